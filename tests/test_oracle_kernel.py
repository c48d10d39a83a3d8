"""Pins the oracle's hyperbolic kernel (particle.cpp:288-310, 686-709) against closed forms."""
import math

import numpy as np
import pytest
from scipy.integrate import quad

PI_REF = 3.1415926


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_normalisation_with_reference_pi(orc, dim):
    # integral of W over R^dim is exactly 1 with the true pi; the reference's truncated
    # PI = 3.1415926 scales it by pi / PI_REF (quirk Q4)
    h = 0.37
    if dim == 1:
        val = 2 * quad(lambda r: orc.W(1, r, h), 0, 2 * h, points=[h])[0]
        expect = 1.0
    elif dim == 2:
        val = quad(lambda r: 2 * math.pi * r * orc.W(2, r, h), 0, 2 * h, points=[h])[0]
        expect = math.pi / PI_REF
    else:
        val = quad(lambda r: 4 * math.pi * r * r * orc.W(3, r, h), 0, 2 * h, points=[h])[0]
        expect = math.pi / PI_REF
    assert abs(val - expect) < 1e-11


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_support_and_continuity(orc, dim):
    h = 0.01
    assert orc.W(dim, 2 * h, h) == 0.0
    assert orc.W(dim, 2.5 * h, h) == 0.0
    C = {1: 1 / 7, 2: 1 / (3 * PI_REF), 3: 15 / (62 * PI_REF)}[dim] / h ** dim
    assert orc.W(dim, 0.0, h) == pytest.approx(6 * C, rel=1e-15)
    eps = 1e-9 * h
    assert orc.W(dim, h - eps, h) == pytest.approx(orc.W(dim, h + eps, h), rel=1e-6)
    assert orc.W(dim, h, h) == pytest.approx(C, rel=1e-14)   # (2-1)^3 branch at s = 1


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_derivative_matches_finite_difference(orc, dim):
    h = 0.02
    for s in np.linspace(0.05, 1.95, 39):
        r = s * h
        d = 1e-7 * h
        fd = (orc.W(dim, r + d, h) - orc.W(dim, r - d, h)) / (2 * d)
        # dW returns (dW/dr) / (r + 1e-20)
        assert orc.dW(dim, r, h) * (r + 1e-20) == pytest.approx(fd, rel=2e-6, abs=1e-9 * abs(fd) + 1e-300)
    assert orc.dW(dim, 2 * h, h) == 0.0
