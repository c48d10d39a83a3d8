"""Pins the oracle's pair terms (Get_kernel_summation, Accumulate_particle_force_real,
Set_timestep) against symmetry, hand cases and an independent long-double evaluation."""
import numpy as np
import pytest


def _cfg(smg_mod, dim):
    cfg = smg_mod.default_config(dim)
    for a in range(3):
        cfg.sym[a] = 0
    return cfg


def test_lattice_has_zero_repulsion(smg_mod, orc):
    cfg = _cfg(smg_mod, 2)
    dx = 0.01
    g = np.arange(20) * dx + 0.3
    X, Y = np.meshgrid(g, g, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel()], 1)
    n = len(x)
    h = np.full(n, dx * cfg.cell_ratio / cfg.cut_off)
    r = orc.pair_terms(cfg, x, np.zeros((n, 2)), h, np.zeros(n), np.zeros(n, np.int32))
    inner = np.all((X.ravel()[:, None] > 0.3 + 2.5 * dx) & (X.ravel()[:, None] < 0.49 - 2.5 * dx) &
                   (Y.ravel()[:, None] > 0.3 + 2.5 * dx) & (Y.ravel()[:, None] < 0.49 - 2.5 * dx), 1)
    scale = np.abs(r["a1"]).max()
    assert scale > 0
    assert np.abs(r["a1"][inner]).max() < 1e-10 * scale
    # every interior particle has the same number density
    assert np.ptp(r["sum"][inner]) < 1e-12 * r["sum"][inner].mean()
    assert np.all(r["a2"] == 0)   # nu = 0 -> no viscous force


def test_two_particles_hand_case(smg_mod, orc):
    cfg = _cfg(smg_mod, 2)
    h = 0.01
    d = 0.013
    x = np.array([[0.5, 0.5], [0.5 + d, 0.5]])
    r = orc.pair_terms(cfg, x, np.zeros((2, 2)), np.array([h, h]), np.zeros(2), np.zeros(2, np.int32))
    w0, wd = orc.W(2, 0.0, h), orc.W(2, d, h)
    s = 1.0 / (w0 + wd + 1e-15)
    assert r["sum"][0] == pytest.approx(s, rel=1e-14)
    dx = h * cfg.cut_off / cfg.cell_ratio
    vol2 = dx ** 4
    # F = -sum * (vol_i^2 + vol_j^2) * dW * r_ij  (P = m = 1)
    f = -s * 2 * vol2 * orc.dW(2, d, h) * (-d)
    assert r["a1"][0, 0] == pytest.approx(f, rel=1e-13)
    assert r["a1"][1, 0] == pytest.approx(-f, rel=1e-13)   # action = reaction for equal h
    assert r["a1"][0, 0] < 0   # repulsive: pushes particle 0 away from 1
    assert r["dt1"] == pytest.approx(0.25 * np.sqrt(h / (abs(f) + 1e-15)), rel=1e-14)


def _ld_pair_terms(x, h, dim, cut, ratio):
    """Independent long-double evaluation of kernel summation + stage-1 force."""
    ld = np.longdouble
    pi = ld("3.1415926")
    C = {2: ld(1) / (ld(3) * pi), 3: ld(15) / (ld(62) * pi)}[dim]
    x = x.astype(ld); h = h.astype(ld)
    n = len(h)
    dxs = h * ld(cut) / ld(ratio)
    vol2 = dxs ** (2 * dim)

    def W(r, hh):
        s = r / hh
        c = C / hh ** dim
        return np.where(s < 1, c * (6 - 6 * s + s ** 3), np.where(s < 2, c * (2 - s) ** 3, ld(0)))

    def dW(r, hh):
        s = r / hh
        c = C / hh ** (dim + 1)
        v = np.where(s < 1, c * (-6 + 3 * s * s), np.where(s < 2, -3 * c * (2 - s) ** 2, ld(0)))
        return v / (r + ld(1e-20))

    sums = np.zeros(n, ld); a = np.zeros((n, dim), ld)
    for i in range(n):
        dr = x[i] - x
        r = np.sqrt((dr * dr).sum(1))
        nb = (r <= h[i] * ld(cut) + ld(1e-10)) | (r <= h * ld(cut) + ld(1e-10))
        sm = W(r[nb], h[i]).sum()
        sums[i] = ld(1) / (sm + ld(1e-15)) if nb.sum() > 1 else sm
        nb[i] = False
        hij = (h[i] + h[nb]) / 2
        aw = dW(r[nb], hij)
        a[i] = (-(sums[i] * (vol2[i] + vol2[nb]))[:, None] * aw[:, None] * dr[nb]).sum(0)
    return sums, a


@pytest.mark.parametrize("dim", [2, 3])
def test_against_long_double(smg_mod, orc, dim):
    rng = np.random.default_rng(5)
    cfg = _cfg(smg_mod, dim)
    n = 300
    x = rng.uniform(0.3, 0.7, size=(n, dim))
    h = rng.uniform(0.03, 0.04, size=n) if dim == 2 else rng.uniform(0.06, 0.07, size=n)
    r = orc.pair_terms(cfg, x, np.zeros((n, dim)), h, np.zeros(n), np.zeros(n, np.int32))
    s_ld, a_ld = _ld_pair_terms(x, h, dim, cfg.cut_off, cfg.cell_ratio)
    assert np.abs(r["sum"] - s_ld.astype(float)).max() <= 1e-12 * np.abs(s_ld.astype(float)).max()
    amax = np.abs(a_ld.astype(float)).max()
    assert np.abs(r["a1"][:, :dim] - a_ld.astype(float)).max() <= 1e-10 * amax
