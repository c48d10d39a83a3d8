"""Pins level-set sampling (particle_mesh_generation.cpp:175-265), cut tags
(lset_package.cpp:664-761), normals and the particle count against closed forms,
and checks the product's host builder against the oracle bit for bit."""
import math

import numpy as np
import pytest


def _linear_lset(smg_mod, orc, dim, coef):
    cfg = smg_mod.default_config(dim)
    for a in range(dim):
        cfg.num_pkg[a] = 8
    X, Y, Z = smg_mod.cell_centres_ext(cfg)
    phi = coef[0] * X + coef[1] * Y + (coef[2] * Z if dim == 3 else 0) + coef[3]
    return cfg, orc.lset_build(cfg, phi, zero_margin=False)


@pytest.mark.parametrize("dim", [2, 3])
def test_sampling_reproduces_linear_phi(smg_mod, orc, dim):
    coef = (0.3, -0.7, 0.45, 0.1)
    cfg, L = _linear_lset(smg_mod, orc, dim, coef)
    rng = np.random.default_rng(11)
    x = rng.uniform(0.02, 0.98, size=(2000, dim))
    phi, n, tag = orc.sample(cfg, L, x)
    exact = x @ np.array(coef[:dim]) + coef[3]
    assert np.abs(phi - exact).max() < 1e-13
    # central-difference normals of a linear field are exact unit vectors
    g = np.array(coef[:dim]) / np.linalg.norm(coef[:dim])
    assert np.abs(n[:, :dim] - g).max() < 1e-6   # FP32 storage of the normal


def test_zero_margin_blends_with_zero(smg_mod, orc):
    # restart-mode margins hold phi = 0 (quirk Q9): at the wall the sample is half the
    # first interior cell value
    cfg, sdf = smg_mod.config_c1()
    X, Y, Z = smg_mod.cell_centres_ext(cfg)
    L = orc.lset_build(cfg, sdf(X, Y, Z), zero_margin=True)
    x = np.array([[0.0, 0.5]])
    phi, _, _ = orc.sample(cfg, L, x)
    dcell = 1.0 / 128
    first = math.hypot(0.5 * dcell - 0.5, 63.5 * dcell + 0.5 * dcell - 0.5) - 0.2
    # y = 0.5 lies between cells 63 and 64 -> average of the two rows
    c63 = math.hypot(0.5 * dcell - 0.5, 63.5 * dcell - 0.5) - 0.2
    c64 = math.hypot(0.5 * dcell - 0.5, 64.5 * dcell - 0.5) - 0.2
    assert phi[0] == pytest.approx(0.5 * (0.5 * c63 + 0.5 * c64), rel=1e-12)
    del first


def test_circle_closed_forms(smg_mod, orc):
    cfg, sdf = smg_mod.config_c1()
    X, Y, Z = smg_mod.cell_centres_ext(cfg)
    L = orc.lset_build(cfg, sdf(X, Y, Z))
    st = L["stats"]
    R = 0.2
    assert st.total_area == pytest.approx(2 * math.pi * R, rel=2e-3)
    # Calculate_volume_mass counts cells with phi > 0 or |phi| < dl: cut cells by their
    # fluid fraction, all other counted cells in full (also those just inside the hole)
    dcell = 1.0 / 128
    c = (np.arange(128) + 0.5) * dcell
    CX, CY = np.meshgrid(c, c, indexing="ij")
    pc = np.hypot(CX - 0.5, CY - 0.5) - R
    g0 = 128 + 2
    tg = L["tag"].reshape(g0, g0)[1:-1, 1:-1].T
    sub = (np.arange(32) + 0.5) / 32 - 0.5
    SX, SY = np.meshgrid(sub, sub, indexing="ij")
    frac = np.zeros_like(pc)
    for i, j in zip(*np.nonzero(tg == 1)):
        frac[i, j] = np.mean(np.hypot(CX[i, j] + SX * dcell - 0.5, CY[i, j] + SY * dcell - 0.5) > R)
    counted = (pc > 0) | (np.abs(pc) < dcell)
    expect = (np.where(tg == 1, frac, 1.0) * counted).sum() * dcell * dcell
    assert st.total_volume == pytest.approx(expect, rel=2e-4)
    assert expect == pytest.approx(1 - math.pi * R * R, rel=1e-2)
    assert st.maximum_phi == pytest.approx(math.hypot(0.5 - 0.5 / 128, 0.5 - 0.5 / 128) - R, rel=1e-12)
    # particle count: V / dx^2 - A / (2 dx) + A / dx  (Get_num_particles)
    dx = 0.01
    assert st.num_particle == pytest.approx(st.total_volume / dx ** 2 + 0.5 * st.total_area / dx, abs=2)
    # cut cells hug the circle: every cell with |phi| < 0.45 dcell is cut, no cut cell is far
    dcell = 1.0 / 128
    g = 128 + 2
    phi = L["phi"].reshape(g, g)[1:-1, 1:-1]
    tag = L["tag"].reshape(g, g)[1:-1, 1:-1]
    assert np.all(tag[np.abs(phi) < 0.45 * dcell] == 1)
    assert np.all(np.abs(phi[tag == 1]) < 1.5 * dcell)
    assert st.n_interface_cell == int(tag.sum())


def test_sphere_closed_forms(smg_mod, orc):
    cfg = smg_mod.default_config(3)
    for a in range(3):
        cfg.num_pkg[a] = 16
    cfg.min_dl = cfg.max_dl = 0.02
    sdf = smg_mod.sdf_circle_hole((0.5, 0.5, 0.5), 0.3)
    X, Y, Z = smg_mod.cell_centres_ext(cfg)
    L = orc.lset_build(cfg, sdf(X, Y, Z))
    st = L["stats"]
    R = 0.3
    assert st.total_area == pytest.approx(4 * math.pi * R * R, rel=1e-2)
    # includes the full cells just inside the hole with |phi| < dl (see the 2D case)
    assert st.total_volume == pytest.approx(1 - 4 / 3 * math.pi * R ** 3, rel=1e-2)


@pytest.mark.parametrize("dim", [2, 3])
def test_product_builder_equals_oracle(smg_mod, orc, dim):
    if dim == 2:
        cfg, sdf = smg_mod.config_c1()
    else:
        cfg, sdf = smg_mod.config_c2(num_pkg=12)
    for zm in (True, False):
        X, Y, Z = smg_mod.cell_centres_ext(cfg)
        ext = sdf(X, Y, Z)
        P = smg_mod.lset_build(cfg, ext, zero_margin=zm)
        O = orc.lset_build(cfg, ext, zero_margin=zm)
        for k in ("phi", "normal", "tag"):
            assert np.array_equal(P[k], O[k]), k
        for f, _ in smg_mod.SmgLsetStats._fields_:
            assert getattr(P["stats"], f) == getattr(O["stats"], f), f


def _clip(poly, a, b, c):
    """Sutherland-Hodgman: the part of polygon poly (k x 2) with a x + b y + c >= 0."""
    out = []
    k = len(poly)
    for q in range(k):
        P, Q = poly[q], poly[(q + 1) % k]
        fp, fq = a * P[0] + b * P[1] + c, a * Q[0] + b * Q[1] + c
        if fp >= 0:
            out.append(P)
        if (fp >= 0) != (fq >= 0):
            t = fp / (fp - fq)
            out.append(P + t * (Q - P))
    return np.array(out)


def _shoelace(poly):
    x, y = poly[:, 0], poly[:, 1]
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


def _plane_cfg(smg_mod, dim, npk):
    cfg = smg_mod.default_config(dim)
    for a in range(dim):
        cfg.num_pkg[a] = npk
    lo = np.array([cfg.box_l[a] for a in range(dim)])
    hi = lo + np.array([cfg.domain[a] for a in range(dim)])
    return cfg, lo, hi


def test_mq_area_oblique_plane_3d(smg_mod, orc):
    """Marching cubes is exact on a linear field: the cut cells' triangles tile the plane's
    section of the box, whose area is the area of its projection on the xy plane over
    |n_z| (the unit square clipped by 0 <= z(x, y) <= 1)."""
    cfg, lo, hi = _plane_cfg(smg_mod, 3, 6)
    n = np.array([0.3, -0.7, 0.45])
    n /= np.linalg.norm(n)
    centre = lo + np.array([0.47, 0.52, 0.61]) * (hi - lo)
    d = -n @ centre
    X, Y, Z = smg_mod.cell_centres_ext(cfg)
    L = orc.lset_build(cfg, n[0] * X + n[1] * Y + n[2] * Z + d, zero_margin=False)
    sq = np.array([[lo[0], lo[1]], [hi[0], lo[1]], [hi[0], hi[1]], [lo[0], hi[1]]])
    # z(x, y) = -(n_x x + n_y y + d) / n_z within [lo_z, hi_z]
    poly = _clip(sq, -n[0] / n[2], -n[1] / n[2], -d / n[2] - lo[2])
    poly = _clip(poly, n[0] / n[2], n[1] / n[2], d / n[2] + hi[2])
    exact = _shoelace(poly) / abs(n[2])
    assert L["stats"].total_area == pytest.approx(exact, rel=1e-12)
    P = smg_mod.lset_build(cfg, n[0] * X + n[1] * Y + n[2] * Z + d, zero_margin=False)
    assert P["stats"].total_area == pytest.approx(exact, rel=1e-12)


def test_mq_area_oblique_line_2d(smg_mod, orc):
    """2D: the cell is measured as three equal layers of unit depth, so the area is the
    length of the line's section of the box."""
    cfg, lo, hi = _plane_cfg(smg_mod, 2, 8)
    n = np.array([0.8, -0.35])
    n /= np.linalg.norm(n)
    centre = lo + np.array([0.43, 0.58]) * (hi - lo)
    d = -n @ centre
    X, Y, Z = smg_mod.cell_centres_ext(cfg)
    L = orc.lset_build(cfg, n[0] * X + n[1] * Y + d, zero_margin=False)
    # the line's ends on the box: y(x) = -(n_x x + d) / n_y at x = lo, hi clipped to [lo_y, hi_y]
    ts = []
    for x in (lo[0], hi[0]):
        ts.append((x, -(n[0] * x + d) / n[1]))
    for y in (lo[1], hi[1]):
        ts.append((-(n[1] * y + d) / n[0], y))
    pts = [p for p in ts if lo[0] - 1e-12 <= p[0] <= hi[0] + 1e-12 and lo[1] - 1e-12 <= p[1] <= hi[1] + 1e-12]
    pts = np.array(sorted(pts))
    exact = np.linalg.norm(pts[-1] - pts[0])
    assert L["stats"].total_area == pytest.approx(exact, rel=1e-12)
