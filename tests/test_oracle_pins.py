"""Pins the oracle functions that the lattice and long-double checks leave open:
the surface-particle branch of the kernel summation and of the stage-1 force
(particle_mesh_generation.cpp:690-765, 883-991, 1122-1126), the stage-2 viscous force
(:995-1118: harmonic-mean eta, quirk Q8, ghost nu = 0), the nu update and stage-2 dt
(:127-145), Levelset::Get_scale (level_set.cpp:311-351), the surface projection and
re-seeding (particle_mesh_generation.cpp:408-467, sph_mesh_generation.cpp:1885-1928)
and seeding (sph_mesh_generation.cpp:1302-1367).

Each check is a closed form, a symmetry, a lattice sum or a property that a plausible
mistake (a dropped term, the wrong kernel dimension, a renormalised normal, a wrong
sign or mean) would break; the kernel W / dW they use is pinned by normalisation and
finite differences in test_oracle_kernel.py."""
import math

import numpy as np
import pytest

REAL, SURFACE = 0, 1


def _cfg(smg_mod, dim, sym=False):
    cfg = smg_mod.default_config(dim)
    for a in range(3):
        cfg.sym[a] = 1 if (sym and a < dim) else 0
    return cfg


def _plane_lattice(dx, m, z0):
    g = (np.arange(m) - m // 2) * dx + 0.5
    X, Y = np.meshgrid(g, g, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), np.full(X.size, z0)], 1)
    inner = (np.abs(X.ravel() - 0.5) < (m // 2 - 3) * dx) & (np.abs(Y.ravel() - 0.5) < (m // 2 - 3) * dx)
    return x, inner


# ---------------------------------------------------------------------------
# surface branch: kernel summation and stage-1 force


def test_surface_plane_lattice_sum_is_the_2d_lattice_sum(smg_mod, orc):
    """Surface particles on a flat square lattice with unit normal: the number density
    uses the (D-1)-dimensional kernel on tangential distances, so an interior particle's
    sum is 1 / (the 2D lattice sum of W_2D within 2h); REAL particles in a layer above
    do not enter a surface particle's sum or force; the force of the symmetric lattice
    is zero and normal-free."""
    cfg = _cfg(smg_mod, 3)
    dx = 0.01
    h = dx * cfg.cell_ratio / cfg.cut_off
    xs, inner = _plane_lattice(dx, 15, 0.5)
    ns = len(xs)
    # lattice points within cut_off h = 1.45 dx of the origin: 0, 4 at dx, 4 at sqrt(2) dx
    pts = [(i, j) for i in range(-2, 3) for j in range(-2, 3) if math.hypot(i, j) * dx <= cfg.cut_off * h + 1e-10]
    assert len(pts) == 9
    lat = sum(orc.W(2, math.hypot(i, j) * dx, h) for i, j in pts)
    expect = 1.0 / (lat + 1e-15)
    nrm = np.tile([0.0, 0.0, 1.0], (ns, 1))
    r = orc.pair_terms(cfg, xs, nrm, np.full(ns, h), np.zeros(ns), np.full(ns, SURFACE, np.int32))
    assert np.abs(r["sum"][inner] - expect).max() <= 1e-13 * expect
    # the 3D kernel would give a different value (the dimension matters)
    lat3 = sum(orc.W(3, math.hypot(i, j) * dx, h) for i, j in pts)
    assert abs(1.0 / lat3 - expect) > 0.1 * expect
    f0 = np.abs(r["a1"][inner]).max()
    # add a REAL layer 0.8 dx above: surface sums and forces do not change at all
    xr = xs.copy(); xr[:, 2] += 0.8 * dx
    x2 = np.vstack([xs, xr])
    n2 = np.vstack([nrm, np.zeros((ns, 3))])
    t2 = np.concatenate([np.full(ns, SURFACE, np.int32), np.full(ns, REAL, np.int32)])
    r2 = orc.pair_terms(cfg, x2, n2, np.full(2 * ns, h), np.zeros(2 * ns), t2)
    assert np.array_equal(r2["sum"][:ns], r["sum"])
    assert np.array_equal(r2["a1"][:ns], r["a1"])
    # the REAL layer feels the surface particles with the full 3D kernel: pushed away (+z)
    assert np.all(r2["a1"][ns:][inner, 2] > 0)
    # interior of the symmetric plane lattice: no tangential force, no normal component
    scale = np.abs(r2["a1"][:ns]).max()
    assert scale > 0
    assert f0 <= 1e-10 * scale
    assert np.all(r2["a1"][:ns, 2] == 0.0)   # Map_surface_particle_force removes a.n exactly here


@pytest.mark.parametrize("theta_deg,counted", [(60.0, True), (120.0, True), (80.0, False), (100.0, False)])
def test_surface_alignment_cut(smg_mod, orc, theta_deg, counted):
    """Two surface particles in the tangent plane of the first one: the neighbour counts
    in the sum (and the force) only if |n_i . n_j| > cos(angle_thres) (angle_thres =
    0.4 pi, cos = 0.309); with only itself counted the sum stays the raw W(0) (Q7)."""
    cfg = _cfg(smg_mod, 3)
    dx = 0.01
    h = dx * cfg.cell_ratio / cfg.cut_off
    d = 1.1 * dx
    th = math.radians(theta_deg)
    x = np.array([[0.5, 0.5, 0.5], [0.5 + d, 0.5, 0.5]])
    nrm = np.array([[0.0, 0.0, 1.0], [math.sin(th), 0.0, math.cos(th)]])
    r = orc.pair_terms(cfg, x, nrm, np.full(2, h), np.zeros(2), np.full(2, SURFACE, np.int32))
    w0, wd = orc.W(2, 0.0, h), orc.W(2, d, h)
    vol2 = dx ** 6
    if counted:
        s = 1.0 / (w0 + wd + 1e-15)
        assert r["sum"][0] == pytest.approx(s, rel=1e-14)
        # F = -sum (vol_i^2 + vol_j^2) dW_2D(d, h) (x_i - x_j), tangential already
        f = -s * 2 * vol2 * orc.dW(2, d, h) * (-d)
        assert r["a1"][0, 0] == pytest.approx(f, rel=1e-13)
        assert r["a1"][0, 0] < 0   # pushed away from the neighbour
    else:
        assert r["sum"][0] == w0   # one contributing neighbour (itself): raw sum
        assert np.all(r["a1"][0] == 0.0)
    assert r["a1"][0, 2] == 0.0


def test_surface_projection_uses_the_raw_normal(smg_mod, orc):
    """The tangential distance is |dr - n (dr . n)| with the interpolated, not
    renormalised normal (Q3): with n = (0, 0, 0.8) and dr = (-d, 0, -e) it is
    sqrt(d^2 + (0.36 e)^2), not d."""
    cfg = _cfg(smg_mod, 3)
    dx = 0.01
    h = dx * cfg.cell_ratio / cfg.cut_off
    d, e = 0.9 * dx, 0.5 * dx
    x = np.array([[0.5, 0.5, 0.5], [0.5 + d, 0.5, 0.5 + e]])
    nrm = np.array([[0.0, 0.0, 0.8], [0.0, 0.0, 0.8]])   # |n_i . n_j| = 0.64: aligned
    r = orc.pair_terms(cfg, x, nrm, np.full(2, h), np.zeros(2), np.full(2, SURFACE, np.int32))
    rt = math.sqrt(d * d + (0.36 * e) ** 2)
    assert r["sum"][0] == pytest.approx(1.0 / (orc.W(2, 0.0, h) + orc.W(2, rt, h) + 1e-15), rel=1e-14)
    assert abs(r["sum"][0] - 1.0 / (orc.W(2, 0.0, h) + orc.W(2, d, h))) > 1e-6 * r["sum"][0]
    # Map_surface_particle_force with the raw normal: a - n (a . n) leaves a.n = 0.36 a_z
    a = r["a1"][0]
    f = -r["sum"][0] * 2 * dx ** 6 * orc.dW(2, rt, h)
    ft = np.array([-d, 0.0, -0.36 * e]) * f            # force along the projected dr
    expect = ft - np.array([0, 0, 0.8]) * (ft @ np.array([0, 0, 0.8]))
    assert np.allclose(a, expect, rtol=1e-12, atol=1e-12 * np.abs(expect).max())


def test_real_sees_surface_with_full_kernel_and_surface_ignores_real(smg_mod, orc):
    """A REAL particle sums every neighbour with the D-dimensional kernel; a surface
    particle ignores its REAL neighbours in the sum and in both force stages."""
    cfg = _cfg(smg_mod, 3)
    dx = 0.01
    h = dx * cfg.cell_ratio / cfg.cut_off
    d = 1.2 * dx
    x = np.array([[0.5, 0.5, 0.5], [0.5, 0.5, 0.5 + d]])
    nrm = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    t = np.array([REAL, SURFACE], np.int32)
    r = orc.pair_terms(cfg, x, nrm, np.full(2, h), np.zeros(2), t)
    s0 = 1.0 / (orc.W(3, 0.0, h) + orc.W(3, d, h) + 1e-15)
    assert r["sum"][0] == pytest.approx(s0, rel=1e-14)
    assert r["a1"][0, 2] == pytest.approx(-s0 * 2 * dx ** 6 * orc.dW(3, d, h) * (-d), rel=1e-13)
    assert r["sum"][1] == orc.W(2, 0.0, h)
    assert np.all(r["a1"][1] == 0.0)
    v = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, -1.0]])
    s2 = orc.stage2(cfg, x, nrm, np.full(2, h), np.full(2, 1e-3), v, t)
    assert np.all(s2["a2"][1] == 0.0)
    assert s2["a2"][0, 2] != 0.0


# ---------------------------------------------------------------------------
# stage 2: viscous force, nu, dt2


def _pair2(smg_mod, orc, nu, v, types=(REAL, REAL), d=None, dim=3, sym=False, x0=0.5):
    cfg = _cfg(smg_mod, dim, sym=sym)
    dx = 0.01
    h = dx * cfg.cell_ratio / cfg.cut_off
    d = 1.1 * dx if d is None else d
    x = np.zeros((2, dim)); x[:, :] = 0.5; x[:, 0] = [x0, x0 + d]
    nrm = np.zeros((2, dim))
    return cfg, h, d, dx, orc.stage2(cfg, x, nrm, np.full(2, h), np.asarray(nu, float), np.asarray(v, float),
                                     np.asarray(types, np.int32))


def test_stage2_two_particle_hand_case(smg_mod, orc):
    u, nu = 0.3, 2e-4
    v = [[u, 0, 0], [-u, 0, 0]]
    cfg, h, d, dx, r = _pair2(smg_mod, orc, [nu, nu], v)
    rho = 1.0 / dx ** 3
    eta = nu * rho
    eta_ij = 2 * eta * eta / (2 * eta + 1e-15)
    f = orc.dW(3, d, h) * eta_ij * (2 / rho ** 2) * (2 * u)
    assert r["a2"][0, 0] == pytest.approx(f, rel=1e-13)
    assert r["a2"][1, 0] == pytest.approx(-f, rel=1e-13)   # momentum conserving
    assert r["a2"][0, 0] < 0                               # opposes the relative motion
    assert np.all(r["a2"][:, 1:] == 0)
    # linear in the velocities
    _, _, _, _, r2 = _pair2(smg_mod, orc, [nu, nu], 2 * np.asarray(v))
    assert r2["a2"][0, 0] == pytest.approx(2 * f, rel=1e-13)


def test_stage2_eta_is_the_harmonic_mean(smg_mod, orc):
    v = [[0.2, 0, 0], [0.0, 0, 0]]
    _, _, _, _, r11 = _pair2(smg_mod, orc, [1e-4, 1e-4], v)
    _, _, _, _, r13 = _pair2(smg_mod, orc, [1e-4, 3e-4], v)
    _, _, _, _, r10 = _pair2(smg_mod, orc, [1e-4, 0.0], v)
    # eta_ij = 2 eta_i eta_j / (eta_i + eta_j): 1.5 eta for nu 1:3 (2.0 for the arithmetic mean)
    assert r13["a2"][0, 0] / r11["a2"][0, 0] == pytest.approx(1.5, rel=1e-9)
    assert np.all(r10["a2"] == 0.0)   # one zero viscosity: no viscous pair force


def test_stage2_real_with_nonreal_neighbour_uses_minus_own_velocity(smg_mod, orc):
    """Q8: for a REAL particle, a non-REAL neighbour's velocity is taken as -v_i, so
    its own velocity does not matter; the result equals a REAL neighbour moving at -v_i."""
    vi = [0.25, -0.1, 0.05]
    _, _, _, _, rs = _pair2(smg_mod, orc, [1e-4, 1e-4], [vi, [3.0, 2.0, -1.0]], types=(REAL, SURFACE))
    _, _, _, _, rs2 = _pair2(smg_mod, orc, [1e-4, 1e-4], [vi, [-7.0, 0.0, 4.0]], types=(REAL, SURFACE))
    _, _, _, _, rr = _pair2(smg_mod, orc, [1e-4, 1e-4], [vi, [-x for x in vi]], types=(REAL, REAL))
    assert np.array_equal(rs["a2"][0], rs2["a2"][0])
    assert np.array_equal(rs["a2"][0], rr["a2"][0])
    # the surface particle sees no REAL neighbour
    assert np.all(rs["a2"][1] == 0.0)


def test_stage2_mirror_ghosts_carry_no_viscosity(smg_mod, orc):
    """Q6: mirror ghosts have nu = 0, so near a symmetric wall the viscous force is the
    same with and without the ghosts, while the stage-1 terms do see them."""
    v = [[0.2, 0.1, 0], [-0.1, 0.3, 0]]
    dx = 0.01
    x0 = 0.3 * dx   # within the cutoff of the low x wall
    _, _, _, _, ron = _pair2(smg_mod, orc, [1e-4, 2e-4], v, sym=True, x0=x0)
    _, _, _, _, roff = _pair2(smg_mod, orc, [1e-4, 2e-4], v, sym=False, x0=x0)
    assert np.array_equal(ron["a2"], roff["a2"])
    cfg_on, cfg_off = _cfg(smg_mod, 3, sym=True), _cfg(smg_mod, 3, sym=False)
    h = dx * cfg_on.cell_ratio / cfg_on.cut_off
    x = np.array([[x0, 0.5, 0.5], [x0 + 1.1 * dx, 0.5, 0.5]])
    p_on = orc.pair_terms(cfg_on, x, np.zeros((2, 3)), np.full(2, h), np.zeros(2), np.zeros(2, np.int32))
    p_off = orc.pair_terms(cfg_off, x, np.zeros((2, 3)), np.full(2, h), np.zeros(2), np.zeros(2, np.int32))
    assert p_on["sum"][0] < p_off["sum"][0]   # the ghost adds to the density sum


def test_nu_update_and_stage2_dt(smg_mod, orc):
    """Set_timestep(2): nu = nu_coeff h |v|; dt2 = min over particles of
    min(0.25 sqrt(h / |a|), h / 40 / |v|, 0.125 h^2 / nu).  With nu_coeff = 0.1 the
    viscous cap (1.25 h / |v|) never binds and h / 40 / |v| does; with nu_coeff = 10 the
    cap (0.0125 h / |v|) does.  Input nu = 0 makes a2 = 0, so the acceleration bound
    is out of play."""
    rng = np.random.default_rng(3)
    cfg = _cfg(smg_mod, 3)
    n = 40
    x = rng.uniform(0.4, 0.6, size=(n, 3))
    h = rng.uniform(0.004, 0.006, size=n)
    v = rng.normal(size=(n, 3)) * 0.1
    U = np.linalg.norm(v, axis=1)
    for coef, expect in ((0.1, (h / 40 / (U + 1e-15)).min()), (10.0, (0.125 * h * h / (10.0 * h * U + 1e-15)).min())):
        cfg.nu_coeff = coef
        r = orc.stage2(cfg, x, np.zeros((n, 3)), h, np.zeros(n), v, np.zeros(n, np.int32))
        assert np.all(r["a2"] == 0.0)
        assert np.allclose(r["nu"], coef * h * U, rtol=1e-15, atol=0)
        assert r["dt2"] == pytest.approx(expect, rel=1e-14)
        assert r["max_v"] == pytest.approx(U.max(), rel=1e-15)


# ---------------------------------------------------------------------------
# Get_scale


def _lset_c1(smg_mod):
    cfg, sdf = smg_mod.config_c1()
    return cfg, smg_mod.make_levelset(cfg, sdf)


def test_get_scale_tanh_profile(smg_mod, orc):
    cfg, lset = _lset_c1(smg_mod)
    cfg.profile = smg_mod.PROFILE_TANH
    cfg.min_dl, cfg.max_dl, cfg.max_phi = 0.01, 0.05, 0.2
    dl = 1.0 / 128
    g = lambda p: orc.get_scale(cfg, lset, p)   # noqa: E731
    assert g(0.0) == 0.01                                   # dx(0) = dl_min
    assert g(0.2) == pytest.approx(0.05, abs=2e-17)         # dx(phi_max) = dl_max
    phi_mid = 0.2 / 2.5 * math.atanh(math.tanh(2.5) / 2)    # tanh term = 1/2: the midpoint
    assert g(phi_mid) == pytest.approx(0.03, rel=1e-14)
    # large phi: the tanh ratio tends to 1 / tanh(2.5), past dl_max
    assert g(5.0) == pytest.approx(0.01 + 0.04 / math.tanh(2.5), rel=1e-14)
    # negative phi: |phi| < dl uses |phi| (symmetric), beyond it dl_min
    assert g(-0.5 * dl) == g(0.5 * dl) > 0.01
    assert g(-1.01 * dl) == 0.01 and g(-0.3) == 0.01
    ps = np.linspace(0, 0.6, 200)
    vals = np.array([g(p) for p in ps])
    assert np.all(np.diff(vals) > 0)


def test_get_scale_linear_profile(smg_mod, orc):
    cfg, lset = _lset_c1(smg_mod)
    cfg.profile = smg_mod.PROFILE_LINEAR
    cfg.lin_dx0, cfg.lin_dx1, cfg.lin_band = 0.008, 0.024, 0.2
    dl = 1.0 / 128
    g = lambda p: orc.get_scale(cfg, lset, p)   # noqa: E731
    assert g(0.0) == 0.008 and g(0.2) == 0.024 and g(0.5) == 0.024
    assert g(0.1) == pytest.approx(0.016, rel=1e-15)
    assert g(-0.5 * dl) == 0.008 and g(-2 * dl) == 0.008


# ---------------------------------------------------------------------------
# projection, re-seeding, seeding


def _annulus_state(orc, cfg, lset, centre, R, n, dim, seed):
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n, dim))
    u /= np.linalg.norm(u, axis=1)[:, None]
    rad = rng.uniform(R - 0.012, R + 0.012, size=n)
    x = centre + u * rad[:, None]
    phi, _, tag = orc.sample(cfg, lset, x)
    s = orc.State.from_arrays(x, np.zeros((n, dim)), np.zeros(n), np.zeros(n), np.zeros(n, np.int32), 0.0)
    return x, phi, tag, s


@pytest.mark.parametrize("dim", [2, 3])
def test_projection_lands_on_the_surface(smg_mod, orc, dim):
    """Pre_run's mapping of particles scattered around a circle (2D) / sphere (3D):
    those starting in a cut cell become SURFACE and end on |x - c| = R to the
    level set's interpolation error, moved along the radius; REAL particles with phi < 0
    are re-seeded to phi > h; the others stay where they were."""
    if dim == 2:
        cfg, sdf = smg_mod.config_c1()
        centre, R, dl = np.array([0.5, 0.5]), 0.2, 1.0 / 128
    else:
        cfg, sdf = smg_mod.config_c2(scale=2.0, num_pkg=16)
        centre, R, dl = np.array([0.5, 0.5, 0.5]), 0.3, 1.0 / 64
    lset = smg_mod.make_levelset(cfg, sdf)
    x, phi, tag, s = _annulus_state(orc, cfg, lset, centre, R, 3000, dim, 7)
    orc.prerun(cfg, lset, s)
    xo = s.x[:, :dim]
    surf = s.type == SURFACE
    assert np.array_equal(surf, tag == 1)          # REAL -> SURFACE exactly in cut cells
    assert surf.sum() > 100
    rad = np.linalg.norm(xo - centre, axis=1)
    err = np.abs(rad[surf] - R)
    # before: up to a cell width off the surface; after: the interpolation error
    assert np.abs(phi[surf]).max() > 0.5 * dl
    assert err.max() < 0.02 * dl
    # moved along the normal (radial) only
    mv = xo[surf] - x[surf]
    radial = (x[surf] - centre) / np.linalg.norm(x[surf] - centre, axis=1)[:, None]
    tang = mv - radial * (mv * radial).sum(1)[:, None]
    assert np.abs(tang).max() < 0.02 * dl
    # re-seeded: REAL in non-cut cells with phi < 0.  The acceptance test compares the
    # new phi with the h the particle had before (Calculate_particle_infor does not
    # update h; Set_particle_info runs after the loop), then h is set at the new place.
    re = (tag != 1) & (phi < 0)
    assert re.sum() > 10
    phin, _, _ = orc.sample(cfg, lset, xo[re])
    h_old = np.array([orc.get_scale(cfg, lset, p) for p in phi[re]]) * cfg.cell_ratio / cfg.cut_off
    h_new = np.array([orc.get_scale(cfg, lset, p) for p in phin]) * cfg.cell_ratio / cfg.cut_off
    assert np.all(phin > h_old)
    assert np.array_equal(s.h[re], h_new)
    assert np.all(s.type[re] == REAL)
    keep = (tag != 1) & (phi >= 0)
    assert np.array_equal(xo[keep], x[keep])


@pytest.mark.parametrize("case", ["2d", "3d"])
def test_seeding_is_inside_and_uniform(smg_mod, orc, case):
    """Random_particle_distribution: every seed has phi > h at its position, lies in the
    level-set box shrunk by one package and in the particle box, and the accepted
    positions are uniform (the domains are symmetric about x = 0.5)."""
    if case == "2d":
        cfg, sdf = smg_mod.config_c1()
        D = 2
    else:
        cfg, sdf = smg_mod.config_c2(scale=2.0, num_pkg=16)
        D = 3
    lset = smg_mod.make_levelset(cfg, sdf)
    n = int(lset["stats"].num_particle)
    s = orc.seed(cfg, lset, n)
    x = s.x[:, :D]
    phi, _, _ = orc.sample(cfg, lset, x)
    assert np.all(phi > s.h)
    pad = cfg.domain[0] / cfg.num_pkg[0]
    for a in range(D):
        assert x[:, a].min() >= cfg.box_l[a] + pad and x[:, a].max() <= cfg.box_l[a] + cfg.domain[a] - pad
        assert x[:, a].min() >= cfg.particle_box_l[a] and x[:, a].max() <= cfg.particle_box_r[a]
        frac = (x[:, a] < 0.5).mean()
        assert abs(frac - 0.5) < 4 * 0.5 / math.sqrt(n)
    assert np.all(s.type == REAL)
    # deterministic per id
    s2 = orc.seed(cfg, lset, n)
    assert np.array_equal(s.x, s2.x)


def test_counter_rng_is_uniform(orc):
    u = orc.rng(20261203, 5, 17, 0, 200000)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 4 * math.sqrt(1 / 12 / len(u))
    assert abs(u.var() - 1 / 12) < 0.002
    h, _ = np.histogram(u, bins=20, range=(0, 1))
    assert (np.abs(h - len(u) / 20) < 5 * math.sqrt(len(u) / 20)).all()
    # streams and ids are independent
    v = orc.rng(20261203, 6, 17, 0, 200000)
    assert abs(np.corrcoef(u, v)[0, 1]) < 0.01
