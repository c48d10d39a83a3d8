"""Multi-resolution level sets (smg_set_levelset_levels).  This is a build extension:
the reference's level set is single-level (level_set.cpp:53, Lmin = Lmax = 0 at
sph_mesh_generation.cpp:308-310).  Pinned by:
* refined levels whose bands keep every package give a run identical, bit for bit, to
  a single-level run on the finest grid (positions, h, types after seeding, Pre_run
  and steps);
* with partial bands every sample equals the oracle's interpolation
  (Calculate_particle_infor) on the finest level that stores all 2^D corner packages
  of the sample, else the base grid.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(smg_mod, case):
    if case == "2d":
        cfg, sdf = smg_mod.config_c1(num_pkg=8, dx=0.02)   # base 32^2, finest 128^2
    else:
        # base 16^3, finest 64^3; h 2:1 over a wide band keeps the lists short
        cfg, sdf = smg_mod.config_c3(num_pkg=4, dx0=0.012, ratio=2.0, band=0.2)
    return cfg, sdf


@pytest.mark.parametrize("case", ["2d", "3d"])
def test_full_bands_equal_single_level_fine(smg_mod, case):
    cfg, sdf = _case(smg_mod, case)
    base = smg_mod.make_levelset(cfg, sdf)
    lv = smg_mod.make_levelset_levels(cfg, sdf, 2, band=None)
    fine = lv[-1]
    n = int(fine["stats"].num_particle)
    print(f"case {case}: n = {n}", flush=True)
    a = smg_mod.Job(cfg, base, levels=lv)
    b = smg_mod.Job(fine["cfg"], fine)
    outs = []
    for job in (a, b):
        job.seed(n)
        s0 = job.state()
        job.step(12)
        outs.append((s0, job.state(), job.stats()))
        job.close()
    (a0, a1, sa), (b0, b1, sb) = outs
    for key in ("x", "h", "type"):
        assert np.array_equal(a0[key], b0[key]), key
        assert np.array_equal(a1[key], b1[key]), key
    assert sa["residual"] == sb["residual"]


def _expected(smg_mod, orc, cfg, base, lv, pts):
    """The oracle's sample on the finest level whose stored packages hold the corners."""
    D = cfg.dim
    out, lev = orc.sample(cfg, base, pts)[0], np.zeros(len(pts), int)
    for l, L in enumerate(lv):
        cl = L["cfg"]
        np_ = [cl.num_pkg[a] for a in range(D)]
        act = L["active"].reshape([cl.num_pkg[a] for a in reversed(range(D))])
        ok = np.ones(len(pts), bool)
        pk = np.zeros((len(pts), D), int)
        m = np.zeros((len(pts), D), int)
        for a in range(D):
            dpkg = cl.domain[a] / float(cl.num_pkg[a])
            dcell = dpkg / 4.0
            t = np.maximum(pts[:, a] - cl.box_l[a], 0.)
            cs = np.minimum(t, cl.domain[a] - 1.e-8)
            p = np.clip(np.floor(cs / dpkg).astype(int), -1, np_[a])
            rr = np.clip((cs - dpkg * p) / dcell - 0.5, -0.5, 3.5)
            pk[:, a], m[:, a] = p, np.floor(rr).astype(int)
        for corner in range(1 << D):
            idx = []
            for a in range(D):
                g = pk[:, a] * 4 + m[:, a] + ((corner >> a) & 1)
                idx.append(np.clip(np.floor_divide(g, 4), 0, np_[a] - 1))   # halo packages: the interior neighbour
            ok &= act[tuple(reversed(idx))] != 0
        phi_l = orc.sample(cl, L, pts)[0]
        out = np.where(ok, phi_l, out)
        lev = np.where(ok, l + 1, lev)
    return out, lev


@pytest.mark.parametrize("case", ["2d", "3d"])
def test_partial_bands_sample_the_finest_stored_level(smg_mod, orc, case):
    cfg, sdf = _case(smg_mod, case)
    base = smg_mod.make_levelset(cfg, sdf)
    lv = smg_mod.make_levelset_levels(cfg, sdf, 2, band=2.0)
    assert all(0 < L["active"].sum() < L["active"].size for L in lv)
    job = smg_mod.Job(cfg, base, levels=lv)
    job.seed(200)
    rng = np.random.default_rng(3)
    pts = rng.uniform(-0.02, 1.02, (20000, 3))
    if cfg.dim == 2:
        pts[:, 2] = 0.
    # the particles' own positions too: near the surface, where the bands are
    pts = np.concatenate([pts, job.state()["x"]])
    got = job.sample_phi(pts)
    want, lev = _expected(smg_mod, orc, cfg, base, lv, pts[:, :cfg.dim])
    print(f"case {case}: samples per level {np.bincount(lev, minlength=3).tolist()}", flush=True)
    assert np.bincount(lev, minlength=3).min() > 0   # every level is exercised
    assert np.array_equal(got, want)
    job.close()


def test_c3_small_relaxes(smg_mod):
    """A reduced multi-resolution case (12 ellipsoids, three-level level set, ~1e5
    particles, h ratio 3:1) seeds, relaxes and lowers its residual without re-seeding
    failures."""
    cfg, sdf = smg_mod.config_c3(num_pkg=4, dx0=0.008, ratio=3.0, band=0.2)
    base = smg_mod.make_levelset(cfg, sdf)
    lv = smg_mod.make_levelset_levels(cfg, sdf, 2)
    n = int(lv[-1]["stats"].num_particle)
    job = smg_mod.Job(cfg, base, levels=lv)
    job.seed(n)
    job.step(50)
    r0 = job.stats()["residual"]
    job.step(300)
    st = job.stats()
    print(f"n = {n}, residual {r0:.3e} -> {st['residual']:.3e}, surface {st['n_surface']}", flush=True)
    assert st["residual"] < r0
    assert st["n_surface"] > 0
    job.close()
