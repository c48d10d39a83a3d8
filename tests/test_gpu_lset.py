"""The GPU level-set builder (smg_lset_build_gpu) equals the host builder
(smg_lset_build), which equals the oracle builder (test_oracle_lset.py): phi, normals,
tags and every statistic bit for bit with the uniform and linear spacing profiles;
with the tanh profile the masses within 1e-12 relative (the device tanh is within
2 ulp of the host's) and the particle count within 1."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXACT = ("maximum_phi", "total_volume", "total_area", "n_interface_cell")
MASS = ("total_mass", "total_mass_surface")
COUNT = ("num_particle_mesh", "num_particle_surface", "num_particle")


def _compare(smg_mod, cfg, ext, zm, exact_mass=True):
    P = smg_mod.lset_build(cfg, ext, zero_margin=zm)
    G = smg_mod.lset_build_gpu(cfg, ext, zero_margin=zm)
    for k in ("phi", "normal", "tag"):
        assert np.array_equal(P[k].view(np.uint8), G[k].view(np.uint8)), k
    for f in EXACT:
        assert getattr(P["stats"], f) == getattr(G["stats"], f), f
    for f in MASS:
        a, b = getattr(P["stats"], f), getattr(G["stats"], f)
        if exact_mass:
            assert a == b, f
        else:
            assert abs(a - b) <= 1e-12 * abs(a), (f, a, b)
    for f in COUNT:
        a, b = getattr(P["stats"], f), getattr(G["stats"], f)
        assert a == b if exact_mass else abs(a - b) <= 1, (f, a, b)
    assert G["kernel_ms"] > 0.0
    return P, G


@pytest.mark.parametrize("zm", [True, False])
def test_circle_uniform_2d(smg_mod, zm):
    cfg, sdf = smg_mod.config_c1()
    X, Y, Z = smg_mod.cell_centres_ext(cfg)
    P, _ = _compare(smg_mod, cfg, sdf(X, Y, Z), zm)
    assert P["stats"].n_interface_cell > 0


@pytest.mark.parametrize("zm", [True, False])
def test_sphere_linear_3d(smg_mod, zm):
    cfg, sdf = smg_mod.config_c2(num_pkg=12)
    X, Y, Z = smg_mod.cell_centres_ext(cfg)
    _compare(smg_mod, cfg, sdf(X, Y, Z), zm)


def test_ellipsoid_union_3d(smg_mod):
    cfg, sdf = smg_mod.config_c3(num_pkg=16)
    X, Y, Z = smg_mod.cell_centres_ext(cfg)
    P, _ = _compare(smg_mod, cfg, sdf(X, Y, Z), True)
    assert P["stats"].n_interface_cell > 1000


@pytest.mark.parametrize("dim", [2, 3])
def test_tanh_profile(smg_mod, dim):
    cfg = smg_mod.default_config(dim)
    for a in range(dim):
        cfg.num_pkg[a] = 32 if dim == 2 else 12
    cfg.min_dl, cfg.max_dl = (0.004, 0.02) if dim == 2 else (0.01, 0.04)
    sdf = smg_mod.sdf_circle_hole((0.5,) * dim, 0.23)
    X, Y, Z = smg_mod.cell_centres_ext(cfg)
    P, G = _compare(smg_mod, cfg, sdf(X, Y, Z), True, exact_mass=False)
    assert P["stats"].maximum_phi > 0.3


@pytest.mark.parametrize("dim", [2, 3])
def test_plane_through_cell_centres(smg_mod, dim):
    """phi = x - (a cell centre): exact zeros along a column of cells."""
    cfg = smg_mod.default_config(dim)
    for a in range(dim):
        cfg.num_pkg[a] = 16 if dim == 2 else 6
    cfg.min_dl = cfg.max_dl = 0.03
    X, Y, Z = smg_mod.cell_centres_ext(cfg)
    x0 = X.ravel()[4 * cfg.num_pkg[0] // 2 + 4]
    ext = X - x0
    assert np.count_nonzero(ext == 0.0) > 0
    _compare(smg_mod, cfg, ext, True)
    _compare(smg_mod, cfg, -ext, False)


@pytest.mark.parametrize("dim", [2, 3])
def test_no_interface(smg_mod, dim):
    """All fluid: no cut cells, the count is the box volume over dx^dim."""
    cfg = smg_mod.default_config(dim)
    for a in range(dim):
        cfg.num_pkg[a] = 8
    cfg.min_dl = cfg.max_dl = 0.05
    X, Y, Z = smg_mod.cell_centres_ext(cfg)
    P, G = _compare(smg_mod, cfg, np.ones_like(X), True)
    assert G["stats"].n_interface_cell == 0
    assert G["stats"].total_volume == pytest.approx(1.0, rel=1e-12)
    assert G["stats"].num_particle == int(1.0 / 0.05 ** dim + 1e-9) or abs(
        G["stats"].num_particle - 1.0 / 0.05 ** dim) <= 1


def test_gpu_builder_equals_oracle(smg_mod, orc):
    cfg, sdf = smg_mod.config_c1()
    X, Y, Z = smg_mod.cell_centres_ext(cfg)
    ext = sdf(X, Y, Z)
    G = smg_mod.lset_build_gpu(cfg, ext, zero_margin=True)
    O = orc.lset_build(cfg, ext, zero_margin=True)
    for k in ("phi", "normal", "tag"):
        assert np.array_equal(G[k], O[k]), k
    for f, _ in smg_mod.SmgLsetStats._fields_:
        assert getattr(G["stats"], f) == getattr(O["stats"], f), f


def test_relaxation_on_gpu_built_levelset(smg_mod):
    """A job on the GPU-built level set equals the job on the host-built one."""
    cfg, sdf = smg_mod.config_c1()
    X, Y, Z = smg_mod.cell_centres_ext(cfg)
    ext = sdf(X, Y, Z)
    out = []
    for build in (smg_mod.lset_build, smg_mod.lset_build_gpu):
        L = build(cfg, ext, zero_margin=True)
        job = smg_mod.Job(cfg, L)
        job.seed(int(L["stats"].num_particle))
        job.step(10)
        out.append(job.state())
        job.close()
    assert np.array_equal(out[0]["x"], out[1]["x"])
    assert np.array_equal(out[0]["type"], out[1]["type"])


def test_short_queue_is_regrown(smg_mod, monkeypatch):
    """A level-2 queue too short for the cut cells (1 entry per cell) is counted, grown
    and the first pass run again: the same bits as with the default size."""
    cfg, sdf = smg_mod.config_c3(num_pkg=8)
    X, Y, Z = smg_mod.cell_centres_ext(cfg)
    ext = sdf(X, Y, Z)
    ref = smg_mod.lset_build_gpu(cfg, ext)
    monkeypatch.setenv("SMG_LSET_QUEUE_PER_CELL", "1")
    G = smg_mod.lset_build_gpu(cfg, ext)
    for k in ("phi", "normal", "tag"):
        assert np.array_equal(ref[k], G[k]), k
    for f, _ in smg_mod.SmgLsetStats._fields_:
        assert getattr(ref["stats"], f) == getattr(G["stats"], f), f
    P = smg_mod.lset_build(cfg, ext)
    assert P["stats"].total_volume == G["stats"].total_volume
    for bad in ("1x", "", "0", "-3"):
        monkeypatch.setenv("SMG_LSET_QUEUE_PER_CELL", bad)
        with pytest.raises(smg_mod.SmgError):
            smg_mod.lset_build_gpu(cfg, ext)
