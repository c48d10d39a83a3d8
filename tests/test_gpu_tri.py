"""2D triangulation on the GPU path: the Voronoi-face edges equal an independent
Qhull construction on the same neighbour sets, the integer stage (3-cliques, centroid
test, rebuild) equals the oracle's on those edges, and the quality statistics equal a
numpy evaluation."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def relaxed(smg_mod, c1):
    cfg, lset = c1
    job = smg_mod.Job(cfg, lset)
    job.seed(int(lset["stats"].num_particle))
    job.step(300)
    return cfg, lset, job


def test_voronoi_edges_equal_qhull(smg_mod, orc, relaxed):
    import tri_oracle
    cfg, lset, job = relaxed
    tris, st = smg_mod.triangulate(job)
    assert st["n_cell_overflow"] == 0
    e_gpu = [tuple(e) for e in smg_mod.tri_edges(job)]
    g = job.state()
    offs, ids, codes = job.neighbors()
    e_ref = tri_oracle.voronoi_edges(cfg, g["x"], g["h"], offs, ids, codes)
    assert e_gpu == e_ref


def test_integer_stage_equals_oracle(smg_mod, orc, relaxed):
    import tri_oracle
    cfg, lset, job = relaxed
    tris, st = smg_mod.triangulate(job)
    edges = [tuple(e) for e in smg_mod.tri_edges(job)]
    g = job.state()
    ref, ndel = tri_oracle.reconstruct(cfg, lset, g["x"], g["type"], edges, orc.sample)
    assert [tuple(t) for t in tris] == ref
    assert st["n_deleted_edge"] == ndel
    q = tri_oracle.quality(g["x"], tris)
    assert st["n_tri"] == q["n_tri"] and st["angle_hist"] == q["angle_hist"]
    for k in ("min_angle", "max_angle", "avg_min_angle", "min_edge", "max_edge", "total_area", "G_min", "G_avg"):
        assert st[k] == pytest.approx(q[k], rel=1e-12, abs=1e-12), k
    # a relaxed uniform 2D set meshes with mostly near-equilateral triangles
    assert st["avg_min_angle"] > 40.0
    assert st["total_area"] == pytest.approx(1.0 - np.pi * 0.2 ** 2, rel=0.05)


def test_triangulate_is_2d_only(smg_mod):
    cfg, sdf = smg_mod.config_c2(scale=3.0, num_pkg=8)
    lset = smg_mod.make_levelset(cfg, sdf)
    job = smg_mod.Job(cfg, lset)
    job.seed(int(lset["stats"].num_particle))
    job.step(1)
    with pytest.raises(smg_mod.SmgError) as e:
        smg_mod.triangulate(job)
    assert e.value.rc == -1
