"""3D tetrahedralisation on the GPU path: the GPU's certified Delaunay tetrahedra equal
Qhull's (scipy.spatial.Delaunay) on the same positions -- every GPU tetrahedron is a
Qhull tetrahedron, and every Qhull tetrahedron whose circumsphere the cell search
certified is a GPU tetrahedron -- their circumspheres are empty (brute force), and the
deletion rule, boundary triangles and quality statistics equal the oracle's on them."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2_small(smg_mod):
    cfg, sdf = smg_mod.config_c2(scale=2.0, num_pkg=16)   # ~24k particles
    lset = smg_mod.make_levelset(cfg, sdf)
    return cfg, lset


def _run(smg_mod, c2_small, steps):
    cfg, lset = c2_small
    job = smg_mod.Job(cfg, lset)
    job.seed(int(lset["stats"].num_particle))
    if steps:
        job.step(steps)
    return job


def _check_against_qhull(smg_mod, job):
    import tri_oracle
    s = job.state()
    x = s["x"]
    n = len(x)
    kept, st = smg_mod.tetrahedralize(job)
    g, status, rcert = smg_mod.tet_delaunay(job, n)
    print(f"n={n} delaunay={len(g)} kept={len(kept)} open={st['n_open']} overflow={st['n_cell_overflow']} "
          f"duplicate={st['n_duplicate']} status={np.bincount(status)}", flush=True)
    assert st["n_cell_overflow"] == 0
    assert st["n_delaunay"] == len(g)
    assert np.all(np.diff(g[:, 0]) >= 0) and np.all(g[:, :-1] < g[:, 1:])   # sorted, ascending ids
    # one vertex per point: coincident particles (the seeding's clamp to the level-set
    # box puts a few on each corner point) keep the smallest id
    _, first = np.unique(x, axis=0, return_index=True)
    first = np.sort(first)
    assert st["n_duplicate"] == n - len(first) == int((status == 6).sum())
    ql, rad = tri_oracle.delaunay_tets(x[first])
    q = np.sort(first[ql], axis=1)
    o = np.lexsort(q.T[::-1])
    q, rad = q[o], rad[o]
    gs = set(map(tuple, g.tolist()))
    qs = set(map(tuple, q.tolist()))
    assert len(gs) == len(g)
    # every GPU tetrahedron is a Delaunay tetrahedron
    extra = gs - qs
    assert not extra, list(extra)[:5]
    # every Delaunay tetrahedron certified by its smallest-id particle's search is found
    i0 = q[:, 0]
    margin = 1e-9
    certified = (status[i0] == 0) | (2.0 * rad < rcert[i0] * (1.0 - margin))
    undecided = (status[i0] != 0) & (np.abs(2.0 * rad - rcert[i0]) <= rcert[i0] * margin)
    need = set(map(tuple, q[certified & ~undecided].tolist()))
    missing = need - gs
    assert not missing, list(missing)[:5]
    # the rest of Qhull's tetrahedra are hull slivers / cavity spanners: far larger
    # circumspheres than the spacing
    rest = np.array([t not in gs for t in map(tuple, q.tolist())])
    if rest.any():
        assert np.all(2.0 * rad[rest] >= rcert[i0[rest]] * (1.0 - margin))
    return x, s, g, kept, st


def test_relaxed_delaunay_equals_qhull(smg_mod, c2_small):
    job = _run(smg_mod, c2_small, 100)
    x, s, g, kept, st = _check_against_qhull(smg_mod, job)
    # cells close inside the search away from the walls and the cavity (there the
    # Voronoi cells of a global tetrahedralisation are unbounded or span the cavity)
    _, status, _ = smg_mod.tet_delaunay(job, len(x))
    wall = np.minimum(x, 1.0 - x).min(axis=1)
    phi = np.linalg.norm(x - 0.5, axis=1) - 0.3
    interior = (wall > 0.15) & (phi > 0.15)
    assert interior.sum() > 1000
    assert np.all(status[interior] == 0)
    # empty circumspheres (brute force over the neighbours of each circumcentre)
    import tri_oracle
    sh = tri_oracle.tet_shapes(x, g)
    tree = cKDTree(x)
    sample = np.random.default_rng(7).choice(len(g), size=min(len(g), 20000), replace=False)
    for k in sample:
        inside = tree.query_ball_point(sh["circumcentre"][k], sh["circumradius"][k] * (1.0 - 1e-9))
        assert set(inside) <= set(g[k].tolist()), (k, g[k], inside)


def test_seeded_random_positions(smg_mod, c2_small):
    job = _run(smg_mod, c2_small, 0)
    _check_against_qhull(smg_mod, job)


def test_host_stage_equals_oracle(smg_mod, c2_small):
    import tri_oracle
    job = _run(smg_mod, c2_small, 200)
    s = job.state()
    kept, st = smg_mod.tetrahedralize(job)
    g, status, rcert = smg_mod.tet_delaunay(job, len(s["x"]))
    rk, rt, rdel, q = tri_oracle.tet_finish(s["x"], s["type"], g)
    assert [tuple(t) for t in kept] == rk
    assert [tuple(t) for t in smg_mod.tet_surface(job)] == rt
    assert st["n_deleted"] == rdel and st["n_surface_tri"] == len(rt) > 0
    for k in ("dihedral_hist", "aspect_hist", "radius_hist", "n_min_dihedral_below"):
        assert st[k] == q[k], k
    for k in ("total_volume", "min_edge", "max_edge", "max_aspect", "avg_aspect", "avg_radius_ratio"):
        assert st[k] == pytest.approx(q[k], rel=1e-9), k
    # acos near 0 and 180 degrees magnifies the rounding of the two formulations (face
    # cross products here, the LU solve in the oracle): slivers agree to 1e-6 degrees
    for k in ("min_dihedral", "max_dihedral", "avg_min_dihedral"):
        assert st[k] == pytest.approx(q[k], rel=1e-9, abs=1e-6), k
    # the kept tetrahedra fill the fluid domain (unit cube minus the sphere r = 0.3 /
    # 2.0 * 0.564 scale does not change the geometry) up to the wall layer
    vol = 1.0 - 4.0 / 3.0 * np.pi * 0.3 ** 3
    assert 0.85 * vol < st["total_volume"] < 1.02 * vol
    # boundary triangles lie on the sphere
    tris = smg_mod.tet_surface(job)
    c = s["x"][tris].mean(axis=1)
    r = np.linalg.norm(c - 0.5, axis=1)
    assert np.median(np.abs(r - 0.3)) < 0.01
    # a second call gives the same mesh
    kept2, st2 = smg_mod.tetrahedralize(job)
    for k in ("cell_ms", "total_ms"):
        st.pop(k), st2.pop(k)
    assert np.array_equal(kept, kept2) and st2 == st


def test_tetrahedralize_is_3d_only(smg_mod, c1):
    cfg, lset = c1
    job = smg_mod.Job(cfg, lset)
    job.seed(int(lset["stats"].num_particle))
    with pytest.raises(smg_mod.SmgError) as e:
        smg_mod.tetrahedralize(job)
    assert e.value.rc == -1
