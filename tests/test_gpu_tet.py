"""3D tetrahedralisation on the GPU path: the GPU's Delaunay tetrahedra equal Qhull's
(scipy.spatial.Delaunay) on the same positions, as a set -- hull slivers and
tetrahedra spanning the cavity included, so they tile the convex hull -- their
circumspheres are empty (brute force), and the deletion rule, boundary triangles and
quality statistics equal the oracle's on them."""
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
    from scipy.spatial import ConvexHull
    s = job.state()
    x = s["x"]
    n = len(x)
    kept, st = smg_mod.tetrahedralize(job)
    g, status, rcert = smg_mod.tet_delaunay(job, n)
    print(f"n={n} delaunay={len(g)} kept={len(kept)} second pass={st['n_second_pass']} open={st['n_open']} "
          f"overflow={st['n_cell_overflow']} duplicate={st['n_duplicate']} status={np.bincount(status)} "
          f"cell_ms={st['cell_ms']:.1f}", flush=True)
    assert st["n_cell_overflow"] == 0 and st["n_open"] == 0
    assert st["n_second_pass"] > 0   # the hull and cavity cells took the verifying pass
    assert st["n_delaunay"] == len(g)
    assert np.all(np.diff(g[:, 0]) >= 0) and np.all(g[:, :-1] < g[:, 1:])   # sorted, ascending ids
    # one vertex per point: coincident particles (the seeding's clamp to the level-set
    # box puts a few on each corner point) keep the smallest id
    _, first = np.unique(x, axis=0, return_index=True)
    first = np.sort(first)
    assert st["n_duplicate"] == n - len(first) == int((status == 6).sum())
    assert np.all(status[first] == 0)
    ql, rad = tri_oracle.delaunay_tets(x[first])
    q = np.sort(first[ql], axis=1)
    gs = set(map(tuple, g.tolist()))
    qs = set(map(tuple, q.tolist()))
    assert len(gs) == len(g)
    extra, missing = gs - qs, qs - gs
    if extra or missing:
        rq = dict(zip(map(tuple, q.tolist()), rad))
        print(f"extra {len(extra)} missing {len(missing)}; missing circumradii "
              f"{sorted(rq[t] for t in missing)[:10]}", flush=True)
    # the whole Delaunay tetrahedralisation of the points, hull slivers and tetrahedra
    # spanning the cavity included
    assert not extra and not missing
    # it tiles the convex hull
    P = x[g]
    vol = np.abs(np.einsum("ij,ij->i", P[:, 1] - P[:, 0], np.cross(P[:, 2] - P[:, 0], P[:, 3] - P[:, 0]))) / 6.0
    assert vol.sum() == pytest.approx(ConvexHull(x[first]).volume, rel=1e-9)
    return x, s, g, kept, st


def test_relaxed_delaunay_equals_qhull(smg_mod, c2_small):
    job = _run(smg_mod, c2_small, 100)
    x, s, g, kept, st = _check_against_qhull(smg_mod, job)
    # the ring search alone closes the cells away from the walls and the cavity (there
    # the Voronoi cells are unbounded or span the cavity): those take the second pass
    _, status, rcert = smg_mod.tet_delaunay(job, len(x))
    wall = np.minimum(x, 1.0 - x).min(axis=1)
    phi = np.linalg.norm(x - 0.5, axis=1) - 0.3
    interior = (wall > 0.15) & (phi > 0.15)
    assert interior.sum() > 1000
    assert np.all(rcert[interior] < 1e300)
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
    # the kept tetrahedra fill the fluid domain (unit cube minus the sphere r = 0.3;
    # the 2.0 * 0.564 scale does not change the geometry) up to the wall layer, and, as
    # the reference's deletion rule keeps the tetrahedra spanning the cavity unless
    # they are flat (all four particles non-REAL with aspect > 10), at most the convex
    # hull of the particles
    from scipy.spatial import ConvexHull
    vol = 1.0 - 4.0 / 3.0 * np.pi * 0.3 ** 3
    assert 0.85 * vol < st["total_volume"] <= ConvexHull(s["x"]).volume * (1 + 1e-9)
    # boundary triangles (the non-REAL face of a kept tetrahedron with one REAL
    # particle) have their corners on the sphere; with the tetrahedra that span the
    # cavity kept, some of them cut across it
    tris = smg_mod.tet_surface(job)
    assert np.all(s["type"][tris] != smg_mod.SMG_REAL)
    r = np.linalg.norm(s["x"][tris] - 0.5, axis=2)
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
