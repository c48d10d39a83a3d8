"""3D tetrahedralisation, host stage and oracle (no GPU): the oracle's shape measures
against closed forms, the library's deletion rule / boundary triangles / quality
statistics (smg_tet_finish) against the oracle on Qhull tetrahedralisations of random
points, and the deletion rule on planted slivers."""
import numpy as np
import pytest

import tri_oracle as T

PI_REF = 3.1415926


def test_oracle_regular_tetrahedron_closed_forms():
    p = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float)   # edge a = 2 sqrt 2
    s = T.tet_shapes(p, [[0, 1, 2, 3]])
    a = 2 * np.sqrt(2)
    assert s["vol"][0] == pytest.approx(a ** 3 / (6 * np.sqrt(2)), rel=1e-14)
    assert s["circumradius"][0] == pytest.approx(a * np.sqrt(3.0 / 8.0), rel=1e-14)
    assert s["aspect"][0] == pytest.approx(np.sqrt(1.5), rel=1e-14)         # a / (a sqrt(2/3))
    assert s["radius_ratio"][0] == pytest.approx(1.0, rel=1e-14)           # 3 r_in / r_circ
    # the reference converts radians with PI = 3.1415926
    assert s["dih"][0] == pytest.approx(np.arccos(1.0 / 3.0) / PI_REF * 180.0, rel=1e-14)
    assert s["dih"][0] != pytest.approx(np.degrees(np.arccos(1.0 / 3.0)), rel=1e-9)


def test_oracle_corner_tetrahedron_closed_forms():
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    s = T.tet_shapes(p, [[0, 1, 2, 3]])
    assert s["vol"][0] == pytest.approx(1.0 / 6.0, rel=1e-14)
    assert s["circumradius"][0] == pytest.approx(np.sqrt(3.0) / 2.0, rel=1e-14)
    # longest edge sqrt 2, smallest height 1/sqrt 3 (onto the slanted face)
    assert s["aspect"][0] == pytest.approx(np.sqrt(6.0), rel=1e-14)
    r_in = 3 * (1.0 / 6.0) / (1.5 + np.sqrt(3.0) / 2.0)
    assert s["radius_ratio"][0] == pytest.approx(3 * r_in / (np.sqrt(3.0) / 2.0), rel=1e-14)
    d = np.sort(s["dih"][0])
    assert d[:3] == pytest.approx([np.arccos(1 / np.sqrt(3)) / PI_REF * 180.0] * 3, rel=1e-13)
    assert d[3:] == pytest.approx([90.0 * np.pi / PI_REF] * 3, rel=1e-13)


def _random_case(seed, n=400):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 3))
    types = (rng.random(n) < 0.4).astype(np.int32)   # 40% SURFACE
    tets, _ = T.delaunay_tets(x)
    return x, types, tets


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_finish_equals_oracle(smg_mod, seed):
    x, types, tets = _random_case(seed)
    kept, tris, st = smg_mod.tet_finish(x, types, tets)
    rk, rt, rdel, q = T.tet_finish(x, types, tets)
    assert st["n_deleted"] == rdel > 0
    assert [tuple(t) for t in kept] == rk
    assert [tuple(t) for t in tris] == rt
    assert st["n_surface_tri"] == len(rt)
    assert st["n_tet"] == q["n_tet"] == len(rk)
    for k in ("dihedral_hist", "aspect_hist", "radius_hist", "n_min_dihedral_below"):
        assert st[k] == q[k], k
    for k in ("total_volume", "min_volume", "max_volume", "min_edge", "max_edge", "min_aspect", "max_aspect",
              "avg_aspect", "min_radius_ratio", "max_radius_ratio", "avg_radius_ratio", "min_dihedral",
              "max_dihedral", "avg_min_dihedral"):
        assert st[k] == pytest.approx(q[k], rel=1e-9, abs=1e-12), k
    # the Delaunay tetrahedra of random points in the unit cube tile their convex hull
    assert T.tet_quality(x, tets)["total_volume"] == pytest.approx(
        __import__("scipy.spatial", fromlist=["ConvexHull"]).ConvexHull(x).volume, rel=1e-12)


def test_deletion_rule_on_planted_slivers(smg_mod):
    # a flat sliver (aspect >> 10) and a regular tetrahedron
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.01],
                  [5, 5, 5], [6, 6, 5], [6, 5, 6], [5, 6, 6]], float)
    tets = [[0, 1, 2, 3], [4, 5, 6, 7]]
    surf = np.ones(8, np.int32)
    kept, tris, st = smg_mod.tet_finish(x, surf, tets)
    assert [tuple(t) for t in kept] == [(4, 5, 6, 7)] and st["n_deleted"] == 1
    assert len(tris) == 0
    # one REAL particle keeps the sliver, and its non-REAL face is a boundary triangle
    mixed = surf.copy()
    mixed[3] = 0
    kept, tris, st = smg_mod.tet_finish(x, mixed, tets)
    assert st["n_deleted"] == 0 and len(kept) == 2
    assert [tuple(t) for t in tris] == [(0, 1, 2)]
    # the threshold is a parameter (the reference's tet_delete_thres)
    kept, tris, st = smg_mod.tet_finish(x, surf, tets, delete_thres=1e6)
    assert st["n_deleted"] == 0


def test_finish_rejects_bad_vertices(smg_mod):
    x = np.zeros((4, 3))
    with pytest.raises(smg_mod.SmgError):
        smg_mod.tet_finish(x, np.zeros(4, np.int32), [[0, 1, 2, 9]])
