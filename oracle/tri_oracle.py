"""Restatement of the 2D triangulation for checking (TEST INFRASTRUCTURE ONLY).

* Voronoi faces of each particle's local cell (Voronoi_construction::Local_VD_generation,
  voronoi_construction.cpp:6-127): built here independently of the engine, from Qhull
  (scipy.spatial.Voronoi) on the particle and its neighbours (mirror images at their
  image positions), each ridge of the particle clipped to the box x +- 1.5 cut_off h
  that the reference gives voro++.  A face with a mirror image counts as its source
  particle; the face with the particle's own image is dropped.
* The integer stage of Graphmeshcls::Reconstruct_tris (graph_mesh.cpp:37-407):
  3-cliques by set intersection, the centroid test of all-non-REAL triangles with the
  oracle's FP64 level-set sampler, the rebuilt graph and its 3-cliques.
* Tri_quality_statistics (graph_mesh.cpp:411-671) in numpy.
"""
import numpy as np

SMG_REAL = 0


def _clip_segment(p0, p1, lo, hi):
    """Liang-Barsky: does the segment p0-p1 meet the box [lo, hi]?"""
    t0, t1 = 0.0, 1.0
    d = p1 - p0
    for a in range(2):
        for p, q in ((-d[a], p0[a] - lo[a]), (d[a], hi[a] - p0[a])):
            if p == 0:
                if q < 0:
                    return False
            else:
                r = q / p
                if p < 0:
                    t0 = max(t0, r)
                else:
                    t1 = min(t1, r)
                if t0 > t1:
                    return False
    return True


def image(cfg, x, code):
    g = np.array(x, float)
    for a in range(2):
        m = (code >> (2 * a)) & 3
        if m == 1:
            g[a] = 2 * cfg.box_l[a] - x[a]
        elif m == 2:
            g[a] = 2 * (cfg.box_l[a] + cfg.domain[a]) - x[a]
    return g


def voronoi_edges(cfg, x, h, offs, ids, codes):
    """Undirected edges (i > j) from the local Voronoi cells of every particle."""
    from scipy.spatial import Voronoi
    n = len(h)
    edges = set()
    for i in range(n):
        sites, src = [np.asarray(x[i, :2], float)], [i]
        for k in range(offs[i], offs[i + 1]):
            j, c = int(ids[k]), int(codes[k])
            if j == i and c == 0:
                continue
            sites.append(image(cfg, x[j, :2], c))
            src.append(j)
        s = 1.5 * cfg.cut_off * h[i]
        lo, hi = x[i, :2] - s, x[i, :2] + s
        P = np.array(sites)
        if len(P) < 3:
            continue
        vor = Voronoi(P)
        center = P.mean(axis=0)
        for (a, b), rv in zip(vor.ridge_points, vor.ridge_vertices):
            if a != 0 and b != 0:
                continue
            other = b if a == 0 else a
            rv = np.asarray(rv)
            if np.all(rv >= 0):
                seg = vor.vertices[rv]
            else:
                v = vor.vertices[rv[rv >= 0][0]]
                t = P[other] - P[0]
                t = t / np.linalg.norm(t)
                nrm = np.array([-t[1], t[0]])
                mid = 0.5 * (P[other] + P[0])
                far = v + np.sign(np.dot(mid - center, nrm)) * nrm * (10.0 * s + np.linalg.norm(v - x[i, :2]))
                seg = np.array([v, far])
            if _clip_segment(seg[0], seg[1], lo, hi):
                j = src[other]
                if j != i:
                    edges.add((max(i, j), min(i, j)))
    return sorted(edges)


def cliques(n, edges):
    adj = [set() for _ in range(n)]
    for a, b in edges:
        if a != b:
            adj[a].add(b)
            adj[b].add(a)
    out = []
    for a in range(n):
        for b in adj[a]:
            if b >= a:
                continue
            for c in adj[a] & adj[b]:
                if c < b:
                    out.append((a, b, c))
    return sorted(out)


def reconstruct(cfg, lset, x, types, edges, sample):
    """The integer stage with the centroid test; sample(cfg, lset, pts) -> (phi, n, tag)."""
    n = len(types)
    t1 = cliques(n, edges)
    phi_p = sample(cfg, lset, x)[0]
    test = [t for t in t1 if all(types[v] != SMG_REAL for v in t)]
    cent = np.array([[1.0 / 3.0 * (x[t[0], a] + x[t[1], a] + x[t[2], a]) for a in range(3)] for t in test]).reshape(-1, 3)
    phi_c = sample(cfg, lset, cent)[0] if len(test) else np.zeros(0)
    dropped, deleted = set(), set()
    for t, pc in zip(test, phi_c):
        if pc > 0.0001 * (1.0 / 3.0 * (phi_p[t[0]] + phi_p[t[1]] + phi_p[t[2]])):
            continue
        dropped.add(t)
        v0, v1, v2 = x[t[0], :2], x[t[1], :2], x[t[2], :2]
        d, e = np.linalg.norm(v1 - v0), (t[0], t[1])
        d2 = np.linalg.norm(v2 - v1)
        if d2 > d:
            d, e = d2, (t[1], t[2])
        d3 = np.linalg.norm(v0 - v2)
        if d3 > d:
            d, e = d3, (t[0], t[2])
        deleted.add(e)
    e2 = set()
    for t in t1:
        if t in dropped:
            continue
        if (t[0], t[1]) in deleted or (t[0], t[2]) in deleted or (t[1], t[2]) in deleted:
            continue
        e2 |= {(t[0], t[1]), (t[1], t[2]), (t[0], t[2])}
    return cliques(n, sorted(e2)), len(deleted)


def quality(x, tris):
    tris = np.asarray(tris, int).reshape(-1, 3)
    v0, v1, v2 = x[tris[:, 0], :2], x[tris[:, 1], :2], x[tris[:, 2], :2]
    a1, a2, a3 = v1 - v0, v2 - v1, v0 - v2
    d1, d2, d3 = (np.sqrt((a * a).sum(1)) for a in (a1, a2, a3))
    rad = 180.0 / np.arccos(-1.0)
    al = np.stack([rad * np.arccos(-(a1 * a3).sum(1) / (d3 * d1 + 1e-20)),
                   rad * np.arccos(-(a1 * a2).sum(1) / (d2 * d1 + 1e-20)),
                   rad * np.arccos(-(a2 * a3).sum(1) / (d3 * d2 + 1e-20))], 1).clip(0, 180)
    area = 0.5 * (d3 * d1) * np.sin(np.arccos(-(a1 * a3).sum(1) / (d3 * d1 + 1e-20)))
    dmax = np.maximum(d1, np.maximum(d2, d3))
    G = 2 * np.sqrt(3.0) * area / (dmax * (d1 + d2 + d3) / 2)
    hist = np.zeros(45, int)
    np.add.at(hist, np.minimum(44, (al / 4.0).astype(int)).ravel(), 1)
    return {"n_tri": len(tris), "min_angle": al.min(), "max_angle": al.max(), "avg_min_angle": al.min(1).mean(),
            "min_edge": np.minimum(d1, np.minimum(d2, d3)).min(), "max_edge": dmax.max(), "total_area": area.sum(),
            "G_min": G.min(), "G_avg": G.mean(), "angle_hist": hist.tolist()}


# ---------------------------------------------------------------------------
# 3D: Graphmeshcls::Reconstruct_tets (graph_mesh.cpp:981-1963)

PI_REF = 3.1415926   # glbparam.h:185; Tet_quality_statistics converts angles with it
TET_DELETE_THRES = 10.0   # sph_mesh_generation.cpp:705


def tet_shapes(x, tets):
    """Shape measures of tetrahedra the way the reference takes them
    (Check_whether_this_tet_should_be_delete, graph_mesh.cpp:850-975;
    Tet_quality_statistics, graph_mesh.cpp:2102-2343): edge vectors V0..V5, the matrix
    A = [V0; V1; V2]; the inward face normals N_j = A^-1 e_j (j < 3), N_3 = -sum; the
    inverse heights H = |N|; the circumcentre A^-1 (|V_j|^2 / 2); dihedral angles
    acos(-N_a.N_b) / PI * 180 over the six face pairs.  Vectorised over tetrahedra."""
    x = np.asarray(x, float)
    t = np.asarray(tets, np.int64).reshape(-1, 4)
    p = x[t]                                   # (m, 4, 3)
    V = np.stack([p[:, 0] - p[:, 3], p[:, 1] - p[:, 3], p[:, 2] - p[:, 3],
                  p[:, 1] - p[:, 0], p[:, 2] - p[:, 1], p[:, 0] - p[:, 2]], axis=1)   # (m, 6, 3)
    el = np.einsum("mek,mek->me", V, V)
    A = V[:, :3, :]
    Ai = np.linalg.inv(A)                      # columns: N_0, N_1, N_2
    N = np.concatenate([np.transpose(Ai, (0, 2, 1)), -Ai.sum(axis=2)[:, None, :]], axis=1)   # (m, 4, 3)
    cc = np.einsum("mij,mj->mi", Ai, 0.5 * el[:, :3])
    cirr = np.sqrt(np.einsum("mi,mi->m", cc, cc))
    H = np.sqrt(np.einsum("mfk,mfk->mf", N, N))
    Nn = N / H[:, :, None]
    insr = 1.0 / H.sum(axis=1)
    dih = []
    for a in range(4):
        for b in range(a + 1, 4):
            c = np.clip(-np.einsum("mk,mk->m", Nn[:, a], Nn[:, b]), -1.0, 1.0)
            dih.append(np.arccos(c) / PI_REF * 180.0)
    # orient3dfast(p1, p0, p2, p3) / 6
    vol = np.linalg.det(np.stack([p[:, 1] - p[:, 3], p[:, 0] - p[:, 3], p[:, 2] - p[:, 3]], axis=1)) / 6.0
    return {"vol": np.abs(vol), "l2min": el.min(axis=1), "l2max": el.max(axis=1),
            "aspect": np.sqrt(el.max(axis=1)) * H.max(axis=1), "radius_ratio": 3.0 * insr / cirr,
            "dih": np.stack(dih, axis=1), "circumradius": cirr, "circumcentre": p[:, 3] + cc}


def tet_finish(x, types, tets, thres=TET_DELETE_THRES):
    """The deletion rule (graph_mesh.cpp:1126-1128 with 675-979: four non-REAL
    particles and aspect > thres; the phi tests there do not decide), the boundary
    triangles (graph_mesh.cpp:1153-1176: the non-REAL face of a tetrahedron with exactly
    one REAL particle, ids ascending) and Tet_quality_statistics over the kept ones."""
    t = np.asarray(tets, np.int64).reshape(-1, 4)
    ty = np.asarray(types)
    nreal = (ty[t] == SMG_REAL).sum(axis=1)
    s = tet_shapes(x, t) if len(t) else None
    dele = (nreal == 0) & (s["aspect"] > thres) if len(t) else np.zeros(0, bool)
    kept = t[~dele]
    one = kept[nreal[~dele] == 1]
    surf = set()
    for q in one:
        surf.add(tuple(sorted(int(v) for v in q if int(ty[v]) != SMG_REAL)))
    return [tuple(int(v) for v in q) for q in kept], sorted(surf), int(dele.sum()), tet_quality(x, kept)


def tet_quality(x, tets):
    """Tet_quality_statistics (graph_mesh.cpp:1967-2460): volumes, edges, aspect ratio
    (histogram bounds 1.5 ... 100), radius ratio (0.02 bins, int(1.1 / 0.02) + 1 of
    them, the last open), dihedral angles (4-degree bins; 180 lands in the last) and
    the counts of tetrahedra whose smallest dihedral angle is below 10/20/30/40."""
    table = [1.5, 2.0, 2.5, 3.0, 4.0, 6.0, 10.0, 15.0, 25.0, 50.0, 100.0]
    nrad = int(1.1 / 0.02) + 1
    t = np.asarray(tets, np.int64).reshape(-1, 4)
    q = {"n_tet": len(t), "dihedral_hist": [0] * 45, "aspect_hist": [0] * 12, "radius_hist": [0] * nrad,
         "n_min_dihedral_below": [0] * 4, "total_volume": 0.0}
    if not len(t):
        return q
    s = tet_shapes(x, t)
    d = s["dih"]
    for b in np.minimum(44, np.floor(d / 4.0).astype(int)).ravel():
        q["dihedral_hist"][b] += 1
    dm = d.min(axis=1)
    q["n_min_dihedral_below"] = [int((dm < lim).sum()) for lim in (10.0, 20.0, 30.0, 40.0)]
    ri = np.floor(s["radius_ratio"] / 0.02).astype(int)
    for b in np.where(ri < nrad - 1, ri, nrad - 1):
        q["radius_hist"][b] += 1
    for a in s["aspect"]:
        ai = 0
        while ai < 11 and a > table[ai]:
            ai += 1
        q["aspect_hist"][ai] += 1
    q.update(total_volume=float(s["vol"].sum()), min_volume=float(s["vol"].min()), max_volume=float(s["vol"].max()),
             min_edge=float(np.sqrt(s["l2min"].min())), max_edge=float(np.sqrt(s["l2max"].max())),
             min_aspect=float(s["aspect"].min()), max_aspect=float(s["aspect"].max()),
             avg_aspect=float(s["aspect"].mean()),
             min_radius_ratio=float(s["radius_ratio"].min()), max_radius_ratio=float(s["radius_ratio"].max()),
             avg_radius_ratio=float(s["radius_ratio"].mean()),
             min_dihedral=float(d.min()), max_dihedral=float(d.max()), avg_min_dihedral=float(dm.mean()))
    return q


def delaunay_tets(x):
    """All Delaunay tetrahedra of the points (Qhull through scipy.spatial.Delaunay),
    ascending ids, sorted, with their circumradii."""
    from scipy.spatial import Delaunay
    x = np.asarray(x, float)
    tri = Delaunay(x)
    s = np.sort(tri.simplices.astype(np.int64), axis=1)
    s = s[np.lexsort(s.T[::-1])]
    return s, tet_shapes(x, s)["circumradius"]
