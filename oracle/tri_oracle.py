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
