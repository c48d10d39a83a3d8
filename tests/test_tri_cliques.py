"""The triangulation's integer stage (3-cliques, i > j > k) and the oracle's Voronoi
restatement, on the CPU: random graphs against a brute-force triple loop, and the
Voronoi faces of a hexagonal lattice (six per interior point) as a closed form."""
import itertools
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def _brute(n, edges):
    E = {(max(a, b), min(a, b)) for a, b in edges if a != b}
    return sorted((i, j, k) for i, j, k in itertools.combinations(range(n - 1, -1, -1), 3)
                  if (i, j) in E and (i, k) in E and (j, k) in E)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_cliques_equal_brute_force(smg_mod, seed):
    rng = np.random.default_rng(seed)
    n = 40
    pairs = [(a, b) for a in range(n) for b in range(n) if rng.random() < 0.15]
    pairs += [(3, 3), (5, 7), (7, 5), (5, 7)]   # a self loop and repeats are ignored
    got = [tuple(t) for t in smg_mod.tri_cliques(n, np.array(pairs))]
    assert got == _brute(n, pairs)
    import tri_oracle
    assert tri_oracle.cliques(n, pairs) == _brute(n, pairs)


def test_cliques_reject_bad_vertex(smg_mod):
    with pytest.raises(smg_mod.SmgError):
        smg_mod.tri_cliques(3, np.array([[0, 5]]))


def test_oracle_voronoi_hex_lattice(smg_mod):
    """Interior points of a hexagonal lattice have exactly their six lattice neighbours
    as Voronoi faces."""
    import tri_oracle
    cfg = smg_mod.default_config(2)
    a = 0.02
    pts = []
    for r in range(30):
        for c in range(30):
            pts.append((0.2 + c * a + (r % 2) * a / 2, 0.2 + r * a * np.sqrt(3) / 2))
    x = np.zeros((len(pts), 3))
    x[:, :2] = pts
    h = np.full(len(pts), a)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    offs, ids, codes = oracle.neighbors(cfg, x, h, brute=True)
    edges = tri_oracle.voronoi_edges(cfg, x, h, offs, ids, codes)
    deg = np.zeros(len(pts), int)
    for i, j in edges:
        deg[i] += 1
        deg[j] += 1
    interior = [r * 30 + c for r in range(2, 28) for c in range(2, 28)]
    assert np.all(deg[interior] == 6)
    for i in interior:
        nb = {j for e in edges for j in e if i in e and j != i}
        d = np.linalg.norm(x[list(nb), :2] - x[i, :2], axis=1)
        assert np.allclose(d, a)
