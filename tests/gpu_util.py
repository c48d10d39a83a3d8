"""Helpers shared by the GPU parity tests."""
import numpy as np
import pytest
from scipy.spatial import cKDTree


def df_round(x):
    x = np.asarray(x, np.float64)
    hi = x.astype(np.float32)
    lo = (x - hi.astype(np.float64)).astype(np.float32)
    return hi.astype(np.float64) + lo.astype(np.float64)


def oracle_state_from_gpu(oracle, job):
    s = job.state()
    st = job.stats()
    dt = st["dt2"] if st["step"] > 0 else 0.0
    return s, oracle.State.from_arrays(s["x"], s["v"], s["h"], s["nu"].astype(np.float64), s["type"], dt)


def coincident(x):
    """Particles that share their exact position with another one.  Seeding clamps
    positions into the level-set box shrunk by one package (sph_mesh_generation.cpp:
    1302-1367), so the few seeds drawn in a corner region land on the same point; from
    there a pair's relative motion is the difference of two sums of the same terms,
    which depends only on the order of the floating-point additions (TBB's in the
    reference), so these particles are left out of spacing comparisons."""
    _, inv, cnt = np.unique(np.asarray(x), axis=0, return_inverse=True, return_counts=True)
    return cnt[inv.ravel()] > 1


def spacing_stats(x, h, types, cfg, dim, exclude=None):
    """d_nn / dx over REAL and SURFACE particles (minus exclude): mean, std, p1, p99, min."""
    keep = (types == 0) | (types == 1)
    if exclude is not None:
        keep = keep & ~exclude
    pts = x[:, :dim]
    tree = cKDTree(pts)
    d, _ = tree.query(pts, k=2)
    dx = h * cfg.cut_off / cfg.cell_ratio
    r = (d[:, 1] / dx)[keep]
    return {"mean": r.mean(), "std": r.std(), "p1": np.percentile(r, 1), "p99": np.percentile(r, 99), "min": r.min()}


def disp_check(g1, st, ref):
    """v dt2 against the oracle: max-norm bar, per-particle relative error percentiles
    over the particles that move (|d| above 1e-3 of the largest)."""
    dg = g1["v"] * st["dt2"]
    do = ref.v * ref.dt
    scale = np.abs(do).max()
    assert np.abs(dg - do).max() <= 1e-5 * scale
    ng = np.linalg.norm(dg - do, axis=1)
    no = np.linalg.norm(do, axis=1)
    mv = no > 1e-3 * no.max()
    rel = ng[mv] / no[mv]
    return {"max_over_max": float(np.abs(dg - do).max() / scale), "rel_p50": float(np.median(rel)),
            "rel_p99": float(np.percentile(rel, 99)), "rel_max": float(rel.max()), "moving": int(mv.sum())}


def step_parity(orc, cfg, lset, job, steps, terms=True, h_ulps=0):
    """steps single steps, each from the engine's state: the oracle's step against the
    engine's.  Returns per-step diagnostics.  h_ulps > 0: h may differ by that many ulp
    (the tanh profile: the device tanh is within 2 ulp of the host's); the neighbour
    sets are then checked against the oracle's search on the engine's own positions
    and h, since a pair at the cutoff follows h's last bit."""
    out = []
    for _ in range(steps):
        g0, ref = oracle_state_from_gpu(orc, job)
        idx = job.stats()["step"] + 1
        d = orc.step(cfg, lset, ref, idx, want_terms=terms, want_nb=True)
        job.step(1)
        g1 = job.state()
        st = job.stats()
        offs, ids, codes = job.neighbors()
        if h_ulps == 0:
            assert np.array_equal(offs, d["offs"]) and np.array_equal(ids, d["ids"]) and np.array_equal(codes, d["codes"])
            assert np.array_equal(g1["h"], ref.h)
        else:
            bo, bi, bc = orc.neighbors(cfg, g1["x"], g1["h"])
            assert np.array_equal(offs, bo) and np.array_equal(ids, bi) and np.array_equal(codes, bc)
            assert np.all(np.abs(g1["h"] - ref.h) <= h_ulps * np.spacing(ref.h))
        assert np.array_equal(g1["type"], ref.type)
        assert np.array_equal(g1["x"], ref.x)
        if terms:
            t = job.pair_terms()
            for key in ("sum", "a1", "a2"):
                scale = np.abs(d[key]).max()
                if scale > 0:
                    assert np.abs(t[key] - d[key]).max() <= 2e-6 * scale, key
        assert st["dt2"] == pytest.approx(d["dt2"], rel=1e-5)
        assert st["n_remapped"] == d["n_remapped"]
        r = disp_check(g1, st, ref)
        r.update(step=idx, n_surface=st["n_surface"], n_remapped=st["n_remapped"])
        print(r, flush=True)
        out.append(r)
    return out
