"""Helpers shared by the GPU parity tests."""
import numpy as np
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


def spacing_stats(x, h, types, cfg, dim):
    """d_nn / dx over REAL and SURFACE particles: mean, std, p1, p99, min."""
    keep = (types == 0) | (types == 1)
    pts = x[:, :dim]
    tree = cKDTree(pts)
    d, _ = tree.query(pts, k=2)
    dx = h * cfg.cut_off / cfg.cell_ratio
    r = (d[:, 1] / dx)[keep]
    return {"mean": r.mean(), "std": r.std(), "p1": np.percentile(r, 1), "p99": np.percentile(r, 99), "min": r.min()}
