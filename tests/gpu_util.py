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
