"""smg_sample_phi_static (the entry point that allocates nothing) against
smg_sample_phi, through the C ABI: bit-identical results on two calls in a row, the
second smaller than the first, and a clear capacity error past its staging array."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pts(n, dim, seed):
    p = np.zeros((n, 3))
    p[:, :dim] = np.random.default_rng(seed).uniform(-0.05, 1.05, size=(n, dim))
    return p


@pytest.mark.parametrize("case", ["2d", "3d"])
def test_static_sampling_equals_allocating_sampling(smg_mod, case):
    if case == "2d":
        cfg, sdf = smg_mod.config_c1()
        dim = 2
    else:
        cfg, sdf = smg_mod.config_c2(scale=2.0, num_pkg=16)
        dim = 3
    lset = smg_mod.make_levelset(cfg, sdf)
    job = smg_mod.Job(cfg, lset)
    job.seed(int(lset["stats"].num_particle))
    L = smg_mod.lib()
    for n, seed in ((200000, 1), (5003, 2)):   # the second call smaller than the first
        p = _pts(n, dim, seed)
        ref = np.zeros(n)
        out = np.full(n, np.nan)
        assert L.smg_sample_phi(job.h, n, p.ctypes.data_as(C.c_void_p), ref.ctypes.data_as(C.c_void_p)) == 0
        rc = L.smg_sample_phi_static(job.h, n, p.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
        assert rc == 0, L.smg_last_error()
        assert np.array_equal(out, ref)
    job.close()


def test_static_sampling_rejects_oversize_input(smg_mod, c1):
    cfg, lset = c1
    job = smg_mod.Job(cfg, lset)
    job.seed(100)
    n = 262145
    p = _pts(n, 2, 3)
    out = np.zeros(n)
    L = smg_mod.lib()
    rc = L.smg_sample_phi_static(job.h, n, p.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
    assert rc == -3 and b"staging array" in L.smg_last_error()
    assert L.smg_sample_phi_static(job.h, 0, None, None) == 0
    job.close()
