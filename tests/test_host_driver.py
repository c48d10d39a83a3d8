"""The C++ host mirror (Initialization / Solver / SPH_mesh_generation over the C ABI,
sph-mesh-generation_amd/host): it builds with the library, fails loudly without a
GPU, and on the GPU gives the same particles as the Python binding of the same case."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "sph-mesh-generation_amd", "lib", "smg_run")


def _driver():
    if not os.path.exists(DRIVER):
        sys.path.insert(0, os.path.join(ROOT, "sph-mesh-generation_amd"))
        import build_lib
        build_lib.build_driver()
    return DRIVER


def test_driver_fails_loudly_without_gpu():
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    out = subprocess.run([_driver(), "--outputs", "1", "--output-dt", "2"], env=env, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 1
    assert "ERROR in smg_seed_particles (-2)" in out.stderr


@pytest.mark.gpu
def test_driver_equals_python_binding(smg_mod, tmp_path):
    csv = tmp_path / "x.csv"
    out = subprocess.run([_driver(), "--outputs", "2", "--output-dt", "10", "--csv", str(csv)], capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert sum(line.startswith("output ") for line in out.stdout.splitlines()) == 2
    data = np.loadtxt(csv, delimiter=",", skiprows=1)
    cfg, sdf = smg_mod.config_c1()
    lset = smg_mod.make_levelset(cfg, sdf)
    job = smg_mod.Job(cfg, lset)
    job.seed(int(lset["stats"].num_particle))
    job.step(20)
    g = job.state()
    assert data.shape[0] == len(g["h"])
    assert np.array_equal(data[:, 1:3], g["x"][:, :2])
    assert np.array_equal(data[:, 4].astype(np.int32), g["type"])
