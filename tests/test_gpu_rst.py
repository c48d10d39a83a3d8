"""A relaxation started from a level-set restart file equals the relaxation started
from the level set the file was written from, bit for bit (2D, the reference's tanh
spacing profile with uniform dx)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_relaxation_from_restart_file_equals_original(smg_mod, c1, tmp_path):
    cfg, lset = c1
    path = tmp_path / "levelset.rst"
    smg_mod.write_levelset_rst(str(path), cfg, lset)
    cfg2, lset2, info, _ = smg_mod.read_levelset_rst(str(path))
    cfg2.seed = cfg.seed
    n = int(lset["stats"].num_particle)
    assert int(lset2["stats"].num_particle) == n
    out = []
    for c, ls in ((cfg, lset), (cfg2, lset2)):
        job = smg_mod.Job(c, ls)
        job.seed(n)
        job.step(20)
        out.append(job.state())
        job.close()
    assert np.array_equal(out[0]["x"], out[1]["x"])
    assert np.array_equal(out[0]["type"], out[1]["type"])
    assert np.array_equal(out[0]["h"], out[1]["h"])
