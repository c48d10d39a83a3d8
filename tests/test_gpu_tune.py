"""The autotuner times every candidate list shape on scratch copies and returns the
fastest; a job built with the tuned shape gives the same particles as with the
default inline depth (only the schedule of list builds depends on the skin)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autotune_picks_the_fastest_candidate(smg_mod, c1):
    cfg, lset = c1
    n = int(lset["stats"].num_particle)
    r = smg_mod.autotune(cfg, lset, n, warmup=10, steps=20)
    assert len(r["tried"]) == 15   # 5 skins x 3 inline depths
    ms = [t[2] for t in r["tried"]]
    assert all(m > 0 for m in ms)
    assert r["ms_per_step"] == min(ms)
    assert (r["skin"], r["max_neighbors"]) == min(r["tried"], key=lambda t: t[2])[:2]
    # the inline depth does not change results
    outs = []
    for K in (64, r["max_neighbors"]):
        c = smg_mod.SmgConfig.from_buffer_copy(cfg)
        c.skin = r["skin"]
        c.max_neighbors = K
        job = smg_mod.Job(c, lset)
        job.seed(n)
        job.step(30)
        outs.append(job.state()["x"])
        job.close()
    assert np.array_equal(outs[0], outs[1])
