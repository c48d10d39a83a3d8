"""Batched relaxation: every job gives the same result alone and in the batch (bit
for bit: the per-job kernels see the same data and reduce per job), and graph
replays equal eager steps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def jobs(smg_mod):
    return smg_mod.batched_jobs(6, 800, 6000, seed=11, max_pkg=12)


def test_batch_equals_single_jobs(smg_mod, jobs):
    b = smg_mod.Batch(jobs)
    b.seed()
    b.step(20)
    for j, (cfg, ls, n) in enumerate(jobs):
        alone = smg_mod.Job(cfg, ls)
        alone.seed(n)
        alone.step(20)
        sa, sb = alone.state(), b.state(j)
        for k in ("x", "v", "type", "h", "nu"):
            assert np.array_equal(sa[k], sb[k]), (j, k)
        ta, tb = alone.stats(), b.stats(j)
        for k in ("dt1", "dt2", "max_v", "residual", "n_neighbor_entries"):
            assert ta[k] == tb[k], (j, k)
        alone.close()
    b.close()


def test_graph_replay_equals_eager(smg_mod, jobs):
    eager = smg_mod.Batch(jobs)
    eager.seed()
    eager.step(12)
    graph = smg_mod.Batch(jobs)
    graph.seed()
    graph.capture(4)
    graph.run(3)
    graph.sync()
    for j in range(len(jobs)):
        a, g = eager.state(j), graph.state(j)
        assert np.array_equal(a["x"], g["x"]), j
        assert np.array_equal(a["v"], g["v"]), j
    eager.close()
    graph.close()
