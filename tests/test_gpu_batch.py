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


@pytest.mark.parametrize("warm", [5, 4])
def test_graph_captured_at_any_parity_equals_eager(smg_mod, c1, warm):
    """A single job's step loop captured after an odd (or even) number of eager steps
    replays the same computation as eager steps, bit for bit."""
    cfg, lset = c1
    n = int(lset["stats"].num_particle)
    eager = smg_mod.Job(cfg, lset)
    eager.seed(n)
    eager.step(warm + 8)
    graph = smg_mod.Job(cfg, lset)
    graph.seed(n)
    graph.step(warm)
    graph.capture(4)
    graph.step_async(8)
    graph.sync()
    a, g = eager.state(), graph.state()
    for k in ("x", "v", "type", "h", "nu"):
        assert np.array_equal(a[k], g[k]), k
    sa, sg = eager.stats(), graph.stats()
    assert sa["step"] == sg["step"] == warm + 8
    for k in ("dt1", "dt2", "max_v", "residual", "n_rebuild"):
        assert sa[k] == sg[k], k
    eager.close()
    graph.close()


def test_graph_replay_at_other_parity_is_refused(smg_mod, c1):
    """After one eager step the step parity differs from the captured one: a replay
    would read the other dt slots, so it is refused with SMG_ERR_STATE."""
    cfg, lset = c1
    job = smg_mod.Job(cfg, lset)
    job.seed(int(lset["stats"].num_particle))
    job.step(3)
    job.capture(2)
    job.step(1)
    with pytest.raises(smg_mod.SmgError) as e:
        job.step_async(2)
    assert e.value.rc == -4 and "parity" in str(e.value)
    job.step(1)
    job.step_async(2)   # same parity again: accepted
    job.sync()
    assert job.stats()["step"] == 7
    job.close()
