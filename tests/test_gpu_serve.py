"""Serving (continuous batching): jobs admitted into a device pool between chunks of
steps give the same bits as the same job alone in a pool and as the plain
single-job path; admission refuses what does not fit and succeeds once retired jobs
free their ranges."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CHUNK = 10


@pytest.fixture(scope="module")
def jobs3(smg_mod):
    return smg_mod.batched_jobs(5, 1500, 9000, seed=21, dims=(3,), max_pkg=10)


def _server(smg_mod, dim, max_jobs=8, slots=1 << 17, table=1 << 22, cells=1 << 20, **kw):
    kw.setdefault("chunk_steps", CHUNK)
    kw.setdefault("min_steps", 2 * CHUNK)
    kw.setdefault("max_steps", 6 * CHUNK)
    kw.setdefault("tol", 2e-3)
    return smg_mod.Server(smg_mod.server_config(dim, max_jobs, slots, table, cells, **kw))


def _run_until_done(srv, tickets, admit_later=()):
    """Advance until every ticket finished; admit_later = [(after_chunk, cfg, ls, n)]."""
    done = {}
    pending = list(admit_later)
    k = 0
    while len(done) < len(tickets) or pending:
        for item in [p for p in pending if p[0] <= k]:
            t = srv.admit(*item[1:])
            assert t is not None
            tickets.append(t)
            pending.remove(item)
        srv.advance()
        k += 1
        for f in srv.finished():
            done[f["ticket"]] = f
        assert k < 1000
    return done


def test_alone_equals_in_the_mix(smg_mod, jobs3):
    cfg, ls, n = jobs3[0]
    alone = _server(smg_mod, 3)
    t = alone.admit(cfg, ls, n)
    info = _run_until_done(alone, [t])[t]
    ra = alone.result(t)
    alone.close()

    mix = _server(smg_mod, 3)
    tickets = [mix.admit(*jobs3[1]), mix.admit(*jobs3[2])]
    # the job of interest arrives after one chunk, another after three
    later = [(1, cfg, ls, n), (3, *jobs3[3])]
    done = _run_until_done(mix, tickets, later)
    tm = tickets[2]
    rm = mix.result(tm)
    assert done[tm]["steps"] == info["steps"] and done[tm]["status"] == info["status"]
    assert done[tm]["window_residual"] == info["window_residual"]
    for k in ("x", "type", "h"):
        assert np.array_equal(ra[k], rm[k]), k
    for k in ("dt1", "dt2", "max_v", "residual", "n_neighbor_entries"):
        assert ra["stats"][k] == rm["stats"][k], k
    mix.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_pool_equals_single_job_path(smg_mod, dim):
    """A job that runs to max_steps in the pool equals smg_seed_particles + smg_step."""
    cfg, ls, n = smg_mod.batched_jobs(2, 2000, 6000, seed=31 + dim, dims=(dim,), max_pkg=10)[0]
    srv = _server(smg_mod, dim, tol=0.0, max_steps=4 * CHUNK)
    t = srv.admit(cfg, ls, n)
    info = _run_until_done(srv, [t])[t]
    assert info["status"] == smg_mod.JOB_MAX_STEPS and info["steps"] == 4 * CHUNK
    r = srv.result(t)
    job = smg_mod.Job(cfg, ls)
    job.seed(n)
    job.step(4 * CHUNK)
    g = job.state()
    for k in ("x", "type", "h"):
        assert np.array_equal(r[k], g[k]), k
    st = job.stats()
    for k in ("dt1", "dt2", "residual", "res_max", "max_disp"):
        assert r["stats"][k] == st[k], k
    job.close()
    srv.close()


def test_admission_by_capacity(smg_mod, jobs3):
    """A pool that holds two of the jobs refuses the third until one retires."""
    sizes = [smg_mod.job_size(c, n)[0] for c, _, n in jobs3[:3]]
    slots = sizes[0] + sizes[1]
    srv = _server(smg_mod, 3, slots=slots, max_steps=2 * CHUNK, min_steps=2 * CHUNK, tol=0.0)
    t0 = srv.admit(*jobs3[0])
    t1 = srv.admit(*jobs3[1])
    assert t0 is not None and t1 is not None
    if sizes[2] > 0:
        assert srv.admit(*jobs3[2]) is None   # no free range now
    srv.advance()
    srv.advance()   # both reach max_steps and retire
    fin = srv.finished()
    assert sorted(f["ticket"] for f in fin) == [t0, t1]
    assert srv.usage()["used_slots"] == 0
    t2 = srv.admit(*jobs3[2])
    assert t2 is not None
    srv.advance()
    srv.advance()
    assert [f["ticket"] for f in srv.finished()] == [t2]
    for t in (t0, t1, t2):
        r = srv.result(t)
        assert np.isfinite(r["x"]).all()
    srv.close()


def test_admission_by_particle_limit(smg_mod, jobs3):
    """With a particle limit the pool admits a job only while the admitted particles
    stay within it (an empty pool admits any job)."""
    n0, n1 = jobs3[0][2], jobs3[1][2]
    srv = _server(smg_mod, 3, max_active_particles=n0 + n1 - 1, max_steps=2 * CHUNK, min_steps=2 * CHUNK,
                  tol=0.0)
    t0 = srv.admit(*jobs3[0])
    assert t0 is not None
    assert srv.admit(*jobs3[1]) is None
    srv.advance()
    srv.advance()
    assert [f["ticket"] for f in srv.finished()] == [t0]
    assert srv.admit(*jobs3[1]) is not None
    srv.close()
