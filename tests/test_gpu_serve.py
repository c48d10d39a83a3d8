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


def _refined_case(smg_mod):
    # 3D multi-resolution job: base 16^3, refined twice in a band (partial bands)
    cfg, sdf = smg_mod.config_c3(num_pkg=4, dx0=0.012, ratio=2.0, band=0.2)
    base = smg_mod.make_levelset(cfg, sdf)
    lv = smg_mod.make_levelset_levels(cfg, sdf, 2, band=2.0)
    n = int(lv[-1]["stats"].num_particle)
    maps = sum(int(np.prod([L["cfg"].num_pkg[a] + 2 for a in range(3)])) for L in lv)
    return cfg, base, lv, n, maps


def test_refined_level_set_job_in_the_pool(smg_mod, jobs3):
    """A job with refined level-set levels, admitted next to single-level jobs (and
    after another refined job retired, so its map range is reused), equals
    smg_set_levelset_levels + smg_seed_particles + smg_step bit for bit; a pool without
    map entries refuses it."""
    cfg, base, lv, n, maps = _refined_case(smg_mod)
    print(f"refined job: n = {n}, map entries = {maps}", flush=True)
    no_maps = _server(smg_mod, 3)
    with pytest.raises(smg_mod.SmgError):
        no_maps.admit(cfg, base, n, levels=lv)
    no_maps.close()

    srv = _server(smg_mod, 3, slots=1 << 18, cells=1 << 22, tol=0.0, max_steps=4 * CHUNK, max_lset_maps=2 * maps)
    first = srv.admit(cfg, base, n, levels=lv)
    other = srv.admit(*jobs3[0])
    assert first is not None and other is not None
    done = _run_until_done(srv, [first, other])
    assert done[first]["steps"] == 4 * CHUNK
    srv.result(first)
    srv.result(other)
    # a second refined job with another seed, mixed with two single-level jobs
    cfg2 = smg_mod.SmgConfig.from_buffer_copy(cfg)
    cfg2.seed = cfg.seed + 17
    tickets = [srv.admit(*jobs3[1]), srv.admit(cfg2, base, n, levels=lv), srv.admit(*jobs3[2])]
    assert None not in tickets
    done = _run_until_done(srv, tickets)
    r = srv.result(tickets[1])
    assert done[tickets[1]]["status"] == smg_mod.JOB_MAX_STEPS
    job = smg_mod.Job(cfg2, base, levels=lv)
    job.seed(n)
    job.step(4 * CHUNK)
    g = job.state()
    for k in ("x", "type", "h"):
        assert np.array_equal(r[k], g[k]), k
    st = job.stats()
    for k in ("dt1", "dt2", "residual", "residual_sum"):
        assert r["stats"][k] == st[k], k
    job.close()
    srv.close()


def test_window_residual_from_residual_sum(smg_mod, jobs3):
    """The server's chunk window residual is the difference of smg_stats.residual_sum
    over the chunk divided by its steps: the plain path can apply the same test."""
    cfg, ls, n = jobs3[4]
    srv = _server(smg_mod, 3, tol=0.0, max_steps=2 * CHUNK)
    t = srv.admit(cfg, ls, n)
    info = _run_until_done(srv, [t])[t]
    srv.result(t)
    srv.close()
    job = smg_mod.Job(cfg, ls)
    job.seed(n)
    job.step(CHUNK)
    s1 = job.stats()["residual_sum"]
    job.step(CHUNK)
    s2 = job.stats()["residual_sum"]
    job.close()
    assert info["window_residual"] == pytest.approx((s2 - s1) / CHUNK, rel=1e-12, abs=0)


def test_odd_chunk_steps_eagerly(smg_mod, jobs3):
    """An odd chunk cannot replay the two-step graph body; it steps eagerly and still
    equals the single-job path."""
    cfg, ls, n = jobs3[0]
    srv = _server(smg_mod, 3, chunk_steps=7, min_steps=7, tol=0.0, max_steps=21)
    t = srv.admit(cfg, ls, n)
    info = _run_until_done(srv, [t])[t]
    assert info["steps"] == 21
    r = srv.result(t)
    srv.close()
    job = smg_mod.Job(cfg, ls)
    job.seed(n)
    job.step(21)
    g = job.state()
    job.close()
    for k in ("x", "type", "h"):
        assert np.array_equal(r[k], g[k]), k
