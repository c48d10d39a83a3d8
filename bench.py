"""Benchmarks of the MI355X relaxation engine.  Each mode prints one JSON line.

relax (default): 3D relaxation throughput in particle-steps/s.  Workload: unit cube
  minus a sphere r = 0.3 at the centre, level set on a 256^3 grid, target spacing
  linear in phi from 0.00451 at the surface to 0.01354 at phi >= 0.2 (~1.0M
  particles), seeded uniform-random positions; each timed step is the full
  relaxation step (mapping, drift test and -- where due -- sort and list rebuild,
  both force stages, both dt reductions, residual), timed as replays of a captured
  hipGraph of 10 steps; per-kernel times come from HIP events over an eager window
  of the same length.
batch: 512 mixed jobs (half 2D circle unions, half 3D sphere unions, N log-uniform
  in [1e3, 2e5]) run for a fixed number of steps, batched (one launch per pass for
  all jobs of a dimension, graph-replayed) against the plain path (one job at a time).
serve: Poisson arrivals of 3D sphere-union jobs (N log-uniform in [5e3, 5e5]) into a
  continuous-batching server, retired when converged; jobs/s and latency percentiles
  against the same arrivals served one job at a time.
c3: the multi-resolution case (12 ellipsoids, three-level level set, ~4M particles)
  run until converged; particle-steps/s and the steps and seconds it took.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sph-mesh-generation_amd"))

METRIC = "particle-steps/s (3D 1M relax); batched jobs/s + p99 latency vs plain GPU"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per particle and launch (3D, 2D), from the per-pass model: each
# particle field counted once per dependent pass, neighbour data assumed served from
# cache (P1 integrate + sample + map + h + key 252/156, P2 sort 156/116, P3 density +
# force1 44/32, P4 force2 60/44, P5 v/nu/residual 40/32; 552/380 per particle-step).
# k_map carries P1 and P5 (the velocity update and residual fold into it); the sort
# passes carry P2 and run only on the steps where a job rebuilds its lists.
STEP_BYTES = (552, 380)
BYTES = {
    "k_map": (252 + 40, 156 + 32),
    "k_force1": (44, 32),
    "k_force2": (60, 44),
    "k_key": (36, 28),
    "k_scatter": (16, 16),
    "k_segsort": (20, 20),
    "k_gather": (60, 40),
    "k_copyback": (60, 40),
    "k_nb": (32, 24),
}
# passes that do work only on the steps where a job rebuilds its lists: their launch
# covers n * (rebuilds / steps) particles on average
REBUILD_ONLY = {"k_key", "k_scatter", "k_segsort", "k_gather", "k_copyback", "k_nb"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(v, q):
    return float(np.percentile(np.asarray(v, float), q)) if len(v) else None


def cpu_baseline(cfg, lset, n, steps):
    """The FP64 oracle on this host's cores, a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    s = oracle.seed(cfg, lset, n)
    oracle.prerun(cfg, lset, s)
    oracle.step(cfg, lset, s, 1)
    t0 = time.perf_counter()
    for k in range(steps):
        oracle.step(cfg, lset, s, 2 + k)
    dt = time.perf_counter() - t0
    return {"value": n * steps / dt, "unit": "particle-steps/s", "cores": oracle.num_threads(), "kind": "port",
            "sample": f"{steps} steps of the same {n}-particle workload after seeding, pre-run and one step "
                      f"(FP64 oracle, OpenMP, {dt:.1f} s)"}


def relax(args, smg):
    t0 = time.perf_counter()
    cfg, sdf = smg.config_c2(scale=args.scale, num_pkg=args.num_pkg)
    if args.skin is not None:
        cfg.skin = args.skin
    lset = smg.make_levelset(cfg, sdf)
    n = int(lset["stats"].num_particle)
    log(f"level set built in {time.perf_counter() - t0:.1f} s, n = {n}")
    tune = None
    if args.skin is None and not args.no_tune:
        # list shape (skin, i.e. the cell-size multiple, and inline depth) from the
        # autotuner, timed on scratch copies over the same window (warmup, steps) as the
        # measurement below; tuning is not part of the timed region
        t0 = time.perf_counter()
        r = smg.autotune(cfg, lset, n, warmup=args.warmup, steps=args.steps)
        cfg.skin = r["skin"]
        cfg.max_neighbors = r["max_neighbors"]
        dflt = [t[2] for t in r["tried"] if t[0] == 0.1 and t[1] == 128]
        tune = {"skin": r["skin"], "max_neighbors": r["max_neighbors"], "seconds": time.perf_counter() - t0,
                "tuned_ms_per_step": r["ms_per_step"],
                "default_shape_ms_per_step": dflt[0] if dflt else None,
                "candidates": len(r["tried"])}
        log(f"autotune: {json.dumps(tune)}")
    job = smg.Job(cfg, lset, profile=False)
    t0 = time.perf_counter()
    job.seed(n)
    job.step(args.warmup)
    log(f"seeded + {args.warmup} warmup steps in {time.perf_counter() - t0:.1f} s")
    rb0 = job.stats()["n_rebuild"]
    G = args.graph_steps
    # whole graphs only, and at least one
    steps = max(G, -(-args.steps // G) * G)
    if steps != args.steps:
        log(f"--steps {args.steps} rounded up to {steps} (whole graphs of {G} steps)")
    # timed region: the step loop replayed as a captured hipGraph of G steps
    job.capture(G)
    job.sync()
    t0 = time.perf_counter()
    job.step_async(steps)
    job.sync()
    elapsed = time.perf_counter() - t0
    st = job.stats()
    rebuild_frac = (st["n_rebuild"] - rb0) / steps
    log(f"timed {steps} steps (graph of {G}) in {elapsed:.3f} s; stats {st}")
    # per-kernel times: HIP events around each kernel on the engine's stream, over an
    # eager window of the same length that continues the same relaxation
    rb1 = st["n_rebuild"]
    smg.lib().smg_profile_enable(job.h, 1)
    t1 = time.perf_counter()
    job.step(steps)
    elapsed_eager = time.perf_counter() - t1
    prof = job.profile()
    rebuild_frac_prof = (job.stats()["n_rebuild"] - rb1) / steps
    log("kernels (avg ms, launches): " + json.dumps({k: [round(v[0], 4), v[1]] for k, v in prof.items()}))
    value = n * steps / elapsed
    D = cfg.dim
    tot = {k: v[0] * v[1] for k, v in prof.items()}
    dom = max(tot, key=tot.get)
    avg_ms = prof[dom][0]
    units = n * (rebuild_frac_prof if dom in REBUILD_ONLY else 1.0)
    bpl = BYTES.get(dom, (0, 0))[0 if D == 3 else 1] * units
    achieved = bpl / (avg_ms * 1e-3) / 1e9
    # HBM bytes per launch of the dominant kernel from the committed rocprofv3 --pmc
    # FETCH_SIZE / WRITE_SIZE summary of this workload (copy-calibrated), if present
    traffic, traffic_src = None, None
    kin = cfg.max_neighbors if cfg.max_neighbors > 0 else 128   # 0: the library's 128

    def profiled_like_this(pr):
        # a committed per-launch count holds for the profiled list shape and step window only
        return (pr.get("skin") == cfg.skin and pr.get("max_neighbors") == kin
                and pr.get("steps") == steps and pr.get("warmup") == args.warmup)

    pmc = os.path.join(ROOT, "profiles", "pmc_traffic_relax3d_1M.json")
    if os.path.exists(pmc) and D == 3:
        with open(pmc) as f:
            pj = json.load(f)
        t = pj.get("kernels", {}).get(dom)
        if t and profiled_like_this(pj.get("profiled_run", {})):
            traffic = t["traffic_bytes"]
            traffic_src = "profiles/pmc_traffic_relax3d_1M.json (FETCH_SIZE + WRITE_SIZE per dispatch, copy-calibrated)"
    # the dominant kernel against the VALU issue rate: its wave64 VALU instructions per
    # launch from the committed rocprofv3 SQ_INSTS_VALU summary of this workload, over
    # its launch time measured here
    issue = None
    sqf = os.path.join(ROOT, "profiles", "sq_counters_relax3d_1M.json")
    if os.path.exists(sqf) and D == 3:
        with open(sqf) as f:
            sq = json.load(f)
        k = sq.get("kernels", {}).get(dom)
        if k and dom not in REBUILD_ONLY and profiled_like_this(sq.get("profiled_run", {})):
            rate = k["valu_instr_per_dispatch"] / (avg_ms * 1e-3)
            issue = {"bound": "valu-issue", "kernel": dom, "achieved": rate, "peak": sq["valu_issue_peak"],
                     "unit": "wave64 VALU instructions/s", "frac": rate / sq["valu_issue_peak"],
                     "source": "profiles/sq_counters_relax3d_1M.json (SQ_INSTS_VALU per dispatch)"}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "3d-relax-1M: unit cube minus sphere r=0.3, level set 256^3, linear dx "
                        f"{cfg.lin_dx0:.5f}->{cfg.lin_dx1:.5f} over phi 0..0.2",
            "n_particles": n,
            "dim": D,
            "num_level": st["num_level"],
            "neighbor_entries_per_particle": st["n_neighbor_entries"] / n,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel_avg_ms": avg_ms,
            "algorithmic_bytes_per_launch": bpl,
        },
        "issue_roofline": issue,
        "step_roofline": {
            "bytes_per_particle_step": STEP_BYTES[0 if D == 3 else 1],
            "achieved_GBs": STEP_BYTES[0 if D == 3 else 1] * value / 1e9,
            "frac": STEP_BYTES[0 if D == 3 else 1] * value / 1e9 / HBM_PEAK_GBS,
        },
        "list_rebuilds_per_step": rebuild_frac,
        "autotune": tune,
        "eager_ms_per_step": elapsed_eager / steps * 1e3,
        "kernels_ms": {k: v[0] for k, v in prof.items()},
    }
    if not args.no_cpu:
        log("accuracy vs the FP64 oracle ...")
        out["accuracy"] = accuracy(cfg, lset, job)
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(cfg, lset, n, args.cpu_steps)
    else:
        out["accuracy"] = None
        out["cpu_baseline"] = None
    job.close()
    if args.full_run:
        out["full_run"] = full_run(smg, cfg, lset, n, args.full_run, G)
    print(json.dumps(out), flush=True)


def full_run(smg, cfg, lset, n, nsteps, G):
    """The whole relaxation of the workload (seeding, pre-run and nsteps steps, the
    step loop as graph replays), timed end to end on the device."""
    job = smg.Job(cfg, lset, profile=False)
    t0 = time.perf_counter()
    job.seed(n)
    job.capture(G)
    job.step_async(-(-nsteps // G) * G)
    job.sync()
    el = time.perf_counter() - t0
    st = job.stats()
    job.close()
    steps = -(-nsteps // G) * G
    return {"steps": steps, "seconds": el, "particle_steps_per_s": n * steps / el,
            "list_rebuilds": st["n_rebuild"], "residual_mean": st["residual"], "residual_max": st["res_max"]}


def accuracy(cfg, lset, job):
    """Single-step error of the timed workload against the FP64 oracle: the oracle
    takes one step from the engine's current state (positions as the engine stores
    them), the engine takes the same step; the displacement each sets up for the next
    step (v dt2) is compared per particle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.set_threads(min(16, os.cpu_count() or 1))
    oracle.set_position_rounding(True)
    s = job.state()
    st = job.stats()
    ref = oracle.State.from_arrays(s["x"], s["v"], s["h"], s["nu"].astype(np.float64), s["type"],
                                   st["dt2"] if st["step"] > 0 else 0.0)
    t0 = time.perf_counter()
    d = oracle.step(cfg, lset, ref, st["step"] + 1)
    t_or = time.perf_counter() - t0
    job.step(1)
    g = job.state()
    st1 = job.stats()
    oracle.set_position_rounding(False)
    dg = g["v"] * st1["dt2"]
    do = ref.v * ref.dt
    ng = np.linalg.norm(dg - do, axis=1)
    no = np.linalg.norm(do, axis=1)
    scale = no.max()
    rel = ng / np.maximum(no, 1e-300)
    moving = no > 1e-3 * scale
    return {
        "what": "one step from the engine's state after the timed window, engine vs FP64 oracle; "
                "displacement = v * dt2 per particle",
        "step": int(st["step"] + 1),
        "positions_bit_exact": bool(np.array_equal(g["x"], ref.x)),
        "types_equal": bool(np.array_equal(g["type"], ref.type)),
        "disp_err_max_over_max_disp": float(ng.max() / scale),
        "disp_rel_err_p50": float(np.median(rel[moving])),
        "disp_rel_err_p99": float(np.percentile(rel[moving], 99)),
        "disp_rel_err_max": float(rel[moving].max()),
        "particles_compared": int(moving.sum()),
        "dt2_rel_err": float(abs(st1["dt2"] - d["dt2"]) / d["dt2"]),
        "n_remapped_oracle": d["n_remapped"],
        "oracle_seconds": t_or,
    }


def batch(args, smg):
    t0 = time.perf_counter()
    jobs = smg.batched_jobs(args.jobs, 1000, 200000, seed=20261203, dims=(2, 3))
    n_tot = sum(n for _, _, n in jobs)
    log(f"{len(jobs)} jobs, {n_tot} particles, level sets in {time.perf_counter() - t0:.1f} s")
    S, G = args.job_steps, args.graph_steps
    # tile shapes from the autotuner (smg_autotune): the list skin, i.e. the cell-size
    # multiple, and the inline list depth, timed on the largest job of each dimension
    # from its seeding on (the phase a 200-step job spends its time in)
    t0 = time.perf_counter()
    tuned = {}
    for dim in (2, 3):
        c, ls, n = max([j for j in jobs if j[0].dim == dim], key=lambda j: j[2])
        tuned[dim] = smg.autotune(c, ls, n, warmup=0, steps=40)
        log(f"autotune {dim}D on n={n}: skin {tuned[dim]['skin']}, max_neighbors {tuned[dim]['max_neighbors']}")
    t_tune = time.perf_counter() - t0
    tjobs = []
    for c, ls, n in jobs:
        cc = smg.SmgConfig.from_buffer_copy(c)
        cc.skin = tuned[c.dim]["skin"]
        cc.max_neighbors = tuned[c.dim]["max_neighbors"]
        tjobs.append((cc, ls, n))

    def batched(js):
        # all jobs of a dimension in one slot space, graph-replayed step loop
        t0 = time.perf_counter()
        b = smg.Batch(js)
        b.seed()
        b.capture(G)
        b.run(S // G)
        b.sync()
        t = time.perf_counter() - t0
        b.close()
        return t

    tu = batched(jobs)
    log(f"batched, default shapes: {tu:.2f} s")
    tb = batched(tjobs)
    log(f"batched, autotuned shapes: {tb:.2f} s (tuning {t_tune:.2f} s)")
    # plain: one job per launch sequence, eager, one after the other
    t0 = time.perf_counter()
    lat = []
    for cfg, ls, n in jobs:
        j = smg.Job(cfg, ls)
        j.seed(n)
        j.step(S)
        j.close()
        lat.append(time.perf_counter() - t0)
    tp = time.perf_counter() - t0
    log(f"plain: {tp:.2f} s")
    out = {
        "metric": "jobs/s (batched mixed 2D/3D relaxation, fixed steps)",
        "value": len(jobs) / tb, "unit": "jobs/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{len(jobs)} jobs, half 2D circle unions, half 3D sphere unions, "
                               f"N log-uniform 1e3..2e5, {S} steps each (seed + pre-run included)",
                   "particles": n_tot, "graph_steps": G},
        "plain_jobs_per_s": len(jobs) / tp,
        "speedup_vs_plain": tp / tb,
        "default_shapes_jobs_per_s": len(jobs) / tu,
        "autotune": {"seconds": t_tune, "2d": {k: tuned[2][k] for k in ("skin", "max_neighbors")},
                     "3d": {k: tuned[3][k] for k in ("skin", "max_neighbors")}},
        "latency_s": {"batched_p50": tb, "batched_p99": tb, "plain_p50": pct(lat, 50), "plain_p99": pct(lat, 99)},
        "particle_steps_per_s": {"batched": n_tot * S / tb, "plain": n_tot * S / tp},
    }
    print(json.dumps(out), flush=True)


def c3(args, smg):
    """The multi-resolution case: 12 seeded ellipsoids, a 64^3 base level set refined
    twice in bands near the surface, h ratio 8:1, run until the step residual (mean
    |v| dt / dx) is below 1e-3."""
    t0 = time.perf_counter()
    cfg, sdf = smg.config_c3()
    # the seeded start is uniform in space, so the particles of the coarse interior
    # (h 8x the surface's) start with lists of up to ~1200 entries: 1024 inline entries
    # per particle (walked in groups of four) and a spill pool for the rest
    cfg.max_neighbors = 1024
    cfg.spill_per_slot = 256
    base = smg.make_levelset(cfg, sdf)
    lv = smg.make_levelset_levels(cfg, sdf, 2)
    n = int(lv[-1]["stats"].num_particle)
    log(f"level sets in {time.perf_counter() - t0:.1f} s; n = {n}; stored packages per level "
        f"{[int(L['active'].sum()) for L in lv]} of {[int(L['active'].size) for L in lv]}")
    job = smg.Job(cfg, base, levels=lv)
    t0 = time.perf_counter()
    job.seed(n)
    log(f"seeded in {time.perf_counter() - t0:.1f} s")
    G, chunk, tol = args.graph_steps, 100, 1e-3
    job.capture(G)
    job.sync()
    steps, hist = 0, []
    t0 = time.perf_counter()
    while steps < args.c3_max_steps:
        job.step_async(chunk)
        job.sync()
        steps += chunk
        st = job.stats()
        hist.append((steps, st["residual"]))
        log(f"step {steps}: residual {st['residual']:.3e}")
        if st["residual"] < tol:
            break
    elapsed = time.perf_counter() - t0
    st = job.stats()
    out = {
        "metric": "particle-steps/s (3D multi-resolution relax, run to residual < 1e-3)",
        "value": n * steps / elapsed, "unit": "particle-steps/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "union of 12 seeded ellipsoids in the unit cube; level set 64^3 base refined "
                               "twice (128^3, 256^3) in bands near the surface; dx linear in phi, 8:1",
                   "n_particles": n, "graph_steps": G},
        "steps": steps, "seconds": elapsed, "converged": st["residual"] < tol, "residual": st["residual"],
        # the residual of this line is the mean of |v| dt2 / dx over the particles; the
        # largest particle's value is pinned near 0.025 by the dt rule
        "residual_max": st["res_max"],
        "ms_per_step": elapsed / steps * 1e3, "list_rebuilds_per_step": st["n_rebuild"] / max(steps, 1),
        "residual_history": hist[::5],
    }
    job.close()
    print(json.dumps(out), flush=True)


def tet(args, smg):
    """3D tetrahedralisation of the relaxed 1M-particle workload: the GPU's Voronoi-cell
    passes (the Delaunay tetrahedra of all particles: ring search, then the verifying
    pass for the hull and cavity cells) plus the host stage (deletion rule, boundary
    triangles, quality statistics); Qhull (scipy.spatial.Delaunay, one host core) on the
    same points beside it, and the two sets compared."""
    cfg, sdf = smg.config_c2(scale=args.scale, num_pkg=args.num_pkg)
    lset = smg.make_levelset(cfg, sdf)
    n = int(lset["stats"].num_particle)
    job = smg.Job(cfg, lset)
    job.seed(n)
    job.step(args.warmup)
    smg.tetrahedralize(job)   # first call: allocation and code-object load
    reps = []
    for _ in range(3):
        tets, st = smg.tetrahedralize(job)
        reps.append((st["cell_ms"], st["total_ms"]))
    cell_ms = min(r[0] for r in reps)
    total_ms = min(r[1] for r in reps)
    log(f"n={n} tets={st['n_tet']} delaunay={st['n_delaunay']} second pass={st['n_second_pass']} "
        f"open={st['n_open']} cell_ms={cell_ms:.1f} total_ms={total_ms:.1f}")
    out = {
        "metric": "tetrahedra/s (3D tetrahedralisation of the relaxed 1M-particle workload, GPU Voronoi-cell passes)",
        "value": st["n_delaunay"] / (cell_ms * 1e-3), "unit": "tetrahedra/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3d-relax-1M after {args.warmup} relaxation steps", "n_particles": n},
        "cell_pass_ms": cell_ms, "whole_call_ms": total_ms,
        "n_delaunay": st["n_delaunay"], "n_tet": st["n_tet"], "n_deleted": st["n_deleted"],
        "n_surface_tri": st["n_surface_tri"], "n_open": st["n_open"], "n_duplicate": st["n_duplicate"],
        "n_cell_overflow": st["n_cell_overflow"], "n_second_pass": st["n_second_pass"],
        "quality": {k: st[k] for k in ("min_dihedral", "avg_min_dihedral", "max_aspect", "avg_aspect",
                                       "avg_radius_ratio", "total_volume")},
    }
    if not args.no_cpu:
        from scipy.spatial import Delaunay
        x = job.state()["x"]
        _, first = np.unique(x, axis=0, return_index=True)
        t0 = time.perf_counter()
        d = Delaunay(x[np.sort(first)])
        dt = time.perf_counter() - t0
        g, _, _ = smg.tet_delaunay(job, len(x))
        q = np.sort(np.sort(first)[d.simplices], axis=1)
        gs, qs = set(map(tuple, g.tolist())), set(map(tuple, q.tolist()))
        out["equal_to_qhull"] = {"gpu_only": len(gs - qs), "qhull_only": len(qs - gs)}
        out["qhull_cpu"] = {"seconds": dt, "tetrahedra": int(len(d.simplices)), "cores": 1,
                            "note": "scipy.spatial.Delaunay on the same points: the whole global tetrahedralisation, "
                                    "hull and cavity tetrahedra included"}
    job.close()
    print(json.dumps(out), flush=True)


def lset(args, smg):
    """Level-set finishing of the 1M-particle workload's 256^3 level set (unit cube minus
    a sphere): the GPU builder (smg_lset_build_gpu) against the host builder
    (smg_lset_build, all host threads) on the same input; both give the same bits."""
    cfg, sdf = smg.config_c2(scale=args.scale, num_pkg=args.num_pkg)
    X, Y, Z = smg.cell_centres_ext(cfg)
    ext = sdf(X, Y, Z)
    del X, Y, Z
    smg.lset_build_gpu(cfg, ext)   # first call: code-object load
    ker, wall = [], []
    for _ in range(5):
        t0 = time.perf_counter()
        G = smg.lset_build_gpu(cfg, ext)
        wall.append((time.perf_counter() - t0) * 1e3)
        ker.append(G["kernel_ms"])
    ncell = 1
    for a in range(3):
        ncell *= 4 * cfg.num_pkg[a]
    kms, wms = min(ker), min(wall)
    st = G["stats"]
    log(f"cells={ncell} cut={st.n_interface_cell} particles={st.num_particle} kernel_ms={kms:.2f} wall_ms={wms:.1f}")
    out = {
        "metric": "level-set cells/s (GPU level-set finishing: normals, cut tags, cut-cell volume and area, "
                  "particle count)",
        "value": ncell / (kms * 1e-3), "unit": "cells/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "3d unit cube minus sphere r=0.3, level set 256^3 cells (the 1M-particle case)",
                   "cells": ncell, "cut_cells": int(st.n_interface_cell), "num_particle": int(st.num_particle)},
        "kernel_ms": kms, "wall_ms_with_transfers": wms,
    }
    if not args.no_cpu:
        t0 = time.perf_counter()
        P = smg.lset_build(cfg, ext)
        dt = time.perf_counter() - t0
        same = all(np.array_equal(P[k], G[k]) for k in ("phi", "normal", "tag")) and all(
            getattr(P["stats"], f) == getattr(st, f) for f, _ in smg.SmgLsetStats._fields_)
        out["host_builder"] = {"ms": dt * 1e3, "threads": min(os.cpu_count() or 1, 32), "bit_identical": bool(same)}
    print(json.dumps(out), flush=True)


def serve_once(smg, specs, lam, duration, max_jobs, budget_gb, rng_seed, label, max_active=0):
    """Poisson arrivals at rate lam for `duration` seconds into a server; returns the
    completions inside the window."""
    dim = specs[0][0].dim
    tbl = max(smg.job_size(c, n)[1] for c, _, n in specs)
    cells = max(smg.job_size(c, n)[2] for c, _, n in specs)
    slots = (1 << 31) - 512   # 32-bit slot indices (neighbour entries are job-relative)
    jobs_cap = max_jobs
    sc = smg.server_config(dim, jobs_cap, slots, min(jobs_cap * tbl, (1 << 31) - 16), jobs_cap * cells,
                           chunk_steps=50, min_steps=100, max_steps=2000, tol=1e-3,
                           max_active_particles=max_active)
    # the footprint is affine in the slot count: take as many slots as the budget holds
    sc.max_slots = 1 << 20
    f0 = smg.server_footprint(sc)
    sc.max_slots = 2 << 20
    per_slot = (smg.server_footprint(sc) - f0) / float(1 << 20)
    sc.max_slots = int(min(slots, (1 << 20) + (budget_gb * 1e9 - f0) / per_slot)) // 256 * 256
    srv = smg.Server(sc)
    rng = np.random.default_rng(rng_seed)
    arrivals = np.cumsum(rng.exponential(1.0 / lam, size=int(lam * duration * 3) + 10))
    arrivals = arrivals[arrivals < duration]
    queue, arrived_at, done = [], {}, []
    k = 0
    t_adv = t_adm = 0.0
    ps_adv = 0   # particle-steps the advances covered
    live = 0     # particles admitted and not yet finished
    t0 = time.perf_counter()
    while True:
        now = time.perf_counter() - t0
        if now >= duration:
            break
        while k < len(arrivals) and arrivals[k] <= now:
            queue.append(k)
            k += 1
        while queue:
            i = queue[0]
            c, ls, n = specs[i % len(specs)]
            cc = smg.SmgConfig.from_buffer_copy(c)
            cc.seed = c.seed + 1000003 * i
            ta = time.perf_counter()
            t = srv.admit(cc, ls, n)
            t_adm += time.perf_counter() - ta
            if t is None:
                break
            arrived_at[t] = (arrivals[i], n)
            live += n
            queue.pop(0)
        if srv.usage()["n_active"] == 0:
            if k < len(arrivals):
                time.sleep(max(0.0, min(arrivals[k] - (time.perf_counter() - t0), duration - now)))
            continue
        ta = time.perf_counter()
        srv.advance()
        t_adv += time.perf_counter() - ta
        ps_adv += live * srv.cfg.chunk_steps
        for f in srv.finished():
            ta, n = arrived_at[f["ticket"]]
            live -= n
            done.append({"latency": time.perf_counter() - t0 - ta, "n": n, "steps": f["steps"], "status": f["status"]})
            srv.result(f["ticket"])
    foot = smg.server_footprint(sc)
    srv.close()
    lat = [d["latency"] for d in done]
    r = {"jobs_per_s": len(done) / duration, "completed": len(done), "arrived": int(len(arrivals)),
         "queued_at_end": len(queue) + (len(arrivals) - k), "p50_s": pct(lat, 50), "p99_s": pct(lat, 99),
         "mean_steps": float(np.mean([d["steps"] for d in done])) if done else None,
         "converged": sum(d["status"] == 0 for d in done),
         "particle_steps_per_s": sum(d["n"] * d["steps"] for d in done) / duration,
         "pool_footprint_GB": foot / 1e9, "max_jobs": jobs_cap,
         # where the window went: stepping (with retirement downloads) and admission
         # (level-set upload, seeding, pre-run); the device rate while stepping
         "advance_s": t_adv, "admit_s": t_adm, "stepping_particle_steps_per_s": ps_adv / max(t_adv, 1e-9)}
    log(f"{label} lam={lam}: {json.dumps(r)}")
    return r


def plain_once(smg, specs, lam, duration, rng_seed, chunk=50, min_steps=100, max_steps=2000, tol=1e-3):
    """The plain GPU path under the same arrivals: one job at a time in arrival order,
    each in an engine of its own (smg_create, smg_set_levelset, smg_seed_particles,
    smg_step in chunks, smg_destroy), retired by the server's rule (mean residual over
    the last chunk, from smg_stats.residual_sum, below tol after min_steps, or
    max_steps).  Returns the completions inside the window."""
    rng = np.random.default_rng(rng_seed)
    arrivals = np.cumsum(rng.exponential(1.0 / lam, size=int(lam * duration * 3) + 10))
    arrivals = arrivals[arrivals < duration]
    done, k = [], 0
    t0 = time.perf_counter()
    while k < len(arrivals):
        now = time.perf_counter() - t0
        if now >= duration:
            break
        if arrivals[k] > now:
            time.sleep(min(arrivals[k], duration) - now)
            continue
        c, ls, n = specs[k % len(specs)]
        cc = smg.SmgConfig.from_buffer_copy(c)
        cc.seed = c.seed + 1000003 * k
        job = smg.Job(cc, ls)
        job.seed(n)
        steps, prev, status = 0, 0.0, None
        while time.perf_counter() - t0 < duration:
            job.step(chunk)
            steps += chunk
            rs = job.stats()["residual_sum"]
            win = (rs - prev) / chunk
            prev = rs
            if steps >= min_steps and win < tol:
                status = 0
            elif steps >= max_steps:
                status = 1
            if status is not None:
                break
        job.close()
        if status is None:
            break   # the window ended inside this job
        done.append({"latency": time.perf_counter() - t0 - arrivals[k], "n": n, "steps": steps, "status": status})
        k += 1
    lat = [d["latency"] for d in done]
    r = {"jobs_per_s": len(done) / duration, "completed": len(done), "arrived": int(len(arrivals)),
         "queued_at_end": int(len(arrivals)) - len(done), "p50_s": pct(lat, 50), "p99_s": pct(lat, 99),
         "mean_steps": float(np.mean([d["steps"] for d in done])) if done else None,
         "converged": sum(d["status"] == 0 for d in done),
         "particle_steps_per_s": sum(d["n"] * d["steps"] for d in done) / duration}
    log(f"plain lam={lam}: {json.dumps(r)}")
    return r


def serve(args, smg):
    t0 = time.perf_counter()
    specs = smg.batched_jobs(args.specs, 5000, 500000, seed=777, dims=(3,), max_pkg=16)
    log(f"{len(specs)} job specs in {time.perf_counter() - t0:.1f} s")
    res = {}
    for lam in args.lam:
        res[str(lam)] = {
            "served": serve_once(smg, specs, lam, args.duration, args.max_jobs, args.budget_gb, 5 + int(lam), "served",
                                 args.max_active),
            "plain": plain_once(smg, specs, lam, args.duration, 5 + int(lam)),
        }
    top = res[str(args.lam[-1])]
    out = {
        "metric": "jobs/s (served 3D sphere unions, Poisson arrivals, retired at residual < 1e-3)",
        "value": top["served"]["jobs_per_s"], "unit": "jobs/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"3D sphere unions, N log-uniform 5e3..5e5 ({len(specs)} level sets, distinct "
                               f"seeds), Poisson lambda {args.lam}/s, {args.duration} s per lambda, chunk 50 steps, "
                               f"at most {args.max_active} particles in the pool, "
                               f"retire at mean residual < 1e-3 over a chunk (>= 100 steps) or 2000 steps"},
        "p50_s": top["served"]["p50_s"], "p99_s": top["served"]["p99_s"],
        "speedup_vs_plain": (top["served"]["jobs_per_s"] / top["plain"]["jobs_per_s"]
                             if top["plain"]["jobs_per_s"] else None),
        "by_lambda": res,
    }
    print(json.dumps(out), flush=True)


def copy(args, smg):
    """A plain device copy of known size (the reference point that scales the HBM
    byte counters of a profiler run): 1 GiB copied args.steps times."""
    import torch
    nbytes = 1 << 30
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.copy_(a)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(json.dumps({"mode": "copy", "bytes_read_per_copy": nbytes, "bytes_written_per_copy": nbytes,
                      "copies": args.steps, "GBs": 2 * nbytes * args.steps / el / 1e9}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--mode", choices=("relax", "batch", "serve", "c3", "copy", "tet", "lset"), default="relax")
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--scale", type=float, default=0.564)
    ap.add_argument("--num-pkg", type=int, default=64)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--jobs", type=int, default=512)
    ap.add_argument("--job-steps", type=int, default=200)
    ap.add_argument("--graph-steps", type=int, default=10)   # steps per captured hipGraph (even)
    ap.add_argument("--specs", type=int, default=48)
    ap.add_argument("--lam", type=float, nargs="+", default=[20.0, 100.0, 200.0])
    ap.add_argument("--duration", type=float, default=15.0)
    ap.add_argument("--max-jobs", type=int, default=256)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--max-active", type=int, default=4000000)
    ap.add_argument("--c3-max-steps", type=int, default=6000)
    ap.add_argument("--skin", type=float, default=None, help="relax mode: list skin (default: the autotuner's)")
    ap.add_argument("--no-tune", action="store_true", help="relax mode: keep the config's list shape")
    ap.add_argument("--full-run", type=int, default=1000, help="relax mode: also time a whole run of this many steps (0: off)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (one GPU per box); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import smg
    {"relax": relax, "batch": batch, "serve": serve, "c3": c3, "copy": copy, "tet": tet,
     "lset": lset}[args.mode](args, smg)


if __name__ == "__main__":
    main()
