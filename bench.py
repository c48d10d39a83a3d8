"""Benchmark: 3D relaxation throughput (particle-steps/s) of the MI355X engine.

Workload (one GPU): unit cube minus a sphere r = 0.3 at the centre, level set on a
256^3 grid, target spacing linear in phi from 0.00451 at the surface to 0.01354 at
phi >= 0.2 (0.564 x the 0.008 -> 0.024 profile, ~1.0M particles), seeded uniform
random positions, the full relaxation step (mapping, sort, neighbour lists, both
force stages, both dt reductions) per timed step.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sph-mesh-generation_amd"))

METRIC = "particle-steps/s (3D 1M relax); batched jobs/s + p99 latency vs plain GPU"
HBM_PEAK_GBS = 8000.0

# algorithmic bytes per particle and launch (3D, 2D): each field counted once per
# dependent pass, neighbour data assumed served from cache
BYTES = {
    "k_map": (292, 188),      # integrate (a1, a2 -> v) + level-set sample + map + h + key
    "k_nb": (32, 24),         # read x, h
    "k_force1": (44, 32),     # read x, h, type, n; write a1
    "k_force2": (60, 44),     # read x, h, type, n, a1, nu; write a2, nu
    "k_scan": (0, 0),
    "k_scatter": (16, 16),
    "k_segsort": (20, 20),
    "k_gather": (120, 80),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--scale", type=float, default=0.564)
    ap.add_argument("--num-pkg", type=int, default=64)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the relaxation of one job has no multi-GPU path in this build "
              "(see the sharding note in DESIGN.md); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import smg
    t0 = time.perf_counter()
    cfg, sdf = smg.config_c2(scale=args.scale, num_pkg=args.num_pkg)
    lset = smg.make_levelset(cfg, sdf)
    n = int(lset["stats"].num_particle)
    log(f"level set built in {time.perf_counter() - t0:.1f} s, n = {n}")
    job = smg.Job(cfg, lset, profile=False)
    t0 = time.perf_counter()
    job.seed(n)
    job.step(args.warmup)
    log(f"seeded + {args.warmup} warmup steps in {time.perf_counter() - t0:.1f} s")
    job.sync = None
    smg.lib().smg_profile_enable(job.h, 1)   # events around each kernel over the timed region
    t0 = time.perf_counter()
    job.step(args.steps)
    elapsed = time.perf_counter() - t0
    prof = job.profile()
    st = job.stats()
    log(f"timed {args.steps} steps in {elapsed:.3f} s; stats {st}")
    log("kernels (avg ms, launches): " + json.dumps({k: [round(v[0], 4), v[1]] for k, v in prof.items()}))
    value = n * args.steps / elapsed
    D = cfg.dim
    tot = {k: v[0] * v[1] for k, v in prof.items()}
    dom = max(tot, key=tot.get)
    avg_ms = prof[dom][0]
    bpl = BYTES.get(dom, (0, 0))[0 if D == 3 else 1] * n
    achieved = bpl / (avg_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
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
            "traffic": None,
            "kernel_avg_ms": avg_ms,
            "algorithmic_bytes_per_launch": bpl,
        },
        "step_roofline": {
            "bytes_per_particle_step": 552 if D == 3 else 380,
            "achieved_GBs": (552 if D == 3 else 380) * value / 1e9,
            "frac": (552 if D == 3 else 380) * value / 1e9 / HBM_PEAK_GBS,
        },
        "kernels_ms": {k: v[0] for k, v in prof.items()},
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(cfg, lset, n, args.cpu_steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
