"""Benchmark of the gauge-field hot path on one MI355X.

Workload (the 32^4 configuration of the headline metric): a 32^4 lattice at
beta = 6.0, thermalised from a Haar-random start (seed 0) with heatbath +
4 overrelaxation updates.  One timed step is one update = 1 Cabibbo-Marinari
heatbath sweep + 4 SU(2)-subgroup overrelaxation sweeps (2-colour checkerboard),
i.e. 5 * 4V link updates.  value = link updates per second over the K timed steps.
The same run then times RK3 Wilson-flow steps (eps = 0.01) and the fused
observables pass, reported as extra fields.

Roofline: the dominant kernel is the overrelaxation colour kernel (4/5 of the
sweeps).  Algorithmic bytes per link update = 1296 B (read the 6 staple links of
the other 3 directions over both parities, the opposite-parity half of the own
direction, read + write the own half: 36 * V * 144 B per sweep over 4V link
updates); one launch updates V/2 links.  Launch durations come from HIP events
recorded around every launch on the library's stream.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size L] [--therm N]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_LINK_UPDATE = 1296.0
BYTES_PER_LINK_STAGE = 480.0
LAYOUT_BYTES_PER_LINK_UPDATE = 864.0   # 8 half-field reads + 1 write of 96-byte (two-row) links


def measured_traffic(path):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    --pmc summary (FETCH_SIZE and WRITE_SIZE in KiB per dispatch, each scaled by the
    factor the same counter reads for a plain copy kernel of known size)."""
    try:
        with open(path) as f:
            t = json.load(f)
        return (t["fetch_kib"] * t["fetch_scale"] + t["write_kib"] * t["write_scale"]) * 1024.0
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(L, beta, budget_s):
    """Oracle (OpenMP C port of the reference algorithms) on a bounded sample of
    the same workload: whole HB+4OR updates of the same L^4 lattice."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lto
    shape = (L, L, L, L)
    U = lto.haar(shape, 0)
    lto.or_sweep(U, shape, 1, 2)       # warm caches and threads
    n_upd, t_tot = 0, 0.0
    counter = 0
    while t_tot < budget_s and n_upd < 8:
        t0 = time.perf_counter()
        counter = lto.hb_sweep(U, shape, beta, 1, 2, 0, counter)
        lto.or_sweep(U, shape, 4, 2)
        t_tot += time.perf_counter() - t0
        n_upd += 1
    links = 5 * 4 * L ** 4 * n_upd
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": links / t_tot, "unit": "link updates/s", "cores": threads, "kind": "port",
            "sample": f"{n_upd} x (1 HB + 4 OR) sweeps of {L}^4 by the OpenMP C oracle, {t_tot:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=32)
    ap.add_argument("--beta", type=float, default=6.0)
    ap.add_argument("--therm", type=int, default=500, help="untimed HB+4OR updates from the Haar start")
    ap.add_argument("--flow-steps", type=int, default=100)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import lettuce_amd as la

    L = args.size
    shape = (L, L, L, L)
    V = L ** 4
    ctx = la.Context(0)
    U = la.GaugeField(ctx, shape)
    U.SetToHaar(0)
    U.set_rng(0, 0)
    hb = la.HeatbathKernel(U, args.beta)
    orx = la.OverrelaxationSubgroupKernel(U)

    def update():
        la.Iterator.Checkerboard(hb, 1)
        la.Iterator.Checkerboard(orx, 4)

    t0 = time.perf_counter()
    for k in range(args.therm):
        update()
        if (k + 1) % 100 == 0:
            ctx.synchronize()
            print(f"# thermalised {k + 1}/{args.therm} updates", file=sys.stderr, flush=True)
    ctx.synchronize()
    t_therm = time.perf_counter() - t0

    for _ in range(args.warmup):
        update()
    ctx.synchronize()

    # timed region: K updates, synchronised on both sides, no per-launch events
    t0 = time.perf_counter()
    for _ in range(args.steps):
        update()
    ctx.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = 1000.0 * dt / args.steps
    link_updates = 5 * 4 * V * args.steps
    value = link_updates / dt

    # kernel durations with HIP events on the library stream (same work, second pass)
    ctx.profile(True)
    for _ in range(max(5, args.steps // 5)):
        update()
    ctx.synchronize()
    n_or, ms_or = ctx.kernel_time("or_cb2")
    n_hb, ms_hb = ctx.kernel_time("hb_cb2")
    ctx.profile(False)
    or_ms = ms_or / max(n_or, 1)
    hb_ms = ms_hb / max(n_hb, 1)
    bytes_per_launch = BYTES_PER_LINK_UPDATE * V / 2
    or_gbs = bytes_per_launch / (or_ms * 1e-3) / 1e9
    hb_gbs = bytes_per_launch / (hb_ms * 1e-3) / 1e9

    obs = la.Observables(U)
    plaq = obs["plaquette_sum"] / (18.0 * V)

    # RK3 Wilson flow from the thermalised configuration
    W = la.GaugeField(ctx, shape)
    fl = la.GradientFlowKernel(U, W, "rk3", 0.01)
    fl(2)
    ctx.synchronize()
    t0 = time.perf_counter()
    fl.Resume(args.flow_steps)
    ctx.synchronize()
    flow_ms = 1000.0 * (time.perf_counter() - t0) / args.flow_steps
    ctx.profile(True)
    fl.Resume(10)
    ctx.synchronize()
    n_fs, ms_fs = ctx.kernel_time("flow_stage")
    ctx.profile(False)
    stage_ms = ms_fs / max(n_fs, 1)
    flow_gbs = BYTES_PER_LINK_STAGE * 4 * V * 3 / (flow_ms * 1e-3) / 1e9
    t0 = time.perf_counter()
    for _ in range(10):
        la.Observables(W)
    obs_ms = 100.0 * (time.perf_counter() - t0)
    fl.close()

    result = {
        "metric": "FP64 link updates/s (HB + 4 OR sweeps, 32^4)",
        "value": value,
        "unit": "link updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"SU(3) Wilson action {L}^4, beta={args.beta}, Haar start seed 0 + "
                               f"{args.therm} HB+4OR thermalisation updates; step = 1 HB + 4 OR sweeps "
                               "(2-colour checkerboard)",
                   "lattice": list(shape), "beta": args.beta, "schedule": "checkerboard-2"},
        "roofline": {"bound": "hbm", "achieved": or_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": or_gbs / HBM_PEAK_GBS,
                     "traffic": measured_traffic(os.path.join(ROOT, "profiles", f"or_cb2_{L}x4_traffic.json")),
                     "kernel": "or_cb2 (one colour of one direction)", "avg_launch_ms": or_ms,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     # links are stored as two rows (96 B): the minimum this layout moves
                     "layout_bytes_per_launch": LAYOUT_BYTES_PER_LINK_UPDATE * V / 2},
        "hb_kernel": {"avg_launch_ms": hb_ms, "achieved_GBs": hb_gbs},
        "hbm_GBs_per_update": BYTES_PER_LINK_UPDATE * 5 * 4 * V / (ms_per_step * 1e-3) / 1e9,
        "flow_rk3": {"ms_per_step": flow_ms, "model_GBs": flow_gbs, "stage_avg_launch_ms": stage_ms,
                     "stage_GBs": BYTES_PER_LINK_STAGE * 4 * V / (stage_ms * 1e-3) / 1e9,
                     "link_stage_updates_per_s": 12 * V / (flow_ms * 1e-3)},
        "observables_pass_ms": obs_ms,
        "plaquette": plaq,
        "thermalisation_s": t_therm,
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(L, args.beta, args.cpu_seconds)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result))


if __name__ == "__main__":
    main()
