"""Benchmark of the gauge-field hot path on one MI355X.

Workload (the 32^4 configuration of the headline metric): a 32^4 lattice at
beta = 6.0, thermalised from a Haar-random start (seed 0) with heatbath +
4 overrelaxation updates.  One timed step is one update = 1 Cabibbo-Marinari
heatbath sweep + 4 SU(2)-subgroup overrelaxation sweeps (2-colour checkerboard),
i.e. 5 * 4V link updates.  value = link updates per second over the K timed steps.
The same run then times RK3 Wilson-flow steps (eps = 0.01) and the fused
observables pass, reported as extra fields, then (unless skipped) the
parallel-tempering MetaD workload (16 replicas of 16^4, batched in one graph vs
one trajectory after another) and the 48^4 workload (HB+OR, RK3 flow to t = 8,
one MetaD trajectory with 8 stout levels and checkpointing, peak memory).

Roofline: the dominant kernel is the overrelaxation colour kernel (4/5 of the
sweeps).  Algorithmic bytes per link update = 1296 B (read the 6 staple links of
the other 3 directions over both parities, the opposite-parity half of the own
direction, read + write the own half: 36 * V * 144 B per sweep over 4V link
updates); one launch updates V/2 links.  The average launch duration comes from two HIP
events on the library's stream around 4 overrelaxation sweeps (32 launches back to back);
the per-launch event pairs of the library's profiler (which add the events' own cost) are
reported beside it.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size L] [--therm N] [--no-pt] [--no-48]
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_LINK_UPDATE = 1296.0
BYTES_PER_LINK_STAGE = 480.0
LAYOUT_BYTES_PER_LINK_UPDATE = 864.0   # 8 half-field reads + 1 write of 96-byte (two-row) links


def block_ms(stream, fn, reps):
    """Milliseconds per call of fn, between two HIP events recorded on the library's stream
    around reps back-to-back calls (HIP runtime C API; one warm-up call first)."""
    hip = ctypes.CDLL("libamdhip64.so")
    ev = [ctypes.c_void_p(), ctypes.c_void_p()]
    for e in ev:
        if hip.hipEventCreate(ctypes.byref(e)) != 0:
            raise RuntimeError("hipEventCreate failed")
    try:
        s = ctypes.c_void_p(stream)
        fn()
        if hip.hipEventRecord(ev[0], s) != 0:
            raise RuntimeError("hipEventRecord failed")
        for _ in range(reps):
            fn()
        if hip.hipEventRecord(ev[1], s) != 0 or hip.hipEventSynchronize(ev[1]) != 0:
            raise RuntimeError("hipEventRecord / hipEventSynchronize failed")
        ms = ctypes.c_float()
        if hip.hipEventElapsedTime(ctypes.byref(ms), ev[0], ev[1]) != 0:
            raise RuntimeError("hipEventElapsedTime failed")
        return ms.value / reps
    finally:
        for e in ev:
            hip.hipEventDestroy(e)


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
    return {"value": links / t_tot, "unit": "link updates/s", "cores": lto.threads(), "kind": "port",
            "host_cpus": os.cpu_count(), "cpus_allowed": len(os.sched_getaffinity(0)),
            "sample": f"{n_upd} x (1 HB + 4 OR) sweeps of {L}^4 by the OpenMP C oracle, {t_tot:.1f} s"}


def reference_cpu_16(beta, nrep=3):
    """The reference's own OpenMP code (oracle/_ref/lettuce_ref, built from its headers;
    extents fixed at 16^4 when it is compiled): seconds per HB sweep, OR sweep and RK3
    step on a Haar start.  None when the harness was not built."""
    import tempfile
    exe = os.path.join(ROOT, "oracle", "_ref", "lettuce_ref")
    if not os.access(exe, os.X_OK):
        return None
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lto
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "u.bin")
        lto.haar((16, 16, 16, 16), 0).tofile(path)
        r = subprocess.run([exe, "time", path, repr(beta), str(nrep)], capture_output=True, text=True, timeout=300)
    if r.returncode:
        return None
    kv = {}
    for line in r.stdout.splitlines():
        parts = line.split()
        if len(parts) == 2 and parts[0] in ("hb_sweep_s", "or_sweep_s", "rk3_step_s", "threads"):
            kv[parts[0]] = float(parts[1])
    if len(kv) < 4:
        return None
    links = 4 * 16 ** 4
    return {"lattice": [16, 16, 16, 16], "threads": int(kv["threads"]),
            "hb_or_link_updates_per_s": 5 * links / (kv["hb_sweep_s"] + 4 * kv["or_sweep_s"]),
            "hb_sweep_ms": kv["hb_sweep_s"] * 1e3, "or_sweep_ms": kv["or_sweep_s"] * 1e3,
            "rk3_step_ms": kv["rk3_step_s"] * 1e3,
            "note": "the reference code itself (checkerboard-2 sweeps, RK3 Wilson flow), timed on the host cores"}


def thermalised(la, ctx, shape, seed, n_update, beta=6.0):
    """Haar(seed) followed by n_update HB + 4 OR updates; returns the AoS host array."""
    U = la.GaugeField(ctx, shape)
    U.SetToHaar(seed)
    U.set_rng(1000 + seed, 0)
    hb = la.HeatbathKernel(U, beta)
    orx = la.OverrelaxationSubgroupKernel(U)
    for _ in range(n_update):
        la.Iterator.Checkerboard(hb, 1)
        la.Iterator.Checkerboard(orx, 4)
    data = U.download()
    U.close()
    return data


def c3_bias(la):
    """The 16^4 MetaD bias: 800 bins on [-8, 8], height 0.05, width 0.05, plain deposits."""
    return la.MetaBiasPotential(-8.0, 8.0, 800, 0.05, 0.05, 10.0, 1000.0, well_tempered=False)


def pt_metad(la, ctx, shapes, n_step=20, steps=2, graph=True):
    """Parallel-tempering MetaD over replicas of the given extents (seeds 0..n-1, bias strength
    r/(n-1), OMF2 tau = 1 with n_step steps, CV = clover Q after 4 stout levels rho = 0.12,
    path-sample deposits, alternating exchanges between equal extents), one trajectory per replica
    per step, each step replayed as one captured graph; against the same trajectories one replica
    after another on the single-lattice path (host-synchronised MetaD-HMC, the reference's call
    structure).  Rates in trajectories/s and site-trajectories/s."""
    n = len(shapes)
    sites = sum(int(np.prod(s)) for s in shapes)
    starts = [thermalised(la, ctx, s, r, 20) for r, s in enumerate(shapes)]
    b = la.ReplicaBatch(ctx, shapes, 4, 0.12)
    for r in range(n):
        b.field(r).upload(starts[r])
        b.set_rng(r, r, 0)
    b.set_bias(c3_bias(la))
    b.use_graph(graph)
    b.metad_hmc(n_step, tau=1.0, update_bias=True, swap="alternate")
    b.results()
    t0 = time.perf_counter()
    b.metad_hmc(n_step, tau=1.0, update_bias=True, swap="alternate", n_traj=steps)
    res = b.results()
    t_batch = (time.perf_counter() - t0) / steps
    b.close()
    sbias = c3_bias(la)
    work = {}   # per extent: U_copy, P, F, CV workspace
    fields, kernels = [], []
    for r, s in enumerate(shapes):
        if s not in work:
            work[s] = (la.GaugeField(ctx, s), la.MatrixField(ctx, s), la.MatrixField(ctx, s),
                       la.CollectiveVariable(ctx, s, 4, 0.12))
        Uc, P, F, cv = work[s]
        U = la.GaugeField(ctx, s)
        U.upload(starts[r])
        U.set_rng(r, 0)
        fields.append(U)
        kernels.append(la.HMCMetaDKernel(U, Uc, P, F, sbias, cv, "omf2", 6.0, 1.0, update_bias=True))
    kernels[0](n_step)
    ctx.synchronize()
    t0 = time.perf_counter()
    for k in kernels:
        k(n_step)
    ctx.synchronize()
    t_serial = time.perf_counter() - t0
    for Uc, P, F, cv in work.values():
        cv.close()
        fields += [Uc, P, F]
    for f in fields:
        f.close()
    return {"replicas": [list(s) for s in sorted(set(shapes))], "n_replicas": n, "sites": sites,
            "config": f"OMF2 tau=1 n_step={n_step}, 4 stout levels rho=0.12, bias x r/{n - 1}",
            "batched_traj_per_s": n / t_batch, "serial_traj_per_s": n / t_serial, "ratio": t_serial / t_batch,
            "batched_site_traj_per_s": sites / t_batch, "accept_rate_last_step": float(res["accepted"].mean())}


def metad16(la, ctx, n_traj=3, n_step=20):
    """The 16^4 metadynamics configuration: MetaD-HMC trajectories (OMF2 tau = 1, n_step
    steps) with the bias force through clover Q after 4 stout levels rho = 0.12, from a
    thermalised start; seconds per trajectory, per force evaluation (2 n_step + 1 per
    trajectory), and the kernel time of each force kernel per evaluation."""
    shape = (16,) * 4
    U = la.GaugeField(ctx, shape)
    U.upload(thermalised(la, ctx, shape, 0, 20))
    U.set_rng(0, 0)
    Uc, P, F = la.GaugeField(ctx, shape), la.MatrixField(ctx, shape), la.MatrixField(ctx, shape)
    cv = la.CollectiveVariable(ctx, shape, 4, 0.12)
    k = la.HMCMetaDKernel(U, Uc, P, F, c3_bias(la), cv, "omf2", 6.0, 1.0, update_bias=True)
    k(n_step)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(n_traj):
        k(n_step)
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / n_traj
    ctx.profile(True)
    k(n_step)
    ctx.synchronize()
    n_eval = 2 * n_step + 1
    kernels = {}
    for name in ("stout_tape", "clover_store", "qdensity", "clover_deriv", "stout_lambda", "stout_sigma",
                 "force_update", "field_update"):
        n, ms = ctx.kernel_time(name)
        if n:
            kernels[name] = {"launches_per_eval": n / n_eval, "us_per_eval": 1e3 * ms / n_eval}
    ctx.profile(False)
    cv.close()
    for f in (U, Uc, P, F):
        f.close()
    return {"lattice": list(shape), "config": f"OMF2 tau=1 n_step={n_step}, 4 stout levels rho=0.12",
            "traj_s": dt, "traj_per_s": 1.0 / dt, "force_evals_per_traj": n_eval,
            "ms_per_force_eval": 1e3 * dt / n_eval, "kernels": kernels}


def lattice8(la, ctx, n_update=200, graph=True):
    """8^4 at beta = 5.7: n_update HB + 4 OR updates from Haar(0), plaquette after each (the
    correctness configuration: launch-latency bound).  Timed twice from the same start:
    one call per sweep with a plaquette read-back per update, and lt_update_sweeps (the
    update captured once as a graph, plaquettes read back after the last update); both
    give the same field bit for bit.  graph=False times the eager calls only."""
    shape = (8,) * 4
    out = {"lattice": list(shape), "beta": 5.7, "updates": n_update}
    fields = []
    for mode in ("eager", "graph") if graph else ("eager",):
        U = la.GaugeField(ctx, shape)
        U.SetToHaar(0)
        U.set_rng(0, 0)
        if mode == "graph":
            la.UpdateSweeps(U, 5.7, 1, 4, 1, la.CB2)     # capture outside the timed loop
            U.SetToHaar(0)
            U.set_rng(0, 0)
        hb = la.HeatbathKernel(U, 5.7)
        orx = la.OverrelaxationSubgroupKernel(U)
        ctx.synchronize()
        t0 = time.perf_counter()
        if mode == "eager":
            plaq = []
            for _ in range(n_update):
                la.Iterator.Checkerboard(hb, 1)
                la.Iterator.Checkerboard(orx, 4)
                plaq.append(la.PlaquetteSum(U))
            plaq = np.array(plaq)
        else:
            plaq = la.UpdateSweeps(U, 5.7, 1, 4, n_update, la.CB2)
        dt = time.perf_counter() - t0
        out[f"ms_per_update_{mode}"] = 1000.0 * dt / n_update
        out[f"mean_plaquette_last_100_{mode}"] = float(np.mean(plaq[-100:]) / (18.0 * 8 ** 4))
        fields.append(U.download())
        U.close()
    if graph:
        out["ms_per_update"] = out["ms_per_update_graph"]
        out["graph_equals_eager"] = bool(np.array_equal(fields[0], fields[1]))
    else:
        out["ms_per_update"] = out["ms_per_update_eager"]
    return out


def lattice48(la, ctx, n_update=200, flow_steps=800, interval=3):
    """48^4 at beta = 6.2: HB + 4 OR updates from Haar(0), RK3 flow to t = 8 with Q, and one
    MetaD-HMC trajectory (OMF2 tau = 1, 20 steps) with 8 stout levels and per-level checkpointing;
    peak device memory of the MetaD workspace with and without checkpointing."""
    L = 48
    shape = (L,) * 4
    V = L ** 4
    U = la.GaugeField(ctx, shape)
    U.SetToHaar(0)
    U.set_rng(0, 0)
    hb = la.HeatbathKernel(U, 6.2)
    orx = la.OverrelaxationSubgroupKernel(U)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(n_update):
        la.Iterator.Checkerboard(hb, 1)
        la.Iterator.Checkerboard(orx, 4)
    ctx.synchronize()
    t_upd = time.perf_counter() - t0
    W = la.GaugeField(ctx, shape)
    fl = la.GradientFlowKernel(U, W, "rk3", 0.01)
    ctx.synchronize()
    t0 = time.perf_counter()
    fl(flow_steps)
    q = la.TopChargeClover(W)
    t_flow = time.perf_counter() - t0
    fl.close()
    W.close()
    fields = [la.GaugeField(ctx, shape), la.MatrixField(ctx, shape), la.MatrixField(ctx, shape)]
    free0, total = ctx.mem_info()
    full = la.CollectiveVariable(ctx, shape, 8, 0.12)
    full_bytes = full.bytes()
    full.close()
    cv = la.CollectiveVariable(ctx, shape, 8, 0.12, checkpoint=interval)
    ck_bytes = cv.bytes()
    k = la.HMCMetaDKernel(U, fields[0], fields[1], fields[2], c3_bias(la), cv, "omf2", 6.2, 1.0, update_bias=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    k(20)
    ctx.synchronize()
    t_md = time.perf_counter() - t0
    free1, _ = ctx.mem_info()
    cv.close()
    for f in fields:
        f.close()
    field_bytes = 72 * (V // 2) * 16
    U.close()
    return {"lattice": list(shape), "beta": 6.2,
            "hb_or_link_updates_per_s": 5 * 4 * V * n_update / t_upd, "hb_or_updates": n_update,
            "flow_rk3_ms_per_step": 1000.0 * t_flow / flow_steps, "flow_steps": flow_steps, "Q_t8": q,
            "metad_traj_s": t_md, "metad_levels": 8, "checkpoint_interval": interval,
            "metad_workspace_GB_full": (4 * field_bytes + full_bytes) / 1e9,
            "metad_workspace_GB_checkpointed": (4 * field_bytes + ck_bytes) / 1e9,
            "device_used_GB_before_workspace": (total - free0) / 1e9,
            "device_used_GB_during_trajectory": (total - free1) / 1e9,
            "reference_model_GB": (10 + 4.33 * 8) * 4 * V * 144 / 1e9}


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
    ap.add_argument("--no-pt", action="store_true", help="skip the 16 x 16^4 PT-MetaD measurement")
    ap.add_argument("--no-48", action="store_true", help="skip the 48^4 measurement")
    ap.add_argument("--pt-16-only", action="store_true",
                    help="of the PT-MetaD measurements run only 16 x 16^4 (kernel profiles of that workload)")
    ap.add_argument("--metad-only", action="store_true",
                    help="run only the 16^4 MetaD-HMC trajectories and print their line (kernel counter runs)")
    ap.add_argument("--eager", action="store_true",
                    help="no hipGraph anywhere: PT-MetaD steps and 8^4 updates as eager launches (kernel tracing)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.metad_only:
        import lettuce_amd as la
        ctx = la.Context(0)
        print(json.dumps({"metad_16": metad16(la, ctx)}))
        ctx.close()
        return

    # the reference's own code runs as a child process before this process touches the GPU
    ref_cpu = None
    if not args.no_cpu:
        try:
            ref_cpu = reference_cpu_16(args.beta)
        except (OSError, ValueError, subprocess.SubprocessError):
            ref_cpu = None

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
    or_ms_events = ms_or / max(n_or, 1)   # an event pair around every launch (adds the events' cost)
    hb_ms_events = ms_hb / max(n_hb, 1)
    # the colour launches back to back, one event pair around each block of sweeps on the
    # library's stream: 4 OR sweeps = 32 launches, 1 HB sweep = 8 launches
    or_ms = block_ms(ctx.stream(), lambda: la.Iterator.Checkerboard(orx, 4), 5) / 32
    hb_ms = block_ms(ctx.stream(), lambda: la.Iterator.Checkerboard(hb, 1), 10) / 8
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
    la.PlaquetteSum(W)
    t0 = time.perf_counter()
    for _ in range(10):
        la.PlaquetteSum(W)
    plaq_sum_ms = 100.0 * (time.perf_counter() - t0)
    # the rest of the reference's Observables(): Symanzik action, plaquette charge,
    # Wilson loops L = 2, 4, 8 on one chain workspace, Polyakov loop
    chain = la.MatrixField(ctx, shape)
    sym = la.RectangularAction(args.beta, 1.0 + 8.0 / 12.0, -1.0 / 12.0)

    def obs_rest():
        sym.ActionNormalized(W)
        la.TopChargePlaquette(W)
        la.WilsonLoop(0, 2, True, W, chain)
        la.WilsonLoop(2, 4, False, W, chain)
        la.WilsonLoop(4, 8, False, W, chain)
        la.PolyakovLoop(W)
    obs_rest()     # first launches of these kernels and first touch of the chain workspace
    t0 = time.perf_counter()
    for _ in range(3):
        obs_rest()
    obs_rest_ms = (time.perf_counter() - t0) / 3 * 1e3
    chain.close()
    fl.close()
    # improved action (Luescher-Weisz) Checkerboard4 sweeps on the flowed copy
    W.assign(U)
    lw_or = la.OverrelaxationSubgroupKernel(W, action=la.LuescherWeiszAction(args.beta))
    la.Iterator.Checkerboard4(lw_or, 1)
    ctx.synchronize()
    t0 = time.perf_counter()
    la.Iterator.Checkerboard4(lw_or, 2)
    ctx.synchronize()
    lw_or_ms = (time.perf_counter() - t0) / 2 * 1e3
    # Wilson HB + 4 OR with the Checkerboard4 schedule the reference's main program uses
    cb4_hb, cb4_or = la.HeatbathKernel(W, args.beta), la.OverrelaxationSubgroupKernel(W)
    la.Iterator.Checkerboard4(cb4_hb, 1)
    la.Iterator.Checkerboard4(cb4_or, 4)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        la.Iterator.Checkerboard4(cb4_hb, 1)
        la.Iterator.Checkerboard4(cb4_or, 4)
    ctx.synchronize()
    cb4_update_ms = (time.perf_counter() - t0) / 3 * 1e3
    # the same loop body as one call over three updates with the plaquette after each
    # (SU3_main.cpp:358-385): the field is put into the four-colour order once per call
    # (a captured graph per update, so not under --eager)
    cb4_update_loop_ms = None
    if not args.eager:
        la.UpdateSweeps(W, args.beta, 1, 4, 1, la.CB4)   # graph capture outside the timed region
        t0 = time.perf_counter()
        la.UpdateSweeps(W, args.beta, 1, 4, 3, la.CB4)
        cb4_update_loop_ms = (time.perf_counter() - t0) / 3 * 1e3
    # BMW files written and read back through the device codec (#18BMW, all rows)
    # read ceilings of this GPU: 16-byte lane loads from a 2 MiB buffer (L2-resident on every
    # XCD) and from a 1 GiB buffer (HBM)
    read_ceiling = {"l2_GBs": ctx.read_rate(1 << 21) / 1e9, "hbm_GBs": ctx.read_rate(1 << 30) / 1e9}
    io = {}
    try:
        import tempfile
        with tempfile.TemporaryDirectory() as tmp:
            path = os.path.join(tmp, "config.conf")
            t0 = time.perf_counter()
            W.save_bmw(path, full=True)
            t_save = time.perf_counter() - t0
            size = os.path.getsize(path)
            t0 = time.perf_counter()
            ok, _ = W.load_bmw(path, full=True)
            t_load = time.perf_counter() - t0
        io = {"format": "#18BMW", "bytes": size, "save_s": t_save, "load_s": t_load, "checksum_ok": ok,
              "save_GBs": size / t_save / 1e9, "load_GBs": size / t_load / 1e9}
    except (OSError, la.LettuceError) as e:
        io = {"error": str(e)}
    W.close()

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
                     "avg_launch_ms_event_pair_per_launch": or_ms_events,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     # links are stored as two rows (96 B): the minimum this layout moves
                     "layout_bytes_per_launch": LAYOUT_BYTES_PER_LINK_UPDATE * V / 2},
        "hb_kernel": {"avg_launch_ms": hb_ms, "achieved_GBs": hb_gbs, "avg_launch_ms_event_pair_per_launch": hb_ms_events},
        "hbm_GBs_per_update": BYTES_PER_LINK_UPDATE * 5 * 4 * V / (ms_per_step * 1e-3) / 1e9,
        "flow_rk3": {"ms_per_step": flow_ms, "model_GBs": flow_gbs, "stage_avg_launch_ms": stage_ms,
                     "stage_GBs": BYTES_PER_LINK_STAGE * 4 * V / (stage_ms * 1e-3) / 1e9,
                     "link_stage_updates_per_s": 12 * V / (flow_ms * 1e-3)},
        "read_ceiling": read_ceiling,
        "observables_pass_ms": obs_ms,
        "plaquette_sum_ms": plaq_sum_ms,
        "observables_rest_ms": obs_rest_ms,
        "lw_or_sweep_ms": lw_or_ms,
        "cb4_update_ms": cb4_update_ms,
        "cb4_update_loop_ms": cb4_update_loop_ms,
        "bmw_io": io,
        "plaquette": plaq,
        "thermalisation_s": t_therm,
    }
    U.close()
    print("# 8^4 correctness configuration", file=sys.stderr, flush=True)
    result["lattice_8"] = lattice8(la, ctx, graph=not args.eager)
    print("# MetaD-HMC 16^4", file=sys.stderr, flush=True)
    result["metad_16"] = metad16(la, ctx)
    if not args.no_pt:
        print("# PT-MetaD 16 x 16^4", file=sys.stderr, flush=True)
        g16 = (16, 16, 16, 16)
        result["pt_metad"] = pt_metad(la, ctx, [g16] * 16, graph=not args.eager)
        result["pt_metad"]["graph"] = not args.eager
    if not args.no_pt and not args.pt_16_only:
        print("# PT-MetaD mixed 16 x 16^4 + 12^4 + 20^4 + 24^4, and 16 x 8^4", file=sys.stderr, flush=True)
        mixed = [g16] * 16 + [(12,) * 4, (20,) * 4, (24,) * 4]
        result["pt_metad_mixed"] = pt_metad(la, ctx, mixed, graph=not args.eager)
        result["pt_metad_8x4"] = pt_metad(la, ctx, [(8,) * 4] * 16, graph=not args.eager)
    if not args.no_48:
        print("# 48^4", file=sys.stderr, flush=True)
        result["lattice_48"] = lattice48(la, ctx)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(L, args.beta, args.cpu_seconds)
    else:
        result["cpu_baseline"] = None
    result["reference_cpu_16x4"] = ref_cpu
    print(json.dumps(result))


if __name__ == "__main__":
    main()
