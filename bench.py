#!/usr/bin/env python3
"""Benchmark of the explicit time-stepping hot path on one MI355X.

Workload: 2D Cartesian coupled acoustic-elastic layered model, 2000 x 1000 m,
800 x 400 quads, order 4 (about 5.1 M GLL points, 9.2 M DOF), FP32; fluid on
top, two elastic layers with a sinusoidal interface; free top, absorbing
elsewhere; explosion source, 101 receivers.  A step is one central-difference
time step of the whole model (element forces, assembly, boundary terms,
coupling, sources, update, receivers), with all fields resident in HBM.

Prints one JSON line with the throughput in GLL DOF-steps/s (timed over the
production path: step chunks replayed as HIP graphs), the roofline of the
dominant kernel (element forces of both blocks, with the update of the
points they own fused in) from HIP events around each of its launches
in a second, directly launched pass of further steps, the HBM traffic of that
launch measured with rocprofv3 counters (read from the committed summary
under profiles/), the CPU oracle timed on the host cores on a bounded
sample (about 20 s) of the same workload, and (under "batched") the solves/s
and latencies of 128 cased-hole dipole models run as batches against a
sample of them run one after another on the plain path, and (under "served") the
throughput and p50/p99 latency of a Poisson stream of mixed requests served
with continuous batching (at 8 requests/s by default; --serve-sweep adds
other arrival rates), against the same stream served one at a time.

`--calibrate-copy` instead runs a plain device-to-device copy of known size
(the reference kernel for scaling the FETCH_SIZE / WRITE_SIZE counters).
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "in-house-sem_amd"))

HBM_PEAK_GBS = 8000.0
OPEN_HOLE_WORKLOAD = ("axisymmetric open-hole monopole, order 4 with GLJ on the axis, 20 kHz AMR, formation radius 4 m, "
                      "13 receivers")
# HBM traffic of the dominant force launch of each workload, measured with
# rocprofv3 --pmc (the L2's memory-side read and write requests resolved by
# request size), from the committed counter summaries.
TRAFFIC_FILES = [os.path.join(ROOT, "profiles", f) for f in
                 ("pmc_forces_pair_layered_800x400.txt", "pmc_open_hole_monopole.txt")]


def progress(msg):
    """A progress line on stderr (stdout carries only the result line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(image, hip, ctx, check):
    """Bytes per step, and per force launch of each block.

    Step: fields 6 c ng + inverse mass ng, plus every block's data: metric
    Rect 3 ne | Affine 5 ne | Curved 5 npe ne | Axisym 8 npe ne, coefficients
    (2 elastic, 1 acoustic) npe ne and l2g npe ne; FP32 values and int32
    indices, 4 bytes each.

    Force work of a block in a time step: the block's data, u at the points it
    gathers (c per point), and per element slot either the update of its
    point (owned points and the first slot of batch-local points: read v
    and 1/m, write u and v: 3 c + 1 values), nothing (the other slots of
    batch-local points, summed in the work-group) or c forces written for
    the tail's assembly."""
    npe = image.ngll ** 2
    step = 0
    blocks = []
    for di, d in enumerate(image.domains):
        c, ng = d["ncomp"], d["n_global"]
        step += (6 * c * ng + ng) * 4
        coef = 2 if d["medium"] == "elastic" else 1
        for bi, b in enumerate(d["blocks"]):
            ne = b["n_elements"]
            metric = {"rect": 3 * ne, "affine": 5 * ne, "curved": 5 * npe * ne, "axisym": 8 * npe * ne}[b["kind"]]
            blk = (metric + coef * npe * ne + npe * ne) * 4
            step += blk
            owned, read, local = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
            check(hip.insem_hip_block_points(ctx, di, bi, ctypes.byref(owned), ctypes.byref(read)))
            check(hip.insem_hip_block_local_slots(ctx, di, bi, ctypes.byref(local)))
            slots = npe * ne
            written = slots - owned.value - local.value
            launch = blk + c * read.value * 4 + owned.value * (3 * c + 1) * 4 + written * c * 4
            blocks.append({"bytes": launch, "medium": d["medium"], "kind": b["kind"], "n_elements": ne,
                           "domain": di, "block": bi, "owned_slots": owned.value, "local_slots": local.value,
                           "points_read": read.value})
    return step, blocks


def measured_traffic(workload):
    """HBM bytes per launch of the dominant force launch from the committed
    counter summary measured on this workload ("workload:" and
    "traffic_bytes_per_launch:" lines), or None."""
    for path in TRAFFIC_FILES:
        try:
            with open(path) as fh:
                text = fh.read()
        except OSError:
            continue
        w = re.search(r"^workload: (.*)$", text, re.M)
        t = re.search(r"^traffic_bytes_per_launch: ([0-9.eE+]+)$", text, re.M)
        if w and t and w.group(1).strip() == workload:
            return float(t.group(1))
    return None


def cpu_baseline(sim, dt, max_seconds=20.0):
    """Times the CPU oracle (all host threads) on the same model, stepping only."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_driver import OracleRun, lib as oracle_lib

    threads = oracle_lib().oracle_set_threads(0)
    img = sim.build_setup()
    o = OracleRun(img, dt=dt)
    o.advance(1)  # initialisation + first step
    t0 = time.perf_counter()
    o.advance(10)  # per-step cost, to size the timed sample
    per_step = (time.perf_counter() - t0) / 10
    n = max(2, min(20000, int(max_seconds / max(per_step, 1e-6))))
    t0 = time.perf_counter()
    o.advance(n)
    el = time.perf_counter() - t0
    dof = img.n_dofs
    o.close()
    return {"value": dof * n / el, "unit": "DOF-steps/s", "cores": int(threads), "kind": "port",
            "sample": f"{n} steps of the same 800x400 order-4 coupled model after 11 warm-up steps "
                      f"({el:.1f} s, OpenMP oracle on {int(threads)} threads, stepping only)"}


def measure_plain(solver, steps, warmup, what="the model"):
    """The plain path's figures of one model: `steps` replayed steps (the
    production path: step chunks replayed as hipGraphs) timed after
    `warmup`, then 200 further steps launched directly with HIP events
    around each force launch on the solver's stream (200 whatever `steps`
    is, so that the per-launch average does not depend on it), and the
    algorithmic bytes of the step and of each block's force work."""
    from insem_amd import _lib

    hip, check = _lib.hip(), _lib.check_hip
    ctx = ctypes.c_void_p(solver.context)
    timing_steps = 200
    need = warmup + steps + timing_steps
    if solver.n_steps < need:
        raise SystemExit(f"{what} has {solver.n_steps} steps, fewer than the {need} this run needs")
    # every step-chunk graph the warmup and the timed advance replay is
    # captured and instantiated here, before the warmup, so that the timed
    # steps follow the warmup's without an idle device in between (a chunk's
    # graph depends on its length and place in the advance, not on its step)
    t0 = time.perf_counter()
    check(hip.insem_hip_prepare(ctx, warmup))
    check(hip.insem_hip_prepare(ctx, steps))
    check(hip.insem_hip_synchronize(ctx))
    capture_s = time.perf_counter() - t0
    solver.advance(warmup)
    check(hip.insem_hip_synchronize(ctx))
    t0 = time.perf_counter()
    solver.advance(steps)
    check(hip.insem_hip_synchronize(ctx))
    el = time.perf_counter() - t0
    check(hip.insem_hip_set_timing(ctx, 1))
    fm, fn, um, un = ctypes.c_double(), ctypes.c_int32(), ctypes.c_double(), ctypes.c_int32()
    check(hip.insem_hip_kernel_time(ctx, ctypes.byref(fm), ctypes.byref(fn), ctypes.byref(um), ctypes.byref(un)))
    solver.advance(timing_steps)
    check(hip.insem_hip_synchronize(ctx))
    check(hip.insem_hip_kernel_time(ctx, ctypes.byref(fm), ctypes.byref(fn), ctypes.byref(um), ctypes.byref(un)))
    check(hip.insem_hip_set_timing(ctx, 0))
    step_bytes, blocks = algorithmic_bytes(solver.image, hip, ctx, check)
    return {"el": el, "capture_s": capture_s, "timing_steps": timing_steps, "force_ms": fm.value, "force_n": fn.value, "tail_ms": um.value,
            "step_bytes": step_bytes, "blocks": blocks, "hip": hip, "ctx": ctx, "check": check}


def plain_leg(sim, workload, steps, warmup):
    """One model on the plain path, measured as the headline is: DOF-steps/s
    and the roofline fractions of the step (its per-step algorithmic bytes) and
    of its force launches (their algorithmic bytes)."""
    solver = sim.solver
    img = solver.image
    m = measure_plain(solver, steps, warmup, workload)
    ms = 1e3 * m["el"] / steps
    force_ms = m["force_ms"] / m["timing_steps"]  # all force launches of a step
    force_bytes = sum(b["bytes"] for b in m["blocks"])
    out = {"workload": workload, "n_dofs": img.n_dofs, "n_elements": sum(d["n_elements"] for d in img.domains),
           "ms_per_step": ms, "value": img.n_dofs * steps / m["el"], "unit": "DOF-steps/s",
           "graph_capture_s": round(m["capture_s"], 4),
           "step_algorithmic_bytes": m["step_bytes"],
           "step_frac": m["step_bytes"] / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9),
           "force_ms_per_step": force_ms, "force_launches_per_step": m["force_n"] / m["timing_steps"],
           "force_algorithmic_bytes": force_bytes,
           "force_frac": force_bytes / (force_ms * 1e-3) / (HBM_PEAK_GBS * 1e9),
           "force_traffic_bytes": measured_traffic(workload),
           "tail_ms_per_step": m["tail_ms"] / m["timing_steps"]}
    solver.close()
    return out


def batched_leg(n_models, duration, plain_sample, batch_size=5, seed=2026):
    """Cased-hole dipole models (azimuthal order 1) drawn from the batched
    workload's distribution, solved end to end (host setup, power-iteration
    Δt, all steps, traces): all n_models on the GPU as batches of up to
    batch_size models of one polynomial order and similar step count, all
    of them side by side on their own streams; and every k-th model of the
    same draw (k = n_models // plain_sample, so that the sample spans the
    draw's orders and sizes) one after another on the plain path.  The two
    paths are compared per model-step (vs_plain) and per solve.  Every solve is
    queued at t = 0; a solve's latency is the time until its traces are
    ready."""
    import numpy as np

    import insem_amd as insem
    from insem_amd.configs import cased_hole

    def draw():
        rng = np.random.default_rng(seed)
        return [cased_hole(rng, duration=duration) for _ in range(n_models)]

    n_plain = min(plain_sample, n_models)
    stride = max(1, n_models // max(1, n_plain))
    t0 = time.perf_counter()
    plain_done, steps = [], 0
    for sim in draw()[::stride][:n_plain]:
        s = sim.solver
        s.run()
        _ = s.traces
        steps += s.n_steps
        s.close()
        plain_done.append(time.perf_counter() - t0)
    plain_s = time.perf_counter() - t0
    progress(f"batched leg: {n_plain} plain solves in {plain_s:.1f} s")

    sims = draw()
    orders = sorted({sim.discretization.order for sim in sims})
    batch_done = [0.0] * n_models
    # insem_amd.solve_many: each order's models sorted by step count and cut
    # into batches of up to batch_size, all side by side on their own streams
    n_batches = sum(-(-sum(1 for s in sims if s.discretization.order == o) // batch_size) for o in orders)
    t0 = time.perf_counter()

    batch_steps = 0

    def done(i, t, n_steps):
        nonlocal batch_steps
        batch_done[i] = t
        batch_steps += n_steps
        if sum(1 for x in batch_done if x > 0.0) % 32 == 0:
            progress(f"batched leg: {sum(1 for x in batch_done if x > 0.0)} models done at {t:.1f} s")

    insem.solve_many(sims, batch_size=batch_size, on_done=done)
    batch_s = time.perf_counter() - t0

    def pct(v, q):
        return float(np.percentile(np.asarray(v) * 1e3, q))

    value, plain_rate = n_models / batch_s, n_plain / plain_s
    # the plain sample's models differ in length from the whole set: compare
    # the two paths per model-step (the rate of steps of one model each)
    batch_step_rate, plain_step_rate = batch_steps / batch_s, steps / plain_s
    return {"metric": "batched solves/s", "value": value, "unit": "solves/s",
            "plain_solves_per_s": plain_rate,
            "batched_model_steps_per_s": batch_step_rate, "plain_model_steps_per_s": plain_step_rate,
            "vs_plain": batch_step_rate / plain_step_rate, "vs_plain_solves": value / plain_rate,
            "latency_ms": {"batched_p50": pct(batch_done, 50), "batched_p99": pct(batch_done, 99),
                           "plain_sequential_p50": pct(plain_done, 50), "plain_sequential_p99": pct(plain_done, 99)},
            "sample": f"{n_models} cased-hole dipole models (seed {seed}, orders {orders}, {duration * 1e3:g} ms each, "
                      f"{batch_steps} model-steps) as {n_batches} batches of up to {batch_size} models of one order "
                      f"and similar step count, side by side, end to end; plain: every {stride}th of them ({n_plain}) "
                      f"({steps} model-steps) one after another, end to end (latencies of that sequential queue)",
            "batches": n_batches}


def served_leg(n_requests, rate, sweep, slots=3, batch=16, seed=7):
    """A Poisson stream of mixed requests (60 % cased-hole dipole, 30 %
    open-hole monopole at 3-8 ms; 10 % 2D layered coupled at 1-4 M nodes for
    0.5-1 s) served by the continuous-batching loop at each arrival rate of
    the sweep (the same requests, inter-arrival times scaled; the layered
    requests in a slot class of their own), and at `rate` also one request
    at a time on the plain path (one slot, no batching, no large slot).
    Latency = arrival to traces (and STC for boreholes) ready; the rate at
    which solves/s stops following the arrival rate is the loop's
    saturation, compared with the one-at-a-time rate at a saturating
    arrival rate (the device never idles there either)."""
    import numpy as np

    from insem_amd import serve
    from insem_amd.configs import served_request as build, served_stream

    unit_arrivals, kinds, seeds = served_stream(n_requests, seed)

    stc = dict(slowness=np.linspace(100e-6, 1500e-6, 141), window=5e-4, times=np.linspace(0.0, 2.5e-3, 51))

    def run(r, slots, batch, large_slots):
        stream = [(float(t) / r, build(k, sd)) for t, k, sd in zip(unit_arrivals, kinds, seeds)]
        srv = serve.Server(max_slots=slots, max_batch=batch, stc=stc, max_large_slots=large_slots)
        t0 = time.perf_counter()
        reqs = srv.serve(stream)
        out = serve.summary(reqs, time.perf_counter() - t0)
        out["loop_thread_s"] = {k: round(v, 3) for k, v in srv.stats.items()}
        for q in reqs:
            if not q.ok:
                progress(f"served leg: request {q.rid} ({kinds[q.rid]}) failed: {q.error}")
        lat = {k: [x.latency * 1e3 for x in reqs if x.ok and kinds[x.rid] == k] for k in ("cased", "open", "layered")}
        out["latency_p50_ms_by_kind"] = {k: float(np.percentile(v, 50)) for k, v in lat.items() if v}
        return out

    out = {"sweep": []}
    for r in sorted(set(sweep) | {rate}):
        res = dict(rate=r, **run(r, slots, batch, 1))
        out["sweep"].append(res)
        progress(f"served leg: {res['solves_per_s']:.3f} solves/s at {r:g}/s "
                 f"(p50 {res['latency_p50_ms']:.0f} ms, p99 {res['latency_p99_ms']:.0f} ms)")
        if r == rate:
            out["served"] = res
    out["one_at_a_time"] = dict(rate=rate, **run(rate, 1, 1, 0))
    progress(f"served leg: one at a time {out['one_at_a_time']['solves_per_s']:.3f} solves/s")
    out["saturation_solves_per_s"] = max(x["solves_per_s"] for x in out["sweep"])
    out["vs_one_at_a_time"] = out["saturation_solves_per_s"] / out["one_at_a_time"]["solves_per_s"]
    out["sample"] = (f"{n_requests} requests, Poisson arrivals (seed {seed}) at rates {sorted(set(sweep) | {rate})}/s: "
                     f"{int((kinds == 'cased').sum())} cased-hole, {int((kinds == 'open').sum())} open-hole, "
                     f"{int((kinds == 'layered').sum())} layered; served: {slots} slots, batches of up to {batch}, "
                     f"plus one slot for layered requests; "
                     f"one at a time at {rate:g}/s")
    return out


def calibrate_copy(n_bytes=1 << 30, reps=5):
    """A plain copy of n_bytes between two device buffers, reps times."""
    import torch

    x = torch.ones(n_bytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n_bytes, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--nx", type=int, default=800)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle leg")
    ap.add_argument("--no-open-hole", action="store_true", help="skip the open-hole monopole leg")
    ap.add_argument("--calibrate-copy", action="store_true", help="run only a 1 GiB device copy (counter scale)")
    ap.add_argument("--only-cased-hole", type=int, default=0, metavar="ORDER",
                    help="run only one cased-hole dipole model of this order (4-6) from the batched workload's "
                         "distribution, alone on the plain path, and print its figures (for profiling its kernels)")
    ap.add_argument("--only-open-hole", action="store_true",
                    help="run only the open-hole monopole leg and print its figures (for profiling it)")
    ap.add_argument("--batch-models", type=int, default=128, help="models of the batched leg (0: skip it)")
    ap.add_argument("--batch-plain-sample", type=int, default=24,
                    help="models of the batched leg's sequential plain comparison (the first of the same draw)")
    ap.add_argument("--batch-duration", type=float, default=5e-3, help="simulated time of each batched model (s)")
    ap.add_argument("--batch-size", type=int, default=5,
                    help="models per batch (each polynomial order's models, sorted by size, in batches of up to this)")
    ap.add_argument("--serve-requests", type=int, default=60, help="requests of the served leg (0: skip it)")
    ap.add_argument("--serve-rate", type=float, default=8.0,
                    help="Poisson arrival rate of the served leg's comparison with one-at-a-time service (1/s)")
    ap.add_argument("--serve-sweep", default="8",
                    help="arrival rates (1/s) the served loop runs at, comma-separated (e.g. 2,4,8)")
    ap.add_argument("--serve-slots", type=int, default=3, help="batches the served loop keeps on the device at once")
    ap.add_argument("--serve-batch", type=int, default=16, help="requests per batch of the served loop")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every run of this build uses a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.only_cased_hole:
        import numpy as np

        from insem_amd.configs import cased_hole

        sim = cased_hole(np.random.default_rng(2026), order=args.only_cased_hole)
        workload = (f"axisymmetric cased-hole dipole (azimuthal order 1), order {args.only_cased_hole}, "
                    f"batched workload draw seed 2026, 8 receivers, alone on the plain path")
        print(json.dumps(plain_leg(sim, workload, args.steps, args.warmup)), flush=True)
        return
    if args.only_open_hole:
        from insem_amd.configs import open_hole_monopole

        print(json.dumps(plain_leg(open_hole_monopole(), OPEN_HOLE_WORKLOAD, args.steps, args.warmup)), flush=True)
        return

    import insem_amd  # noqa: F401
    from insem_amd.configs import layered_coupled

    nx = args.nx
    nzf, nzs = nx // 10, (nx * 2) // 10
    sim = layered_coupled(nx=nx, nz_fluid=nzf, nz_solid=(nzs, nzs), duration=1.0)
    t0 = time.perf_counter()
    solver = sim.solver
    setup_s = time.perf_counter() - t0
    img = solver.image
    m = measure_plain(solver, args.steps, args.warmup)
    hip, ctx, check = m["hip"], m["ctx"], m["check"]
    el, timing_steps, step_bytes, blocks = m["el"], m["timing_steps"], m["step_bytes"], m["blocks"]
    dof = img.n_dofs
    value = dof * args.steps / el
    launches_per_step = m["force_n"] / timing_steps
    if round(launches_per_step) == 1:
        # one force launch per step covers every block of both domains
        kbytes = sum(b["bytes"] for b in blocks)
        k_ms = m["force_ms"] / max(m["force_n"], 1)
        kname = "element forces + owned-point update, all blocks of both domains + previous-step receivers"
        kblocks = blocks
    else:
        dom = max(blocks, key=lambda b: b["bytes"])
        bms, bn = ctypes.c_double(), ctypes.c_int32()
        check(hip.insem_hip_block_time(ctx, dom["domain"], dom["block"], ctypes.byref(bms), ctypes.byref(bn)))
        kbytes, k_ms = dom["bytes"], bms.value / max(bn.value, 1)
        kname = f"element forces + owned-point update, {dom['medium']} {dom['kind']} block ({dom['n_elements']} elements)"
        kblocks = [dom]
    achieved = kbytes / (k_ms * 1e-3) / 1e9
    workload = (f"2D coupled acoustic-elastic layered, {nx}x{nzf + 2 * nzs} quads, order 4, "
                f"curved elastic blocks, 101 receivers")
    result = {
        "metric": "GLL DOF-steps/s at 1 GPU",
        "value": value,
        "unit": "DOF-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload, "n_points": img.n_points, "n_dofs": dof, "dt": solver.dt,
                   "setup_s": round(setup_s, 2), "graph_capture_s": round(m["capture_s"], 4)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(workload),
                     "kernel": kname, "algorithmic_bytes": kbytes, "avg_launch_ms": k_ms,
                     "launches_per_step": launches_per_step,
                     "blocks": [{k: b[k] for k in ("medium", "kind", "n_elements", "bytes", "owned_slots",
                                                   "local_slots", "points_read")} for b in kblocks],
                     "tail_ms_per_step": m["tail_ms"] / timing_steps, "timing_steps": timing_steps,
                     "step_algorithmic_bytes": step_bytes,
                     "step_frac": step_bytes / (el / args.steps) / (HBM_PEAK_GBS * 1e9)},
    }
    dt = solver.dt
    solver.close()  # free the model's device memory before the other legs
    progress(f"layered model: {result['ms_per_step']:.4f} ms/step")
    if not args.no_open_hole:
        from insem_amd.configs import open_hole_monopole

        result["open_hole"] = plain_leg(open_hole_monopole(), OPEN_HOLE_WORKLOAD, args.steps, args.warmup)
        progress(f"open-hole model: {result['open_hole']['ms_per_step']:.4f} ms/step, "
                 f"step at {result['open_hole']['step_frac']:.3f} of the HBM roofline")
    if args.batch_models > 0:
        result["batched"] = batched_leg(args.batch_models, args.batch_duration, args.batch_plain_sample,
                                        args.batch_size)
        progress(f"batched leg: {result['batched']['value']:.3f} solves/s")
    if args.serve_requests > 0:
        sweep = [float(x) for x in args.serve_sweep.split(",") if x]
        result["served"] = served_leg(args.serve_requests, args.serve_rate, sweep, args.serve_slots, args.serve_batch)
    if not args.no_cpu:
        progress("CPU oracle leg")
    result["cpu_baseline"] = None if args.no_cpu else cpu_baseline(sim, dt)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
