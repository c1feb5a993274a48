#!/usr/bin/env python3
"""Benchmark of the explicit time-stepping hot path on one MI355X.

Workload: 2D Cartesian coupled acoustic-elastic layered model, 2000 x 1000 m,
800 x 400 quads, order 4 (about 5.1 M GLL points, 9.2 M DOF), FP32; fluid on
top, two elastic layers with a sinusoidal interface; free top, absorbing
elsewhere; explosion source, 101 receivers.  A step is one central-difference
time step of the whole model (element forces, assembly, boundary terms,
coupling, sources, update, receivers), with all fields resident in HBM.

Prints one JSON line with the throughput in GLL DOF-steps/s, the roofline of
the dominant kernel (element forces of the elastic block) measured with HIP
events over the timed region, and the CPU oracle timed on the host cores on a
bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "in-house-sem_amd"))

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(image):
    """Bytes per step and per force launch of the dominant block, by the
    count of the survey's measurement plan: fields 6 c ng + inverse mass ng,
    metric Rect 3 ne | Affine 5 ne | Curved 5 npe ne | Axisym 8 npe ne,
    coefficients (2 elastic, 1 acoustic) npe ne, l2g npe ne; FP32 values and
    int32 indices (4 bytes each)."""
    npe = image.ngll ** 2
    step = 0
    dominant = None
    for di, d in enumerate(image.domains):
        c, ng = d["ncomp"], d["n_global"]
        step += (6 * c * ng + ng) * 4
        coef = 2 if d["medium"] == "elastic" else 1
        for bi, b in enumerate(d["blocks"]):
            ne = b["n_elements"]
            metric = {"rect": 3 * ne, "affine": 5 * ne, "curved": 5 * npe * ne, "axisym": 8 * npe * ne}[b["kind"]]
            blk = (metric + coef * npe * ne + npe * ne) * 4
            step += blk
            # the force launch reads the block data and the field, writes c * npe * ne forces
            launch = blk + c * ng * 4 + c * npe * ne * 4
            if dominant is None or launch > dominant[0]:
                dominant = (launch, d["medium"], b["kind"], ne, di, bi)
    return step, dominant


def cpu_baseline(sim, dt, max_seconds=20.0):
    """Times the CPU oracle (all host threads) on the same model, stepping only."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_driver import OracleRun, lib as oracle_lib

    threads = oracle_lib().oracle_set_threads(0)
    img = sim.build_setup()
    o = OracleRun(img, dt=dt)
    t0 = time.perf_counter()
    o.advance(1)  # initialisation + first step
    first = time.perf_counter() - t0
    n = max(2, min(200, int(max_seconds / max(first, 1e-6))))
    t0 = time.perf_counter()
    o.advance(n)
    el = time.perf_counter() - t0
    dof = img.n_dofs
    o.close()
    return {"value": dof * n / el, "unit": "DOF-steps/s", "cores": int(threads), "kind": "port",
            "sample": f"{n} steps of the same 800x400 order-4 coupled model after 1 warm-up step "
                      f"({el:.1f} s, OpenMP oracle, stepping only)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--nx", type=int, default=800)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every run of this build uses a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        sys.exit(2)

    import insem_amd  # noqa: F401
    from insem_amd import _lib
    from insem_amd.configs import layered_coupled

    nx = args.nx
    nzf, nzs = nx // 10, (nx * 2) // 10
    sim = layered_coupled(nx=nx, nz_fluid=nzf, nz_solid=(nzs, nzs), duration=1.0)
    t0 = time.perf_counter()
    solver = sim.solver
    setup_s = time.perf_counter() - t0
    img = solver.image
    need = args.warmup + args.steps
    if solver.n_steps < need:
        raise SystemExit(f"the model has {solver.n_steps} steps, fewer than warmup + steps = {need}")
    hip = _lib.hip()
    ctx = ctypes.c_void_p(solver.context)

    solver.advance(args.warmup)
    _lib.check_hip(hip.insem_hip_synchronize(ctx))
    _lib.check_hip(hip.insem_hip_set_timing(ctx, 1))
    fm, fn, um, un = ctypes.c_double(), ctypes.c_int32(), ctypes.c_double(), ctypes.c_int32()
    _lib.check_hip(hip.insem_hip_kernel_time(ctx, ctypes.byref(fm), ctypes.byref(fn), ctypes.byref(um),
                                             ctypes.byref(un)))  # reset
    _lib.check_hip(hip.insem_hip_synchronize(ctx))
    t0 = time.perf_counter()
    solver.advance(args.steps)
    _lib.check_hip(hip.insem_hip_synchronize(ctx))
    el = time.perf_counter() - t0
    _lib.check_hip(hip.insem_hip_kernel_time(ctx, ctypes.byref(fm), ctypes.byref(fn), ctypes.byref(um),
                                             ctypes.byref(un)))
    _lib.check_hip(hip.insem_hip_set_timing(ctx, 0))

    dof = img.n_dofs
    value = dof * args.steps / el
    step_bytes, dominant = algorithmic_bytes(img)
    force_avg_ms = fm.value / max(fn.value, 1)
    bms, bn = ctypes.c_double(), ctypes.c_int32()
    _lib.check_hip(hip.insem_hip_block_time(ctx, dominant[4], dominant[5], ctypes.byref(bms), ctypes.byref(bn)))
    dom_ms = bms.value / max(bn.value, 1)
    achieved = dominant[0] / (dom_ms * 1e-3) / 1e9
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
        "config": {"workload": f"2D coupled acoustic-elastic layered, {nx}x{nzf + 2 * nzs} quads, order 4, "
                               f"curved elastic blocks, 101 receivers",
                   "n_points": img.n_points, "n_dofs": dof, "dt": solver.dt, "setup_s": round(setup_s, 2)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": f"element forces, {dominant[1]} {dominant[2]} block ({dominant[3]} elements)",
                     "algorithmic_bytes": dominant[0], "avg_launch_ms": dom_ms,
                     "launches": bn.value, "force_avg_ms_all_blocks": force_avg_ms,
                     "update_ms_per_step": um.value / args.steps,
                     "step_algorithmic_bytes": step_bytes,
                     "step_frac": step_bytes / (el / args.steps) / (HBM_PEAK_GBS * 1e9)},
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(sim, solver.dt)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
