"""Autotuner of the paired element-force launch (one GPU).

For each polynomial order 4, 5, 6 and each element kind of the coupled
models (Cartesian: curved elastic + rectangular acoustic blocks, as in the
layered workload; axisymmetric m = 0 and m = 1 blocks, as in the open-hole
and cased-hole workloads) it times every variant of the launch on a
representative model at its real size -- level 0/1/2 (registers budgeted
for the default waves per SIMD, at most 6, at most 4, with as many, 6 or 4
work-groups per CU) under LLVM's max-ilp and default scheduling strategies
(insem_hip_set_force_variant) -- as whole time steps replayed from captured
chunks, interleaving the variants over several rounds; then, with the
winning variant, every length of the captured step chunks (16 to 256 steps
per hipGraph, insem_hip_set_chunk).  For each (order, kind) it records the
variant and chunk length with the lowest median step time when they beat
the defaults (variant 0, 32 steps) by more than the noise margin.

    python -m insem_amd.autotune [--table PATH] [--report PATH] [--rounds N]

writes the C table that the library's launches read
(csrc/hip/tuned_forces.inc; rebuild the library afterwards) and a report
with every measurement.  The variants change no result bit (each runs the
same arithmetic in the same order), so the table only moves time.
"""

from __future__ import annotations

import argparse
import ctypes
import os
import statistics
import sys
import time

import numpy as np

from . import _lib
from .configs import cased_hole, layered_coupled, open_hole, open_hole_monopole

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(os.path.dirname(HERE), "csrc", "hip", "tuned_forces.inc")
LEVELS, SCHEDS = 3, 2
CHUNKS = (16, 32, 64, 128, 256)
DEFAULT_CHUNK = 32
VARIANTS = LEVELS * SCHEDS
SCHED_NAMES = ("max-ilp", "default")
# physics codes of the C table (device.hpp: Phys)
ELASTIC_CART, ACOUSTIC_CART, ELASTIC_AXI0, ACOUSTIC_AXI0, ELASTIC_AXI1, ACOUSTIC_AXI1 = 1, 0, 3, 2, 5, 4


def workloads(orders):
    """(name, ngll, key, builder): key = (elastic phys, rect, acoustic phys, rect)."""
    out = []
    for order in orders:
        nx = {4: 800, 5: 640, 6: 540}.get(order, 400)
        out.append((f"layered {nx}x{nx // 2} order {order}", order + 1, (ELASTIC_CART, 0, ACOUSTIC_CART, 1),
                    lambda o=order, n=nx: layered_coupled(nx=n, nz_fluid=n // 10, nz_solid=(n // 5, n // 5), order=o,
                                                          n_receivers=101)))
        if order == 4:
            out.append(("open-hole monopole order 4", 5, (ELASTIC_AXI0, 0, ACOUSTIC_AXI0, 0),
                        lambda: open_hole_monopole()))
        else:
            out.append((f"open-hole draw order {order}", order + 1, (ELASTIC_AXI0, 0, ACOUSTIC_AXI0, 0),
                        lambda o=order: open_hole(np.random.default_rng(11), order=o, f_max=15e3)))
        out.append((f"cased-hole draw order {order}", order + 1, (ELASTIC_AXI1, 0, ACOUSTIC_AXI1, 0),
                    lambda o=order: cased_hole(np.random.default_rng(5), order=o, f_max=8e3)))
    return out


def time_variants(solver, steps, rounds):
    """Median ms per step of each variant, interleaved over rounds."""
    hip, check = _lib.hip(), _lib.check_hip
    ctx = ctypes.c_void_p(solver.context)
    var, paired = ctypes.c_int32(), ctypes.c_int32()
    check(hip.insem_hip_force_variant(ctx, ctypes.byref(var), ctypes.byref(paired)))
    if not paired.value:
        raise RuntimeError("the model has no paired force launch")
    need = 20 + rounds * VARIANTS * (steps + 5)
    if solver.n_steps - solver.step < need:
        raise RuntimeError(f"the model has {solver.n_steps} steps; the sweep needs {need}")
    solver.advance(20)
    times = {v: [] for v in range(VARIANTS)}
    for _ in range(rounds):
        for v in range(VARIANTS):
            check(hip.insem_hip_set_force_variant(ctx, v))
            solver.advance(5)  # capture the chunk shapes of this variant's first steps
            check(hip.insem_hip_prepare(ctx, steps))
            check(hip.insem_hip_synchronize(ctx))
            t0 = time.perf_counter()
            solver.advance(steps)
            check(hip.insem_hip_synchronize(ctx))
            times[v].append(1e3 * (time.perf_counter() - t0) / steps)
    check(hip.insem_hip_set_force_variant(ctx, -1))
    return {v: statistics.median(t) for v, t in times.items()}, times


def time_chunks(solver, variant, steps, rounds):
    """Median ms per step of each chunk length (steps per captured graph)
    with the given force variant, interleaved over rounds."""
    hip, check = _lib.hip(), _lib.check_hip
    ctx = ctypes.c_void_p(solver.context)
    check(hip.insem_hip_set_force_variant(ctx, variant))
    times = {k: [] for k in CHUNKS}
    for _ in range(rounds):
        for k in CHUNKS:
            check(hip.insem_hip_set_chunk(ctx, k))
            check(hip.insem_hip_prepare(ctx, steps))
            check(hip.insem_hip_synchronize(ctx))
            t0 = time.perf_counter()
            solver.advance(steps)
            check(hip.insem_hip_synchronize(ctx))
            times[k].append(1e3 * (time.perf_counter() - t0) / steps)
    check(hip.insem_hip_set_chunk(ctx, 0))
    check(hip.insem_hip_set_force_variant(ctx, -1))
    return {k: statistics.median(t) for k, t in times.items()}


def write_table(path, rows):
    with open(path, "w") as fh:
        fh.write("// Autotuned paired force launches: {ngll, elastic physics, rect, acoustic physics, rect, level, sched,\n"
                 "// steps per captured chunk (0: 32)} (written by in-house-sem_amd/insem_amd/autotune.py from\n"
                 "// measurements on one MI355X).\n")
        for r in rows:
            fh.write("    {%d, %d, %d, %d, %d, %d, %d, %d},  // %s\n" % (*r[:8], r[8]))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--table", default=TABLE)
    ap.add_argument("--report", default=None, help="write every measurement here (text)")
    ap.add_argument("--orders", default="4,5,6")
    ap.add_argument("--steps", type=int, default=100, help="steps per timing")
    ap.add_argument("--rounds", type=int, default=3)
    ap.add_argument("--margin", type=float, default=0.02, help="a variant must beat variant 0 by this fraction")
    args = ap.parse_args(argv)
    orders = [int(x) for x in args.orders.split(",")]
    rows, lines = [], []
    for name, ngll, key, build in workloads(orders):
        t0 = time.perf_counter()
        sim = build()
        chunk_steps = 256
        sim.set_time(n_steps=20 + args.rounds * VARIANTS * (args.steps + 5) + args.rounds * len(CHUNKS) * chunk_steps + 10)
        solver = sim.solver
        med, all_t = time_variants(solver, args.steps, args.rounds)
        best = min(med, key=med.get)
        chosen = best if med[best] < (1.0 - args.margin) * med[0] else 0
        cmed = time_chunks(solver, chosen, chunk_steps, args.rounds)
        solver.close()
        kbest = min(cmed, key=cmed.get)
        chunk = kbest if cmed[kbest] < (1.0 - args.margin) * cmed[DEFAULT_CHUNK] else DEFAULT_CHUNK
        desc = ", ".join(f"v{v} {med[v]:.4f}" for v in range(VARIANTS))
        cdesc = ", ".join(f"{k} {cmed[k]:.4f}" for k in CHUNKS)
        line = (f"{name}: ngll {ngll}, ms/step median of {args.rounds} -> {desc}; "
                f"chosen v{chosen} (level {chosen % LEVELS}, {SCHED_NAMES[chosen // LEVELS]}); "
                f"steps per chunk -> {cdesc}; chosen {chunk} [{time.perf_counter() - t0:.0f} s]")
        print(line, flush=True)
        lines.append(line)
        lines.append("    all: " + "; ".join(f"v{v} " + " ".join(f"{x:.4f}" for x in all_t[v]) for v in range(VARIANTS)))
        if chosen or chunk != DEFAULT_CHUNK:
            rows.append((ngll, *key, chosen % LEVELS, chosen // LEVELS, 0 if chunk == DEFAULT_CHUNK else chunk, name))
    write_table(args.table, rows)
    if args.report:
        with open(args.report, "w") as fh:
            fh.write("\n".join(lines) + "\n")
    print(f"table: {args.table} ({len(rows)} tuned entries)")


if __name__ == "__main__":
    sys.exit(main())
