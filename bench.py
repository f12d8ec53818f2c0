"""Benchmark of the MI355X FDTD core on its headline workload.

Workload (one "step" = one time step of every shot of the batch):
batched 2-D elastic, 96 shots over three layered snow / firn / ice models
(512x256, 1024x512, 2048x768 cells + C-PML 20, dx = 0.5 m), 150 Hz, FP32,
all shots advanced together in the same kernel launches, hipGraph blocks of
50 steps.  Beside it, in the same run:
  * the plain FP64 GPU path (one shot per plan, no batching) on one shot of
    each model, for the FP32-batched / FP64-plain speed ratio and the FP32
    trace error against FP64;
  * the oracle (FP64 C restatement of the legacy loops, OpenMP) on a bounded
    sample of the same workload on the host cores (cpu_baseline).

Prints one JSON line.  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s & %HBM roofline, FP32 batched vs FP64 plain; FWI grad s/shot"
HBM_PEAK_GBS = 8000.0
PROFILE_STEPS = 20


def algorithmic_bytes(problems, nshots_by_model, itemsize):
    """Compulsory HBM bytes per launch of the two field kernels of the
    2-D elastic step, over the whole batch.

    stress  : read vx vz, read+write sxx szz sxz (per shot); read the 8
              staggered-node Cij averages (per model, shared by its shots);
              read+write the 2+2 velocity-derivative psi slabs (per shot).
    velocity: read sxx szz sxz, read+write vx vz (per shot); read dt/rho x2
              and 1-gamma x2 (per model); read+write the 2+2 stress-derivative
              psi slabs (per shot).
    1-D C-PML profiles and source samples are counted too (negligible)."""
    st = vel = 0
    for p, S in zip(problems, nshots_by_model):
        n = p.nx * p.nz
        lx, _, lz = p.slabs()
        psi = 2 * (p.nz * 2 * lx) + 2 * (2 * lz * p.nx)     # elements, one kernel's half
        prof = 6 * p.nx + 7 * p.nz
        st += S * (2 * n + 6 * n + 2 * psi) + 8 * n + prof
        vel += S * (3 * n + 4 * n + 2 * psi + 2) + 4 * n + prof
    return st * itemsize, vel * itemsize


def cpu_baseline(problems, budget_s):
    """Oracle on one shot of the middle model, on the host cores, for a
    bounded number of steps (~budget_s seconds)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.build()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    if os.environ.get("OMP_NUM_THREADS"):
        cores = min(cores, int(os.environ["OMP_NUM_THREADS"]))
    p = problems[1]
    shot = len(p.shots) // 2
    o = pyoracle.OracleRun(p, shot=shot, threads=cores)
    t0 = time.perf_counter()
    o.run(0, 10)
    probe = time.perf_counter() - t0
    n = int(max(10, min(p.nt - 10, budget_s / max(probe / 10, 1e-6))))
    t0 = time.perf_counter()
    o.run(10, n)
    el = time.perf_counter() - t0
    cells = p.nx * p.nz * n
    return {"value": cells / el / 1e9, "unit": "Gcell-updates/s", "cores": cores, "kind": "port",
            "sample": f"oracle (FP64 C/OpenMP restatement of seismic2), 1 shot of the "
                      f"{p.nx}x{p.nz} padded model, {n} steps, {el:.1f} s"}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--graph-block", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(the batch shards by shot across GPUs by design; this build measures one GPU)", file=sys.stderr)
        return 2

    import torch
    from seidart_amd import configs
    from seidart_amd.problem import Problem
    from seidart_amd.solver import Simulation

    K, W = a.steps, a.warmup
    nt = W + K + PROFILE_STEPS
    t0 = time.perf_counter()
    problems = configs.batched_layered(nt=nt)
    setup_s = time.perf_counter() - t0
    shots = [len(p.shots) for p in problems]
    cells_per_step = sum(p.nx * p.nz * s for p, s in zip(problems, shots))

    sim = Simulation(problems, precision="fp32")
    gb = a.graph_block if W >= a.graph_block else 0
    sim.run(W, graph_block=gb)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    sim.run(K, graph_block=a.graph_block)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    sim.check()
    value = cells_per_step * K / wall / 1e9
    # per-kernel durations (HIP events around each launch, same stream)
    k_ms = sim.profile(PROFILE_STEPS)
    sim.check()
    b_st, b_vel = algorithmic_bytes(problems, shots, 4)
    dom = 0 if k_ms[0] >= k_ms[1] else 1
    dom_bytes = (b_st, b_vel)[dom]
    achieved = dom_bytes / (k_ms[dom] * 1e-3) / 1e9
    step_bytes = b_st + b_vel
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": ("k_el2_stress<float>", "k_el2_vel<float>")[dom],
                "kernel_ms": round(k_ms[dom], 4),
                "alg_bytes_per_launch": int(dom_bytes),
                "other_kernels_ms": {"stress": round(k_ms[0], 4), "velocity": round(k_ms[1], 4),
                                     "gather": round(k_ms[2], 4)},
                "step_alg_GBs": round(step_bytes / (sum(k_ms[:2]) * 1e-3) / 1e9, 1)}
    traces32 = sim.traces()
    sim.close()
    del sim
    torch.cuda.empty_cache()

    plain = None
    if not a.no_plain:
        # plain FP64 GPU path: one shot per plan, sample shot = middle shot of each model
        tot_cells, tot_s, errs = 0, 0.0, []
        off = 0
        for m, p in enumerate(problems):
            s = len(p.shots) // 2
            single = Problem(p.kind, p.nx, p.ny, p.nz, p.dx, p.dy, p.dz, p.dt, p.nt, p.coef,
                             p.prof, [p.shots[s]], p.gamma)
            sp = Simulation(single, precision="fp64")
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            sp.run(nt, graph_block=a.graph_block)
            torch.cuda.synchronize()
            tot_s += time.perf_counter() - t1
            sp.check()
            tot_cells += p.nx * p.nz * nt
            t64 = sp.traces()[0]
            t32 = traces32[off + s]
            errs.append(float(np.linalg.norm(t32 - t64) / np.linalg.norm(t64)))
            sp.close()
            off += len(p.shots)
        pv = tot_cells / tot_s / 1e9
        plain = {"value": round(pv, 3), "unit": "Gcell-updates/s", "dtype": "f64",
                 "sample": "middle shot of each model, one plan per shot, full run",
                 "speedup_fp32_batched": round(value / pv, 3),
                 "fp32_trace_rel_l2_vs_fp64": [float(f"{e:.3e}") for e in errs]}

    cpu = None if a.no_cpu else cpu_baseline(problems, a.cpu_seconds)
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": round(wall / K * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched_layered: 2-D elastic, 96 shots (32 per model) over 3 "
                               "layered snow/firn/ice models 512x256, 1024x512, 2048x768 "
                               "(+C-PML 20), dx 0.5 m, 150 Hz, one launch set per step, "
                               "hipGraph per 50 steps",
                   "cells_per_step": cells_per_step, "shots": sum(shots),
                   "graph_block": a.graph_block},
        "gpu_ms_events": round(gpu_ms, 3),
        "roofline": roofline,
        "fp64_plain": plain,
        "cpu_baseline": cpu,
        "host_setup_s": round(setup_s, 2),
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
