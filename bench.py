"""Benchmark of the MI355X FDTD core on its headline workload.

Workload (one "step" = one time step of every shot of the batch):
batched 2-D elastic, 96 shots over three layered snow / firn / ice models
(512x256, 1024x512, 2048x768 cells + C-PML 20, dx = 0.5 m), 150 Hz, FP32,
all shots advanced together in the same kernel launches, hipGraph blocks of
up to 50 steps captured before the timed region.  Beside it, in the same run:
  * the full 3000-step FP32 batched run and the plain FP64 GPU path (every
    one of the 96 shots alone in its own plan, no batching, 3000 steps), for
    the FP32-batched / FP64-plain speed ratio and the FP32 trace error of
    every shot against its FP64 run;
  * the other reference configurations at their full lengths (plain_psv,
    gpr, aniso25), FP32 and FP64, with the FP32 trace error;
  * the oracle (FP64 C restatement of the legacy loops, OpenMP) on a bounded
    sample of the same workload on the host cores (cpu_baseline);
  * the FWI gradient workload (fwi_vti: 2-D VTI elastic 1024x512 + C-PML 20,
    16 shots, 6000 steps, L2 trace misfit, gradients for C11 C13 C33 C55 rho,
    FP32 with FP64 accumulation, checkpoints every 64 steps): seconds per
    shot of one full gradient (forward with checkpoints + adjoint with
    recomputation) and its ratio to a plain forward of the same shots.

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
    """Compulsory HBM bytes of one launch of the single-pass 2-D elastic step
    kernel over the whole batch: per shot, read and write vx vz sxx szz sxz
    once and read and write each C-PML memory array on its slab cells once
    (4 + 4 arrays with the legacy sharing, 6 + 6 with the split C-PML); per
    model (shared by its shots), read its coefficient arrays once (up to 12:
    8 staggered-node Cij averages, dt/rho x2, 1-gamma x2; the c15/c35
    averages and 1-gamma are absent for isotropic, lossless models) plus the
    1-D C-PML profiles."""
    from seidart_amd import _hip
    from seidart_amd.problem import kernel_coefs
    tot = 0
    for p, S in zip(problems, nshots_by_model):
        n = p.nx * p.nz
        lx, _, lz = p.slabs()
        kid = _hip.SDT_ELASTIC2 | (_hip.SDT_SPLIT_CPML if getattr(p, "cpml_split", False) else 0)
        psi = _hip.psi_elems(kid, p.nx, p.ny, p.nz, lx, 0, lz)
        prof = 6 * p.nx + 7 * p.nz
        K = kernel_coefs(p)
        ncoef = 12 - (4 if all(not np.any(K[q]) for q in (2, 4, 5, 6)) else 0) \
            - (2 if all(np.all(K[q] == 1.0) for q in (10, 11)) else 0)
        tot += S * (10 * n + 2 * psi) + ncoef * n + prof
    return tot * itemsize


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(budget_s):
    """Oracle (FP64 C restatement of the legacy loops, OpenMP) on one shot of
    the largest model of the workload, on the host cores, for about budget_s
    seconds of stepping from the start of the run; plus a shorter
    single-thread sample for the per-core rate."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from seidart_amd import configs
    pyoracle.build()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    if os.environ.get("OMP_NUM_THREADS"):
        cores = min(cores, int(os.environ["OMP_NUM_THREADS"]))
    q = configs.batched_layered()[-1]
    p = q.with_shots([q.shots[len(q.shots) // 2]])

    def timed(threads, budget):
        o = pyoracle.OracleRun(p, threads=threads)
        o.run(0, 2)  # first touch of every array
        t0 = time.perf_counter()
        steps = 2
        while time.perf_counter() - t0 < budget and steps < p.nt:
            n = min(25, p.nt - steps)
            o.run(steps, n)
            steps += n
        return steps - 2, time.perf_counter() - t0

    n, el = timed(cores, budget_s)
    n1, el1 = timed(1, min(4.0, budget_s / 4))
    cells = p.nx * p.nz
    return {"value": round(cells * n / el / 1e9, 4), "unit": "Gcell-updates/s", "cores": cores,
            "kind": "port", "cpu_model": cpu_model(),
            "sample": f"oracle (FP64 C/OpenMP restatement of seismic2, split C-PML) on the middle shot of the "
                      f"{p.nx}x{p.nz} padded model (2048x768 + C-PML 20), steps 2..{n + 1} of its "
                      f"{p.nt} in {el:.1f} s on {cores} threads; 1 thread: "
                      f"{cells * n1 / el1 / 1e9:.4f} Gcell-updates/s ({n1} steps in {el1:.1f} s)",
            "value_1core": round(cells * n1 / el1 / 1e9, 4)}


FWI_PARAMS = ("c11", "c13", "c33", "c55", "rho")


def fwi_gradient(nt: int = 6000, nshots: int = 16, every: int = 64, fd_cells: int = 0, tiles=None):
    """Time one full FWI gradient of the fwi_vti workload (FP32, FP64
    accumulation) against a plain FP32 forward of the same shots; optionally
    check it against central differences of the FP64 forward at fd_cells
    seeded cells per parameter (relative step 1e-4)."""
    import torch
    from seidart_amd.adjoint import Elastic2Engine, elastic2_traces
    from seidart_amd.configs import fwi_vti
    from seidart_amd.solver import Simulation

    true, start = fwi_vti(nt=nt, nshots=nshots)
    sim = Simulation(true, precision="fp32")
    sim.run(graph_block=every).check()
    obs = torch.as_tensor(np.stack(sim.traces()), device="cuda")
    sim.close()
    sim = Simulation(start, precision="fp32")
    if tiles:
        sim.set_tiles(*tiles)
    else:
        sim.autotune(steps=10)
    sim.capture_graph(every)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.run(graph_block=every)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    sim.check()
    sim.close()
    del sim
    eng = Elastic2Engine(start, "fp32", every=every)
    if tiles:
        eng.sim.set_tiles(*tiles)
    else:
        eng.sim.autotune(steps=10)
    prm = {k: torch.tensor(start.coef[k], device="cuda", requires_grad=k in FWI_PARAMS)
           for k in ("c11", "c13", "c15", "c33", "c35", "c55", "rho")}

    def gradient():
        for v in prm.values():
            v.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tr = elastic2_traces(eng, [prm])
        J = 0.5 * ((tr - obs) ** 2).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        J.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return float(J.detach()), t1 - t0, t2 - t1

    gradient()  # captures the segment graphs
    J, t_fwd, t_bwd = gradient()
    mem = sum(x.numel() * x.element_size() for x in eng.ckpt + eng.trail)
    out = {"workload": f"fwi_vti: 2-D VTI elastic {start.nx}x{start.nz} padded, {nshots} shots, "
                       f"{nt} steps, 256 receivers, L2 misfit, grad of C11 C13 C33 C55 rho, "
                       f"checkpoint every {every} steps, split C-PML",
           "s_per_shot": round((t_fwd + t_bwd) / nshots, 4), "gradient_s": round(t_fwd + t_bwd, 3),
           "forward_with_checkpoints_s": round(t_fwd, 3), "adjoint_s": round(t_bwd, 3),
           "plain_forward_s": round(t_plain, 3),
           "gradient_over_forward": round((t_fwd + t_bwd) / t_plain, 3),
           "misfit": J, "dtype": "f32 (gradient accumulated in f64)",
           "checkpoint_and_trail_GB": round(mem / 1e9, 2)}
    g32 = {k: prm[k].grad.double().cpu().numpy() for k in FWI_PARAMS}
    eng.close()
    del eng
    torch.cuda.empty_cache()
    if fd_cells > 0:
        # FD of the FP64 forward; gradients of the FP64 and FP32 adjoints
        eng = Elastic2Engine(start, "fp64", every=every)
        prm64 = {k: v.detach().double().clone().requires_grad_(k in FWI_PARAMS) for k, v in prm.items()}
        tr = elastic2_traces(eng, [prm64])
        (0.5 * ((tr - obs.double()) ** 2).sum()).backward()
        g64 = {k: prm64[k].grad.cpu().numpy() for k in FWI_PARAMS}
        base = {k: v.detach() for k, v in prm64.items()}

        def Jof(p):
            with torch.no_grad():
                return float(0.5 * ((elastic2_traces(eng, [p]) - obs.double()) ** 2).sum())

        rng = np.random.default_rng(2025)
        fd = {}
        for k in FWI_PARAMS:
            g = g64[k]
            cand = np.argwhere(np.abs(g) >= 0.05 * np.abs(g).max())
            cells = cand[rng.choice(len(cand), size=min(fd_cells, len(cand)), replace=False)]
            e64, e32 = [], []
            for j, i in cells:
                h = 1e-4 * abs(float(start.coef[k][j, i]))
                up = dict(base)
                dn = dict(base)
                up[k] = base[k].clone()
                dn[k] = base[k].clone()
                up[k][j, i] += h
                dn[k][j, i] -= h
                d = (Jof(up) - Jof(dn)) / (2 * h)
                e64.append(abs(g[j, i] - d) / abs(d))
                e32.append(abs(g32[k][j, i] - d) / abs(d))
            fd[k] = {"cells": len(cells), "max_rel_err_fp64_adjoint": float(f"{max(e64):.3e}"),
                     "max_rel_err_fp32_adjoint": float(f"{max(e32):.3e}")}
            print(f"fd check {k}: {fd[k]}", file=sys.stderr, flush=True)
        out["fd_check"] = {"step": "1e-4 relative, central, FP64 forward",
                           "cells": "seeded among those with |grad| >= 5% of the field's max",
                           "per_param": fd}
        eng.close()
    return out


def alg_bytes_per_step(problems, itemsize):
    """Compulsory HBM bytes of one time step of a plan over these problems,
    for any kind: per shot every field and every slab memory array read and
    written once; per model every coefficient array it passes to the device
    read once (2.5-D coefficients are y-invariant planes)."""
    from seidart_amd.problem import FIELDS, kernel_coefs
    from seidart_amd import _hip
    from seidart_amd.problem import KINDS
    tot = 0
    for p in problems:
        lx, ly, lz = p.slabs()
        kid = KINDS[p.kind] | (_hip.SDT_SPLIT_CPML if getattr(p, "cpml_split", False) else 0)
        npsi = _hip.psi_elems(kid, p.nx, p.ny, p.nz, lx, ly, lz)
        n = p.nx * p.ny * p.nz
        K = kernel_coefs(p)
        ncoef = len(K)
        if p.kind == "elastic2":
            ncoef -= (4 if all(not np.any(K[q]) for q in (2, 4, 5, 6)) else 0) + \
                (2 if all(np.all(K[q] == 1.0) for q in (10, 11)) else 0)
        tot += len(p.shots) * (2 * len(FIELDS[p.kind]) * n + 2 * npsi) + ncoef * p.nx * p.nz
    return tot * itemsize


def config_lines(survey_steps: int = 200, ref_steps: int = 0, only=None):
    """Throughput and FP32 accuracy of the other configurations at one GPU.
    The reference configurations run their full length (plain_psv 2000
    steps, gpr 4000, aniso25 3000) in FP32 and FP64, after the hipGraph of
    their block is captured; the two measured lines that are not reference
    configurations (gpr25, gpr_survey) run `survey_steps` steps.  Each line
    gives Gcell-updates/s, the algorithmic bytes per cell-update, the
    fraction of the HBM peak and the FP32 trace error against the FP64 run
    over the same steps.  For profiling runs, `ref_steps` > 0 shortens the
    reference configurations to that many steps and `only` (names) keeps
    only those lines."""
    import torch
    from seidart_amd import configs
    from seidart_amd.solver import Simulation
    out = {}
    cases = [("plain_psv", configs.plain_psv, None),
             ("gpr", configs.gpr, None),
             ("aniso25", configs.aniso25, None),
             ("gpr25", configs.gpr25, survey_steps),
             ("gpr_survey", configs.gpr_survey, survey_steps)]
    for name, make, K in cases:
        if only and name not in only:
            continue
        if K is None and ref_steps > 0:
            K = ref_steps
        p = make() if K is None else make(nt=K)
        K = p.nt
        cells = p.nx * p.ny * p.nz * len(p.shots)
        rec = {"grid": f"{p.nx}x{p.ny}x{p.nz}" if p.is3d else f"{p.nx}x{p.nz}", "kind": p.kind,
               "shots": len(p.shots), "steps": K}
        traces = {}
        for prec, size in (("fp32", 4), ("fp64", 8)):
            sim = Simulation(p, precision=prec)
            if p.kind in ("elastic2", "em2"):
                rec.setdefault("tiles", {})[prec] = list(min((r := sim.autotune(steps=10)), key=r.get))
            sim.capture_graph(50)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sim.run(K, graph_block=50)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            sim.check()
            traces[prec] = sim.traces()[0]
            sim.close()
            gcs = cells * K / el / 1e9
            bpc = alg_bytes_per_step([p], size) / cells
            rec[prec] = {"Gcell_updates_s": round(gcs, 3), "alg_B_per_cell_update": round(bpc, 2),
                         "alg_GBs": round(gcs * bpc, 1), "frac_hbm": round(gcs * bpc / HBM_PEAK_GBS, 4)}
        t32, t64 = traces["fp32"], traces["fp64"]
        rec["fp32_trace_rel_l2_vs_fp64"] = float(f"{np.linalg.norm(t32 - t64) / np.linalg.norm(t64):.3e}")
        if p.kind == "elastic25" and K > 1000:
            # the legacy 2.5-D elastic scheme grows without bound (energy x1e6
            # and more between steps 1000 and 3000, in FP64 as in FP32), so
            # FP32 and FP64 part ways late in the run; the bar holds early
            a, b = t32[:1000], t64[:1000]
            rec["fp32_trace_rel_l2_vs_fp64_first_1000_steps"] = float(
                f"{np.linalg.norm(a - b) / np.linalg.norm(b):.3e}")
        out[name] = rec
        print(f"bench: {name} done", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    return out


def measured_traffic(alg_bytes: int):
    """HBM bytes per launch of the headline kernel measured by rocprofv3 byte
    counters (profiles/el2_step_batched_layered_fp32_hbm_bytes.md, scaled by
    the copy calibration there), if that profile was taken on this workload
    (same algorithmic bytes per launch); else None."""
    path = os.path.join(ROOT, "profiles", "el2_step_batched_layered_fp32_hbm_bytes.md")
    try:
        kv = dict(line.split(":", 1) for line in open(path) if line.count(":") == 1
                  and line.split(":")[0] in ("traffic_bytes_per_launch", "alg_bytes_per_launch"))
        if int(kv["alg_bytes_per_launch"]) == int(alg_bytes):
            return int(kv["traffic_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        pass
    return None


def copy_rate(gib: float = 2.0, reps: int = 10) -> float:
    """GB/s of torch's device-to-device copy of `gib` GiB (read + write): what a
    plain streaming kernel reaches on this device, the practical ceiling a
    stencil's measured traffic is compared with."""
    import torch
    n = int(gib * 2**30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * n * 4 / (ms * 1e-3) / 1e9


def copy_calibration(gib: float) -> int:
    """Time torch's device-to-device copy of `gib` GiB (read gib + write gib):
    the achievable-bandwidth reference and the known-size kernel the HBM
    byte counters are scaled by."""
    import torch
    n = int(gib * 2**30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 10
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "ms": round(ms, 4),
                      "GBs": round(2 * n * 4 / (ms * 1e-3) / 1e9, 1)}), flush=True)
    return 0


def probe_calibration(gib: float) -> int:
    """Run the library's HBM probe (a copy of `gib` GiB with one 4-byte
    element per lane, the stencils' access width) 10 times: the known-traffic
    kernel the byte counters of the stencil kernels are calibrated against,
    beside torch's vectorised copy (--copy-gib)."""
    import ctypes as C
    import torch
    from seidart_amd import _hip
    n = int(gib * 2**30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for _ in range(3):
        _hip.check(_hip.lib().sdt_hbm_probe(C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), n, st))
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 10
    e0.record()
    for _ in range(reps):
        _hip.check(_hip.lib().sdt_hbm_probe(C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), n, st))
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    assert torch.equal(a, b)
    print(json.dumps({"probe_bytes_read": n * 4, "probe_bytes_written": n * 4, "ms": round(ms, 4),
                      "GBs": round(2 * n * 4 / (ms * 1e-3) / 1e9, 1)}), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--graph-block", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-legacy", action="store_true",
                    help="skip the legacy-C-PML leg (throughput and FP32 error of the workload "
                         "with the reference's shared C-PML memory arrays)")
    ap.add_argument("--no-fwi", action="store_true")
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--tiles", default="",
                    help="fixed tile shape WxR of the 2-D elastic step kernel (e.g. 4x16) instead "
                         "of the autotuner's choice (profiling runs)")
    ap.add_argument("--config-steps", type=int, default=200,
                    help="steps of the measured lines that are not reference configurations "
                         "(gpr25, gpr_survey); the reference configurations run their full length")
    ap.add_argument("--ref-config-steps", type=int, default=0,
                    help="profiling runs: shorten the reference configurations to this many steps "
                         "(0: their full length)")
    ap.add_argument("--only-configs", default="",
                    help="profiling runs: comma-separated names of the configuration lines to run")
    ap.add_argument("--fwi-steps", type=int, default=6000)
    ap.add_argument("--fwi-fd", type=int, default=0,
                    help="also check the FWI gradient against central differences at this "
                         "many seeded cells per parameter (slow)")
    ap.add_argument("--copy-gib", type=float, default=0.0,
                    help="only time a plain device copy of this many GiB (HBM calibration)")
    ap.add_argument("--probe-gib", type=float, default=0.0,
                    help="only run the library's 4-byte-per-lane HBM probe copy of this many GiB")
    a = ap.parse_args(argv)
    if a.copy_gib > 0:
        return copy_calibration(a.copy_gib)
    if a.probe_gib > 0:
        return probe_calibration(a.probe_gib)
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(the batch shards by shot across GPUs by design; this build measures one GPU)", file=sys.stderr)
        return 2

    import torch
    from seidart_amd import configs
    from seidart_amd.solver import Simulation

    K, W = a.steps, a.warmup
    FULL = 3000                       # the workload's own length
    nt = max(W + K + PROFILE_STEPS, FULL)
    # hipGraph block of the timed run: at most --graph-block and at most K
    # steps, even (two steps per launch); captured before the timed region
    gb = min(a.graph_block, K) // 2 * 2
    gb = gb if gb >= 2 else 0
    t0 = time.perf_counter()
    problems = configs.batched_layered(nt=nt)
    setup_s = time.perf_counter() - t0
    shots = [len(p.shots) for p in problems]
    cells_per_step = sum(p.nx * p.nz * s for p, s in zip(problems, shots))

    sim = Simulation(problems, precision="fp32")
    fixed = tuple(int(v) for v in a.tiles.split("x")) if a.tiles else None
    if fixed:
        sim.set_tiles(*fixed)
        tune = {(sim.steps_per_launch, *fixed, sim.tile_order): float("nan")}
    else:
        tune = sim.autotune(steps=10)   # tile shape of the step kernel (setup, untimed)
    sim_spl, sim_tiles, sim_order = sim.steps_per_launch, sim.tiles(), sim.tile_order
    captured = sim.capture_graph(gb) if gb else False
    sim.run(W, graph_block=gb if W >= gb else 0)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    sim.run(K, graph_block=gb)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    sim.check()
    value = cells_per_step * K / wall / 1e9
    # per-kernel durations (HIP events around each launch, same stream)
    k_ms = sim.profile(PROFILE_STEPS)
    sim.check()
    # one launch of the step kernel advances `spl` steps (two by default: the
    # state of the first stays in registers), so its algorithmic bytes are
    # spl times those of one step
    spl = sim.steps_per_launch
    dom_bytes = spl * algorithmic_bytes(problems, shots, 4)
    achieved = dom_bytes / (k_ms[0] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(dom_bytes),
                "kernel": "k_el2_step2<float>" if spl == 2 else "k_el2_step<float>",
                "kernel_ms": round(k_ms[0], 4), "steps_per_launch": spl,
                "alg_bytes_per_launch": int(dom_bytes),
                "alg_bytes_per_cell_update": round(dom_bytes / spl / cells_per_step, 2),
                "gather_ms": round(k_ms[2], 4)}
    if roofline["traffic"]:
        # measured bytes per launch over the launch time, against a plain copy
        cgbs = copy_rate()
        tgbs = roofline["traffic"] / (k_ms[0] * 1e-3) / 1e9
        roofline.update({"traffic_GBs": round(tgbs, 1), "copy_GBs": round(cgbs, 1),
                         "traffic_frac_of_copy": round(tgbs / cgbs, 4)})
    sim.close()
    del sim
    torch.cuda.empty_cache()
    print("bench: headline timed", file=sys.stderr, flush=True)

    plain = None
    if not a.no_plain:
        # Accuracy and the plain FP64 path over the workload's full length,
        # independent of --steps: one FP32 batched run of all 96 shots, then
        # every shot alone in its own FP64 plan (tile shape autotuned on the
        # first shot of each model), timed.
        full = configs.batched_layered(nt=FULL)
        sb = Simulation(full, precision="fp32")
        if fixed:
            sb.set_tiles(*fixed)
        else:
            sb.autotune(steps=10)
        sb.capture_graph(50)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        sb.run(graph_block=50)
        torch.cuda.synchronize()
        full_s = time.perf_counter() - t1
        sb.check()
        traces32 = sb.traces()
        sb.close()
        del sb
        torch.cuda.empty_cache()
        tot_cells, tot_s, errs = 0, 0.0, []
        k = 0
        for p in full:
            best = None
            for sh in p.shots:
                single = p.with_shots([sh])
                sp = Simulation(single, precision="fp64")
                if best is None:
                    r = sp.autotune(steps=12)   # tile shape and steps per launch (setup, untimed)
                    best = min(r, key=r.get) if r else None
                elif best:
                    sp.use(best)
                sp.capture_graph(50)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                sp.run(graph_block=50)
                torch.cuda.synchronize()
                tot_s += time.perf_counter() - t1
                sp.check()
                tot_cells += p.nx * p.nz * p.nt
                t64 = sp.traces()[0]
                errs.append(float(np.linalg.norm(traces32[k] - t64) / np.linalg.norm(t64)))
                sp.close()
                k += 1
        pv = tot_cells / tot_s / 1e9
        plain = {"value": round(pv, 3), "unit": "Gcell-updates/s", "dtype": "f64",
                 "sample": f"all {k} shots, each alone in its own FP64 plan (tiles autotuned per "
                           f"model), {FULL} steps each",
                 "speedup_fp32_batched": round(value / pv, 3),
                 "fp32_batched_full_run_Gcell_updates_s": round(cells_per_step * FULL / full_s / 1e9, 3),
                 "fp32_trace_rel_l2_vs_fp64": {
                     "shots": len(errs), "steps": FULL, "max": float(f"{max(errs):.3e}"),
                     "median": float(f"{np.median(errs):.3e}"),
                     "per_model_max": [float(f"{max(errs[o:o + n]):.3e}") for o, n in
                                       zip(np.cumsum([0] + shots[:-1]), shots)]}}
        print("bench: fp64 plain and accuracy done", file=sys.stderr, flush=True)

    legacy = None
    if not a.no_legacy:
        # The same workload with the legacy sharing of the C-PML memory
        # arrays (the reference's arithmetic): throughput over the same K
        # timed steps, and the FP32 error of the middle shot of each model
        # against its FP64 single-shot run over the full length, where the
        # legacy C-PML's growing mode amplifies rounding differences.
        lp = configs.batched_layered(nt=nt, cpml_split=False)
        sl = Simulation(lp, precision="fp32")
        if fixed:
            sl.set_tiles(*fixed)
        else:
            sl.autotune(steps=10)
        if gb:
            sl.capture_graph(gb)
        sl.run(W, graph_block=gb if W >= gb else 0)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        sl.run(K, graph_block=gb)
        torch.cuda.synchronize()
        lv = cells_per_step * K / (time.perf_counter() - t1) / 1e9
        sl.check()
        sl.close()
        del sl
        lfull = configs.batched_layered(nt=FULL, cpml_split=False)
        sb = Simulation(lfull, precision="fp32")
        sb.run(graph_block=50).check()
        l32 = sb.traces()
        sb.close()
        del sb
        lerr, off = [], 0
        for p in lfull:
            s = len(p.shots) // 2
            single = p.with_shots([p.shots[s]])
            sp = Simulation(single, precision="fp64")
            sp.run(graph_block=50).check()
            t64 = sp.traces()[0]
            lerr.append(float(f"{np.linalg.norm(l32[off + s] - t64) / np.linalg.norm(t64):.3e}"))
            sp.close()
            off += len(p.shots)
        legacy = {"value": round(lv, 3), "unit": "Gcell-updates/s", "dtype": "f32",
                  "fp32_trace_rel_l2_vs_fp64_middle_shots": lerr, "steps": FULL}
        torch.cuda.empty_cache()
        print("bench: legacy C-PML leg done", file=sys.stderr, flush=True)

    fwi = None if a.no_fwi else fwi_gradient(nt=a.fwi_steps, fd_cells=a.fwi_fd, tiles=fixed)
    others = None if a.no_configs else config_lines(
        a.config_steps, a.ref_config_steps, [n for n in a.only_configs.split(",") if n] or None)
    cpu = None if a.no_cpu else cpu_baseline(a.cpu_seconds)
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": round(wall / K * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched_layered: 2-D elastic, 96 shots (32 per model) over 3 "
                               "layered snow/firn/ice models 512x256, 1024x512, 2048x768 "
                               f"(+C-PML 20, split memory arrays), dx 0.5 m, 150 Hz, FP32, all shots in the same "
                               f"launches, {K} timed steps after {W} warmup steps" +
                               (f", hipGraph blocks of {gb} steps (captured before timing)"
                                if gb else ", direct launches"),
                   "cells_per_step": cells_per_step, "shots": sum(shots),
                   "graph_block": gb if captured else 0,
                   "tiles": {"chosen": {"steps_per_launch": sim_spl, "waves": sim_tiles[0],
                                        "rows": sim_tiles[1], "order": sim_order},
                             "ms_per_step": {} if fixed else
                             {f"{n}-step {w}x{r} order {o}": round(v, 4)
                              for (n, w, r, o), v in tune.items()}}},
        "gpu_ms_events": round(gpu_ms, 3),
        "roofline": roofline,
        "fp64_plain": plain,
        "legacy_cpml": legacy,
        "fwi_grad": fwi,
        "other_configs": others,
        "cpu_baseline": cpu,
        "host_setup_s": round(setup_s, 2),
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
