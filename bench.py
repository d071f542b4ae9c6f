#!/usr/bin/env python3
"""Throughput of the fused FDTD step on one MI355X.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 20] [--workload t512]

Workloads: t512 (default: 512^3 voxel mix, FP32 storage), c5 (640^3 voxel mix
with 8 Ez DFT planes x 64 frequencies, FP32; about 21 GB of device memory),
dispersive256 (FP32), dispersive256-fp64 and vacuum64.

A step is one full FDTD time step (magnetic then electric half step) over the
whole grid, with inputs resident in HBM.  The default workload is the 512^3
voxel mix: 512^3 Yee cells, 16-cell CPML, a planar Gaussian Ez current sheet
normal to x, 8^3-cell voxels of which 30% hold a three-pole Lorentz dielectric
and 10% a Drude metal, fields and polarizations stored in FP32 with FP64
arithmetic and coefficients.

Prints one JSON line.  `roofline` prices the dominant kernel's algorithmic bytes
(ffip_sim_traffic) against its HIP-event duration over the timed region;
`cpu_baseline` times the CPU oracle (the FP64 restatement of the reference's
hot loop) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import contextlib
import json
import os
import platform
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def workload(name):
    from ffip_amd import workloads as W
    if name == "dispersive256":
        return W.dispersive_spheres(n=256, precision="fp32"), None, (
            "dispersive scattering 256^3: 12-cell CPML, Ez Gaussian1 sheet, 32 seeded spheres "
            "(16 two-pole Lorentz, 16 Drude), 2 Ez DFT planes x 16 freqs")
    if name == "dispersive256-fp64":
        return W.dispersive_spheres(n=256, precision="fp64"), None, "dispersive scattering 256^3, FP64 storage"
    if name == "vacuum64":
        return W.vacuum_dipole(), None, "vacuum Ez dipole 64^3, 8-cell CPML, 1 Ez DFT plane x 3 freqs"
    if name == "t512":
        return W.voxel_mix(n=512, precision="fp32"), None, (
            "voxel mix 512^3 FP32: 16-cell CPML, Ez Gaussian1 sheet, 8^3-cell voxels, 30% 3-pole Lorentz "
            "(palette of 8), 10% Drude (palette of 4), no DFT")
    if name == "c5":
        return W.voxel_mix(n=640, precision="fp32", nfreq=64, nplanes=8), None, (
            "voxel mix 640^3 FP32: 16-cell CPML, Ez Gaussian1 sheet, 8^3-cell voxels, 30% 3-pole Lorentz "
            "(palette of 8), 10% Drude (palette of 4), 8 Ez DFT planes at half-integer z x 64 freqs")
    raise SystemExit("unknown workload %r" % name)


# workloads too large for the CPU oracle's set-up are sampled on a smaller
# instance of the same recipe (per-cell rate reported)
CPU_SAMPLE = {"t512": lambda: __import__("ffip_amd.workloads", fromlist=["voxel_mix"]).voxel_mix(n=128),
              "c5": lambda: __import__("ffip_amd.workloads", fromlist=["voxel_mix"]).voxel_mix(n=128, nfreq=64)}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_threads():
    """Host threads for the CPU baseline: the cores this process may run on."""
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except (AttributeError, OSError):
        return max(1, os.cpu_count() or 1)


# the reference's own 1-core rate on the 64^3 vacuum dipole (8-cell CPML, one
# Ez plane x 3 freqs), g++ -O3 -march=native on an Intel Xeon @ 2.7 GHz
REFERENCE_VACUUM64_1CORE = 12.08


def cpu_baseline(cfg, ds, steps, cells, threads):
    """Time the oracle's C hot loop (FP64, the -O3 timing build) for `steps`
    steps after its set-up, on the same configuration, first on one thread and
    then on `threads` (Mcell-updates/s for each)."""
    from oracle.oracle import OracleSim, set_threads, use_timing_build
    use_timing_build()
    cfg = dict(cfg)
    cfg["precision"] = "fp64"
    sim = OracleSim(cfg, ds, strict_pool=False)
    rates = []
    for n in (1, threads):
        set_threads(n)
        sim.run_steps(1)  # first touch of the thread team's pages
        t0 = time.perf_counter()
        sim.run_steps(steps)
        rates.append(cells * steps / (time.perf_counter() - t0) / 1e6)
    return rates


def cpu_calibration(steps=100):
    """1-thread rate of the timing build on the 64^3 vacuum dipole, the case the
    reference's own 1-core rate is quoted on."""
    from ffip_amd import workloads as W
    return cpu_baseline(W.vacuum_dipole(64), None, steps, 64 ** 3, 1)[0]


def design_gradient(steps):
    """Design gradients per second on the inverse-design workload: 200x200x100
    cells, 10-cell CPML, Ex sheet at z = -35, a 100x100x20 two-medium design box
    (eps 1..12, blurred seeded density), F = sum |DFT Ez|^2 over a 20x20 patch at
    z = +30, f = 0.1, `steps` forward + `steps` adjoint steps.  One gradient is
    the whole forward solve, the adjoint solve and the device gather, set-ups
    included; FP32 storage is compared with FP64 storage on the same density."""
    import numpy as np
    from ffip_amd import adjoint as A
    from ffip_amd.workloads import blurred_density
    out, grads = {}, {}
    for prec in ("fp64", "fp32"):
        p = A.c3_problem(steps=steps, precision=prec)
        rho = blurred_density(p.shape, seed=3, radius=3, sigma=1.5)
        p.value_and_grad(rho)  # warm-up: first launches, phasor tables
        t0 = time.perf_counter()
        F, g = p.value_and_grad(rho)
        out[prec] = time.perf_counter() - t0
        grads[prec] = (F, g)
        print("[bench] design gradient %s: %.3f s (F = %.6e)" % (prec, out[prec], F), file=sys.stderr, flush=True)
    g64, g32 = grads["fp64"][1], grads["fp32"][1]
    err = float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64))
    return {"value": round(1.0 / out["fp32"], 4), "unit": "design-gradients/s", "storage": "fp32",
            "fp64_value": round(1.0 / out["fp64"], 4), "speedup_vs_fp64": round(out["fp64"] / out["fp32"], 3),
            "rel_err_vs_fp64": err,
            "objective_rel_err_vs_fp64": abs(grads["fp32"][0] - grads["fp64"][0]) / abs(grads["fp64"][0]),
            "workload": "inverse design 200x200x100, 100x100x20 design box, eps 1..12, 20x20 Ez patch at f = 0.1, "
                        "%d + %d steps, set-ups included" % (steps, steps)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="t512",
                    choices=["dispersive256", "dispersive256-fp64", "vacuum64", "t512", "c5"])
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--gradient-steps", type=int, default=3000)
    ap.add_argument("--no-gradient", action="store_true")
    ap.add_argument("--no-stream", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine holds the whole grid on a single MI355X and has no "
              "multi-GPU path (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from ffip_amd._lib import LibSim

    cfg, ds, desc = workload(args.workload)
    cells = 1
    for s in cfg["size"]:
        cells *= int(round(s / (1 / cfg["resolution"])))
    t_setup = time.perf_counter()
    sim = LibSim(cfg, ds)
    t_setup = time.perf_counter() - t_setup
    print("[bench] set-up %.1f s, device memory %.2f GB" % (t_setup, sim.info()["device_bytes"] / 1e9),
          file=sys.stderr, flush=True)
    sim.step(args.warmup)
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass
    t0 = time.perf_counter()
    tm = sim.step_timed(args.steps, per_kernel=True)
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1000 / args.steps
    value = cells * args.steps / wall / 1e6
    print("[bench] %d steps: wall %.3f s, events %.3f ms total, E kernels %.3f ms, H kernels %.3f ms"
          % (args.steps, wall, tm["total_ms"], tm["e_ms"], tm["h_ms"]), file=sys.stderr, flush=True)

    te, th = sim.traffic(0), sim.traffic(1)
    e_avg = tm["e_ms"] / args.steps / 1000
    h_avg = tm["h_ms"] / args.steps / 1000
    dom_e = e_avg >= h_avg
    tr, avg = (te, e_avg) if dom_e else (th, h_avg)
    achieved = tr["total"] / avg / 1e9
    kname = ("k_half<%s,%s>" % ("float" if cfg.get("precision") == "fp32" else "double", "E" if dom_e else "H"))
    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on this
    # workload (calibrated against a plain copy), when a measurement is on file
    traffic = None
    try:
        with open(os.path.join(HERE, "profiles", "traffic.json")) as f:
            traffic = json.load(f).get(args.workload, {}).get(kname)
    except (OSError, ValueError):
        traffic = None
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": kname, "avg_launch_us": round(avg * 1e6, 2), "bytes_per_launch": tr["total"]}
    step_bytes = te["total"] + th["total"]
    if not args.no_stream:
        # measured ceiling: a streaming kernel that reads two arrays for every one it
        # writes (the half steps' mix) and a copy, 1 GiB per array
        from ffip_amd._lib import stream_bandwidth
        s21 = stream_bandwidth(2, 1, 1 << 30, 10)
        s11 = stream_bandwidth(1, 1, 1 << 30, 10)
        roof["measured_ceiling"] = {"read2_write1_gbs": round(s21[0], 1), "copy_gbs": round(s11[0], 1),
                                    "frac_of_read2_write1": round(achieved / s21[0], 4)}
    detail = {"electric": {"avg_us": round(e_avg * 1e6, 2), "traffic": te},
              "magnetic": {"avg_us": round(h_avg * 1e6, 2), "traffic": th},
              "step_hbm_frac": round(step_bytes / (ms_per_step / 1000) / 1e9 / HBM_PEAK_GBS, 4),
              "setup_s": round(t_setup, 2)}
    sim.close()

    # the front end announces each Simulation on stdout, as the reference does;
    # this program's stdout carries only the result line
    with contextlib.redirect_stdout(sys.stderr):
        grad = None if args.no_gradient else design_gradient(args.gradient_steps)

    cpu = None
    if not args.no_cpu and int(os.environ.get("RANK", "0")) == 0:
        nth = cpu_threads()
        if args.workload in CPU_SAMPLE:
            scfg = CPU_SAMPLE[args.workload]()
            scells = 1
            for s in scfg["size"]:
                scells *= int(round(s * scfg["resolution"]))
            v1, vn = cpu_baseline(scfg, None, args.cpu_steps, scells, nth)
            what = "the same recipe at %d^3 (the full grid exceeds the oracle's set-up memory)" % round(
                scells ** (1 / 3))
        else:
            v1, vn = cpu_baseline(cfg, ds, args.cpu_steps, cells, nth)
            what = "same workload and grid"
        c1 = cpu_calibration()
        cpu = {"value": round(vn, 3), "unit": "Mcell-updates/s", "cores": nth, "kind": "port",
               "value_1_thread": round(v1, 3),
               "calibration": {"vacuum64_1_thread": round(c1, 3),
                               "reference_vacuum64_1_core": REFERENCE_VACUUM64_1CORE,
                               "ratio": round(c1 / REFERENCE_VACUUM64_1CORE, 3),
                               "note": "the reference's rate was measured on another host (Xeon @ 2.7 GHz)"},
               "sample": "%s, %d steps after set-up on 1 thread and then %d on %d threads (all cores this "
                         "process may use), FP64 oracle (C restatement of the reference hot loop, gcc -O3 "
                         "-march=x86-64-v3, OpenMP over each phase's points) on %s"
                         % (what, args.cpu_steps, args.cpu_steps, nth, cpu_model())}

    line = {"metric": "Mcell-updates/s", "value": round(value, 1), "unit": "Mcell-updates/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": desc, "grid": [int(s) for s in cfg["size"]], "storage": cfg.get("precision"),
                       "cells": cells},
            "roofline": roof, "cpu_baseline": cpu, "design_gradient": grad, "detail": detail}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
