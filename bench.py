#!/usr/bin/env python3
"""Benchmark of the solver step on one MI355X.

Workload: 2D resistive MHD, HLLD + constrained transport + SOR self-gravity (two 600-sweep Poisson
solves per Heun step), Gaussian collapse on a 2048 x 2048 grid over [-0.5,0.5]^2 with clamped
boundaries, gamma 5/3, G 1, eta 1e-4, CFL 0.4 (the 2D MHD program's step with the collapse IC).
The same K steps run once in plain FP64 and once in the mixed form (FP32 reconstruction, Riemann
solver and SOR sweeps; FP64 update, B/CT and Poisson residual), from the same seeded IC.

A zone-update is one cell advanced one full Heun step including its two Poisson solves.
`value` is the mixed-precision zone-updates/s with all state resident in HBM; the FP64 figure,
the ratio, the mixed-vs-FP64 relative L1 density error and max|div B| ride along.

roofline: dominant kernel = the tiled SOR kernel (k_sor_sflow_* / k_sor_flow_*: one dataflow launch
per 600-sweep solve on sheared or rectangular tiles; k_sor_(s)tile_*, one launch per tile level, when
CC_SOR_LEVELS is set), timed with HIP events on the handle's stream around every launch inside the
timed region.  Algorithmic bytes per launch = (cell-sweeps in the launch) x 24 B (FP64: read phi,
read rhs, write phi) or x 12 B (mixed: FP32 phi and rhs), the compulsory traffic of one sweep.  The kernel keeps its tile in LDS for
16 sweeps, so its algorithmic rate can exceed the HBM peak; `traffic` (measured HBM bytes per
launch, from rocprofv3 FETCH_SIZE/WRITE_SIZE in profiles/hbm_traffic.json) shows the real HBM load.

cpu_baseline: the reference program's own step functions (compiled from its sources in
oracle/_ref, single thread) on a bounded sample of the same problem (see "sample").

The line also carries "turb3d": the 3D resistive MHD step (HLLD + CT, no gravity) on the 256^3
periodic turbulence problem (ideal gas gamma 5/3, beta 1 field along z, k^-2 solenoidal velocity at
rms Mach 1, eta 5e-5, CFL 0.4), FP64 and mixed, with the roofline of its dominant kernel (the face
kernel, one launch per axis: per face it reads one cell's 8 primitives and the face B and writes 5
fluxes + 2 EMFs, 128 B in FP64 and 68 B with FP32 primitives and fluxes)
and the CPU restatement's rate on a 64^3 sample (the reference has no ideal-gas EOS, so the CPU
leg is the oracle's port of its step with that EOS).  --workload selects one of the two problems.
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "zone-updates/s mixed vs FP64 on 1 MI355X at ≤1e-4 L1 err; % HBM roofline"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--N", type=int, default=2048)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="bound on the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("all", "collapse2d", "turb3d", "batch2d", "sphere3d"), default="all")
    ap.add_argument("--N5", type=int, default=512, help="grid of the 3D Euler sphere problem")
    ap.add_argument("--steps5", type=int, default=10)
    ap.add_argument("--batch", type=int, default=64, help="instances of the batched collapse")
    ap.add_argument("--steps-batch", type=int, default=200)
    ap.add_argument("--sor-autotune", type=int, default=1,
                    help="1: time the compiled SOR tile shapes once per 2D grid size and use the fastest")
    ap.add_argument("--seq-steps", type=int, default=10,
                    help="steps per instance of the one-after-another comparison (0: skip)")
    ap.add_argument("--N3", type=int, default=256, help="grid of the 3D turbulence problem")
    ap.add_argument("--steps3", type=int, default=10)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (for counter calibration) and exit")
    return ap.parse_args()


def run_gpu(N, precision, warmup, steps):
    from core_collapse_amd import ic
    from core_collapse_amd import mhd2d as M
    p = ic.collapse2d_params(N, precision=precision)
    st = ic.collapse2d_ic(p)
    h = M.Mhd2D(p)
    h.upload(st)
    t = h.run(warmup) if warmup > 0 else 0.0
    h.set_profiling(True)
    h.reset_profiling()
    h.sync()
    t0 = time.perf_counter()
    t = h.run(steps, t)          # synchronises at the end
    elapsed = time.perf_counter() - t0
    prof = {k: h.kernel_time(k) for k in ("sor", "flux", "corner", "ct", "grad", "mean", "dt")}
    h.set_profiling(False)
    out = h.download()
    divb = h.divB_max()
    h.close()
    tile = M.sor_tile_choice(N, N, p.sor_iters, precision)
    return {"elapsed": elapsed, "t": t, "rho": out["rho"], "divb": divb, "prof": prof, "params": p, "tile": tile}


def sor_cell_sweeps_per_solve(N, iters=600):
    return float(N) * N * iters


def cpu_baseline(N, budget_s):
    """Time the reference's own compiled step (driver loop M2:1111-1117) on a sample grid."""
    from core_collapse_amd import ic
    from oracle import bind as OB
    kind = "reference"
    try:
        lib = OB.Lib("ref_m2")
    except FileNotFoundError:
        lib = OB.Lib("oracle")
        kind = "port"
    n = 512 if N >= 512 else N
    p = ic.collapse2d_params(n)
    st = ic.collapse2d_ic(p)
    op = OB.m2_params(gamma=p.gamma, cfl=p.cfl, x0=p.x0, x1=p.x1, y0=p.y0, y1=p.y1, Nx=n, Ny=n,
                      floor_rho=p.floor_rho, floor_P=p.floor_P, periodic_x=0, periodic_y=0, eta=p.eta,
                      gravity_on=1, G=p.G, sor_iters=p.sor_iters, sor_omega=p.sor_omega)
    S = OB.M2Arrays(n, n, {k: v.reshape(-1) for k, v in st.items()})
    steps, t, tsum = 0, 0.0, 0.0
    while steps < 20 and (steps == 0 or tsum < budget_s):
        t0 = time.perf_counter()
        t = lib.or_m2_run(C.byref(op), C.byref(S.cstate()), C.c_double(p.dx), C.c_double(p.dy), 1, C.c_double(t),
                          C.c_double(1e300), None)
        tsum += time.perf_counter() - t0
        steps += 1
    zups = n * n * steps / tsum
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": zups, "unit": "zone-updates/s", "cores": 1, "kind": kind,
            "sample": "%d Heun step(s) of the same collapse problem at %dx%d (FP64, 600-sweep SOR x2 per step, "
                      "%.1f s), single thread, %s, nproc=%d" % (steps, n, n, tsum, cpu, os.cpu_count() or 0)}


def run_gpu3d(N, precision, warmup, steps):
    from core_collapse_amd import ic
    from core_collapse_amd import mhd3d as M3
    p = ic.turbulence3d_params(N, precision=precision)
    h = M3.Mhd3D(p)
    h.upload(ic.turbulence3d_ic(p))
    t = h.run(warmup) if warmup > 0 else 0.0
    h.set_profiling(True)
    h.reset_profiling()
    h.sync()
    t0 = time.perf_counter()
    t = h.run(steps, t)
    elapsed = time.perf_counter() - t0
    prof = {k: h.kernel_time(k) for k in ("prim", "face", "update", "edge", "ct", "joule", "dt")}
    h.set_profiling(False)
    out = h.download()
    divb = h.divB_max()
    h.close()
    return {"elapsed": elapsed, "rho": out["rho"], "divb": divb, "prof": prof}


def cpu_baseline3d(n=128, budget_s=12.0):
    """The oracle's restatement of the 3D step (ideal-gas EOS mode) on a 128^3 sample."""
    from core_collapse_amd import ic
    from oracle import bind as OB
    lib = OB.Lib("oracle")
    p = ic.turbulence3d_params(n)
    st = ic.turbulence3d_ic(p)
    op = OB.m3_params(Nx=n, Ny=n, Nz=n, cfl=p.cfl, eta=p.eta, ideal_eos=1, gamma_ideal=p.gamma)
    S = OB.M3Arrays(n, n, n, {k: v.reshape(-1) for k, v in st.items()})
    steps, tsum = 0, 0.0
    while steps < 20 and (steps == 0 or tsum < budget_s):
        t0 = time.perf_counter()
        lib.or_m3_run(C.byref(op), C.byref(S.cstate()), C.c_double(p.dx), C.c_double(p.dy), C.c_double(p.dz), 1,
                      C.c_double(0.0), C.c_double(1e300), None)
        tsum += time.perf_counter() - t0
        steps += 1
    return {"value": n ** 3 * steps / tsum, "unit": "zone-updates/s", "cores": 1, "kind": "port",
            "sample": "%d Heun step(s) of the same turbulence problem at %d^3 (FP64, %.1f s), single thread, "
                      "oracle restatement (the reference program has no ideal-gas EOS)" % (steps, n, tsum)}


def bench3d(N, warmup, steps, with_cpu):
    from core_collapse_amd import CC_FP64, CC_MIXED
    res = {prec: run_gpu3d(N, prec, warmup, steps) for prec in (CC_FP64, CC_MIXED)}
    zu = float(N) ** 3 * steps

    def roof(r, word):
        n, ms = r["prof"]["face"]
        if n == 0 or ms <= 0:
            return None
        # per face: read the 8 primitives of one cell and the face B (FP64), write 5 fluxes + 2 EMFs
        bpl = float(N) ** 3 * (8 * word + 8 + 7 * word)
        avg = ms / n * 1e-3
        return {"bound": "hbm", "achieved": round(bpl / avg / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(bpl / avg / 1e9 / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_face3d",
                "launches": n, "avg_launch_us": round(avg * 1e6, 2), "alg_bytes_per_launch": bpl}

    a, b = res[CC_MIXED]["rho"], res[CC_FP64]["rho"]
    r64, rmx = roof(res[CC_FP64], 8), roof(res[CC_MIXED], 4)
    if r64:
        r64["traffic"] = load_traffic("cc::k_face3d<0, false")
    if rmx:
        rmx["traffic"] = load_traffic("cc::k_face3d<0, true")
    out = {
        "workload": "3D resistive MHD HLLD+CT, periodic turbulence %d^3, ideal gas gamma 5/3, beta 1 along z, "
                    "rms Mach 1, eta 5e-5, CFL 0.4" % N,
        "steps": steps, "warmup": warmup,
        "value": zu / res[CC_MIXED]["elapsed"], "unit": "zone-updates/s",
        "ms_per_step": 1e3 * res[CC_MIXED]["elapsed"] / steps,
        "fp64": {"value": zu / res[CC_FP64]["elapsed"], "ms_per_step": 1e3 * res[CC_FP64]["elapsed"] / steps,
                 "roofline": r64, "max_divB": res[CC_FP64]["divb"]},
        "mixed_over_fp64": res[CC_FP64]["elapsed"] / res[CC_MIXED]["elapsed"],
        "l1_rho_rel_err": float(np.abs(a - b).sum() / np.abs(b).sum()),
        "linf_rho_rel_err": float(np.abs(a - b).max() / np.abs(b).max()),
        "max_divB": res[CC_MIXED]["divb"],
        "roofline": rmx,
        "zone_update_alg_bytes": 320,
        "kernel_times": {tag: {k: {"launches": v[0], "total_ms": round(v[1], 3)} for k, v in res[prec]["prof"].items()
                               if v[0] > 0} for prec, tag in ((CC_FP64, "fp64"), (CC_MIXED, "mixed"))},
    }
    out["hbm_frac_alg"] = round(out["value"] * 320 / 1e9 / HBM_PEAK_GBS, 4)
    for prec, tag, R, mx in ((CC_FP64, "fp64", "double", "false"), (CC_MIXED, "mixed", "float", "true")):
        groups = {"prim": ["cc::k_prim3d<%s>" % R],
                  "face": ["cc::k_face3d<%d, %s," % (a, mx) for a in range(3)],
                  "update": ["cc::k_update3d<%s>" % R], "edge": ["cc::k_edge3d<%s>" % R], "ct": ["cc::k_ct3d"],
                  "joule": ["cc::k_joule3d<true>", "cc::k_joule3d<false>"]}
        out.setdefault("hbm_measured", {})[tag] = measured_hbm(res[prec]["prof"], groups)
        out.setdefault("fusion_factor", {})[tag] = fusion_factor(out["hbm_measured"][tag], zu, 320)
    if with_cpu:
        out["cpu_baseline"] = cpu_baseline3d()
    return out


def bench_batch2d(n_inst, steps, warmup, nmin, nmax, seq_steps):
    """The batched collapse: n_inst independent instances from the batch sampler (grids log-uniform in
    [nmin, nmax]^2, contrast, eta, B0 and IC seed per instance), mixed precision, all advanced together
    by one graph launch per step (per-instance dt; the SOR tile levels of all instances merged into one
    launch per level, the rest of each instance's step a parallel branch).  Also times each instance run
    alone for seq_steps steps, to set the batch against running the instances one after another."""
    from core_collapse_amd import CC_MIXED
    from core_collapse_amd import ic
    from core_collapse_amd import mhd2d as M
    sample = ic.collapse2d_batch_sample(n_inst, 42, nmin, nmax)
    t_setup = time.perf_counter()
    sims = []
    for e in sample:
        p = ic.collapse2d_params(e["N"], eta=e["eta"], precision=CC_MIXED)
        h = M.Mhd2D(p)
        h.upload(ic.collapse2d_ic(p, **e["ic"]))
        sims.append(h)
    cells = float(sum(e["N"] ** 2 for e in sample))
    t_setup = time.perf_counter() - t_setup
    tiles = {str(e["N"]): M.sor_tile_choice(e["N"], e["N"], 600, CC_MIXED) for e in sample}
    b = M.Mhd2DBatch(sims)
    binfo = b.info()
    t = b.run(warmup) if warmup > 0 else None
    t0 = time.perf_counter()
    t = b.run(steps, t)
    elapsed = time.perf_counter() - t0
    b.close()
    seq = 0.0
    if seq_steps > 0:
        for h, ti in zip(sims, t):
            h.run(1, ti)
            h.sync()
            t1 = time.perf_counter()
            h.run(seq_steps, ti)
            seq += (time.perf_counter() - t1) * steps / seq_steps
    divb = max(h.divB_max() for h in sims)
    for h in sims:
        h.close()
    zu = cells * steps
    return {
        "workload": "batch of %d 2D collapses (HLLD+CT+SOR gravity, clamped boundaries), grids log-uniform in "
                    "[%d, %d]^2, contrast U[5,50], eta U[1e-5,1e-3], B0 U[0.01,0.1] along x, IC seed 42+i, "
                    "mixed precision, per-instance dt" % (n_inst, nmin, nmax),
        "instances": n_inst, "cells": cells, "grids": [e["N"] for e in sample],
        "steps": steps, "warmup": warmup,
        "value": zu / elapsed, "unit": "zone-updates/s", "ms_per_step": 1e3 * elapsed / steps,
        "sequential_estimate": {"value": zu / seq if seq > 0 else None, "steps_timed": seq_steps,
                                "ms_per_step": 1e3 * seq / steps if seq > 0 else None},
        "batch_over_sequential": seq / elapsed if seq > 0 else None,
        "sor_alg_GBps": cells * steps * 2 * 600 * 12 / elapsed / 1e9,
        "setup_s": round(t_setup, 2), "sor_merged": binfo["merged"], "sor_merged_tile": binfo["tile"],
        "sor_tiles_alone": tiles,
        "max_divB": divb, "t_min": float(np.min(t)), "t_max": float(np.max(t)),
    }


def run_gpu_sphere(N, precision, warmup, steps):
    from core_collapse_amd import euler3d as E3
    from core_collapse_amd import ic
    p = ic.sphere3d_params(N, precision=precision)
    h = E3.Euler3D(p)
    h.upload(ic.sphere3d_ic(p))
    t = h.run(warmup) if warmup > 0 else 0.0
    h.set_profiling(True)
    h.reset_profiling()
    h.sync()
    t0 = time.perf_counter()
    t = h.run(steps, t)
    elapsed = time.perf_counter() - t0
    prof = {k: h.kernel_time(k) for k in ("prim", "face", "update", "mean", "fft_x", "fft_y", "fft_z", "check",
                                          "grad", "dt")}
    h.set_profiling(False)
    stats = h.poisson_stats()
    out = h.download()
    tot = h.totals()
    h.close()
    return {"elapsed": elapsed, "rho": out["rho"], "prof": prof, "poisson": stats, "mass": tot["mass"]}


def cpu_baseline_sphere(n=128, budget_s=12.0):
    """The oracle's restatement of the 3D Euler + spectral gravity step on a 128^3 sample."""
    from core_collapse_amd import ic
    from oracle import bind as OB
    lib = OB.Lib("oracle")
    p = ic.sphere3d_params(n)
    st = ic.sphere3d_ic(p)
    op = OB.e3_params(Nx=n, Ny=n, Nz=n, cfl=p.cfl, gamma=p.gamma, G=p.G, gravity_on=1)
    S = OB.E3Arrays(n, n, n, st)
    steps, t, tsum = 0, 0.0, 0.0
    while steps < 20 and (steps == 0 or tsum < budget_s):
        t0 = time.perf_counter()
        t = lib.or_e3_run(C.byref(op), C.byref(S.cstate()), 1, C.c_double(t), C.c_double(1e300), None)
        tsum += time.perf_counter() - t0
        steps += 1
    return {"value": n ** 3 * steps / tsum, "unit": "zone-updates/s", "cores": 1, "kind": "port",
            "sample": "%d Heun step(s) of the same sphere problem at %d^3 (FP64, exact spectral gravity, %.1f s), "
                      "single thread, oracle restatement (no reference program has this step)" % (steps, n, tsum)}


def bench_sphere3d(N, warmup, steps, with_cpu):
    """3D Euler HLLC + FFT self-gravity, uniform sphere, FP64 and mixed (FP32 primitives, fluxes and FFT;
    FP64 state, update and Poisson residual)."""
    from core_collapse_amd import CC_FP64, CC_MIXED
    res = {prec: run_gpu_sphere(N, prec, warmup, steps) for prec in (CC_FP64, CC_MIXED)}
    zu = float(N) ** 3 * steps
    cells = float(N) ** 3

    def roof(r, word, tname):
        n, ms = r["prof"]["update"]
        if n == 0 or ms <= 0:
            return None
        # the stage update, one launch per stage (half of them the final stage): per cell it reads 15 face
        # fluxes (3 axes x 5, in the flux precision), the state (5 FP64), the potential (1 FP64, differenced
        # for g), on the final stage also U^n (5 FP64), and writes the state: 15 w + 88 B or 15 w + 128 B,
        # 15 w + 108 B on average
        bpl = cells * (15 * word + 108)
        avg = ms / n * 1e-3
        t1, t2 = load_traffic("cc::k_e3_update<1, %s>" % tname), load_traffic("cc::k_e3_update<2, %s>" % tname)
        return {"bound": "hbm", "achieved": round(bpl / avg / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(bpl / avg / 1e9 / HBM_PEAK_GBS, 4),
                "traffic": 0.5 * (t1 + t2) if t1 and t2 else None, "kernel": "k_e3_update",
                "launches": n, "avg_launch_us": round(avg * 1e6, 2), "alg_bytes_per_launch": bpl}

    def balg(r):
        k = max(r["poisson"]["iters_last"], 1)
        return 392 + 192 * k

    a, b = res[CC_MIXED]["rho"], res[CC_FP64]["rho"]
    out = {
        "workload": "3D Euler HLLC + FFT Poisson self-gravity, periodic %d^3, gamma 5/3, G 1, uniform sphere rho 10 "
                    "(r < 0.2) in 0.1, 1%% seeded noise, p 0.01, at rest; FFT refined in FP64 to 1e-10" % N,
        "steps": steps, "warmup": warmup,
        "value": zu / res[CC_MIXED]["elapsed"], "unit": "zone-updates/s",
        "ms_per_step": 1e3 * res[CC_MIXED]["elapsed"] / steps,
        "fp64": {"value": zu / res[CC_FP64]["elapsed"], "ms_per_step": 1e3 * res[CC_FP64]["elapsed"] / steps,
                 "roofline": roof(res[CC_FP64], 8, "double"), "poisson": res[CC_FP64]["poisson"]},
        "mixed_over_fp64": res[CC_FP64]["elapsed"] / res[CC_MIXED]["elapsed"],
        "l1_rho_rel_err": float(np.abs(a - b).sum() / np.abs(b).sum()),
        "linf_rho_rel_err": float(np.abs(a - b).max() / np.abs(b).max()),
        "roofline": roof(res[CC_MIXED], 4, "float"),
        "poisson": res[CC_MIXED]["poisson"],
        "zone_update_alg_bytes": balg(res[CC_MIXED]),
        "kernel_times": {tag: {k: {"launches": v[0], "total_ms": round(v[1], 3)} for k, v in res[prec]["prof"].items()
                               if v[0] > 0} for prec, tag in ((CC_FP64, "fp64"), (CC_MIXED, "mixed"))},
    }
    out["hbm_frac_alg"] = round(out["value"] * balg(res[CC_MIXED]) / 1e9 / HBM_PEAK_GBS, 4)
    for prec, tag, R in ((CC_FP64, "fp64", "double"), (CC_MIXED, "mixed", "float")):
        groups = {"prim": ["cc::k_e3_prim<%s>" % R], "face": tuple("cc::k_e3_face<%s, %d>" % (R, a) for a in range(3)),
                  "update": ["cc::k_e3_update<1, %s>" % R, "cc::k_e3_update<2, %s>" % R], "grad": ["cc::k_p3_grad"]}
        out.setdefault("hbm_measured", {})[tag] = measured_hbm(res[prec]["prof"], groups)
        # hydro kernels only: B_alg's hydro share is 5 state passes of 40 B plus g, 224 B per zone-update
        out.setdefault("fusion_factor_hydro", {})[tag] = fusion_factor(out["hbm_measured"][tag], zu, 224)
    if with_cpu:
        out["cpu_baseline"] = cpu_baseline_sphere()
    return out


def measured_hbm(prof, groups):
    """Per profiler scope: rocprof-measured HBM bytes per launch (profiles/hbm_traffic.json, averaged over
    the kernels the scope launches) divided by the scope's average launch time from the HIP events of
    this run (a tuple of kernels: one scope that launches each of them once); plus the same ratio over all
    listed scopes together (bytes and time summed)."""
    out, tb, tt = {}, 0.0, 0.0
    for scope, names in groups.items():
        n, ms = prof.get(scope, (0, 0.0))
        together = isinstance(names, tuple)  # a tuple: one timed scope launches all of them
        tr = [load_traffic(k) for k in names]
        if n == 0 or ms <= 0 or any(t is None for t in tr):
            continue
        b = sum(tr) if together else sum(tr) / len(tr)
        avg = ms / n * 1e-3
        out[scope] = {"bytes_per_launch": b, "avg_launch_us": round(avg * 1e6, 2),
                      "GBps": round(b / avg / 1e9, 1), "frac": round(b / avg / 1e9 / HBM_PEAK_GBS, 4)}
        tb += b * n
        tt += ms * 1e-3
    if tt > 0:
        out["all"] = {"GBps": round(tb / tt / 1e9, 1), "frac": round(tb / tt / 1e9 / HBM_PEAK_GBS, 4),
                      "bytes_total": tb}
    return out


def fusion_factor(hbm, zone_updates, balg):
    """Measured HBM bytes per zone-update of the listed kernels over the algorithmic bytes B_alg."""
    a = hbm.get("all")
    return round(a["bytes_total"] / zone_updates / balg, 3) if a and zone_updates > 0 else None


def load_traffic(prefix):
    """Measured HBM bytes per launch of the first kernel whose name starts with `prefix`, from the
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary in profiles/hbm_traffic.json (None if absent)."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    for name, b in k.items():
        if name.startswith(prefix):
            return b
    return None


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: runs on one GPU only (the solver step is built and measured on one GPU; sharding "
              "across GPUs is not built); got --gpus %d" % a.gpus, file=sys.stderr)
        sys.exit(2)
    from core_collapse_amd import CC_FP64, CC_MIXED
    if a.copy_calibration:
        from core_collapse_amd import _lib
        ms = C.c_double()
        nbytes = a.copy_calibration << 20
        _lib.check(_lib.lib().cc_copy_calibration(nbytes, 5, C.byref(ms)))
        print(json.dumps({"copy_bytes": nbytes, "read_bytes": nbytes, "write_bytes": nbytes, "avg_ms": ms.value,
                          "GBps": 2 * nbytes / ms.value / 1e6}))
        return
    N, K, W = a.N, a.steps, a.warmup
    if a.sor_autotune:
        from core_collapse_amd import mhd2d as M2
        M2.sor_autotune(True)
    t3d = None
    if a.workload in ("all", "turb3d"):
        print("bench: %d^3 turbulence, warmup %d, timed %d steps" % (a.N3, 2, a.steps3), file=sys.stderr, flush=True)
        t3d = bench3d(a.N3, 2, a.steps3, not a.no_cpu)
        print("bench: turb3d mixed %.3f ms/step, fp64 %.3f ms/step" % (t3d["ms_per_step"], t3d["fp64"]["ms_per_step"]),
              file=sys.stderr, flush=True)
        if a.workload == "turb3d":
            print(json.dumps({"turb3d": t3d}))
            return
    sp = None
    if a.workload in ("all", "sphere3d"):
        print("bench: %d^3 sphere, warmup %d, timed %d steps" % (a.N5, 2, a.steps5), file=sys.stderr, flush=True)
        sp = bench_sphere3d(a.N5, 2, a.steps5, not a.no_cpu)
        print("bench: sphere3d mixed %.3f ms/step, fp64 %.3f ms/step" % (sp["ms_per_step"], sp["fp64"]["ms_per_step"]),
              file=sys.stderr, flush=True)
        if a.workload == "sphere3d":
            print(json.dumps({"sphere3d": sp}))
            return
    bt = None
    if a.workload in ("all", "batch2d"):
        print("bench: batch of %d collapses, timed %d steps" % (a.batch, a.steps_batch), file=sys.stderr, flush=True)
        bt = bench_batch2d(a.batch, a.steps_batch, 3, 128, 1024, a.seq_steps)
        print("bench: batch2d %.3f ms/step (sequential %s ms/step)" % (bt["ms_per_step"],
              bt["sequential_estimate"]["ms_per_step"]), file=sys.stderr, flush=True)
        if a.workload == "batch2d":
            print(json.dumps({"batch2d": bt}))
            return
    print("bench: %dx%d collapse, warmup %d, timed %d steps" % (N, N, W, K), file=sys.stderr, flush=True)
    res = {}
    for prec in (CC_FP64, CC_MIXED):
        res[prec] = run_gpu(N, prec, W, K)
        r = res[prec]
        print("bench: %s %.3f ms/step" % ("fp64" if prec == CC_FP64 else "mixed", 1e3 * r["elapsed"] / K),
              file=sys.stderr, flush=True)
    zone_updates = float(N) * N * K
    v64 = zone_updates / res[CC_FP64]["elapsed"]
    vmx = zone_updates / res[CC_MIXED]["elapsed"]
    rho64, rhomx = res[CC_FP64]["rho"], res[CC_MIXED]["rho"]
    l1 = float(np.abs(rhomx - rho64).sum() / np.abs(rho64).sum())
    linf = float(np.abs(rhomx - rho64).max() / np.abs(rho64).max())

    def roof(r, bytes_per_sweep):
        n, ms = r["prof"]["sor"]
        if n == 0 or ms <= 0:
            return None
        sweeps = sor_cell_sweeps_per_solve(N) * 2 * K  # two solves per step
        bytes_per_launch = sweeps * bytes_per_sweep / n
        avg_s = ms / n * 1e-3
        ach = bytes_per_launch / avg_s / 1e9
        return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": None, "launches": n, "avg_launch_us": round(avg_s * 1e6, 2),
                "alg_bytes_per_launch": bytes_per_launch}

    def sor_kernel(tile):
        S, WJ, sheared = tile
        return ("k_sor_s" if sheared else "k_sor_") + ("tile" if os.environ.get("CC_SOR_LEVELS") else "flow")

    roof_mx = roof(res[CC_MIXED], 12.0)
    roof_64 = roof(res[CC_FP64], 24.0)
    if roof_mx:
        tl = res[CC_MIXED]["tile"]
        roof_mx["kernel"] = sor_kernel(tl)
        roof_mx["traffic"] = load_traffic("cc::%s_mixed<%d, %d>" % (sor_kernel(tl), tl[0], tl[1]))
    if roof_64:
        tl = res[CC_FP64]["tile"]
        roof_64["kernel"] = sor_kernel(tl)
        roof_64["traffic"] = load_traffic("cc::%s_fp64<%d, %d%s" % (sor_kernel(tl), tl[0], tl[1], ">" if tl[2] else ","))
    kernels = {}
    for prec, tag in ((CC_FP64, "fp64"), (CC_MIXED, "mixed")):
        kernels[tag] = {k: {"launches": v[0], "total_ms": round(v[1], 3)} for k, v in res[prec]["prof"].items()
                        if v[0] > 0}
    cpu = None if a.no_cpu else cpu_baseline(N, a.cpu_seconds)
    line = {
        "metric": METRIC,
        "value": vmx,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * res[CC_MIXED]["elapsed"] / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fluxes/SOR sweeps + f64 update/CT/residual",
        "data": "synthetic",
        "config": {"workload": "2D resistive MHD HLLD+CT+SOR self-gravity, Gaussian collapse %dx%d, clamped "
                               "boundaries, 600 SOR sweeps x 2 solves per Heun step" % (N, N),
                   "grid": [N, N], "gamma": 5.0 / 3.0, "G": 1.0, "eta": 1e-4, "cfl": 0.4, "sor_iters": 600,
                   "sor_omega": 1.8, "seed": 42},
        "fp64": {"value": v64, "ms_per_step": 1e3 * res[CC_FP64]["elapsed"] / K, "roofline": roof_64,
                 "max_divB": res[CC_FP64]["divb"]},
        "mixed_over_fp64": vmx / v64,
        "l1_rho_rel_err": l1,
        "linf_rho_rel_err": linf,
        "max_divB": res[CC_MIXED]["divb"],
        "roofline": roof_mx,
        "cpu_baseline": cpu,
        "sor_tile": {"fp64": res[CC_FP64]["tile"], "mixed": res[CC_MIXED]["tile"]},
        "kernel_times": kernels,
        "turb3d": t3d,
        "batch2d": bt,
        "sphere3d": sp,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
