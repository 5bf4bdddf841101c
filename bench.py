#!/usr/bin/env python3
"""Benchmark of the solver step on one MI355X.

Headline workload (config 2): 2D resistive MHD, HLLD + constrained transport + SOR self-gravity (two
600-sweep Poisson solves per Heun step), Gaussian collapse on a 2048 x 2048 grid over [-0.5,0.5]^2 with
clamped boundaries, gamma 5/3, G 1, eta 1e-4, CFL 0.4, compensated face update.  --steps steps run once in
plain FP64 and once in the mixed form (FP32 reconstruction, Riemann solver and SOR sweeps; FP64 update,
B/CT and Poisson residual) from the same seeded IC, each timed on the captured-graph path with no
per-launch events.  A zone-update is one cell advanced one full Heun step including its two Poisson
solves; `value` is the mixed zone-updates/s with all state resident in HBM.

After the timed steps, each precision runs --profile-steps more with HIP events around every launch on
the handle's stream: `roofline` (dominant kernel: the SOR dataflow launch, one per 600-sweep solve) and
`kernel_times` come from that pass.  The SOR's algorithmic bytes per launch are cell-sweeps x 24 B (FP64:
read phi, read rhs, write phi) or 12 B (mixed); the tile keeps 16 sweeps in LDS, so this rate can exceed
the HBM peak -- `traffic` (rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, profiles/hbm_traffic.json) and
`traffic_frac` give the HBM load itself.

`accuracy`: both precisions run --accuracy-steps (1000) steps from the IC (timed too); per frame field
the relative L1 and L-inf error of mixed against FP64 and max|div B| of both.

cpu_baseline: the oracle's C restatement of the reference step (bit-identical to the reference's step
functions on the golden vectors), single thread, on a bounded sample of the same problem.

The line also carries "turb3d" (config 3: 3D HLLD+CT turbulence at 256^3, 500 steps per precision,
per-field errors, face-kernel roofline without the FP64 pass over queued faces), "batch2d" (config 4:
64 collapse instances in one graph per step) and "sphere3d" (config 5: 3D Euler + FFT gravity at 512^3,
300 steps per precision).  --workload selects one problem.
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
    ap.add_argument("--steps", type=int, default=20, help="timed steps of the headline (2048^2 collapse)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--N", type=int, default=2048)
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="length of the FP64 and mixed 2048^2 runs compared field by field (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="bound on each CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("all", "collapse2d", "turb3d", "batch2d", "sphere3d"), default="all")
    ap.add_argument("--N5", type=int, default=512, help="grid of the 3D Euler sphere problem")
    ap.add_argument("--steps5", type=int, default=300)
    ap.add_argument("--batch", type=int, default=64, help="instances of the batched collapse")
    ap.add_argument("--steps-batch", type=int, default=200)
    ap.add_argument("--sor-autotune", type=int, default=1,
                    help="1: time the compiled SOR tile shapes once per 2D grid size and use the fastest")
    ap.add_argument("--seq-steps", type=int, default=10,
                    help="steps per instance of the one-after-another comparison (0: skip)")
    ap.add_argument("--N3", type=int, default=256, help="grid of the 3D turbulence problem")
    ap.add_argument("--steps3", type=int, default=500)
    ap.add_argument("--profile-steps", type=int, default=5,
                    help="steps of the separate profiled pass (HIP events per launch) that gives kernel times")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (for counter calibration) and exit")
    return ap.parse_args()


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def timed_run(h, steps, t=0.0):
    """steps of the handle's driver loop (captured HIP graphs, no per-launch events), wall time around them."""
    h.sync()
    t0 = time.perf_counter()
    t = h.run(steps, t)          # synchronises at the end
    return t, time.perf_counter() - t0


def profiled_pass(h, steps, t, names):
    """steps more with per-launch HIP events on the handle's stream: {kernel scope: (launches, total ms)}."""
    h.set_profiling(True)
    h.reset_profiling()
    t = h.run(steps, t)
    prof = {k: h.kernel_time(k) for k in names}
    h.set_profiling(False)
    return t, prof


def field_errors(cmp, ref_frame, names):
    """Per frame field: relative L1 (sum|a-b| / sum|b|) and L-inf relative to the field's max |b|."""
    out = {}
    for f in names:
        mean, rl1, mx = cmp[f]
        scale = float(np.abs(ref_frame[f]).max())
        out[f] = {"l1_rel": rl1, "linf_rel": mx / scale if scale > 0 else mx, "linf_abs": mx}
    return out


def bench_collapse2d(N, warmup, steps, acc_steps, prof_steps):
    """Config 2: the 2D collapse at N^2, plain FP64 and mixed.  Headline: `steps` steps per precision timed
    on the graph path; a profiled pass of prof_steps gives the kernel times; then both precisions run
    acc_steps from the IC and are compared field by field on the device."""
    from core_collapse_amd import CC_FP64, CC_MIXED
    from core_collapse_amd import ic
    from core_collapse_amd import mhd2d as M
    res = {}
    for prec in (CC_FP64, CC_MIXED):
        p = ic.collapse2d_params(N, precision=prec)
        h = M.Mhd2D(p)
        h.upload(ic.collapse2d_ic(p))
        t = h.run(warmup) if warmup > 0 else 0.0
        t, el = timed_run(h, steps, t)
        t, prof = profiled_pass(h, prof_steps, t, ("sor", "flux", "ct", "grad", "mean", "dt"))
        res[prec] = {"elapsed": el, "prof": prof, "tile": M.sor_tile_choice(N, N, p.sor_iters, prec, with_kernels=True)}
        h.close()
        log("collapse2d %s %.3f ms/step" % ("fp64" if prec == CC_FP64 else "mixed", 1e3 * el / steps))
    acc = None
    if acc_steps > 0:
        hs, tacc = {}, {}
        for prec in (CC_FP64, CC_MIXED):
            p = ic.collapse2d_params(N, precision=prec)
            hs[prec] = M.Mhd2D(p)
            hs[prec].upload(ic.collapse2d_ic(p))
            _, tacc[prec] = timed_run(hs[prec], acc_steps)
            log("collapse2d %d-step %s run %.1f s" % (acc_steps, "fp64" if prec == CC_FP64 else "mixed", tacc[prec]))
        cmp = hs[CC_MIXED].compare(hs[CC_FP64])
        ref = hs[CC_FP64].frame()
        zu = float(N) * N * acc_steps
        acc = {"steps": acc_steps,
               "fp64": {"value": zu / tacc[CC_FP64], "ms_per_step": 1e3 * tacc[CC_FP64] / acc_steps,
                        "max_divB": hs[CC_FP64].divB_max()},
               "mixed": {"value": zu / tacc[CC_MIXED], "ms_per_step": 1e3 * tacc[CC_MIXED] / acc_steps,
                         "max_divB": hs[CC_MIXED].divB_max()},
               "mixed_over_fp64": tacc[CC_FP64] / tacc[CC_MIXED],
               "errors": field_errors(cmp, ref, ("rho", "P", "u", "v", "vabs", "Bx", "By", "Bz"))}
        for h in hs.values():
            h.close()
    return res, acc


def sor_cell_sweeps_per_solve(N, iters=600):
    return float(N) * N * iters


def cpu_info():
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return "%s, nproc=%d" % (cpu, os.cpu_count() or 0)


def cpu_baseline(N, budget_s):
    """The oracle's restatement of the 2D step (bit-identical to the reference's step functions on the
    golden vectors) on a sample grid of the same problem, single thread."""
    from core_collapse_amd import ic
    from oracle import bind as OB
    lib = OB.Lib("oracle")
    n = 512 if N >= 512 else N
    p = ic.collapse2d_params(n, ct_compensated=False)
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
    return {"value": n * n * steps / tsum, "unit": "zone-updates/s", "cores": 1, "kind": "port",
            "sample": "%d Heun step(s) of the same collapse problem at %dx%d (FP64 reference arithmetic, 600-sweep "
                      "SOR x2 per step, %.1f s), single thread, the oracle's C restatement of the reference step "
                      "(bit-identical to it on the golden vectors), %s" % (steps, n, n, tsum, cpu_info())}


def bench_turb3d(N, warmup, steps, prof_steps):
    """Config 3: 3D resistive HLLD+CT, turbulence at N^3, FP64 and mixed, `steps` steps each from the IC
    timed on the graph path, compared field by field on the device; then a profiled pass."""
    from core_collapse_amd import CC_FP64, CC_MIXED
    from core_collapse_amd import ic
    from core_collapse_amd import mhd3d as M3
    hs, el, prof = {}, {}, {}
    names = ("prim", "face", "face_fp64", "update", "zupd", "ect", "edge", "ct", "joule", "dt")
    for prec in (CC_FP64, CC_MIXED):
        p = ic.turbulence3d_params(N, precision=prec)
        h = M3.Mhd3D(p)
        h.upload(ic.turbulence3d_ic(p))
        if warmup > 0:  # warm the graphs on a copy of the IC, then restart from it
            h.run(warmup)
            h.upload(ic.turbulence3d_ic(p))
        _, el[prec] = timed_run(h, steps)
        hs[prec] = h
        log("turb3d %s %.3f ms/step over %d steps" % ("fp64" if prec == CC_FP64 else "mixed", 1e3 * el[prec] / steps,
                                                        steps))
    cmp = hs[CC_MIXED].compare(hs[CC_FP64])
    ref = hs[CC_FP64].frame()
    errs = field_errors(cmp, ref, ("rho", "P", "u", "v", "w", "vabs", "Bx", "By", "Bz"))
    divb = {prec: hs[prec].divB_max() for prec in hs}
    emax = float(ref["P"].max())
    for prec in (CC_FP64, CC_MIXED):
        _, prof[prec] = profiled_pass(hs[prec], prof_steps, 0.0, names)
        hs[prec].close()
    return el, prof, errs, divb, emax


def cpu_baseline3d(n=128, budget_s=12.0):
    """The oracle's restatement of the 3D step (ideal-gas EOS mode) on a 128^3 sample."""
    from core_collapse_amd import ic
    from oracle import bind as OB
    lib = OB.Lib("oracle")
    p = ic.turbulence3d_params(n)
    st = ic.turbulence3d_ic(p)
    op = OB.m3_params(Nx=n, Ny=n, Nz=n, cfl=p.cfl, eta=p.eta, ideal_eos=1, gamma_ideal=p.gamma, hlld_guard=p.hlld_guard,
                      ct_comp=int(p.ct_compensated))
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
                      "oracle restatement (the reference program has no ideal-gas EOS), %s" % (steps, n, tsum, cpu_info())}


def bench3d(N, warmup, steps, prof_steps, with_cpu):
    from core_collapse_amd import CC_FP64, CC_MIXED
    el, prof, errs, divb, pmax = bench_turb3d(N, warmup, steps, prof_steps)
    zu = float(N) ** 3 * steps

    def roof(pr, word):
        n, ms = pr["face"]
        if n == 0 or ms <= 0:
            return None
        # per face: read the 8 primitives of one cell and the face B (FP64), write 5 fluxes + 2 EMFs; the
        # mixed path's FP64 pass over queued faces is timed apart ("face_fp64") and not included here
        bpl = float(N) ** 3 * (8 * word + 8 + 7 * word)
        avg = ms / n * 1e-3
        return {"bound": "hbm", "achieved": round(bpl / avg / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(bpl / avg / 1e9 / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_face3d",
                "launches": n, "avg_launch_us": round(avg * 1e6, 2), "alg_bytes_per_launch": bpl}

    r64, rmx = roof(prof[CC_FP64], 8), roof(prof[CC_MIXED], 4)
    if r64:
        r64["traffic"] = load_traffic("cc::k_face3d<0, false")
    if rmx:
        rmx["traffic"] = load_traffic("cc::k_face3d<0, true")
    out = {
        "workload": "3D resistive MHD HLLD+CT, periodic turbulence %d^3, ideal gas gamma 5/3, beta 1 along z, "
                    "rms Mach 1, eta 5e-5, CFL 0.4, HLLD degeneracy guard 5e-2, compensated face update" % N,
        "steps": steps, "warmup": warmup,
        "value": zu / el[CC_MIXED], "unit": "zone-updates/s",
        "ms_per_step": 1e3 * el[CC_MIXED] / steps,
        "fp64": {"value": zu / el[CC_FP64], "ms_per_step": 1e3 * el[CC_FP64] / steps, "roofline": r64,
                 "max_divB": divb[CC_FP64]},
        "mixed_over_fp64": el[CC_FP64] / el[CC_MIXED],
        "errors": errs,
        "l1_rho_rel_err": errs["rho"]["l1_rel"],
        "linf_rho_rel_err": errs["rho"]["linf_rel"],
        "max_divB": divb[CC_MIXED],
        "max_P_fp64": pmax,
        "roofline": rmx,
        "zone_update_alg_bytes": 320,
        "kernel_times": {tag: {k: {"launches": v[0], "total_ms": round(v[1], 3)} for k, v in prof[prec].items()
                               if v[0] > 0} for prec, tag in ((CC_FP64, "fp64"), (CC_MIXED, "mixed"))},
        "profiled_steps": prof_steps,
    }
    out["hbm_frac_alg"] = round(out["value"] * 320 / 1e9 / HBM_PEAK_GBS, 4)
    for prec, tag, R, mx in ((CC_FP64, "fp64", "double", "false"), (CC_MIXED, "mixed", "float", "true")):
        # FP64 solves the z faces inside the column pass (k_zupd3d) that also updates the cells; the mixed path
        # runs the three face kernels and the update
        face_axes = 2 if prec == CC_FP64 else 3
        groups = {"prim": ["cc::k_prim3d<%s>" % R],
                  "face": ["cc::k_face3d<%d, %s," % (a, mx) for a in range(face_axes)],
                  "update": ["cc::k_update3d<%s>" % R], "zupd": ["cc::k_zupd3d<%s," % mx],
                  "ect": ["cc::k_ect3d<%s, 32, 12," % R],
                  "edge": ["cc::k_edge3d<%s>" % R], "ct": ["cc::k_ct3d"],
                  "joule": ["cc::k_joule3d<%s, true, true>" % R, "cc::k_joule3d<%s, false, true>" % R]}
        hbm = measured_hbm(prof[prec], groups)
        out.setdefault("hbm_measured", {})[tag] = hbm
        zu_p = float(N) ** 3 * prof_steps
        out.setdefault("fusion_factor", {})[tag] = fusion_factor(hbm, zu_p, 320)
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
        "sor_merged_fp64_kernels": binfo["fp64_kernels"],
        "sor_tiles_alone": tiles,
        "max_divB": divb, "t_min": float(np.min(t)), "t_max": float(np.max(t)),
    }


def run_gpu_sphere(N, precision, warmup, steps, prof_steps):
    from core_collapse_amd import euler3d as E3
    from core_collapse_amd import ic
    p = ic.sphere3d_params(N, precision=precision)
    h = E3.Euler3D(p)
    h.upload(ic.sphere3d_ic(p))
    if warmup > 0:  # warm the graphs, then restart from the IC
        h.run(warmup)
        h.upload(ic.sphere3d_ic(p))
    t, elapsed = timed_run(h, steps)
    stats = h.poisson_stats()
    out = h.download()
    tot = h.totals()
    _, prof = profiled_pass(h, prof_steps, t, ("prim", "face", "update", "fused", "mean", "fft_x", "fft_y", "fft_z",
                                               "check", "grad", "dt"))
    h.close()
    return {"elapsed": elapsed, "state": out, "prof": prof, "poisson": stats, "mass": tot["mass"]}


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


def bench_sphere3d(N, warmup, steps, prof_steps, with_cpu):
    """3D Euler HLLC + FFT self-gravity, uniform sphere, FP64 and mixed (FP32 primitives, fluxes and FFT;
    FP64 state, update and Poisson residual), `steps` steps each from the IC."""
    from core_collapse_amd import CC_FP64, CC_MIXED
    res = {}
    for prec in (CC_FP64, CC_MIXED):
        res[prec] = run_gpu_sphere(N, prec, warmup, steps, prof_steps)
        log("sphere3d %s %.3f ms/step over %d steps" % ("fp64" if prec == CC_FP64 else "mixed",
                                                         1e3 * res[prec]["elapsed"] / steps, steps))
    zu = float(N) ** 3 * steps
    cells = float(N) ** 3

    def roof(r, word, tname):
        # the faces + update pass (k_e3_fused), or with CC_E3_FUSED=0 the stage update (k_e3_update); one launch
        # per stage, half of them the final stage.  Per cell the fused pass reads the 5 primitives (in the flux
        # precision) and the update reads 15 face fluxes instead; both read the state (5 FP64), the potential
        # (1 FP64, differenced for g), on the final stage also U^n (5 FP64), and write the state: 5 (or 15) w +
        # 88 B or + 128 B, + 108 B on average
        kern = "k_e3_fused" if r["prof"].get("fused", (0, 0.0))[0] > 0 else "k_e3_update"
        n, ms = r["prof"]["fused" if kern == "k_e3_fused" else "update"]
        if n == 0 or ms <= 0:
            return None
        bpl = cells * ((5 if kern == "k_e3_fused" else 15) * word + 108)
        avg = ms / n * 1e-3
        t1, t2 = load_traffic("cc::%s<1, %s>" % (kern, tname)), load_traffic("cc::%s<2, %s>" % (kern, tname))
        return {"bound": "hbm", "achieved": round(bpl / avg / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(bpl / avg / 1e9 / HBM_PEAK_GBS, 4),
                "traffic": 0.5 * (t1 + t2) if t1 and t2 else None, "kernel": kern,
                "launches": n, "avg_launch_us": round(avg * 1e6, 2), "alg_bytes_per_launch": bpl}

    def balg(r):
        k = max(r["poisson"]["iters_last"], 1)
        return 392 + 192 * k

    errs = {}
    for f in ("rho", "mx", "my", "mz", "E"):
        a, b = res[CC_MIXED]["state"][f], res[CC_FP64]["state"][f]
        d = np.abs(a - b)
        errs[f] = {"l1_rel": float(d.sum() / max(np.abs(b).sum(), 1e-300)),
                   "linf_rel": float(d.max() / max(np.abs(b).max(), 1e-300))}
    out = {
        "workload": "3D Euler HLLC + FFT Poisson self-gravity, periodic %d^3, gamma 5/3, G 1, uniform sphere rho 10 "
                    "(r < 0.2) in 0.1, 1%% seeded noise, p 0.01, at rest; FFT refined in FP64 to 1e-10" % N,
        "steps": steps, "warmup": warmup,
        "value": zu / res[CC_MIXED]["elapsed"], "unit": "zone-updates/s",
        "ms_per_step": 1e3 * res[CC_MIXED]["elapsed"] / steps,
        "fp64": {"value": zu / res[CC_FP64]["elapsed"], "ms_per_step": 1e3 * res[CC_FP64]["elapsed"] / steps,
                 "roofline": roof(res[CC_FP64], 8, "double"), "poisson": res[CC_FP64]["poisson"]},
        "mixed_over_fp64": res[CC_FP64]["elapsed"] / res[CC_MIXED]["elapsed"],
        "errors": errs,
        "l1_rho_rel_err": errs["rho"]["l1_rel"],
        "linf_rho_rel_err": errs["rho"]["linf_rel"],
        "mass_fp64": res[CC_FP64]["mass"], "mass_mixed": res[CC_MIXED]["mass"],
        "roofline": roof(res[CC_MIXED], 4, "float"),
        "poisson": res[CC_MIXED]["poisson"],
        "zone_update_alg_bytes": balg(res[CC_MIXED]),
        "kernel_times": {tag: {k: {"launches": v[0], "total_ms": round(v[1], 3)} for k, v in res[prec]["prof"].items()
                               if v[0] > 0} for prec, tag in ((CC_FP64, "fp64"), (CC_MIXED, "mixed"))},
        "profiled_steps": prof_steps,
    }
    out["hbm_frac_alg"] = round(out["value"] * balg(res[CC_MIXED]) / 1e9 / HBM_PEAK_GBS, 4)
    for prec, tag, R in ((CC_FP64, "fp64", "double"), (CC_MIXED, "mixed", "float")):
        groups = {"prim": ["cc::k_e3_prim<%s>" % R], "face": tuple("cc::k_e3_face<%s, %d>" % (R, a) for a in range(3)),
                  "update": ["cc::k_e3_update<1, %s>" % R, "cc::k_e3_update<2, %s>" % R],
                  "fused": ["cc::k_e3_fused<1, %s>" % R, "cc::k_e3_fused<2, %s>" % R], "grad": ["cc::k_p3_grad"]}
        hbm = measured_hbm(res[prec]["prof"], groups)
        out.setdefault("hbm_measured", {})[tag] = hbm
        # hydro kernels only: B_alg's hydro share is 5 state passes of 40 B plus g, 224 B per zone-update
        out.setdefault("fusion_factor_hydro", {})[tag] = fusion_factor(hbm, cells * prof_steps, 224)
    if with_cpu:
        out["cpu_baseline"] = cpu_baseline_sphere()
    for prec in res:
        res[prec]["state"] = None
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
        log("%d^3 turbulence, %d steps per precision" % (a.N3, a.steps3))
        t3d = bench3d(a.N3, 2, a.steps3, a.profile_steps, not a.no_cpu)
        if a.workload == "turb3d":
            print(json.dumps({"turb3d": t3d}))
            return
    sp = None
    if a.workload in ("all", "sphere3d"):
        log("%d^3 sphere, %d steps per precision" % (a.N5, a.steps5))
        sp = bench_sphere3d(a.N5, 2, a.steps5, a.profile_steps, not a.no_cpu)
        if a.workload == "sphere3d":
            print(json.dumps({"sphere3d": sp}))
            return
    bt = None
    if a.workload in ("all", "batch2d"):
        log("batch of %d collapses, timed %d steps" % (a.batch, a.steps_batch))
        bt = bench_batch2d(a.batch, a.steps_batch, 3, 128, 1024, a.seq_steps)
        log("batch2d %.3f ms/step (sequential %s ms/step)" % (bt["ms_per_step"], bt["sequential_estimate"]["ms_per_step"]))
        if a.workload == "batch2d":
            print(json.dumps({"batch2d": bt}))
            return
    log("%dx%d collapse, warmup %d, timed %d steps per precision, then %d-step runs compared" % (N, N, W, K,
                                                                                                 a.accuracy_steps))
    res, acc = bench_collapse2d(N, W, K, a.accuracy_steps, a.profile_steps)
    zone_updates = float(N) * N * K
    v64 = zone_updates / res[CC_FP64]["elapsed"]
    vmx = zone_updates / res[CC_MIXED]["elapsed"]

    def roof(r, bytes_per_sweep, traffic_name):
        n, ms = r["prof"]["sor"]
        if n == 0 or ms <= 0:
            return None
        sweeps = sor_cell_sweeps_per_solve(N) * 2 * a.profile_steps  # two solves per step
        bytes_per_launch = sweeps * bytes_per_sweep / n
        avg_s = ms / n * 1e-3
        ach = bytes_per_launch / avg_s / 1e9
        tr = load_traffic(traffic_name)
        return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": tr,
                "traffic_frac": round(tr / avg_s / 1e9 / HBM_PEAK_GBS, 4) if tr else None,
                "kernel": traffic_name.replace("cc::", ""), "launches": n, "avg_launch_us": round(avg_s * 1e6, 2),
                "alg_bytes_per_launch": bytes_per_launch,
                "note": "algorithmic rate of a temporally blocked solver (16 sweeps per tile pass through LDS): "
                        "achieved can exceed the HBM peak; traffic_frac is the counter-measured HBM fraction"}

    def sor_kernel(tile, prec):
        """The SOR kernel a tile choice runs and its algorithmic bytes per cell-sweep (the mixed kernels sweep
        an FP32 correction: 12 B; the FP64 kernels, also the choice of a mixed grid where they are faster: 24 B)."""
        S, WJ, sheared, fp64k = tile
        form = ("k_sor_s" if sheared else "k_sor_") + ("tile" if os.environ.get("CC_SOR_LEVELS") else "flow")
        if prec == CC_MIXED and not fp64k:
            return "cc::%s_mixed<%d, %d>" % (form, S, WJ), 12.0
        return "cc::%s_fp64<%d, %d%s" % (form, S, WJ, ">" if sheared else ","), 24.0

    kmx, bmx = sor_kernel(res[CC_MIXED]["tile"], CC_MIXED)
    k64, b64 = sor_kernel(res[CC_FP64]["tile"], CC_FP64)
    roof_mx = roof(res[CC_MIXED], bmx, kmx)
    roof_64 = roof(res[CC_FP64], b64, k64)
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
        "dtype": ("f32 reconstruction/Riemann fluxes + f64 SOR sweeps/update/CT" if res[CC_MIXED]["tile"][3] else
                  "f32 fluxes/SOR sweeps + f64 update/CT/residual"),
        "data": "synthetic",
        "config": {"workload": "2D resistive MHD HLLD+CT+SOR self-gravity, Gaussian collapse %dx%d, clamped "
                               "boundaries, 600 SOR sweeps x 2 solves per Heun step, compensated face update" % (N, N),
                   "grid": [N, N], "gamma": 5.0 / 3.0, "G": 1.0, "eta": 1e-4, "cfl": 0.4, "sor_iters": 600,
                   "sor_omega": 1.8, "seed": 42},
        "fp64": {"value": v64, "ms_per_step": 1e3 * res[CC_FP64]["elapsed"] / K, "roofline": roof_64},
        "mixed_over_fp64": vmx / v64,
        "accuracy": acc,
        "l1_rho_rel_err": acc["errors"]["rho"]["l1_rel"] if acc else None,
        "linf_rho_rel_err": acc["errors"]["rho"]["linf_rel"] if acc else None,
        "max_divB": acc["mixed"]["max_divB"] if acc else None,
        "roofline": roof_mx,
        "cpu_baseline": cpu,
        "sor_tile": {"fp64": res[CC_FP64]["tile"], "mixed": res[CC_MIXED]["tile"]},
        "kernel_times": kernels,
        "profiled_steps": a.profile_steps,
        "turb3d": t3d,
        "batch2d": bt,
        "sphere3d": sp,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
