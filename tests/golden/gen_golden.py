#!/usr/bin/env python3
"""Generate the golden vectors of tests/golden/*.npz from the reference programs' own step functions.

The reference has no tests and no known-answer vectors, so these are made here: oracle/_ref/
holds the reference sources compiled in place (oracle/Makefile `ref`, this container only), and this
script calls them through the oracle C API on seeded inputs and stores inputs and outputs.  Only
the vectors are committed.  Run:  make -C oracle ref && python tests/golden/gen_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import bind as OB  # noqa: E402


def _prims(rng, n, rho_lo=0.1, rho_hi=5.0, p_lo=0.01, p_hi=3.0):
    W = np.empty((n, 8))
    W[:, 0] = rng.uniform(rho_lo, rho_hi, n)
    W[:, 1:4] = rng.normal(0, 1.0, (n, 3))
    W[:, 4] = rng.uniform(p_lo, p_hi, n)
    W[:, 5:8] = rng.normal(0, 1.0, (n, 3))
    return W


def m2_vectors(ref, orc):
    rng = np.random.default_rng(2024)
    p = OB.m2_params()
    # HLLD, stratified over the 7 branches (classified with the restatement's branch log)
    want = 300
    bins = {b: [] for b in range(7)}
    tries = 0
    while min(len(v) for v in bins.values()) < want and tries < 400000:
        tries += 1
        WL, WR = _prims(rng, 1)[0], _prims(rng, 1)[0]
        mode = tries % 6
        Bn = 0.0 if mode == 0 else rng.normal(0, 1.0)
        if mode == 1:  # push toward supersonic
            WL[1] += 4.0 * np.sign(rng.normal())
            WR[1] = WL[1]
        if mode == 2:  # uniform state
            WR = WL.copy()
        out = np.zeros(8)
        orc.or_m2_hlld_x(C.byref(p), OB.ptr(WL), OB.ptr(WR), C.c_double(Bn), OB.ptr(out))
        b = orc.or_m2_last_branch()
        if len(bins[b]) < want:
            bins[b].append((WL, WR, Bn))
    items = [x for b in range(7) for x in bins[b]]
    WL = np.array([x[0] for x in items]); WR = np.array([x[1] for x in items]); Bn = np.array([x[2] for x in items])
    br = np.repeat(np.arange(7), [len(bins[b]) for b in range(7)])
    Fx = np.zeros((len(items), 8)); Fy = np.zeros((len(items), 8))
    for k in range(len(items)):
        a, bb = np.ascontiguousarray(WL[k]), np.ascontiguousarray(WR[k])
        ref.or_m2_hlld_x(C.byref(p), OB.ptr(a), OB.ptr(bb), C.c_double(Bn[k]), OB.ptr(Fx[k]))
        ref.or_m2_hlld_y(C.byref(p), OB.ptr(a), OB.ptr(bb), C.c_double(Bn[k]), OB.ptr(Fy[k]))
    np.savez_compressed(os.path.join(HERE, "m2_hlld.npz"), WL=WL, WR=WR, Bn=Bn, branch=br, Fx=Fx, Fy=Fy)

    # cons_to_prim (incl. floor cases), limiter, fast speeds
    c8 = np.concatenate([_prims(rng, 200), _prims(rng, 50, 1e-12, 1e-11, 1e-12, 1e-11)])
    c8[:, 4] = np.abs(rng.normal(1, 1, len(c8)))  # total energy, sometimes below kin+mag
    W8 = np.zeros_like(c8)
    for k in range(len(c8)):
        ref.or_m2_cons_to_prim(C.byref(p), OB.ptr(np.ascontiguousarray(c8[k])), OB.ptr(W8[k]))
    lim_in = rng.normal(0, 1, (400, 2)); lim_in[::5, 1] = 0.0; lim_in[::7, 0] = -lim_in[::7, 1]
    lim = np.array([ref.or_m2_mc_limiter(C.c_double(a), C.c_double(b)) for a, b in lim_in])
    Wf = _prims(rng, 200)
    cfx = np.array([ref.or_m2_fast_speed_x(C.byref(p), OB.ptr(np.ascontiguousarray(w))) for w in Wf])
    cfy = np.array([ref.or_m2_fast_speed_y(C.byref(p), OB.ptr(np.ascontiguousarray(w))) for w in Wf])
    np.savez_compressed(os.path.join(HERE, "m2_prims.npz"), c8=c8, W8=W8, lim_in=lim_in, lim=lim, Wf=Wf, cfx=cfx,
                        cfy=cfy)

    # Poisson (SOR, 600 sweeps) periodic and clamped
    out = {}
    for name, Nx, Ny, per in (("p16", 16, 16, 1), ("c16", 16, 16, 0), ("c24x20", 24, 20, 0), ("p20x12", 20, 12, 1)):
        pp = OB.m2_params(Nx=Nx, Ny=Ny, periodic_x=per, periodic_y=per, G=1.0)
        rho = 1.0 + rng.uniform(0, 2, Nx * Ny)
        gx, gy = np.zeros(Nx * Ny), np.zeros(Nx * Ny)
        ref.or_m2_poisson(C.byref(pp), OB.ptr(rho), C.c_double(1.0 / Nx), C.c_double(1.0 / Ny), OB.ptr(gx), OB.ptr(gy))
        out[name + "_rho"] = rho; out[name + "_gx"] = gx; out[name + "_gy"] = gy
    np.savez_compressed(os.path.join(HERE, "m2_poisson.npz"), **out)

    # advance_stage / rk2_step at 16x12 and a 12-step driver run at 32^2 (collapse IC, clamped)
    from core_collapse_amd import ic
    from tests.helpers import random_mhd2d_state
    res = {}
    for per in (1, 0):
        pm = ic.collapse2d_params(16) if False else None
        from core_collapse_amd.mhd2d import Params
        pp = Params(Nx=16, Ny=12, periodic_x=bool(per), periodic_y=bool(per), eta=2e-3, G=1.0, sor_iters=60)
        st = random_mhd2d_state(pp, 11 + per)
        op = OB.m2_params(Nx=16, Ny=12, periodic_x=per, periodic_y=per, eta=2e-3, G=1.0, sor_iters=60)
        S0 = OB.M2Arrays(16, 12, {k: v.reshape(-1) for k, v in st.items()})
        gx = rng.normal(0, 0.3, 16 * 12); gy = rng.normal(0, 0.3, 16 * 12)
        S1 = S0.copy()
        ref.or_m2_advance_stage(C.byref(op), C.byref(S0.cstate()), C.byref(S1.cstate()), OB.ptr(gx), OB.ptr(gy),
                                C.c_double(pp.dx), C.c_double(pp.dy), C.c_double(2e-3))
        S2 = S0.copy()
        ref.or_m2_rk2_step(C.byref(op), C.byref(S2.cstate()), OB.ptr(gx), OB.ptr(gy), C.c_double(pp.dx),
                           C.c_double(pp.dy), C.c_double(2e-3))
        for f in S0.FIELDS:
            res["per%d_in_%s" % (per, f)] = S0.a[f]
            res["per%d_stage_%s" % (per, f)] = S1.a[f]
            res["per%d_rk2_%s" % (per, f)] = S2.a[f]
        res["per%d_gx" % per] = gx; res["per%d_gy" % per] = gy
    pc = ic.collapse2d_params(32)
    st = ic.collapse2d_ic(pc)
    op = OB.m2_params(gamma=pc.gamma, cfl=pc.cfl, x0=pc.x0, x1=pc.x1, y0=pc.y0, y1=pc.y1, Nx=32, Ny=32, periodic_x=0,
                      periodic_y=0, eta=pc.eta, G=pc.G)
    S = OB.M2Arrays(32, 32, {k: v.reshape(-1) for k, v in st.items()})
    for f in S.FIELDS:
        res["run_in_" + f] = S.a[f].copy()
    dts = np.zeros(12)
    t = ref.or_m2_run(C.byref(op), C.byref(S.cstate()), C.c_double(pc.dx), C.c_double(pc.dy), 12, C.c_double(0.0),
                      C.c_double(1e300), OB.ptr(dts))
    for f in S.FIELDS:
        res["run_out_" + f] = S.a[f]
    res["run_dts"] = dts
    res["run_t"] = np.array([t])
    np.savez_compressed(os.path.join(HERE, "m2_step.npz"), **res)


def m3_vectors(ref):
    rng = np.random.default_rng(77)
    p = OB.m3_params()
    n = 1500
    WL = _prims(rng, n, 0.5, 200.0, 0.01, 300.0)
    WR = _prims(rng, n, 0.5, 200.0, 0.01, 300.0)
    Bn = rng.normal(0, 1.0, n); Bn[::6] = 0.0
    F = np.zeros((n, 7))
    for k in range(n):
        ref.or_m3_hlld_x(C.byref(p), OB.ptr(np.ascontiguousarray(WL[k])), OB.ptr(np.ascontiguousarray(WR[k])),
                         C.c_double(Bn[k]), OB.ptr(F[k]))
    rho = np.concatenate([np.geomspace(1e-11, 79.9, 60), np.linspace(80, 119.9, 40), np.linspace(120, 400, 40)])
    cold = np.zeros((len(rho), 3))
    for k, r in enumerate(rho):
        ref.or_m3_eos_cold(C.byref(p), C.c_double(r), OB.ptr(cold[k]))
    c8 = _prims(rng, 300, 0.5, 200.0, 0.0, 0.0)
    c8[:, 4] = np.abs(rng.normal(50, 60, 300))
    W8 = np.zeros_like(c8)
    for k in range(len(c8)):
        ref.or_m3_cons_to_prim(C.byref(p), OB.ptr(np.ascontiguousarray(c8[k])), OB.ptr(W8[k]))
    res = {"WL": WL, "WR": WR, "Bn": Bn, "F": F, "rho": rho, "cold": cold, "c8": c8, "W8": W8}
    from core_collapse_amd import ic
    from core_collapse_amd.mhd3d import Params
    for per in (1, 0):
        pp = Params(Nx=10, Ny=9, Nz=8, periodic_x=bool(per), periodic_y=bool(per), periodic_z=bool(per))
        st = ic.mhd3d_reference_state(pp)
        op = OB.m3_params(Nx=10, Ny=9, Nz=8, periodic_x=per, periodic_y=per, periodic_z=per)
        S = OB.M3Arrays(10, 9, 8, {k: v.reshape(-1) for k, v in st.items()})
        for f in S.FIELDS:
            res["per%d_in_%s" % (per, f)] = S.a[f].copy()
        dts = np.zeros(3)
        ref.or_m3_run(C.byref(op), C.byref(S.cstate()), C.c_double(pp.dx), C.c_double(pp.dy), C.c_double(pp.dz), 3,
                      C.c_double(0.0), C.c_double(0.15), OB.ptr(dts))
        for f in S.FIELDS:
            res["per%d_out_%s" % (per, f)] = S.a[f]
        res["per%d_dts" % per] = dts
    np.savez_compressed(os.path.join(HERE, "m3_step.npz"), **res)


def _hllc_branch(W, V, gamma, floor_rho, axis, guard):
    """0: F_L, 1: F_R, 2: star L, 3: star R (classification only; the vectors hold the reference output)."""
    un, vn = (W[1], V[1]) if axis == 0 else (W[2], V[2])
    aL = np.sqrt(gamma * W[3] / max(W[0], floor_rho))
    aR = np.sqrt(gamma * V[3] / max(V[0], floor_rho))
    SL, SR = min(un - aL, vn - aR), max(un + aL, vn + aR)
    rL, rR = max(W[0], floor_rho), max(V[0], floor_rho)
    num = (V[3] - W[3]) + rL * un * (SL - un) - rR * vn * (SR - vn)
    den = rL * (SL - un) - rR * (SR - vn)
    SM = num / den if abs(den) > guard else 0.0
    return 0 if SL >= 0 else 1 if SR <= 0 else 2 if SM > 0 else 3


def _hllc_vectors(ref, prefix, params, guard, fx, fy, rng, want=250):
    bins = {b: [] for b in range(4)}
    tries = 0
    while min(len(v) for v in bins.values()) < want and tries < 200000:
        tries += 1
        W = np.array([rng.uniform(0.1, 4), *rng.normal(0, 1, 2), rng.uniform(0.01, 3)])
        V = np.array([rng.uniform(0.1, 4), *rng.normal(0, 1, 2), rng.uniform(0.01, 3)])
        mode = tries % 5
        if mode == 1:
            W[1:3] += 5.0 * np.sign(rng.normal())
            V[1:3] = W[1:3] + rng.normal(0, 0.1, 2)
        if mode == 2:
            V = W.copy()
        b = _hllc_branch(W, V, params.gamma, params.floor_rho, tries % 2, guard)
        if len(bins[b]) < want:
            bins[b].append((W, V))
    pairs = [x for b in range(4) for x in bins[b]]
    WL = np.array([a for a, _ in pairs])
    WR = np.array([b for _, b in pairs])
    Fx, Fy = np.zeros_like(WL), np.zeros_like(WL)
    for k in range(len(WL)):
        getattr(ref, fx)(C.byref(params), OB.ptr(WL[k]), OB.ptr(WR[k]), OB.ptr(Fx[k]))
        getattr(ref, fy)(C.byref(params), OB.ptr(WL[k]), OB.ptr(WR[k]), OB.ptr(Fy[k]))
    return dict(WL=WL, WR=WR, Fx=Fx, Fy=Fy)


def eh_vectors(ref):
    rng = np.random.default_rng(77)
    res = {}
    for k, v in _hllc_vectors(ref, "eh", OB.eh_params(), 1e-14, "or_eh_hllc_x", "or_eh_hllc_y", rng).items():
        res["hllc_" + k] = v
    # Poisson, rhs, rk2 and short runs on the program's own IC (init_random_blob, EH:574-642)
    for tag, (Nx, Ny, per) in {"a": (16, 16, 1), "b": (24, 20, 0)}.items():
        p = OB.eh_params(Nx=Nx, Ny=Ny, periodic_x=per, periodic_y=per)
        S = OB.EHArrays(Nx, Ny)
        ref.or_eh_init_random_blob(C.byref(p), C.byref(S.cstate()))
        for f in S.FIELDS:
            res[tag + "_ic_" + f] = S.a[f].copy()
        gx, gy = np.zeros(Nx * Ny), np.zeros(Nx * Ny)
        ref.or_eh_poisson(C.byref(p), C.byref(S.cstate()), OB.ptr(gx), OB.ptr(gy), None)
        res[tag + "_gx"], res[tag + "_gy"] = gx, gy
        res[tag + "_dt"] = np.array([ref.or_eh_compute_dt(C.byref(p), C.byref(S.cstate()))])
        dU = OB.EHArrays(Nx, Ny)
        ref.or_eh_rhs(C.byref(p), C.byref(S.cstate()), C.byref(dU.cstate()))
        for f in S.FIELDS:
            res[tag + "_rhs_" + f] = dU.a[f]
        T = S.copy()
        ref.or_eh_rk2_step(C.byref(p), C.byref(T.cstate()), C.c_double(2e-3))
        for f in S.FIELDS:
            res[tag + "_rk2_" + f] = T.a[f]
        R = S.copy()
        dts = np.zeros(10)
        t = ref.or_eh_run(C.byref(p), C.byref(R.cstate()), 10, C.c_double(0.0), C.c_double(5.0), OB.ptr(dts))
        res[tag + "_run_dts"], res[tag + "_run_t"] = dts, np.array([t])
        for f in S.FIELDS:
            res[tag + "_run_" + f] = R.a[f]
    np.savez_compressed(os.path.join(HERE, "eh_step.npz"), **res)


def ns_vectors(ref):
    rng = np.random.default_rng(78)
    res = {}
    for k, v in _hllc_vectors(ref, "ns", OB.ns_params(), 0.0, "or_ns_hllc_x", "or_ns_hllc_y", rng).items():
        res["hllc_" + k] = v
    n = 4000
    for name, fn, args in (("normal", "or_std_normal", ()), ("uniform", "or_std_uniform", (C.c_double(-1.0),
                                                                                           C.c_double(1.0)))):
        a = np.zeros(n)
        getattr(ref, fn)(42, n, *args, OB.ptr(a))
        res["std_" + name] = a
    for (Nx, Ny) in ((16, 16), (32, 8)):
        d = rng.normal(0, 1, (Ny, Nx, 2))
        res["fft_in_%dx%d" % (Nx, Ny)] = d.copy()
        for inv in (0, 1):
            e = d.reshape(-1).copy()
            ref.or_ns_fft2d(Nx, Ny, inv, OB.ptr(e))
            res["fft_out%d_%dx%d" % (inv, Nx, Ny)] = e.reshape(Ny, Nx, 2)
    from core_collapse_amd import ic
    from core_collapse_amd.ns2d import Params as NSP
    for tag, (Nx, Ny, vis) in {"a": (32, 32, 1), "b": (16, 32, 0)}.items():
        p = OB.ns_params(Nx=Nx, Ny=Ny, viscosity_on=vis, mu=1e-3)
        U = ic.ns2d_reference_ic(NSP(Nx=Nx, Ny=Ny)).reshape(-1).copy()
        res[tag + "_ic"] = U.copy()
        rho = np.maximum(U[0::4], p.floor_rho).copy()
        phi = np.zeros(Nx * Ny)
        ref.or_ns_poisson_fft(C.byref(p), OB.ptr(rho), OB.ptr(phi))
        res[tag + "_phi"] = phi
        res[tag + "_dt"] = np.array([ref.or_ns_compute_dt(C.byref(p), OB.ptr(U))])
        dU = np.zeros(4 * Nx * Ny)
        ref.or_ns_compute_rhs(C.byref(p), OB.ptr(U), OB.ptr(dU))
        res[tag + "_rhs"] = dU
        T = U.copy()
        ref.or_ns_heun_step(C.byref(p), OB.ptr(T), C.c_double(2e-3))
        res[tag + "_heun"] = T
        R = U.copy()
        dts = np.zeros(20)
        t = ref.or_ns_run(C.byref(p), OB.ptr(R), 20, C.c_double(0.0), C.c_double(10.0), OB.ptr(dts))
        res[tag + "_run"], res[tag + "_run_dts"], res[tag + "_run_t"] = R, dts, np.array([t])
    np.savez_compressed(os.path.join(HERE, "ns_step.npz"), **res)


def config1_vectors(ref):
    """The 2D collapse benchmark's plain FP64 run (128^2, 200 steps, reference arithmetic) through the
    reference's own step loop: the final state, every dt and t (m2_config1.npz)."""
    from core_collapse_amd import ic
    p = ic.collapse2d_params(128, ct_compensated=False)
    st = ic.collapse2d_ic(p)
    op = OB.m2_params(gamma=p.gamma, cfl=p.cfl, x0=p.x0, x1=p.x1, y0=p.y0, y1=p.y1, Nx=128, Ny=128, periodic_x=0,
                      periodic_y=0, eta=p.eta, G=p.G, sor_iters=p.sor_iters, sor_omega=p.sor_omega)
    S = OB.M2Arrays(128, 128, {k: v.reshape(-1) for k, v in st.items()})
    res = {"in_rho_sum": np.array([S.a["rho"].sum()])}
    dts = np.zeros(200)
    t = ref.or_m2_run(C.byref(op), C.byref(S.cstate()), C.c_double(p.dx), C.c_double(p.dy), 200, C.c_double(0.0),
                      C.c_double(1e300), OB.ptr(dts))
    for f in S.FIELDS:
        res["out_" + f] = S.a[f]
    res["dts"] = dts
    res["t"] = np.array([t])
    np.savez_compressed(os.path.join(HERE, "m2_config1.npz"), **res)


if __name__ == "__main__":
    import sys
    if sys.argv[1:] == ["config1"]:
        config1_vectors(OB.Lib("ref_m2"))
        sys.exit(0)
    orc = OB.Lib("oracle")
    m2_vectors(OB.Lib("ref_m2"), orc)
    m3_vectors(OB.Lib("ref_m3"))
    eh_vectors(OB.Lib("ref_eh"))
    ns_vectors(OB.Lib("ref_ns"))
    config1_vectors(OB.Lib("ref_m2"))
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
