"""Shared helpers for the parity tests: oracle access and state conversion."""
import ctypes as C
import os

import numpy as np

from oracle import bind as OB

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        path = os.path.join(ROOT, "oracle", "liboracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"])
        _oracle = OB.Lib("oracle")
    return _oracle


def ref(kind):
    """The reference harness library oracle/_ref/lib<kind>.so, or None where it is not built."""
    try:
        return OB.Lib(kind)
    except FileNotFoundError:
        return None


def m2_oparams(p):
    """Oracle params from a core_collapse_amd.mhd2d.Params."""
    return OB.m2_params(gamma=p.gamma, cfl=p.cfl, x0=p.x0, x1=p.x1, y0=p.y0, y1=p.y1, Nx=p.Nx, Ny=p.Ny,
                        floor_rho=p.floor_rho, floor_P=p.floor_P, periodic_x=int(p.periodic_x),
                        periodic_y=int(p.periodic_y), eta=p.eta, gravity_on=int(p.gravity_on), G=p.G,
                        sor_iters=p.sor_iters, sor_omega=p.sor_omega, hlld_guard=p.hlld_guard,
                        ct_comp=int(p.ct_compensated))


def m2_arrays(p, state):
    return OB.M2Arrays(p.Nx, p.Ny, {k: np.asarray(v).reshape(-1) for k, v in state.items()})


def m2_dict(p, arr):
    Nx, Ny = p.Nx, p.Ny
    sh = {"Bx_f": (Ny, Nx + 1), "By_f": (Ny + 1, Nx)}
    return {k: arr.a[k].reshape(sh.get(k, (Ny, Nx))).copy() for k in arr.FIELDS}


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


def rel_l1(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


def random_mhd2d_state(p, seed, periodic_div_free=True):
    """A smooth random state with a divergence-free face field from a corner potential."""
    rng = np.random.default_rng(seed)
    Nx, Ny = p.Nx, p.Ny
    x = (np.arange(Nx) + 0.5) / Nx
    y = (np.arange(Ny) + 0.5) / Ny
    XC, YC = np.meshgrid(x, y)
    def smooth(a, k=3):
        ph = rng.uniform(0, 2 * np.pi, 4)
        return a * (np.sin(2 * np.pi * XC + ph[0]) * np.cos(2 * np.pi * YC + ph[1]) +
                    0.5 * np.sin(4 * np.pi * XC + ph[2]) * np.sin(2 * np.pi * YC + ph[3]))
    rho = 1.0 + 0.3 * smooth(1.0) + 0.02 * rng.standard_normal((Ny, Nx))
    rho = np.abs(rho) + 0.2
    u = smooth(0.4) + 0.05 * rng.standard_normal((Ny, Nx))
    v = smooth(0.4) + 0.05 * rng.standard_normal((Ny, Nx))
    w = smooth(0.2)
    P = 0.6 + 0.2 * smooth(1.0) + 0.01 * rng.standard_normal((Ny, Nx))
    P = np.abs(P) + 0.05
    # face field from a periodic corner potential (exactly divergence free)
    Xc = np.arange(Nx + 1) / Nx
    Yc = np.arange(Ny + 1) / Ny
    XX, YY = np.meshgrid(Xc, Yc)
    Az = 0.15 * YY + 0.05 * np.sin(2 * np.pi * XX) * np.sin(2 * np.pi * YY)
    Bx_f = (Az[1:, :] - Az[:-1, :]) * Ny
    By_f = -(Az[:, 1:] - Az[:, :-1]) * Nx
    Bz = smooth(0.1)
    Bx = 0.5 * (Bx_f[:, :-1] + Bx_f[:, 1:])
    By = 0.5 * (By_f[:-1, :] + By_f[1:, :])
    E = P / (p.gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * (Bx * Bx + By * By + Bz * Bz)
    return {"rho": rho, "mx": rho * u, "my": rho * v, "mz": rho * w, "E": E, "Bz": Bz, "Bx_f": Bx_f,
            "By_f": By_f}


def m3_oparams(p):
    """Oracle params from a core_collapse_amd.mhd3d.Params."""
    return OB.m3_params(gamma1=p.gamma1, gamma2=p.gamma2, K1=p.K1, rho_nuc=p.rho_nuc, gamma_th=p.gamma_th,
                        rho_bounce=p.rho_bounce, gamma_bounce=p.gamma_bounce, Kbounce=p.Kbounce, cfl=p.cfl,
                        x0=p.x0, x1=p.x1, y0=p.y0, y1=p.y1, z0=p.z0, z1=p.z1, Nx=p.Nx, Ny=p.Ny, Nz=p.Nz,
                        floor_rho=p.floor_rho, floor_P=p.floor_P, periodic_x=int(p.periodic_x),
                        periodic_y=int(p.periodic_y), periodic_z=int(p.periodic_z), eta=p.eta,
                        ideal_eos=int(p.eos == 1), gamma_ideal=p.gamma, hlld_guard=p.hlld_guard,
                        ct_comp=int(p.ct_compensated))


def m3_arrays(p, state):
    return OB.M3Arrays(p.Nx, p.Ny, p.Nz, {k: np.asarray(v).reshape(-1) for k, v in state.items()})


def m3_dict(p, arr):
    Nx, Ny, Nz = p.Nx, p.Ny, p.Nz
    sh = {"Bx_f": (Nz, Ny, Nx + 1), "By_f": (Nz, Ny + 1, Nx), "Bz_f": (Nz + 1, Ny, Nx)}
    return {k: arr.a[k].reshape(sh.get(k, (Nz, Ny, Nx))).copy() for k in arr.FIELDS}


def random_mhd3d_state(p, seed):
    """Smooth random 3D state with a divergence-free face field (from a corner vector potential)."""
    from tests import ic_restated as ic
    rng = np.random.default_rng(seed)
    Nx, Ny, Nz = p.Nx, p.Ny, p.Nz
    x = (np.arange(Nx) + 0.5) / Nx
    y = (np.arange(Ny) + 0.5) / Ny
    z = (np.arange(Nz) + 0.5) / Nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    def smooth(a):
        ph = rng.uniform(0, 2 * np.pi, 3)
        return a * (np.sin(2 * np.pi * X + ph[0]) * np.cos(2 * np.pi * Y + ph[1]) * np.cos(2 * np.pi * Z + ph[2]))
    rho = 1.0 + 0.4 * smooth(1.0) + 0.02 * rng.standard_normal((Nz, Ny, Nx))
    if p.eos == 0:
        rho = rho * 60.0   # straddle the hybrid EOS branches (rho_nuc = 80, rho_bounce = 120)
    u, v, w = smooth(0.5), smooth(0.5), smooth(0.5)
    P = (0.6 + 0.2 * smooth(1.0)) * (30.0 if p.eos == 0 else 1.0)
    Xc = np.arange(Nx + 1) / Nx
    Yc = np.arange(Ny + 1) / Ny
    Zc = np.arange(Nz + 1) / Nz
    ZZ, YY, XX = np.meshgrid(Zc, Yc, Xc, indexing="ij")
    Az = 0.05 * np.sin(2 * np.pi * XX) * np.sin(2 * np.pi * YY) * np.cos(2 * np.pi * ZZ)
    Bx_f = (Az[:Nz, 1:, :] - Az[:Nz, :-1, :]) * Ny + 0.15
    By_f = -(Az[:Nz, :, 1:] - Az[:Nz, :, :-1]) * Nx
    Bz_f = np.full((Nz + 1, Ny, Nx), 0.08)
    Bx = 0.5 * (Bx_f[:, :, :-1] + Bx_f[:, :, 1:])
    By = 0.5 * (By_f[:, :-1, :] + By_f[:, 1:, :])
    Bz = 0.5 * (Bz_f[:-1] + Bz_f[1:])
    rho, mx, my, mz, E = ic.mhd3d_prim_to_cons(p, rho, u, v, w, P, Bx, By, Bz)
    return {"rho": rho, "mx": mx, "my": my, "mz": mz, "E": E, "Bx_f": Bx_f, "By_f": By_f, "Bz_f": Bz_f}
