"""Initial conditions (host side): the reference programs' default ICs and the benchmark-config ICs.

Random draws use std::mt19937 semantics (numpy's MT19937 with legacy seeding produces the same
32-bit stream as std::mt19937(seed)) and a documented Box-Muller transform, so the C++ drop-in
programs and the Python tests build identical inputs.
"""
import numpy as np

from .mhd2d import Params as Mhd2dParams


def mt19937_raw(seed, n):
    """The first n outputs of std::mt19937(seed) as uint64."""
    bg = np.random.MT19937()
    bg._legacy_seeding(int(seed))
    return bg.random_raw(int(n)).astype(np.uint64)


def box_muller_pairs(seed, npairs):
    """npairs standard-normal pairs (z0, z1) from std::mt19937(seed):
    u1 = (r1 + 0.5) / 2^32, u2 = (r2 + 0.5) / 2^32, rad = sqrt(-2 ln u1),
    z0 = rad cos(2 pi u2), z1 = rad sin(2 pi u2)."""
    r = mt19937_raw(seed, 2 * npairs).astype(np.float64)
    u1 = (r[0::2] + 0.5) / 4294967296.0
    u2 = (r[1::2] + 0.5) / 4294967296.0
    rad = np.sqrt(-2.0 * np.log(u1))
    return rad * np.cos(2.0 * np.pi * u2), rad * np.sin(2.0 * np.pi * u2)


def _floors(p, rho, mx, my, mz, E, Bx, By, Bz):
    """enforce_floors (M2:114-124), vectorised."""
    rho = np.maximum(rho, p.floor_rho)
    u, v, w = mx / rho, my / rho, mz / rho
    kin = 0.5 * rho * (u * u + v * v + w * w)
    mag = 0.5 * (Bx * Bx + By * By + Bz * Bz)
    P = (p.gamma - 1.0) * (E - kin - mag)
    E = np.where(P < p.floor_P, kin + mag + p.floor_P / (p.gamma - 1.0), E)
    return rho, E


def _prim_to_cons(p, rho, u, v, w, P, Bx, By, Bz):
    """prim_to_cons (M2:137-146) followed by enforce_floors, as the IC loop of M2:1044-1046 does."""
    rho = np.maximum(rho, p.floor_rho)
    mx, my, mz = rho * u, rho * v, rho * w
    kin = 0.5 * rho * (u * u + v * v + w * w)
    mag = 0.5 * (Bx * Bx + By * By + Bz * Bz)
    E = kin + mag + P / (p.gamma - 1.0)
    rho, E = _floors(p, rho, mx, my, mz, E, Bx, By, Bz)
    return rho, mx, my, mz, E


def mhd2d_reference_ic(p: Mhd2dParams = None, rho_bg=1.0, rho_core=20.0, r_core=0.12, r_edge=0.02, P0=0.15,
                       vin=1.25, rot_omega=1.0, B0x=0.15, Bshear=0.05, gamma_eff=1.05):
    """The default IC of mhd2d_resistive_hlld_ct_gravity.cpp (M2:954-1048): face B from a corner
    vector potential Az (M2:957-1006), tanh-core overdensity with infall and rotation (M2:1011-1048)."""
    p = p or Mhd2dParams()
    Nx, Ny, dx, dy = p.Nx, p.Ny, p.dx, p.dy
    cx, cy = 0.5 * (p.x0 + p.x1), 0.5 * (p.y0 + p.y1)
    X = p.x0 + np.arange(Nx + 1) * dx
    Y = p.y0 + np.arange(Ny + 1) * dy
    XX, YY = np.meshgrid(X, Y)
    rx, ry = XX - cx, YY - cy
    r = np.sqrt(rx * rx + ry * ry) + 1e-14
    s = (r - r_core) / max(r_edge, 1e-12)
    core = 0.5 * (1.0 - np.tanh(s))
    Lx = p.x1 - p.x0
    Az = B0x * (YY - p.y0) + (Bshear * r_core) * core * np.sin(2.0 * np.pi * (XX - p.x0) / Lx)
    Bx_f = (Az[1:, :] - Az[:-1, :]) / dy          # (Ny, Nx+1)
    By_f = -(Az[:, 1:] - Az[:, :-1]) / dx         # (Ny+1, Nx)
    x = p.x0 + (np.arange(Nx) + 0.5) * dx
    y = p.y0 + (np.arange(Ny) + 0.5) * dy
    XC, YC = np.meshgrid(x, y)
    rx, ry = XC - cx, YC - cy
    r = np.sqrt(rx * rx + ry * ry) + 1e-14
    s = (r - r_core) / max(r_edge, 1e-12)
    core = 0.5 * (1.0 - np.tanh(s))
    rho = rho_bg + (rho_core - rho_bg) * core
    P = P0 * np.power(rho / max(rho_bg, 1e-12), gamma_eff)
    vr = -vin * np.minimum(r / r_core, 1.0) * core
    vphi = rot_omega * r * core
    erx, ery = rx / r, ry / r
    u = vr * erx + vphi * (-ery)
    v = vr * ery + vphi * erx
    w = np.zeros_like(u)
    Bx = 0.5 * (Bx_f[:, :-1] + Bx_f[:, 1:])
    By = 0.5 * (By_f[:-1, :] + By_f[1:, :])
    Bz = np.zeros_like(u)
    rho, mx, my, mz, E = _prim_to_cons(p, rho, u, v, w, P, Bx, By, Bz)
    return {"rho": rho, "mx": mx, "my": my, "mz": mz, "E": E, "Bz": Bz, "Bx_f": Bx_f, "By_f": By_f}


def collapse2d_params(N=128, precision=0, **kw):
    """Parameters of the 2D MHD Gaussian-collapse benchmark: box [-0.5,0.5]^2, clamped ("outflow")
    boundaries, gamma 5/3, G 1, eta 1e-4, CFL 0.4, SOR 600 sweeps at omega 1.8."""
    d = dict(gamma=5.0 / 3.0, cfl=0.4, x0=-0.5, x1=0.5, y0=-0.5, y1=0.5, Nx=N, Ny=N, periodic_x=False,
             periodic_y=False, eta=1e-4, gravity_on=True, G=1.0, sor_iters=600, sor_omega=1.8, precision=precision)
    d.update(kw)
    return Mhd2dParams(**d)


def collapse2d_ic(p: Mhd2dParams, seed=42, contrast=10.0, width=0.1, P0=0.1, vamp=1e-3, B0=0.05):
    """Gaussian collapse: rho = 1 + (contrast-1) exp(-r^2/width^2), p = P0, (u, v) = vamp * (z0, z1)
    with one Box-Muller pair per cell in row-major order from std::mt19937(seed), w = 0,
    Bx_f = B0, By_f = 0, Bz = 0; conserved variables by prim_to_cons + floors (M2:1044-1046)."""
    Nx, Ny = p.Nx, p.Ny
    x = p.x0 + (np.arange(Nx) + 0.5) * p.dx
    y = p.y0 + (np.arange(Ny) + 0.5) * p.dy
    XC, YC = np.meshgrid(x, y)
    r2 = XC * XC + YC * YC
    rho = 1.0 + (contrast - 1.0) * np.exp(-r2 / (width * width))
    z0, z1 = box_muller_pairs(seed, Nx * Ny)
    u = vamp * z0.reshape(Ny, Nx)
    v = vamp * z1.reshape(Ny, Nx)
    w = np.zeros_like(u)
    P = np.full_like(u, P0)
    Bx_f = np.full((Ny, Nx + 1), B0)
    By_f = np.zeros((Ny + 1, Nx))
    Bx = np.full_like(u, B0)
    By = np.zeros_like(u)
    Bz = np.zeros_like(u)
    rho, mx, my, mz, E = _prim_to_cons(p, rho, u, v, w, P, Bx, By, Bz)
    return {"rho": rho, "mx": mx, "my": my, "mz": mz, "E": E, "Bz": Bz, "Bx_f": Bx_f, "By_f": By_f}
