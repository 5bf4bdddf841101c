"""Independent restatements of the initial conditions (test infrastructure).

numpy / pure-Python versions of the ICs that core_collapse_amd.ic builds in C++ (libccamd's
cc_*_ic_* entry points), kept here as checkers: the EH and NS restatements reproduce libstdc++'s
distributions draw by draw (StdRandom) and match the reference harness bit for bit; the MHD ones
use numpy's transcendental functions and agree with the C++ ICs to rounding.
"""
import numpy as np

from core_collapse_amd.mhd2d import Params as Mhd2dParams

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


# ----------------------------------------------------------------------------- 3D
def mhd3d_reference_ic(p=None, rho_bg=1.0, rho_core=20.0, r_core=0.12, r_edge=0.02, P0=0.15, vin=1.25,
                       rot_omega=1.0, B0x=0.15, B0z=0.05, Bshear=0.05):
    """The default IC of mhd3d_resistive_hlld_ct.cpp (M3:1159-1272): Az on corners -> Bx, By
    (M3:1162-1219), uniform Bz (M3:1221-1228), tanh core with infall and rotation (M3:1233-1272),
    prim_to_cons through the hybrid EOS and floors (done by the oracle's cons path: here via
    total_energy of the EOS in numpy)."""
    from core_collapse_amd.mhd3d import Params as P3
    p = p or P3()
    Nx, Ny, Nz = p.Nx, p.Ny, p.Nz
    dx, dy, dz = p.dx, p.dy, p.dz
    cx, cy, cz = 0.5 * (p.x0 + p.x1), 0.5 * (p.y0 + p.y1), 0.5 * (p.z0 + p.z1)
    Xc = p.x0 + np.arange(Nx + 1) * dx
    Yc = p.y0 + np.arange(Ny + 1) * dy
    Zc = p.z0 + np.arange(Nz + 1) * dz
    ZZ, YY, XX = np.meshgrid(Zc, Yc, Xc, indexing="ij")
    r = np.sqrt((XX - cx) ** 2 + (YY - cy) ** 2 + (ZZ - cz) ** 2) + 1e-14
    core = 0.5 * (1.0 - np.tanh((r - r_core) / max(r_edge, 1e-12)))
    Az = Bshear * core * np.sin(2.0 * np.pi * XX) * np.sin(2.0 * np.pi * YY)   # (Nz+1, Ny+1, Nx+1)
    Bx_f = (Az[:Nz, 1:, :] - Az[:Nz, :-1, :]) / dy + B0x                        # (Nz, Ny, Nx+1)
    By_f = -(Az[:Nz, :, 1:] - Az[:Nz, :, :-1]) / dx                             # (Nz, Ny+1, Nx)
    Bz_f = np.full((Nz + 1, Ny, Nx), B0z)
    x = p.x0 + (np.arange(Nx) + 0.5) * dx
    y = p.y0 + (np.arange(Ny) + 0.5) * dy
    z = p.z0 + (np.arange(Nz) + 0.5) * dz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    rx, ry, rz = X - cx, Y - cy, Z - cz
    r = np.sqrt(rx * rx + ry * ry + rz * rz) + 1e-14
    core = 0.5 * (1.0 - np.tanh((r - r_core) / max(r_edge, 1e-12)))
    rho = rho_bg + (rho_core - rho_bg) * core
    P = P0 * np.power(rho / rho_bg, 5.0 / 3.0)   # gamma_init (M3:53)
    vinc = vin * core
    u = -vinc * rx / r
    v = -vinc * ry / r
    w = -vinc * rz / r
    if rot_omega != 0.0:
        u = u + (-rot_omega * (Y - cy))
        v = v + rot_omega * (X - cx)
    Bx = 0.5 * (Bx_f[:, :, :-1] + Bx_f[:, :, 1:])
    By = 0.5 * (By_f[:, :-1, :] + By_f[:, 1:, :])
    Bz = 0.5 * (Bz_f[:-1] + Bz_f[1:])
    return {"prim": (rho, u, v, w, P, Bx, By, Bz), "Bx_f": Bx_f, "By_f": By_f, "Bz_f": Bz_f}


def turbulence3d_ic(p, seed=42, kmin=1, kmax=8, mach=1.0, beta=1.0):
    """rho = 1, p = 1/gamma (sound speed 1), uniform Bz with plasma beta `beta` (B0 = sqrt(2 p /
    beta)); velocity: random-phase Fourier modes on integer wavevectors 1 <= |k| <= kmax with
    |v(k)| ~ k^-2 (so E(k) ~ k^-2 in 3D shells), projected solenoidal, inverse FFT, scaled to rms
    Mach `mach`.  Amplitudes and phases: std::mt19937(seed) uniform draws in numpy's FFT order."""
    Nx, Ny, Nz = p.Nx, p.Ny, p.Nz
    g = p.gamma
    P0 = 1.0 / g
    B0 = np.sqrt(2.0 * P0 / beta)
    kx = np.fft.fftfreq(Nx, 1.0 / Nx)
    ky = np.fft.fftfreq(Ny, 1.0 / Ny)
    kz = np.fft.fftfreq(Nz, 1.0 / Nz)
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    k2 = KX * KX + KY * KY + KZ * KZ
    kk = np.sqrt(k2)
    mask = (kk >= kmin) & (kk <= kmax)
    nm = int(mask.sum())
    raw = mt19937_raw(seed, 6 * nm).astype(np.float64) / 4294967296.0
    vh = []
    for c in range(3):
        amp = raw[2 * c * nm:(2 * c + 1) * nm]
        ph = raw[(2 * c + 1) * nm:(2 * c + 2) * nm]
        a = np.zeros(k2.shape, dtype=np.complex128)
        a[mask] = (0.5 + amp) * np.power(kk[mask], -2.0) * np.exp(2j * np.pi * ph)
        vh.append(a)
    # solenoidal projection v_hat -= k (k . v_hat) / k^2
    with np.errstate(invalid="ignore", divide="ignore"):
        kdv = (KX * vh[0] + KY * vh[1] + KZ * vh[2]) / np.where(k2 > 0, k2, 1.0)
    vh = [vh[0] - KX * kdv, vh[1] - KY * kdv, vh[2] - KZ * kdv]
    vel = [np.real(np.fft.ifftn(a)) for a in vh]
    rms = np.sqrt(np.mean(vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2))
    scale = mach * 1.0 / rms   # sound speed 1
    u, v, w = (c * scale for c in vel)
    rho = np.ones((Nz, Ny, Nx))
    mx, my, mz = rho * u, rho * v, rho * w
    E = P0 / (g - 1.0) + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * B0 * B0
    return {"rho": rho, "mx": mx, "my": my, "mz": mz, "E": E, "Bx_f": np.zeros((Nz, Ny, Nx + 1)),
            "By_f": np.zeros((Nz, Ny + 1, Nx)), "Bz_f": np.full((Nz + 1, Ny, Nx), B0)}


def _hybrid_cold(p, rho):
    """eos_cold (M3:118-144), vectorised."""
    rho = np.maximum(rho, p.floor_rho)
    K2 = p.K1 * p.rho_nuc ** (p.gamma1 - p.gamma2)
    lo = rho < p.rho_nuc
    Pc = np.where(lo, p.K1 * np.power(rho, p.gamma1), K2 * np.power(rho, p.gamma2))
    ec1_n = p.K1 * p.rho_nuc ** (p.gamma1 - 1.0) / (p.gamma1 - 1.0)
    ec2_n = K2 * p.rho_nuc ** (p.gamma2 - 1.0) / (p.gamma2 - 1.0)
    ec = np.where(lo, p.K1 * np.power(rho, p.gamma1 - 1.0) / (p.gamma1 - 1.0),
                  ec1_n + (K2 * np.power(rho, p.gamma2 - 1.0) / (p.gamma2 - 1.0) - ec2_n))
    b = (~lo) & (rho >= p.rho_bounce)
    pb = p.Kbounce * (np.power(rho, p.gamma_bounce) - p.rho_bounce ** p.gamma_bounce)
    eb = p.Kbounce * (np.power(rho, p.gamma_bounce - 1.0) - p.rho_bounce ** (p.gamma_bounce - 1.0)) / \
        (p.gamma_bounce - 1.0)
    Pc = np.where(b, Pc + pb, Pc)
    ec = np.where(b, ec + eb, ec)
    return Pc, ec


def mhd3d_prim_to_cons(p, rho, u, v, w, P, Bx, By, Bz):
    """prim_to_cons (M3:199-204) for the hybrid (eos=0) or ideal (eos=1) EOS, without floors."""
    rho = np.maximum(rho, p.floor_rho)
    kin = 0.5 * rho * (u * u + v * v + w * w)
    mag = 0.5 * (Bx * Bx + By * By + Bz * Bz)
    if p.eos == 1:
        E = kin + mag + P / (p.gamma - 1.0)
    else:
        Pc, ec = _hybrid_cold(p, rho)
        Pth = np.maximum(0.0, P - Pc)
        eth = Pth / ((p.gamma_th - 1.0) * rho) if p.gamma_th > 1.0 else 0.0 * Pth
        E = kin + mag + rho * (ec + eth)
    return rho, rho * u, rho * v, rho * w, E


def mhd3d_reference_state(p=None, **kw):
    """Conserved state of the 3D reference default IC (M3:1233-1272)."""
    from core_collapse_amd.mhd3d import Params as P3
    p = p or P3()
    d = mhd3d_reference_ic(p, **kw)
    rho, mx, my, mz, E = mhd3d_prim_to_cons(p, *d["prim"])
    return {"rho": rho, "mx": mx, "my": my, "mz": mz, "E": E, "Bx_f": d["Bx_f"], "By_f": d["By_f"],
            "Bz_f": d["Bz_f"]}


# ------------------------------------------------------------------ libstdc++ distributions
class StdRandom:
    """std::mt19937(seed) with libstdc++'s (GCC 11) distributions, as the EH and NS ICs draw them.

    canonical(): generate_canonical<double, 53> -- two 32-bit draws, (r1 + r2 * 2^32) / 2^64 in
    double arithmetic, clamped below 1.  uniform(a, b): canonical() * (b - a) + a.  normal():
    Marsaglia's polar method, x, y = 2 canonical() - 1 until 0 < x^2 + y^2 <= 1, mult =
    sqrt(-2 log(r2) / r2), returns y * mult and keeps x * mult for the next call.
    """

    def __init__(self, seed, chunk=1 << 16):
        self._bg = np.random.MT19937()
        self._bg._legacy_seeding(int(seed))
        self._buf = []
        self._chunk = chunk
        self._saved = None

    def _raw(self):
        if not self._buf:
            self._buf = [float(v) for v in self._bg.random_raw(self._chunk)][::-1]
        return self._buf.pop()

    def canonical(self):
        s = self._raw() + self._raw() * 4294967296.0
        r = s / 18446744073709551616.0
        return r if r < 1.0 else np.nextafter(1.0, 0.0)

    def uniform(self, a, b):
        return self.canonical() * (b - a) + a

    def normal(self):
        import math
        if self._saved is not None:
            v, self._saved = self._saved, None
            return v
        while True:
            x = 2.0 * self.canonical() - 1.0
            y = 2.0 * self.canonical() - 1.0
            r2 = x * x + y * y
            if not (r2 > 1.0 or r2 == 0.0):
                break
        mult = math.sqrt(-2 * math.log(r2) / r2)
        self._saved = x * mult
        return y * mult


def euler2d_reference_ic(p=None, rho0=1.0, P0=1.0, amp_rho=0.10, amp_P=0.10, vmax=0.05, sigma=0.08, bump_amp=0.5,
                         seed=42):
    """init_random_blob of euler2d_gravity_hllc.cpp (EH:574-642): uniform background with 10 %
    random density and pressure fluctuations, a central Gaussian bump, random velocities in
    [-vmax, vmax], floors.  Returns SoA float64 arrays of shape (Ny, Nx)."""
    import math
    from core_collapse_amd.euler2d import Params as EHParams
    p = p or EHParams()
    Nx, Ny, dx, dy = p.Nx, p.Ny, p.dx, p.dy
    rng = StdRandom(seed)
    xc, yc = 0.5 * (p.x0 + p.x1), 0.5 * (p.y0 + p.y1)
    out = {f: np.empty((Ny, Nx)) for f in ("rho", "mx", "my", "E")}
    g1 = p.gamma - 1.0
    for j in range(Ny):
        y = p.y0 + (j + 0.5) * dy
        for i in range(Nx):
            x = p.x0 + (i + 0.5) * dx
            fr = 2.0 * rng.uniform(0.0, 1.0) - 1.0
            fp = 2.0 * rng.uniform(0.0, 1.0) - 1.0
            rho = rho0 * (1.0 + amp_rho * fr)
            P = P0 * (1.0 + amp_P * fp)
            rx, ry = x - xc, y - yc
            r2 = rx * rx + ry * ry
            rho += bump_amp * rho0 * math.exp(-r2 / (2.0 * sigma * sigma))
            u = vmax * rng.uniform(-1.0, 1.0)
            v = vmax * rng.uniform(-1.0, 1.0)
            rho = p.floor_rho if rho < p.floor_rho else rho  # std::max (EH:628-629)
            P = p.floor_P if P < p.floor_P else P
            mx, my = rho * u, rho * v
            E = P / g1 + 0.5 * rho * (u * u + v * v)
            # enforce_floors (EH:62-72)
            rho = rho if not (rho < p.floor_rho) else p.floor_rho
            uu, vv = mx / rho, my / rho
            kin = 0.5 * rho * (uu * uu + vv * vv)
            if g1 * (E - kin) < p.floor_P:
                E = kin + p.floor_P / g1
            out["rho"][j, i], out["mx"][j, i], out["my"][j, i], out["E"][j, i] = rho, mx, my, E
    return out


def ns2d_reference_ic(p=None, seed=42, rho0=1.0, P0=0.05, A=1.0, sigma_frac=0.10, Arho=0.01, Mach0=0.05):
    """The collapse IC in main() of navier_stokes_2d_hllc_poisson.cpp (NS:638-688): Gaussian
    overdensity rho0 (1 + A exp(-r^2 / 2 sigma^2) + Arho N(0,1)), uniform P0, velocities
    Mach0 a0 N(0,1), drawn per cell in the order noise, u, v.  Returns the AoS state (Ny, Nx, 4)."""
    import math
    from core_collapse_amd.ns2d import Params as NSParams
    p = p or NSParams()
    Nx, Ny = p.Nx, p.Ny
    Lx, Ly = p.x1 - p.x0, p.y1 - p.y0
    dx, dy = Lx / Nx, Ly / Ny
    rng = StdRandom(seed)
    sigma = sigma_frac * min(Lx, Ly)
    a0 = math.sqrt(p.gamma * P0 / rho0)
    vamp = Mach0 * a0
    xc, yc = 0.5 * (p.x0 + p.x1), 0.5 * (p.y0 + p.y1)
    U = np.empty((Ny, Nx, 4))
    clamp = lambda v, lo: lo if v < lo else v  # noqa: E731  clamp_min, NS:51
    for j in range(Ny):
        y = p.y0 + (j + 0.5) * dy
        for i in range(Nx):
            x = p.x0 + (i + 0.5) * dx
            rx, ry = x - xc, y - yc
            r2 = rx * rx + ry * ry
            bump = A * math.exp(-0.5 * r2 / (sigma * sigma))
            noise = Arho * rng.normal()
            d = clamp(rho0 * (1.0 + bump + noise), p.floor_rho)
            P = clamp(P0, p.floor_P)
            u = vamp * rng.normal()
            v = vamp * rng.normal()
            # prim_to_cons NS:134-145
            d = clamp(d, p.floor_rho)
            P = clamp(P, p.floor_P)
            U[j, i] = (d, d * u, d * v, P / (p.gamma - 1.0) + 0.5 * d * (u * u + v * v))
    return U
