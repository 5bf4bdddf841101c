// ic.cpp -- initial conditions on the host: the reference programs' default ICs (built inline in
// their main() functions) and the benchmark-problem ICs.  Compiled without FP contraction, with the
// same libm calls (tanh, pow, sin, exp, sqrt) and libstdc++ <random> distributions as the
// reference, so the reference ICs are reproduced bit for bit.
//
// Files cited: M2 = mhd2d_resistive_hlld_ct_gravity.cpp, M3 = mhd3d_resistive_hlld_ct.cpp,
//              EH = euler2d_gravity_hllc.cpp, NS = navier_stokes_2d_hllc_poisson.cpp.
#include <algorithm>
#include <cmath>
#include <complex>
#include <random>
#include <thread>
#include <vector>

#include "common.hpp"
#include "riemann.hpp"

namespace {

using cc::set_error;

constexpr double kPi = 3.14159265358979323846;  // M_PI

cc_status bad(const char* msg) {
    set_error(CC_ERR_INVALID_ARG, msg);
    return CC_ERR_INVALID_ARG;
}

// ------------------------------------------------------------------ M2 helpers (ideal gas)
// prim_to_cons (M2:137-146) followed by enforce_floors (M2:113-123), as the IC loop M2:1044-1045.
void m2_cons(const cc_mhd2d_params* p, double rho, double u, double v, double w, double P, double Bx, double By,
             double Bz, double* out5) {
    const double gm1 = p->gamma - 1.0;
    double r = std::max(rho, p->floor_rho);
    double mx = r * u, my = r * v, mz = r * w;
    double kin = 0.5 * r * (u * u + v * v + w * w);
    double mag = 0.5 * (Bx * Bx + By * By + Bz * Bz);
    double E = kin + mag + P / gm1;
    r = std::max(r, p->floor_rho);
    const double uu = mx / r, vv = my / r, ww = mz / r;
    kin = 0.5 * r * (uu * uu + vv * vv + ww * ww);
    mag = 0.5 * (Bx * Bx + By * By + Bz * Bz);
    const double Pf = gm1 * (E - kin - mag);
    if (Pf < p->floor_P) E = kin + mag + p->floor_P / gm1;
    out5[0] = r; out5[1] = mx; out5[2] = my; out5[3] = mz; out5[4] = E;
}

bool check2d(const cc_mhd2d_params* p) { return p && p->Nx > 0 && p->Ny > 0 && p->gamma > 1.0; }

// ------------------------------------------------------------------ M3 helpers
struct M3Eos {
    const cc_mhd3d_params* p;
    cc::HybridEOS<double> h;
    bool ideal() const { return p->eos == 1; }
    // total_energy_from_prim (M3:154-163) or the ideal-gas form.
    double total_energy(const cc::Prim<double>& W) const {
        if (ideal()) {
            const double rho = std::max(W.rho, p->floor_rho);
            const double kin = 0.5 * rho * (W.u * W.u + W.v * W.v + W.w * W.w);
            const double mag = 0.5 * (W.Bx * W.Bx + W.By * W.By + W.Bz * W.Bz);
            return kin + mag + W.P / (p->gamma - 1.0);
        }
        return h.total_energy(W, h.aux(W));
    }
    // enforce_floors (M3:165-183; M2:113-123 for the ideal gas).
    void floors(double& rho, double& mx, double& my, double& mz, double& E, double Bx, double By, double Bz) const {
        rho = std::max(rho, p->floor_rho);
        const double u = mx / rho, v = my / rho, w = mz / rho;
        const double kin = 0.5 * rho * (u * u + v * v + w * w);
        const double mag = 0.5 * (Bx * Bx + By * By + Bz * Bz);
        if (ideal()) {
            const double P = (p->gamma - 1.0) * (E - kin - mag);
            if (P < p->floor_P) E = kin + mag + p->floor_P / (p->gamma - 1.0);
            return;
        }
        double eint_tot = E - kin - mag;
        if (eint_tot < 0.0) eint_tot = 0.0;
        const double e_spec = eint_tot / rho;
        const auto c0 = h.cold(rho);  // pressure_from_rho_espec, M3:146-152
        const double eth0 = std::max(0.0, e_spec - c0.ec);
        const double Pth0 = (p->gamma_th - 1.0) * rho * eth0;
        double P = std::max(c0.Pc + Pth0, p->floor_P);
        if (P < p->floor_P) P = p->floor_P;
        const auto c = h.cold(rho);
        const double Pth = std::max(0.0, P - c.Pc);
        const double eth = (p->gamma_th > 1.0) ? (Pth / ((p->gamma_th - 1.0) * rho)) : 0.0;
        E = kin + mag + rho * (c.ec + eth);
    }
};

M3Eos m3_eos(const cc_mhd3d_params* p) {
    M3Eos e{p, {}};
    cc::HybridEOS<double>& h = e.h;
    const double K2 = p->K1 * std::pow(p->rho_nuc, p->gamma1 - p->gamma2);  // M3:120
    h.gamma1 = p->gamma1; h.gamma2 = p->gamma2; h.K1 = p->K1; h.K2 = K2; h.rho_nuc = p->rho_nuc;
    h.gamma_th = p->gamma_th; h.rho_bounce = p->rho_bounce; h.gamma_bounce = p->gamma_bounce; h.Kbounce = p->Kbounce;
    h.ec1_n = p->K1 * std::pow(p->rho_nuc, p->gamma1 - 1.0) / (p->gamma1 - 1.0);
    h.ec2_n = K2 * std::pow(p->rho_nuc, p->gamma2 - 1.0) / (p->gamma2 - 1.0);
    h.rb_g = std::pow(p->rho_bounce, p->gamma_bounce);
    h.rb_g1 = std::pow(p->rho_bounce, p->gamma_bounce - 1.0);
    h.floor_rho = p->floor_rho;
    h.floor_P = p->floor_P;
    return e;
}

bool check3d(const cc_mhd3d_params* p) { return p && p->Nx > 0 && p->Ny > 0 && p->Nz > 0; }

// ------------------------------------------------------------------ host FFT (turbulence IC)
bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// In-place radix-2 inverse DFT (sign +, unscaled) of `n` points at stride `st`, on `cnt` lines whose
// starts are given by line(l).  Twiddles are evaluated directly from cos / sin.
template <typename LineStart>
void ifft_lines(std::complex<double>* a, int n, size_t st, size_t cnt, LineStart line) {
    std::vector<std::complex<double>> tw((size_t)n / 2 + 1);
    for (int k = 0; k <= n / 2; k++) tw[k] = std::polar(1.0, 2.0 * kPi * k / n);
    int lg = 0;
    while ((1 << lg) < n) lg++;
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto work = [&](size_t l0, size_t l1) {
        std::vector<std::complex<double>> buf(n);
        for (size_t l = l0; l < l1; l++) {
            std::complex<double>* b = a + line(l);
            for (int i = 0; i < n; i++) {
                unsigned r = 0;
                for (int s = 0; s < lg; s++) r |= ((unsigned)(i >> s) & 1u) << (lg - 1 - s);
                buf[r] = b[(size_t)i * st];
            }
            for (int len = 2; len <= n; len <<= 1) {
                const int step = n / len;
                for (int s = 0; s < n; s += len)
                    for (int k = 0; k < len / 2; k++) {
                        const std::complex<double> t = tw[(size_t)k * step] * buf[s + k + len / 2];
                        const std::complex<double> u = buf[s + k];
                        buf[s + k] = u + t;
                        buf[s + k + len / 2] = u - t;
                    }
            }
            for (int i = 0; i < n; i++) b[(size_t)i * st] = buf[i];
        }
    };
    std::vector<std::thread> th;
    const size_t per = (cnt + nt - 1) / nt;
    for (unsigned t = 0; t < nt; t++) {
        const size_t l0 = t * per, l1 = std::min(cnt, l0 + per);
        if (l0 < l1) th.emplace_back(work, l0, l1);
    }
    for (auto& x : th) x.join();
}

// Integer wavenumber of FFT index i on n points (numpy.fft.fftfreq(n, 1/n) order).
inline int kfreq(int i, int n) { return (i < (n + 1) / 2) ? i : i - n; }

}  // namespace

extern "C" {

// ================================================================== M2
void cc_mhd2d_ic_defaults(cc_mhd2d_ic* k) {
    if (!k) return;
    *k = cc_mhd2d_ic{1.0, 20.0, 0.12, 0.02, 0.15, 1.25, 1.0, 0.15, 0.05, 1.05};  // M2:74-85, M2:1011
}

cc_status cc_mhd2d_ic_reference(const cc_mhd2d_params* p, const cc_mhd2d_ic* kn, double* rho, double* mx, double* my,
                                double* mz, double* E, double* Bz, double* Bx_f, double* By_f) {
    if (!check2d(p) || !kn || !rho || !mx || !my || !mz || !E || !Bz || !Bx_f || !By_f)
        return bad("cc_mhd2d_ic_reference: bad parameters or null output");
    const int Nx = p->Nx, Ny = p->Ny;
    const double dx = (p->x1 - p->x0) / Nx, dy = (p->y1 - p->y0) / Ny;
    const double cx = 0.5 * (p->x0 + p->x1), cy = 0.5 * (p->y0 + p->y1);
    const double redge = std::max(kn->r_edge, 1e-12);
    // Az on the (Nx+1) x (Ny+1) corners (M2:957-989); the periodic wrap modulo N+1 of M2:967-968
    // maps every corner index 0..N to itself.
    std::vector<double> Az((size_t)(Nx + 1) * (Ny + 1));
    for (int jc = 0; jc <= Ny; ++jc)
        for (int ic = 0; ic <= Nx; ++ic) {
            const double X = p->x0 + ic * dx, Y = p->y0 + jc * dy;
            const double rx = X - cx, ry = Y - cy;
            const double r = std::sqrt(rx * rx + ry * ry) + 1e-14;
            const double s = (r - kn->r_core) / redge;
            const double core = 0.5 * (1.0 - std::tanh(s));
            const double Az_base = kn->B0x * (Y - p->y0);
            const double Lx = p->x1 - p->x0;
            const double Az_pert = (kn->Bshear * kn->r_core) * core * std::sin(2.0 * kPi * (X - p->x0) / Lx);
            Az[(size_t)jc * (Nx + 1) + ic] = Az_base + Az_pert;
        }
    for (int j = 0; j < Ny; ++j)  // M2:992-998
        for (int i = 0; i <= Nx; ++i)
            Bx_f[(size_t)j * (Nx + 1) + i] = (Az[(size_t)(j + 1) * (Nx + 1) + i] - Az[(size_t)j * (Nx + 1) + i]) / dy;
    for (int j = 0; j <= Ny; ++j)  // M2:999-1005
        for (int i = 0; i < Nx; ++i)
            By_f[(size_t)j * Nx + i] = -(Az[(size_t)j * (Nx + 1) + i + 1] - Az[(size_t)j * (Nx + 1) + i]) / dx;
    for (int j = 0; j < Ny; j++)  // M2:1011-1048
        for (int i = 0; i < Nx; i++) {
            const double X = p->x0 + (i + 0.5) * dx, Y = p->y0 + (j + 0.5) * dy;
            const double rx = X - cx, ry = Y - cy;
            const double r = std::sqrt(rx * rx + ry * ry) + 1e-14;
            const double s = (r - kn->r_core) / redge;
            const double core = 0.5 * (1.0 - std::tanh(s));
            const double d = kn->rho_bg + (kn->rho_core - kn->rho_bg) * core;
            const double P = kn->P0 * std::pow(d / std::max(kn->rho_bg, 1e-12), kn->gamma_eff);
            const double vr = -kn->vin * (std::min(r / kn->r_core, 1.0)) * core;
            const double vphi = kn->rot_omega * r * core;
            const double erx = rx / r, ery = ry / r;
            const double epx = -ery, epy = erx;
            const double u = vr * erx + vphi * epx;
            const double v = vr * ery + vphi * epy;
            // cell-centred B from faces (M2:262-286): faces i, i+1 and j, j+1 lie in range
            const double bx = 0.5 * (Bx_f[(size_t)j * (Nx + 1) + i] + Bx_f[(size_t)j * (Nx + 1) + i + 1]);
            const double by = 0.5 * (By_f[(size_t)j * Nx + i] + By_f[(size_t)(j + 1) * Nx + i]);
            double c[5];
            m2_cons(p, d, u, v, 0.0, P, bx, by, 0.0, c);
            const size_t q = (size_t)j * Nx + i;
            rho[q] = c[0]; mx[q] = c[1]; my[q] = c[2]; mz[q] = c[3]; E[q] = c[4]; Bz[q] = 0.0;
        }
    return CC_OK;
}

void cc_collapse2d_ic_defaults(cc_collapse2d_ic* k) {
    if (!k) return;
    *k = cc_collapse2d_ic{42u, 10.0, 0.1, 0.1, 1e-3, 0.05};
}

cc_status cc_mhd2d_ic_collapse(const cc_mhd2d_params* p, const cc_collapse2d_ic* kn, double* rho, double* mx, double* my,
                               double* mz, double* E, double* Bz, double* Bx_f, double* By_f) {
    if (!check2d(p) || !kn || !rho || !mx || !my || !mz || !E || !Bz || !Bx_f || !By_f)
        return bad("cc_mhd2d_ic_collapse: bad parameters or null output");
    const int Nx = p->Nx, Ny = p->Ny;
    const double dx = (p->x1 - p->x0) / Nx, dy = (p->y1 - p->y0) / Ny;
    const double cx = 0.5 * (p->x0 + p->x1), cy = 0.5 * (p->y0 + p->y1);
    std::mt19937 rng(kn->seed);
    for (size_t f = 0; f < (size_t)(Nx + 1) * Ny; f++) Bx_f[f] = kn->B0;
    for (size_t f = 0; f < (size_t)Nx * (Ny + 1); f++) By_f[f] = 0.0;
    for (int j = 0; j < Ny; j++)
        for (int i = 0; i < Nx; i++) {
            const double X = p->x0 + (i + 0.5) * dx - cx, Y = p->y0 + (j + 0.5) * dy - cy;
            const double r2 = X * X + Y * Y;
            const double d = 1.0 + (kn->contrast - 1.0) * std::exp(-r2 / (kn->width * kn->width));
            const double u1 = ((double)rng() + 0.5) / 4294967296.0;
            const double u2 = ((double)rng() + 0.5) / 4294967296.0;
            const double rad = std::sqrt(-2.0 * std::log(u1));
            const double u = kn->vamp * (rad * std::cos(2.0 * kPi * u2));
            const double v = kn->vamp * (rad * std::sin(2.0 * kPi * u2));
            double c[5];
            m2_cons(p, d, u, v, 0.0, kn->P0, kn->B0, 0.0, 0.0, c);
            const size_t q = (size_t)j * Nx + i;
            rho[q] = c[0]; mx[q] = c[1]; my[q] = c[2]; mz[q] = c[3]; E[q] = c[4]; Bz[q] = 0.0;
        }
    return CC_OK;
}

cc_status cc_collapse2d_batch_sample(unsigned seed, int n, int nmin, int nmax, int* N, cc_collapse2d_ic* ic,
                                     double* eta) {
    if (n < 0 || nmin < 1 || nmax < nmin || !N || !ic || !eta) return bad("cc_collapse2d_batch_sample: bad arguments");
    std::mt19937 rng(seed);
    auto U = [&]() { return ((double)rng() + 0.5) / 4294967296.0; };
    for (int i = 0; i < n; i++) {
        const double uN = U(), uc = U(), ue = U(), uB = U();
        N[i] = (int)std::lround(nmin * std::pow((double)nmax / nmin, uN));
        cc_collapse2d_ic_defaults(&ic[i]);
        ic[i].seed = 42u + (unsigned)i;
        ic[i].contrast = 5.0 + 45.0 * uc;
        ic[i].B0 = 0.01 + 0.09 * uB;
        eta[i] = 1e-5 + (1e-3 - 1e-5) * ue;
    }
    return CC_OK;
}

// ================================================================== 3D Euler uniform sphere
void cc_sphere3d_ic_defaults(cc_sphere3d_ic* k) {
    if (!k) return;
    *k = cc_sphere3d_ic{42u, 10.0, 0.1, 0.2, 0.01, 0.01};
}

cc_status cc_euler3d_ic_sphere(const cc_euler3d_params* p, const cc_sphere3d_ic* kn, double* rho, double* mx,
                               double* my, double* mz, double* E) {
    if (!p || !kn || !rho || !mx || !my || !mz || !E || p->Nx < 1 || p->Ny < 1 || p->Nz < 1 || !(p->gamma > 1.0))
        return bad("cc_euler3d_ic_sphere: bad parameters or null output");
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
    const double dx = (p->x1 - p->x0) / Nx, dy = (p->y1 - p->y0) / Ny, dz = (p->z1 - p->z0) / Nz;
    const double cx = 0.5 * (p->x0 + p->x1), cy = 0.5 * (p->y0 + p->y1), cz = 0.5 * (p->z0 + p->z1);
    std::mt19937 rng(kn->seed);
    double z1 = 0.0;
    const size_t n = (size_t)Nx * Ny * Nz;
    for (size_t q = 0; q < n; q++) {
        const int i = (int)(q % Nx), j = (int)((q / Nx) % Ny), k = (int)(q / ((size_t)Nx * Ny));
        double z;
        if ((q & 1) == 0) {  // a new Box-Muller pair: z0 now, z1 for the next cell
            const double u1 = ((double)rng() + 0.5) / 4294967296.0;
            const double u2 = ((double)rng() + 0.5) / 4294967296.0;
            const double rad = std::sqrt(-2.0 * std::log(u1));
            z = rad * std::cos(2.0 * kPi * u2);
            z1 = rad * std::sin(2.0 * kPi * u2);
        } else {
            z = z1;
        }
        const double X = p->x0 + (i + 0.5) * dx - cx, Y = p->y0 + (j + 0.5) * dy - cy, Z = p->z0 + (k + 0.5) * dz - cz;
        const double r = std::sqrt(X * X + Y * Y + Z * Z);
        rho[q] = (r < kn->radius ? kn->rho_in : kn->rho_out) * (1.0 + kn->noise * z);
        mx[q] = 0.0; my[q] = 0.0; mz[q] = 0.0;
        E[q] = kn->P0 / (p->gamma - 1.0);
    }
    return CC_OK;
}

// ================================================================== M3
void cc_mhd3d_ic_defaults(cc_mhd3d_ic* k) {
    if (!k) return;
    *k = cc_mhd3d_ic{1.0, 20.0, 0.12, 0.02, 0.15, 1.25, 1.0, 0.15, 0.05, 0.05, 5.0 / 3.0};  // M3:53, 76-88
}

cc_status cc_mhd3d_ic_reference(const cc_mhd3d_params* p, const cc_mhd3d_ic* kn, double* rho, double* mx, double* my,
                                double* mz, double* E, double* Bx_f, double* By_f, double* Bz_f) {
    if (!check3d(p) || !kn || !rho || !mx || !my || !mz || !E || !Bx_f || !By_f || !Bz_f)
        return bad("cc_mhd3d_ic_reference: bad parameters or null output");
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
    const double dx = (p->x1 - p->x0) / Nx, dy = (p->y1 - p->y0) / Ny, dz = (p->z1 - p->z0) / Nz;
    const double cx = 0.5 * (p->x0 + p->x1), cy = 0.5 * (p->y0 + p->y1), cz = 0.5 * (p->z0 + p->z1);
    const double redge = std::max(kn->r_edge, 1e-12);
    const size_t NX1 = (size_t)Nx + 1, NY1 = (size_t)Ny + 1;
    // Az on corners (M3:1162-1200); corner indices 0..N map to themselves.
    std::vector<double> Az(NX1 * NY1 * ((size_t)Nz + 1));
    for (int kc = 0; kc <= Nz; ++kc)
        for (int jc = 0; jc <= Ny; ++jc)
            for (int ic = 0; ic <= Nx; ++ic) {
                const double X = p->x0 + ic * dx, Y = p->y0 + jc * dy, Z = p->z0 + kc * dz;
                const double rx = X - cx, ry = Y - cy, rz = Z - cz;
                const double r = std::sqrt(rx * rx + ry * ry + rz * rz) + 1e-14;
                const double s = (r - kn->r_core) / redge;
                const double core = 0.5 * (1.0 - std::tanh(s));
                Az[((size_t)kc * NY1 + jc) * NX1 + ic] =
                    kn->Bshear * core * std::sin(2.0 * kPi * X) * std::sin(2.0 * kPi * Y);
            }
    auto az = [&](int i, int j, int k) { return Az[((size_t)k * NY1 + j) * NX1 + i]; };
    for (int k = 0; k < Nz; k++)  // Bx = dAz/dy + B0x (M3:1203-1211)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i <= Nx; i++)
                Bx_f[((size_t)k * Ny + j) * NX1 + i] = (az(i, j + 1, k) - az(i, j, k)) / dy + kn->B0x;
    for (int k = 0; k < Nz; k++)  // By = -dAz/dx (M3:1213-1220)
        for (int j = 0; j <= Ny; j++)
            for (int i = 0; i < Nx; i++) By_f[((size_t)k * NY1 + j) * Nx + i] = -(az(i + 1, j, k) - az(i, j, k)) / dx;
    for (size_t f = 0; f < (size_t)Nx * Ny * ((size_t)Nz + 1); f++) Bz_f[f] = kn->B0z;  // M3:1222-1228
    const M3Eos eos = m3_eos(p);
    for (int k = 0; k < Nz; k++)  // M3:1233-1272
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                const double X = p->x0 + (i + 0.5) * dx, Y = p->y0 + (j + 0.5) * dy, Z = p->z0 + (k + 0.5) * dz;
                const double rx = X - cx, ry = Y - cy, rz = Z - cz;
                const double r = std::sqrt(rx * rx + ry * ry + rz * rz) + 1e-14;
                const double s = (r - kn->r_core) / redge;
                const double core = 0.5 * (1.0 - std::tanh(s));
                const double d = kn->rho_bg + (kn->rho_core - kn->rho_bg) * core;
                const double P = kn->P0 * std::pow(d / kn->rho_bg, kn->gamma_init);
                double u = 0.0, v = 0.0, w = 0.0;
                const double vin = kn->vin * core;
                u -= vin * rx / r;
                v -= vin * ry / r;
                w -= vin * rz / r;
                if (kn->rot_omega != 0.0) {
                    u += -kn->rot_omega * (Y - cy);
                    v += kn->rot_omega * (X - cx);
                }
                const double bx = 0.5 * (Bx_f[((size_t)k * Ny + j) * NX1 + i] + Bx_f[((size_t)k * Ny + j) * NX1 + i + 1]);
                const double by = 0.5 * (By_f[((size_t)k * NY1 + j) * Nx + i] + By_f[((size_t)k * NY1 + j + 1) * Nx + i]);
                const double bz = 0.5 * (Bz_f[((size_t)k * Ny + j) * Nx + i] + Bz_f[(((size_t)k + 1) * Ny + j) * Nx + i]);
                const cc::Prim<double> W{d, u, v, w, P, bx, by, bz};
                double r0 = std::max(W.rho, p->floor_rho);  // prim_to_cons, M3:199-204
                double m0 = r0 * W.u, m1 = r0 * W.v, m2 = r0 * W.w;
                double e0 = eos.total_energy(W);
                eos.floors(r0, m0, m1, m2, e0, bx, by, bz);
                const size_t q = ((size_t)k * Ny + j) * Nx + i;
                rho[q] = r0; mx[q] = m0; my[q] = m1; mz[q] = m2; E[q] = e0;
            }
    return CC_OK;
}

void cc_turbulence3d_ic_defaults(cc_turbulence3d_ic* k) {
    if (!k) return;
    *k = cc_turbulence3d_ic{42u, 1, 8, 1.0, 1.0};
}

cc_status cc_mhd3d_ic_turbulence(const cc_mhd3d_params* p, const cc_turbulence3d_ic* kn, double* rho, double* mx,
                                 double* my, double* mz, double* E, double* Bx_f, double* By_f, double* Bz_f) {
    if (!check3d(p) || !kn || !rho || !mx || !my || !mz || !E || !Bx_f || !By_f || !Bz_f)
        return bad("cc_mhd3d_ic_turbulence: bad parameters or null output");
    if (p->eos != 1 || !(p->gamma > 1.0)) return bad("cc_mhd3d_ic_turbulence: needs the ideal EOS (eos = 1, gamma > 1)");
    if (!pow2(p->Nx) || !pow2(p->Ny) || !pow2(p->Nz)) {
        set_error(CC_ERR_NON_POW2, "cc_mhd3d_ic_turbulence: grid sizes must be powers of two");
        return CC_ERR_NON_POW2;
    }
    if (!(kn->beta > 0.0) || !(kn->mach >= 0.0) || kn->kmax < kn->kmin)
        return bad("cc_mhd3d_ic_turbulence: bad beta / mach / k range");
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
    const size_t N = (size_t)Nx * Ny * Nz;
    const double g = p->gamma, P0 = 1.0 / g, B0 = std::sqrt(2.0 * P0 / kn->beta);
    // Modes in the band, in C order over (kz, ky, kx) FFT indices.
    std::vector<size_t> modes;
    for (int k = 0; k < Nz; k++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                const double KX = kfreq(i, Nx), KY = kfreq(j, Ny), KZ = kfreq(k, Nz);
                const double kk = std::sqrt(KX * KX + KY * KY + KZ * KZ);
                if (kk >= kn->kmin && kk <= kn->kmax) modes.push_back(((size_t)k * Ny + j) * Nx + i);
            }
    const size_t nm = modes.size();
    std::mt19937 rng(kn->seed);
    std::vector<double> raw(6 * nm);
    for (auto& r : raw) r = (double)rng() / 4294967296.0;
    std::vector<std::complex<double>> vh[3];
    for (int c = 0; c < 3; c++) {
        vh[c].assign(N, std::complex<double>(0.0, 0.0));
        const double* amp = raw.data() + 2 * c * nm;
        const double* ph = raw.data() + (2 * c + 1) * nm;
        for (size_t m = 0; m < nm; m++) {
            const size_t q = modes[m];
            const int i = (int)(q % Nx), j = (int)((q / Nx) % Ny), k = (int)(q / ((size_t)Nx * Ny));
            const double KX = kfreq(i, Nx), KY = kfreq(j, Ny), KZ = kfreq(k, Nz);
            const double kk = std::sqrt(KX * KX + KY * KY + KZ * KZ);
            vh[c][q] = (0.5 + amp[m]) * std::pow(kk, -2.0) * std::polar(1.0, 2.0 * kPi * ph[m]);
        }
    }
    for (size_t m = 0; m < nm; m++) {  // solenoidal projection v -= k (k . v) / k^2
        const size_t q = modes[m];
        const int i = (int)(q % Nx), j = (int)((q / Nx) % Ny), k = (int)(q / ((size_t)Nx * Ny));
        const double KX = kfreq(i, Nx), KY = kfreq(j, Ny), KZ = kfreq(k, Nz);
        const double k2 = KX * KX + KY * KY + KZ * KZ;
        if (k2 == 0.0) continue;  // the mean mode carries no compressive part
        const std::complex<double> kdv = (KX * vh[0][q] + KY * vh[1][q] + KZ * vh[2][q]) / k2;
        vh[0][q] -= KX * kdv;
        vh[1][q] -= KY * kdv;
        vh[2][q] -= KZ * kdv;
    }
    for (int c = 0; c < 3; c++) {  // inverse 3D FFT: x lines, y lines, z lines
        std::complex<double>* a = vh[c].data();
        ifft_lines(a, Nx, 1, (size_t)Ny * Nz, [&](size_t l) { return l * Nx; });
        ifft_lines(a, Ny, Nx, (size_t)Nx * Nz, [&](size_t l) { return (l / Nx) * Nx * Ny + (l % Nx); });
        ifft_lines(a, Nz, (size_t)Nx * Ny, (size_t)Nx * Ny, [&](size_t l) { return l; });
    }
    double ss = 0.0;
    for (size_t q = 0; q < N; q++) {
        const double u = vh[0][q].real(), v = vh[1][q].real(), w = vh[2][q].real();
        ss += u * u + v * v + w * w;
    }
    const double rms = std::sqrt(ss / (double)N);
    const double scale = (rms > 0.0) ? kn->mach / rms : 0.0;  // sound speed 1
    for (size_t q = 0; q < N; q++) {
        const double u = vh[0][q].real() * scale, v = vh[1][q].real() * scale, w = vh[2][q].real() * scale;
        rho[q] = 1.0;
        mx[q] = u; my[q] = v; mz[q] = w;
        E[q] = P0 / (g - 1.0) + 0.5 * (u * u + v * v + w * w) + 0.5 * B0 * B0;
    }
    for (size_t f = 0; f < ((size_t)Nx + 1) * Ny * Nz; f++) Bx_f[f] = 0.0;
    for (size_t f = 0; f < (size_t)Nx * ((size_t)Ny + 1) * Nz; f++) By_f[f] = 0.0;
    for (size_t f = 0; f < (size_t)Nx * Ny * ((size_t)Nz + 1); f++) Bz_f[f] = B0;
    return CC_OK;
}

// ================================================================== EH
void cc_euler2d_ic_defaults(cc_euler2d_ic* k) {
    if (!k) return;
    *k = cc_euler2d_ic{42u, 1.0, 1.0, 0.10, 0.10, 0.05, 0.08, 0.5};  // EH:585-603
}

cc_status cc_euler2d_ic_reference(const cc_euler2d_params* p, const cc_euler2d_ic* kn, double* rho, double* mx,
                                  double* my, double* E) {
    if (!p || p->Nx <= 0 || p->Ny <= 0 || !(p->gamma > 1.0) || !kn || !rho || !mx || !my || !E)
        return bad("cc_euler2d_ic_reference: bad parameters or null output");
    const int Nx = p->Nx, Ny = p->Ny;
    const double dx = (p->x1 - p->x0) / Nx, dy = (p->y1 - p->y0) / Ny;
    std::mt19937 rng(kn->seed);
    std::uniform_real_distribution<double> uni01(0.0, 1.0);
    std::uniform_real_distribution<double> uniV(-1.0, 1.0);
    const double xc = 0.5 * (p->x0 + p->x1), yc = 0.5 * (p->y0 + p->y1);
    for (int j = 0; j < Ny; ++j) {
        const double y = p->y0 + (j + 0.5) * dy;
        for (int i = 0; i < Nx; ++i) {
            const double x = p->x0 + (i + 0.5) * dx;
            const size_t q = (size_t)j * Nx + i;
            const double fr = 2.0 * uni01(rng) - 1.0;
            const double fp = 2.0 * uni01(rng) - 1.0;
            double d = kn->rho0 * (1.0 + kn->amp_rho * fr);
            double P = kn->P0 * (1.0 + kn->amp_P * fp);
            const double rx = x - xc, ry = y - yc;
            const double r2 = rx * rx + ry * ry;
            d += kn->bump_amp * kn->rho0 * std::exp(-r2 / (2.0 * kn->sigma * kn->sigma));
            const double u = kn->vmax * uniV(rng);
            const double v = kn->vmax * uniV(rng);
            d = std::max(d, p->floor_rho);
            P = std::max(P, p->floor_P);
            double r0 = d, m0 = d * u, m1 = d * v;
            double e0 = P / (p->gamma - 1.0) + 0.5 * d * (u * u + v * v);
            r0 = std::max(r0, p->floor_rho);  // enforce_floors, EH:62-72
            const double uu = m0 / r0, vv = m1 / r0;
            const double kin = 0.5 * r0 * (uu * uu + vv * vv);
            if ((p->gamma - 1.0) * (e0 - kin) < p->floor_P) e0 = kin + p->floor_P / (p->gamma - 1.0);
            rho[q] = r0; mx[q] = m0; my[q] = m1; E[q] = e0;
        }
    }
    return CC_OK;
}

// ================================================================== NS
void cc_ns2d_ic_defaults(cc_ns2d_ic* k) {
    if (!k) return;
    *k = cc_ns2d_ic{42u, 1.0, 0.05, 1.0, 0.10, 0.01, 0.05};  // NS:69, NS:642-660
}

cc_status cc_ns2d_ic_reference(const cc_ns2d_params* p, const cc_ns2d_ic* kn, double* U) {
    if (!p || p->Nx <= 0 || p->Ny <= 0 || !(p->gamma > 1.0) || !kn || !U)
        return bad("cc_ns2d_ic_reference: bad parameters or null output");
    const int Nx = p->Nx, Ny = p->Ny;
    const double Lx = p->x1 - p->x0, Ly = p->y1 - p->y0;
    const double dx = Lx / Nx, dy = Ly / Ny;
    std::mt19937 rng(kn->seed);
    std::normal_distribution<double> gauss(0.0, 1.0);
    const double sigma = kn->sigma_frac * std::min(Lx, Ly);
    const double a0 = std::sqrt(p->gamma * kn->P0 / kn->rho0);
    const double vamp = kn->Mach0 * a0;
    const double xc = 0.5 * (p->x0 + p->x1), yc = 0.5 * (p->y0 + p->y1);
    auto clamp_min = [](double v, double lo) { return (v < lo) ? lo : v; };  // NS:51
    for (int j = 0; j < Ny; j++)
        for (int i = 0; i < Nx; i++) {
            const size_t q = (size_t)i + (size_t)Nx * j;
            const double x = p->x0 + (i + 0.5) * dx, y = p->y0 + (j + 0.5) * dy;
            const double rx = x - xc, ry = y - yc;
            const double r2 = rx * rx + ry * ry;
            const double bump = kn->A * std::exp(-0.5 * r2 / (sigma * sigma));
            const double noise = kn->Arho * gauss(rng);
            double d = clamp_min(kn->rho0 * (1.0 + bump + noise), p->floor_rho);
            double P = clamp_min(kn->P0, p->floor_P);
            const double u = vamp * gauss(rng);
            const double v = vamp * gauss(rng);
            d = clamp_min(d, p->floor_rho);  // prim_to_cons, NS:134-145
            P = clamp_min(P, p->floor_P);
            U[4 * q + 0] = d;
            U[4 * q + 1] = d * u;
            U[4 * q + 2] = d * v;
            U[4 * q + 3] = P / (p->gamma - 1.0) + 0.5 * d * (u * u + v * v);
        }
    return CC_OK;
}

}  // extern "C"
