// Drives the host-only code of libccamd under AddressSanitizer + UndefinedBehaviorSanitizer: every
// initial condition at several grid sizes (odd, one-cell-wide, power-of-two), the batch sampler, and
// the SOR tile-schedule builder over tile shapes and grids (edge tiles, partial sweep blocks), with the
// size query and the fill of its two-call protocol.  Exits non-zero on a failed call or check.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ccamd.h"

#define CHECK(x)                                                                      \
    do {                                                                              \
        if ((x) != CC_OK) {                                                           \
            std::fprintf(stderr, "%s:%d: %s failed: %s\n", __FILE__, __LINE__, #x, cc_last_error()); \
            std::exit(1);                                                             \
        }                                                                             \
    } while (0)

static void mhd2d(int Nx, int Ny) {
    cc_mhd2d_params p{};
    p.gamma = 5.0 / 3.0; p.cfl = 0.4; p.x0 = -0.5; p.x1 = 0.5; p.y0 = -0.5; p.y1 = 0.5; p.Nx = Nx; p.Ny = Ny;
    p.floor_rho = 1e-10; p.floor_P = 1e-10; p.sor_iters = 10; p.sor_omega = 1.8;
    const size_t n = (size_t)Nx * Ny;
    std::vector<double> r(n), mx(n), my(n), mz(n), E(n), Bz(n), bx((size_t)(Nx + 1) * Ny), by((size_t)Nx * (Ny + 1));
    cc_mhd2d_ic k;
    cc_mhd2d_ic_defaults(&k);
    CHECK(cc_mhd2d_ic_reference(&p, &k, r.data(), mx.data(), my.data(), mz.data(), E.data(), Bz.data(), bx.data(), by.data()));
    cc_collapse2d_ic c;
    cc_collapse2d_ic_defaults(&c);
    CHECK(cc_mhd2d_ic_collapse(&p, &c, r.data(), mx.data(), my.data(), mz.data(), E.data(), Bz.data(), bx.data(), by.data()));
}

static void mhd3d(int N, int eos) {
    cc_mhd3d_params p{};
    p.gamma1 = 1.3; p.gamma2 = 2.5; p.K1 = 1.0; p.rho_nuc = 2e2; p.gamma_th = 1.5; p.rho_bounce = 2e2;
    p.gamma_bounce = 2.5; p.Kbounce = 1.0;
    p.cfl = 0.4; p.x0 = -0.5; p.x1 = 0.5; p.y0 = -0.5; p.y1 = 0.5; p.z0 = -0.5; p.z1 = 0.5;
    p.Nx = N; p.Ny = N; p.Nz = N; p.floor_rho = 1e-10; p.floor_P = 1e-10; p.eos = eos; p.gamma = 5.0 / 3.0;
    const size_t n = (size_t)N * N * N, nf = (size_t)(N + 1) * N * N;
    std::vector<double> r(n), mx(n), my(n), mz(n), E(n), bx(nf), by(nf), bz(nf);
    cc_mhd3d_ic k;
    cc_mhd3d_ic_defaults(&k);
    CHECK(cc_mhd3d_ic_reference(&p, &k, r.data(), mx.data(), my.data(), mz.data(), E.data(), bx.data(), by.data(), bz.data()));
    if (eos == 1 && (N & (N - 1)) == 0) {
        cc_turbulence3d_ic t;
        cc_turbulence3d_ic_defaults(&t);
        t.kmax = N / 4 > 1 ? N / 4 : 1;
        CHECK(cc_mhd3d_ic_turbulence(&p, &t, r.data(), mx.data(), my.data(), mz.data(), E.data(), bx.data(), by.data(), bz.data()));
    }
}

static void hydro(int N) {
    cc_euler2d_params e{};
    e.gamma = 1.4; e.cfl = 0.4; e.x0 = 0; e.x1 = 1; e.y0 = 0; e.y1 = 1; e.Nx = N; e.Ny = N + 3;
    e.floor_rho = 1e-10; e.floor_P = 1e-10;
    const size_t n = (size_t)e.Nx * e.Ny;
    std::vector<double> r(n), mx(n), my(n), E(n);
    cc_euler2d_ic k;
    cc_euler2d_ic_defaults(&k);
    CHECK(cc_euler2d_ic_reference(&e, &k, r.data(), mx.data(), my.data(), E.data()));
    cc_ns2d_params s{};
    s.gamma = 1.4; s.cfl = 0.4; s.x0 = 0; s.x1 = 1; s.y0 = 0; s.y1 = 1; s.Nx = 32; s.Ny = 16; s.floor_rho = 1e-10; s.floor_P = 1e-10;
    std::vector<double> U((size_t)4 * s.Nx * s.Ny);
    cc_ns2d_ic kn;
    cc_ns2d_ic_defaults(&kn);
    CHECK(cc_ns2d_ic_reference(&s, &kn, U.data()));
    cc_euler3d_params g{};
    g.Nx = 16; g.Ny = 8; g.Nz = 32; g.x0 = 0; g.x1 = 1; g.y0 = 0; g.y1 = 1; g.z0 = 0; g.z1 = 1; g.gamma = 5.0 / 3.0; g.cfl = 0.4;
    g.floor_rho = 1e-10; g.floor_P = 1e-10;
    const size_t m = (size_t)16 * 8 * 32;
    std::vector<double> a(m), b(m), c(m), d(m), f(m);
    cc_sphere3d_ic ks;
    cc_sphere3d_ic_defaults(&ks);
    CHECK(cc_euler3d_ic_sphere(&g, &ks, a.data(), b.data(), c.data(), d.data(), f.data()));
}

static void schedule(int Nx, int Ny, int iters, int W, int WJ, int S, int sheared) {
    int nt = 0, nd = 0;
    CHECK(cc_sor_flow_schedule(Nx, Ny, iters, W, WJ, S, sheared, &nt, &nd, nullptr, nullptr, nullptr));
    std::vector<int> tiles((size_t)3 * nt), off((size_t)nt + 1), idx((size_t)(nd > 0 ? nd : 1));
    CHECK(cc_sor_flow_schedule(Nx, Ny, iters, W, WJ, S, sheared, &nt, &nd, tiles.data(), off.data(), idx.data()));
    for (int t = 0; t < nt; ++t)
        for (int d = off[(size_t)t]; d < off[(size_t)t + 1]; ++d)
            if (idx[(size_t)d] < 0 || idx[(size_t)d] >= t) {
                std::fprintf(stderr, "schedule %dx%d: tile %d waits for %d\n", Nx, Ny, t, idx[(size_t)d]);
                std::exit(1);
            }
}

int main() {
    for (int N : {1, 2, 3, 17, 64}) mhd2d(N, N + 1);
    for (int N : {4, 5, 8, 16}) mhd3d(N, 0), mhd3d(N, 1);
    hydro(9);
    std::vector<int> Ns(8);
    std::vector<cc_collapse2d_ic> ics(8);
    std::vector<double> eta(8);
    CHECK(cc_collapse2d_batch_sample(42, 8, 16, 256, Ns.data(), ics.data(), eta.data()));
    for (int sh : {0, 1})
        for (int S : {16, 32})
            for (int WJ : {32, 64, 80, 128}) {
                schedule(1, 1, 1, 64, WJ, S, sh);
                schedule(130, 67, 37, 64, WJ, S, sh);
                schedule(300, 170, 600, 64, WJ, S, sh);
                schedule(64, 2, 33, 64, WJ, S, sh);
            }
    std::printf("host code under ASan/UBSan: ok\n");
    return 0;
}
