// mhd3d_main.cpp -- drop-in for mhd3d_resistive_hlld_ct (M3): 3D resistive MHD with HLLD +
// constrained transport and the hybrid cold + thermal + bounce EOS, the step on the GPU.
//
// With no arguments it runs the reference configuration (Params M3:43-89, IC M3:1159-1272) and
// writes out3d_mhd.h5: /x /y /z /t and /rho /P /Mach /u /v /w /vabs /Bx /By /Bz /divB of shape
// (Nt, Nz, Ny, Nx).  As in the reference there is no t = 0 frame: a frame is stored after a step
// whenever t >= tnext - 1e-12 (M3:1280-1327), announced by "Stored frame k at t=...".
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "cli.hpp"
#include "common.hpp"
#include "snapshot_writer.hpp"

using namespace ccprog;

int main(int argc, char** argv) {
    cc_mhd3d_params p{};
    p.gamma1 = 1.30; p.gamma2 = 2.50; p.K1 = 0.05; p.rho_nuc = 80.0; p.gamma_th = 1.50;  // M3:45-52
    p.rho_bounce = 120.0; p.gamma_bounce = 3.00; p.Kbounce = 0.02;
    p.cfl = 0.4;
    p.x0 = 0.0; p.x1 = 1.0; p.y0 = 0.0; p.y1 = 1.0; p.z0 = 0.0; p.z1 = 1.0;
    p.Nx = 48; p.Ny = 48; p.Nz = 48;
    double t_end = 0.15, dt_store = 1e-3;  // M3:63-64
    p.floor_rho = 1e-10; p.floor_P = 1e-10;
    p.periodic_x = 1; p.periodic_y = 1; p.periodic_z = 1;
    p.eta = 5e-4;
    p.eos = 0;
    p.gamma = 5.0 / 3.0;
    p.degen_tau = 0.0;
    cc_mhd3d_ic ic;
    cc_mhd3d_ic_defaults(&ic);
    cc_turbulence3d_ic tb;
    cc_turbulence3d_ic_defaults(&tb);
    std::string precision = "fp64", out = "out3d_mhd.h5", ic_name = "reference", eos = "hybrid";

    Cli cli("mhd3d_resistive_hlld_ct",
            "3D resistive MHD (HLLD + CT) on the GPU; defaults are the reference run.");
    cli.add("gamma1", &p.gamma1, "hybrid EOS: cold index below rho_nuc");
    cli.add("gamma2", &p.gamma2, "hybrid EOS: cold index above rho_nuc");
    cli.add("K1", &p.K1, "hybrid EOS: cold constant");
    cli.add("rho_nuc", &p.rho_nuc, "hybrid EOS: transition density");
    cli.add("gamma_th", &p.gamma_th, "hybrid EOS: thermal index");
    cli.add("rho_bounce", &p.rho_bounce, "hybrid EOS: bounce density");
    cli.add("gamma_bounce", &p.gamma_bounce, "hybrid EOS: bounce index");
    cli.add("Kbounce", &p.Kbounce, "hybrid EOS: bounce constant");
    cli.add("eos", &eos, "hybrid (reference) | ideal (gamma-law gas with --gamma)");
    cli.add("gamma", &p.gamma, "ideal EOS: adiabatic index");
    cli.add("cfl", &p.cfl, "CFL number");
    cli.add("x0", &p.x0, "domain"); cli.add("x1", &p.x1, "domain");
    cli.add("y0", &p.y0, "domain"); cli.add("y1", &p.y1, "domain");
    cli.add("z0", &p.z0, "domain"); cli.add("z1", &p.z1, "domain");
    cli.add("Nx", &p.Nx, "cells in x"); cli.add("Ny", &p.Ny, "cells in y"); cli.add("Nz", &p.Nz, "cells in z");
    cli.add("t_end", &t_end, "end time"); cli.add("dt_store", &dt_store, "snapshot interval");
    cli.add("floor_rho", &p.floor_rho, "density floor"); cli.add("floor_P", &p.floor_P, "pressure floor");
    cli.add_bool("periodic_x", &p.periodic_x, "periodic in x (else clamped outflow)");
    cli.add_bool("periodic_y", &p.periodic_y, "periodic in y (else clamped outflow)");
    cli.add_bool("periodic_z", &p.periodic_z, "periodic in z (else clamped outflow)");
    cli.add("eta", &p.eta, "resistivity");
    cli.add("ic", &ic_name, "initial condition: reference | turbulence (needs --eos ideal)");
    cli.add("rho_bg", &ic.rho_bg, "reference IC: background density");
    cli.add("rho_core", &ic.rho_core, "reference IC: core density");
    cli.add("r_core", &ic.r_core, "reference IC: core radius");
    cli.add("r_edge", &ic.r_edge, "reference IC: core edge thickness");
    cli.add("P0", &ic.P0, "reference IC: base pressure scale");
    cli.add("vin", &ic.vin, "reference IC: infall speed scale");
    cli.add("rot_omega", &ic.rot_omega, "reference IC: rotation strength");
    cli.add("B0x", &ic.B0x, "reference IC: uniform Bx");
    cli.add("B0z", &ic.B0z, "reference IC: uniform Bz");
    cli.add("Bshear", &ic.Bshear, "reference IC: Az perturbation strength");
    cli.add("gamma_init", &ic.gamma_init, "reference IC: P ~ rho^gamma_init");
    cli.add("seed", &tb.seed, "turbulence IC: mt19937 seed");
    cli.add("kmin", &tb.kmin, "turbulence IC: lowest driven |k|");
    cli.add("kmax", &tb.kmax, "turbulence IC: highest driven |k|");
    cli.add("mach", &tb.mach, "turbulence IC: rms Mach number");
    cli.add("beta", &tb.beta, "turbulence IC: plasma beta of the uniform Bz");
    cli.add("precision", &precision, "fp64 (reference arithmetic) | mixed");
    cli.add("degen_tau", &p.degen_tau, "mixed: FP64 fallback margin near the Alfven degeneracy (<= 0: 0.02)");
    cli.add("out", &out, "output file");
    cli.add("hlld_guard", &p.hlld_guard,
            "HLL for faces whose HLLD star-state denominator is below this fraction of its scale (0: the reference)");
    cli.add_bool("ct_compensated", &p.ct_compensated, "carry the faces' rounding errors through CT (0: the reference)");
    DtGuard guard;
    cli.add("dt_guard", &guard.ratio,
            "stop with status 3 when dt falls below this fraction of the last 64 steps' largest dt (0: never)");
    if (int r = cli.parse(argc, argv)) return r == 1 ? 0 : 2;
    p.precision = parse_precision(precision);
    p.eos = (eos == "hybrid") ? 0 : (eos == "ideal") ? 1 : -1;
    if (p.precision < 0 || p.eos < 0 || (ic_name != "reference" && ic_name != "turbulence") ||
        (ic_name == "turbulence" && p.eos != 1) || !(dt_store > 0.0)) {
        std::fprintf(stderr, "bad --precision, --eos, --ic or --dt_store (see --help)\n");
        return 2;
    }

    const int Nx = p.Nx, Ny = p.Ny, Nz = p.Nz;
    const size_t N = (size_t)Nx * Ny * Nz;
    std::vector<double> rho(N), mx(N), my(N), mz(N), E(N);
    std::vector<double> Bx_f((size_t)(Nx + 1) * Ny * Nz), By_f((size_t)Nx * (Ny + 1) * Nz), Bz_f((size_t)Nx * Ny * (Nz + 1));
    if (ic_name == "reference")
        check(cc_mhd3d_ic_reference(&p, &ic, rho.data(), mx.data(), my.data(), mz.data(), E.data(), Bx_f.data(),
                                    By_f.data(), Bz_f.data()), "initial condition");
    else
        check(cc_mhd3d_ic_turbulence(&p, &tb, rho.data(), mx.data(), my.data(), mz.data(), E.data(), Bx_f.data(),
                                     By_f.data(), Bz_f.data()), "initial condition");

    cc_mhd3d* h = nullptr;
    check(cc_mhd3d_create(&p, &h), "cc_mhd3d_create");
    check(cc_mhd3d_upload(h, rho.data(), mx.data(), my.data(), mz.data(), E.data(), Bx_f.data(), By_f.data(),
                          Bz_f.data()), "cc_mhd3d_upload");

    SnapshotWriter w;
    if (!w.open(out, {"/rho", "/P", "/Mach", "/u", "/v", "/w", "/vabs", "/Bx", "/By", "/Bz", "/divB"},
                {(hsize_t)Nz, (hsize_t)Ny, (hsize_t)Nx})) {
        std::fprintf(stderr, "%s\n", w.error().c_str());
        return 1;
    }
    double t = 0.0, tnext = 0.0;  // M3:1280-1327
    while (t < t_end) {
        double dt = 0.0;
        check(cc_mhd3d_run(h, 1, &t, t_end, &dt), "cc_mhd3d_run");
        if (guard.trip(dt, t, t_end)) {
            std::vector<double> f(w.buffer().size());
            check(cc_mhd3d_frame(h, f.data()), "cc_mhd3d_frame");
            const size_t n = (size_t)Nx * Ny * Nz;
            const double gam = p.eos == 1 ? p.gamma : p.gamma2;
            const size_t q = fastest_cell(&f[0], &f[n], &f[6 * n], &f[7 * n], &f[8 * n], &f[9 * n], n, gam);
            std::fprintf(stderr,
                         "dt collapsed at t=%.9g: dt=%.4g after steps of up to %.4g (--dt_guard %g); fastest cell "
                         "i=%zu j=%zu k=%zu: rho=%.6g P=%.6g -- an HLLD star state near the Alfven degeneracy (the "
                         "reference guards its denominator only by 1e-14); stopping\n",
                         t, dt, guard.scale(), guard.ratio, q % Nx, (q / Nx) % Ny, q / ((size_t)Nx * Ny), f[q], f[n + q]);
            if (!w.finish({{"/x", centres(p.x0, p.x1, Nx)}, {"/y", centres(p.y0, p.y1, Ny)},
                           {"/z", centres(p.z0, p.z1, Nz)}}))
                std::fprintf(stderr, "%s\n", w.error().c_str());
            cc_mhd3d_destroy(h);
            return 3;
        }
        if (t >= tnext - 1e-12) {
            std::vector<double> f = w.buffer();
            check(cc_mhd3d_frame(h, f.data()), "cc_mhd3d_frame");
            if (!w.append(t, std::move(f))) {
                std::fprintf(stderr, "%s\n", w.error().c_str());
                return 1;
            }
            std::cout << "Stored frame " << w.frames() - 1 << " at t=" << t << "\n" << std::flush;
            tnext += dt_store;
        }
    }
    if (!w.finish({{"/x", centres(p.x0, p.x1, Nx)}, {"/y", centres(p.y0, p.y1, Ny)},
                   {"/z", centres(p.z0, p.z1, Nz)}})) {
        std::fprintf(stderr, "%s\n", w.error().c_str());
        return 1;
    }
    cc_mhd3d_destroy(h);
    std::cout << "Wrote " << out << "\n";
    return 0;
}
