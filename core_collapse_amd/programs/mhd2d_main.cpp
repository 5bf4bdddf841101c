// mhd2d_main.cpp -- drop-in for mhd2d_resistive_hlld_ct_gravity (M2): 2D resistive MHD with
// HLLD + constrained transport + SOR self-gravity, the step on the GPU through libccamd.
//
// With no arguments it runs the reference configuration (Params M2:44-86, IC M2:954-1048) and
// writes out2d_mhd.h5 with the reference layout: /x /y /t and /rho /P /Mach /u /v /vabs /Bx /By
// /Bz /divB of shape (Nt, Ny, Nx), frame 0 at t = 0, then a frame whenever t >= next_store - 1e-15
// (M2:1108-1130).  Standard output carries the reference's lines: "init max|divB|=...", one
// "t=... dt=... max|divB|=..." per stored frame and "Wrote out2d_mhd.h5".
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "cli.hpp"
#include "common.hpp"
#include "snapshot_writer.hpp"

using namespace ccprog;

int main(int argc, char** argv) {
    cc_mhd2d_params p{};
    p.gamma = 5.0 / 3.0; p.cfl = 0.45;  // M2:45-51
    p.x0 = 0.0; p.x1 = 1.0; p.y0 = 0.0; p.y1 = 1.0;
    p.Nx = 192; p.Ny = 192;
    double t_end = 0.35, dt_store = 1e-3;  // M2:53-54
    p.floor_rho = 1e-10; p.floor_P = 1e-10;
    p.periodic_x = 1; p.periodic_y = 1;
    p.eta = 5e-4;
    p.gravity_on = 1; p.G = 8.0;
    p.sor_iters = 600; p.sor_omega = 1.8;
    cc_mhd2d_ic ic;
    cc_mhd2d_ic_defaults(&ic);
    cc_collapse2d_ic col;
    cc_collapse2d_ic_defaults(&col);
    std::string precision = "fp64", out = "out2d_mhd.h5", ic_name = "reference";

    Cli cli("mhd2d_resistive_hlld_ct_gravity",
            "2D resistive MHD (HLLD + CT) with SOR self-gravity on the GPU; defaults are the reference run.");
    cli.add("gamma", &p.gamma, "adiabatic index");
    cli.add("cfl", &p.cfl, "CFL number");
    cli.add("x0", &p.x0, "domain"); cli.add("x1", &p.x1, "domain");
    cli.add("y0", &p.y0, "domain"); cli.add("y1", &p.y1, "domain");
    cli.add("Nx", &p.Nx, "cells in x"); cli.add("Ny", &p.Ny, "cells in y");
    cli.add("t_end", &t_end, "end time"); cli.add("dt_store", &dt_store, "snapshot interval");
    cli.add("floor_rho", &p.floor_rho, "density floor"); cli.add("floor_P", &p.floor_P, "pressure floor");
    cli.add_bool("periodic_x", &p.periodic_x, "periodic in x (else clamped outflow)");
    cli.add_bool("periodic_y", &p.periodic_y, "periodic in y (else clamped outflow)");
    cli.add("eta", &p.eta, "resistivity");
    cli.add_bool("gravity_on", &p.gravity_on, "self-gravity");
    cli.add("G", &p.G, "gravitational constant");
    cli.add("sor_iters", &p.sor_iters, "SOR sweeps per Poisson solve");
    cli.add("sor_omega", &p.sor_omega, "SOR relaxation factor");
    cli.add("ic", &ic_name, "initial condition: reference | collapse (Gaussian collapse)");
    cli.add("rho_bg", &ic.rho_bg, "reference IC: background density");
    cli.add("rho_core", &ic.rho_core, "reference IC: core density");
    cli.add("r_core", &ic.r_core, "reference IC: core radius");
    cli.add("r_edge", &ic.r_edge, "reference IC: core edge thickness");
    cli.add("P0", &ic.P0, "reference IC: base pressure scale");
    cli.add("vin", &ic.vin, "reference IC: infall speed scale");
    cli.add("rot_omega", &ic.rot_omega, "reference IC: rotation strength");
    cli.add("B0x", &ic.B0x, "reference IC: uniform Bx");
    cli.add("Bshear", &ic.Bshear, "reference IC: Az perturbation strength");
    cli.add("gamma_eff", &ic.gamma_eff, "reference IC: P ~ rho^gamma_eff");
    cli.add("seed", &col.seed, "collapse IC: mt19937 seed");
    cli.add("contrast", &col.contrast, "collapse IC: central density contrast");
    cli.add("width", &col.width, "collapse IC: Gaussian width");
    cli.add("collapse_P0", &col.P0, "collapse IC: uniform pressure");
    cli.add("vamp", &col.vamp, "collapse IC: velocity noise amplitude");
    cli.add("B0", &col.B0, "collapse IC: uniform Bx");
    cli.add("precision", &precision, "fp64 (reference arithmetic) | mixed");
    cli.add("out", &out, "output file");
    cli.add("hlld_guard", &p.hlld_guard,
            "HLL for faces whose HLLD star-state denominator is below this fraction of its scale (0: the reference)");
    cli.add_bool("ct_compensated", &p.ct_compensated, "carry the faces' rounding errors through CT (0: the reference)");
    DtGuard guard;
    cli.add("dt_guard", &guard.ratio,
            "stop with status 3 when dt falls below this fraction of the last 64 steps' largest dt (0: never)");
    if (int r = cli.parse(argc, argv)) return r == 1 ? 0 : 2;
    p.precision = parse_precision(precision);
    if (p.precision < 0 || (ic_name != "reference" && ic_name != "collapse") || !(dt_store > 0.0)) {
        std::fprintf(stderr, "bad --precision, --ic or --dt_store (see --help)\n");
        return 2;
    }

    const int Nx = p.Nx, Ny = p.Ny;
    const size_t N = (size_t)Nx * Ny;
    std::vector<double> rho(N), mx(N), my(N), mz(N), E(N), Bz(N), Bx_f((size_t)(Nx + 1) * Ny), By_f((size_t)Nx * (Ny + 1));
    if (ic_name == "reference")
        check(cc_mhd2d_ic_reference(&p, &ic, rho.data(), mx.data(), my.data(), mz.data(), E.data(), Bz.data(),
                                    Bx_f.data(), By_f.data()), "initial condition");
    else
        check(cc_mhd2d_ic_collapse(&p, &col, rho.data(), mx.data(), my.data(), mz.data(), E.data(), Bz.data(),
                                   Bx_f.data(), By_f.data()), "initial condition");

    cc_mhd2d* h = nullptr;
    check(cc_mhd2d_create(&p, &h), "cc_mhd2d_create");
    check(cc_mhd2d_upload(h, rho.data(), mx.data(), my.data(), mz.data(), E.data(), Bz.data(), Bx_f.data(),
                          By_f.data()), "cc_mhd2d_upload");

    double maxdiv0 = 0.0;  // M2:1051-1059
    check(cc_mhd2d_divb_max(h, &maxdiv0), "cc_mhd2d_divb_max");
    std::cout << "init max|divB|=" << std::setprecision(12) << maxdiv0 << "\n";

    SnapshotWriter w;
    if (!w.open(out, {"/rho", "/P", "/Mach", "/u", "/v", "/vabs", "/Bx", "/By", "/Bz", "/divB"},
                {(hsize_t)Ny, (hsize_t)Nx})) {
        std::fprintf(stderr, "%s\n", w.error().c_str());
        return 1;
    }
    auto store = [&](double t) {
        std::vector<double> f = w.buffer();
        check(cc_mhd2d_frame(h, f.data()), "cc_mhd2d_frame");
        if (!w.append(t, std::move(f))) {
            std::fprintf(stderr, "%s\n", w.error().c_str());
            std::exit(1);
        }
    };

    double t = 0.0, next_store = 0.0;  // M2:1108-1130
    store(t);
    while (t < t_end) {
        double dt = 0.0;
        check(cc_mhd2d_run(h, 1, &t, t_end, &dt), "cc_mhd2d_run");
        if (guard.trip(dt, t, t_end)) {
            std::vector<double> f(w.buffer().size());
            check(cc_mhd2d_frame(h, f.data()), "cc_mhd2d_frame");
            const size_t n = (size_t)Nx * Ny;
            const size_t q = fastest_cell(&f[0], &f[n], &f[5 * n], &f[6 * n], &f[7 * n], &f[8 * n], n, p.gamma);
            std::fprintf(stderr,
                         "dt collapsed at t=%.9g: dt=%.4g after steps of up to %.4g (--dt_guard %g); fastest cell "
                         "i=%zu j=%zu: rho=%.6g P=%.6g u=%.6g v=%.6g Bx=%.6g By=%.6g Bz=%.6g -- an HLLD star state "
                         "near the Alfven degeneracy (the reference guards its denominator only by 1e-14); stopping\n",
                         t, dt, guard.scale(), guard.ratio, q % Nx, q / Nx, f[q], f[n + q], f[3 * n + q], f[4 * n + q],
                         f[6 * n + q], f[7 * n + q], f[8 * n + q]);
            if (!w.finish({{"/x", centres(p.x0, p.x1, Nx)}, {"/y", centres(p.y0, p.y1, Ny)}}))
                std::fprintf(stderr, "%s\n", w.error().c_str());
            cc_mhd2d_destroy(h);
            return 3;
        }
        if (t >= next_store - 1e-15) {
            double maxdiv = 0.0;
            check(cc_mhd2d_divb_max(h, &maxdiv), "cc_mhd2d_divb_max");
            std::cout << "t=" << std::setprecision(7) << t << " dt=" << dt << " max|divB|=" << std::setprecision(12)
                      << maxdiv << "\n" << std::flush;
            store(t);
            next_store += dt_store;
        }
    }
    if (!w.finish({{"/x", centres(p.x0, p.x1, Nx)}, {"/y", centres(p.y0, p.y1, Ny)}})) {
        std::fprintf(stderr, "%s\n", w.error().c_str());
        return 1;
    }
    cc_mhd2d_destroy(h);
    std::cout << "Wrote " << out << "\n";
    return 0;
}
