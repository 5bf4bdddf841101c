// ns2d_main.cpp -- drop-in for navier_stokes_2d_hllc_poisson (NS): 2D compressible Navier-Stokes
// with first-order HLLC, explicit viscosity and FFT self-gravity, the step on the GPU.
//
// With no arguments it runs the reference configuration (Params NS:54-86, collapse IC NS:638-688)
// and writes out_ns2d_poisson.h5 (/x /y /t /rho /P /Mach /u /v /momx /momy /Phi, (Nt, Ny, Nx)),
// Phi solved for each frame's density (NS:740-746), with the cadence of NS:751-789 and the
// reference's stderr lines.  Non-power-of-two grids end the run with status 1 (NS:614-617).
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
    cc_ns2d_params p{};
    p.gamma = 1.4; p.cfl = 0.35;  // NS:54-86
    p.x0 = 0.0; p.x1 = 1.0; p.y0 = 0.0; p.y1 = 1.0;
    p.Nx = 128; p.Ny = 128;
    double t_end = 10.0, dt_store = 5e-3;
    p.floor_rho = 1e-10; p.floor_P = 1e-10;
    p.viscosity_on = 1; p.mu = 1e-6; p.cfl_visc = 0.05;
    p.gravity_on = 1; p.G = 10.0;
    cc_ns2d_ic ic;
    cc_ns2d_ic_defaults(&ic);
    std::string precision = "fp64", out = "out_ns2d_poisson.h5";

    Cli cli("navier_stokes_2d_hllc_poisson",
            "2D Navier-Stokes (HLLC + viscosity) with FFT self-gravity on the GPU; defaults are the reference run.");
    cli.add("gamma", &p.gamma, "adiabatic index");
    cli.add("cfl", &p.cfl, "CFL number");
    cli.add("x0", &p.x0, "domain"); cli.add("x1", &p.x1, "domain");
    cli.add("y0", &p.y0, "domain"); cli.add("y1", &p.y1, "domain");
    cli.add("Nx", &p.Nx, "cells in x (power of two)"); cli.add("Ny", &p.Ny, "cells in y (power of two)");
    cli.add("t_end", &t_end, "end time"); cli.add("dt_store", &dt_store, "snapshot interval");
    cli.add("seed", &ic.seed, "IC: mt19937 seed");
    cli.add("floor_rho", &p.floor_rho, "density floor"); cli.add("floor_P", &p.floor_P, "pressure floor");
    cli.add_bool("viscosity_on", &p.viscosity_on, "viscous fluxes");
    cli.add("mu", &p.mu, "dynamic viscosity");
    cli.add("cfl_visc", &p.cfl_visc, "viscous stability number");
    cli.add_bool("gravity_on", &p.gravity_on, "self-gravity");
    cli.add("G", &p.G, "gravitational constant");
    cli.add("out_h5", &out, "output file");
    cli.add("rho0", &ic.rho0, "IC: background density");
    cli.add("P0", &ic.P0, "IC: uniform pressure");
    cli.add("A", &ic.A, "IC: central overdensity");
    cli.add("sigma_frac", &ic.sigma_frac, "IC: Gaussian width / box size");
    cli.add("Arho", &ic.Arho, "IC: relative density noise");
    cli.add("Mach0", &ic.Mach0, "IC: velocity noise Mach number");
    cli.add("precision", &precision, "fp64 (reference arithmetic) | mixed");
    if (int r = cli.parse(argc, argv)) return r == 1 ? 0 : 2;
    p.precision = parse_precision(precision);
    if (p.precision < 0 || !(dt_store > 0.0)) {
        std::fprintf(stderr, "bad --precision or --dt_store (see --help)\n");
        return 2;
    }

    auto pow2 = [](int n) { return n > 0 && (n & (n - 1)) == 0; };
    if (!pow2(p.Nx) || !pow2(p.Ny)) {  // NS:614-617
        std::cerr << "Nx and Ny must be powers of two for the FFT Poisson solver.\n";
        return 1;
    }
    cc_ns2d* h = nullptr;
    check(cc_ns2d_create(&p, &h), "cc_ns2d_create");
    const size_t N = (size_t)p.Nx * p.Ny;
    std::vector<double> U(4 * N);
    check(cc_ns2d_ic_reference(&p, &ic, U.data()), "initial condition");
    check(cc_ns2d_upload(h, U.data()), "cc_ns2d_upload");

    SnapshotWriter w;
    if (!w.open(out, {"/rho", "/P", "/Mach", "/u", "/v", "/momx", "/momy", "/Phi"}, {(hsize_t)p.Ny, (hsize_t)p.Nx})) {
        std::fprintf(stderr, "%s\n", w.error().c_str());
        return 1;
    }
    auto store = [&](double t) {
        std::vector<double> f = w.buffer();
        check(cc_ns2d_frame(h, f.data()), "cc_ns2d_frame");
        if (!w.append(t, std::move(f))) {
            std::fprintf(stderr, "%s\n", w.error().c_str());
            std::exit(1);
        }
    };

    double t = 0.0, next_store = 0.0;  // NS:748-789
    store(t);
    next_store += dt_store;
    while (t < t_end - 1e-14) {
        double dt = 0.0;
        check(cc_ns2d_compute_dt(h, &dt), "cc_ns2d_compute_dt");
        if (t + dt > t_end) dt = t_end - t;
        check(cc_ns2d_heun_step(h, dt), "cc_ns2d_heun_step");
        t += dt;
        if (t + 1e-14 >= next_store || t >= t_end - 1e-14) {
            store(t);
            next_store += dt_store;
            std::cerr << "t=" << std::setprecision(10) << t << " dt=" << dt << " snapshots=" << w.frames() << "\n" << std::flush;
        }
    }
    if (!w.finish({{"/x", centres(p.x0, p.x1, p.Nx)}, {"/y", centres(p.y0, p.y1, p.Ny)}})) {
        std::fprintf(stderr, "%s\n", w.error().c_str());
        return 1;
    }
    cc_ns2d_destroy(h);
    std::cout << "Wrote " << out << " with datasets /x /y /t /rho /P /Mach /u /v /momx /momy /Phi\n";
    return 0;
}
