// euler2d_main.cpp -- drop-in for euler2d_gravity_hllc (EH): 2D Euler with MUSCL-MC + HLLC and
// SOR-to-tolerance self-gravity, the step on the GPU through libccamd.
//
// With no arguments it runs the reference configuration (Params EH:29-56, init_random_blob
// EH:574-642) and writes out2d_hllc.h5 (/x /y /t /rho /P /Mach, (Nt, Ny, Nx)) with the reference's
// cadence (EH:692-712): frame 0 at t = 0, then one whenever t + 1e-14 >= next_store or at t_end,
// each announced on stderr as "t=... dt=... snapshots=...".  A non-finite or non-positive dt ends
// the run with "Bad dt=... at t=..." and status 1 (EH:699-702).
#include <cmath>
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
    cc_euler2d_params p{};
    p.gamma = 1.2; p.cfl = 0.45;  // EH:29-56
    p.x0 = 0.0; p.x1 = 1.0; p.y0 = 0.0; p.y1 = 1.0;
    p.Nx = 50; p.Ny = 50;
    double t_end = 5.0, dt_store = 1e-3;
    p.floor_rho = 1e-12; p.floor_P = 1e-12;
    p.periodic_x = 1; p.periodic_y = 1;
    p.gravity_on = 1; p.G = 1.0;
    p.max_iter = 50000; p.tol = 1e-10; p.omega = 1.7;  // EH:329-331
    cc_euler2d_ic ic;
    cc_euler2d_ic_defaults(&ic);
    std::string precision = "fp64", out = "out2d_hllc.h5";

    Cli cli("euler2d_gravity_hllc", "2D Euler (HLLC) with SOR self-gravity on the GPU; defaults are the reference run.");
    cli.add("gamma", &p.gamma, "adiabatic index");
    cli.add("cfl", &p.cfl, "CFL number");
    cli.add("x0", &p.x0, "domain"); cli.add("x1", &p.x1, "domain");
    cli.add("y0", &p.y0, "domain"); cli.add("y1", &p.y1, "domain");
    cli.add("Nx", &p.Nx, "cells in x"); cli.add("Ny", &p.Ny, "cells in y");
    cli.add("t_end", &t_end, "end time"); cli.add("dt_store", &dt_store, "snapshot interval");
    cli.add("floor_rho", &p.floor_rho, "density floor"); cli.add("floor_P", &p.floor_P, "pressure floor");
    cli.add_bool("periodic_x", &p.periodic_x, "periodic hydro boundary in x (else clamped outflow)");
    cli.add_bool("periodic_y", &p.periodic_y, "periodic hydro boundary in y (else clamped outflow)");
    cli.add_bool("gravity_on", &p.gravity_on, "self-gravity");
    cli.add("G", &p.G, "gravitational constant");
    cli.add("max_iter", &p.max_iter, "SOR sweep limit");
    cli.add("tol", &p.tol, "SOR residual tolerance");
    cli.add("omega", &p.omega, "SOR relaxation factor");
    cli.add("seed", &ic.seed, "IC: mt19937 seed");
    cli.add("rho0", &ic.rho0, "IC: background density");
    cli.add("P0", &ic.P0, "IC: background pressure");
    cli.add("amp_rho", &ic.amp_rho, "IC: relative density noise");
    cli.add("amp_P", &ic.amp_P, "IC: relative pressure noise");
    cli.add("vmax", &ic.vmax, "IC: velocity noise amplitude");
    cli.add("sigma", &ic.sigma, "IC: Gaussian bump width");
    cli.add("bump_amp", &ic.bump_amp, "IC: Gaussian bump amplitude");
    cli.add("precision", &precision, "fp64 (reference arithmetic) | mixed");
    cli.add("out", &out, "output file");
    if (int r = cli.parse(argc, argv)) return r == 1 ? 0 : 2;
    p.precision = parse_precision(precision);
    if (p.precision < 0 || !(dt_store > 0.0)) {
        std::fprintf(stderr, "bad --precision or --dt_store (see --help)\n");
        return 2;
    }

    const size_t N = (size_t)p.Nx * p.Ny;
    std::vector<double> rho(N), mx(N), my(N), E(N);
    check(cc_euler2d_ic_reference(&p, &ic, rho.data(), mx.data(), my.data(), E.data()), "initial condition");
    cc_euler2d* h = nullptr;
    check(cc_euler2d_create(&p, &h), "cc_euler2d_create");
    check(cc_euler2d_upload(h, rho.data(), mx.data(), my.data(), E.data()), "cc_euler2d_upload");

    SnapshotWriter w;
    if (!w.open(out, {"/rho", "/P", "/Mach"}, {(hsize_t)p.Ny, (hsize_t)p.Nx})) {
        std::fprintf(stderr, "%s\n", w.error().c_str());
        return 1;
    }
    auto store = [&](double t) {
        std::vector<double> f = w.buffer();
        check(cc_euler2d_frame(h, f.data()), "cc_euler2d_frame");
        if (!w.append(t, std::move(f))) {
            std::fprintf(stderr, "%s\n", w.error().c_str());
            std::exit(1);
        }
    };

    double t = 0.0, next_store = 0.0;  // EH:692-712
    store(t);
    next_store += dt_store;
    while (t < t_end - 1e-14) {
        double dt = 0.0;
        check(cc_euler2d_compute_dt(h, &dt), "cc_euler2d_compute_dt");
        if (!std::isfinite(dt) || dt <= 0.0) {
            std::cerr << "Bad dt=" << dt << " at t=" << t << "\n";
            return 1;
        }
        if (t + dt > t_end) dt = t_end - t;
        check(cc_euler2d_rk2_step(h, dt), "cc_euler2d_rk2_step");
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
    cc_euler2d_destroy(h);
    std::cout << "Wrote " << out << " datasets: /x /y /t /rho /P /Mach\n";
    return 0;
}
