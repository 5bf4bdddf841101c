/*
 * ccamd.h -- C ABI of the MI355X (gfx950) solver-step library libccamd.so.
 *
 * The reference (MohhamedRafi/core_collapse) has no plugin or FFI API: each program is a
 * zero-argument executable whose step is a set of internal-linkage C++ functions.  Every
 * entry point below replaces one of those functions (cited as file:line) with plain
 * pointers and sizes; state crosses the boundary in the reference's own SoA layouts
 * (cells i-fastest, N+1 faces per face axis, FP64).  The library owns all device buffers;
 * host buffers stay caller-owned and are copied in and out.
 *
 * Threading: one handle = one HIP stream; a handle is not thread-safe, independent handles
 * are.  Step calls are asynchronous on the handle's stream; download / dt / diagnostics
 * calls synchronise it.  No call exits the process; errors are returned as cc_status and
 * described by cc_last_error() (thread-local).
 *
 * Files cited: M2 = mhd2d_resistive_hlld_ct_gravity.cpp, M3 = mhd3d_resistive_hlld_ct.cpp,
 *              EH = euler2d_gravity_hllc.cpp, NS = navier_stokes_2d_hllc_poisson.cpp.
 */
#ifndef CCAMD_H
#define CCAMD_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    CC_OK = 0,
    CC_ERR_INVALID_ARG = 1, /* bad size / null pointer / unsupported parameter combination */
    CC_ERR_NON_POW2 = 2,    /* FFT Poisson needs power-of-two grids (NS:614-617 returns 1) */
    CC_ERR_BAD_DT = 3,      /* non-finite or non-positive dt (EH:699-702 returns 1)         */
    CC_ERR_HIP = 4,         /* HIP runtime error                                            */
    CC_ERR_OOM = 5,         /* device allocation failed                                     */
    CC_ERR_NO_DEVICE = 6    /* no gfx950 device visible                                     */
} cc_status;

typedef enum {
    CC_FP64 = 0, /* every step in FP64: the reference arithmetic                                   */
    CC_MIXED = 1 /* FP32 reconstruction, Riemann fluxes and Poisson sweeps; FP64 conserved update,
                    pressure recovery, face B / CT and the Poisson residual                      */
} cc_precision;

/* Last error message of the calling thread ("" if none). */
const char* cc_last_error(void);
/* Library version (major*10000 + minor*100 + patch). */
int cc_version(void);
/* Name and compute-unit count of the current device (for logs). */
cc_status cc_device_info(char* name, int name_len, int* cu_count);

/* ============================================================================================
 * 2D resistive MHD + SOR self-gravity (M2).
 * ========================================================================================== */

/* The hot-path subset of struct Params (M2:44-72); the IC knobs (M2:74-85) live with the ICs. */
typedef struct cc_mhd2d_params {
    double gamma, cfl, x0, x1, y0, y1;
    int Nx, Ny;
    double floor_rho, floor_P;
    int periodic_x, periodic_y;
    double eta;
    int gravity_on;
    double G;
    int sor_iters;    /* M2:71 */
    double sor_omega; /* M2:72 */
    int precision;    /* cc_precision */
    /* No reference counterpart (0 keeps the reference's arithmetic):
     * hlld_guard: an HLLD face whose star-state denominator rho (S - u)(S - S_M) - Bn^2 (M2:419-442) on
     *   either side is below hlld_guard * max(Bn^2, |rho (S - u)(S - S_M)|) takes the reference's own HLL
     *   flux (its |Bn| < 1e-12 branch, M2:449) instead of star states divided by a near-zero d (the
     *   reference guards d only by 1e-14).  In [0, 1).
     * ct_compensated: the face update of CT (M2:807-830) and the Heun average of the faces (M2:850-866)
     *   carry each face's rounding error in an FP32 low word, so max|div B| of the stored field stays at
     *   one rounding instead of growing with the number of steps. */
    double hlld_guard;
    int ct_compensated;
} cc_mhd2d_params;

typedef struct cc_mhd2d cc_mhd2d;

/* Allocates device state, temporaries and the SOR schedule for one grid. */
cc_status cc_mhd2d_create(const cc_mhd2d_params* p, cc_mhd2d** out);
cc_status cc_mhd2d_destroy(cc_mhd2d* h);

/* State in the reference layout (M2:88-95): cells q = j*Nx+i; Bx_f[(Nx+1)*Ny] at j*(Nx+1)+i_f;
 * By_f[Nx*(Ny+1)] at j_f*Nx+i.  upload copies host -> device, download device -> host (syncs). */
cc_status cc_mhd2d_upload(cc_mhd2d* h, const double* rho, const double* mx, const double* my, const double* mz,
                          const double* E, const double* Bz, const double* Bx_f, const double* By_f);
cc_status cc_mhd2d_download(cc_mhd2d* h, double* rho, double* mx, double* my, double* mz, double* E, double* Bz,
                            double* Bx_f, double* By_f);

/* g = -grad(phi), phi from 4*pi*G*(rho - mean rho) by lexicographic SOR on the current state's rho
 * (compute_poisson_gravity_periodic, M2:178-237; called at M2:1112).  The result is kept on the
 * device for the next cc_mhd2d_rk2_step / cc_mhd2d_stage. */
cc_status cc_mhd2d_gravity(cc_mhd2d* h);
cc_status cc_mhd2d_get_gravity(cc_mhd2d* h, double* gx, double* gy);
cc_status cc_mhd2d_set_gravity(cc_mhd2d* h, const double* gx, const double* gy);

/* CFL / resistive time step of the current state (compute_dt, M2:579-598).  Syncs. */
cc_status cc_mhd2d_compute_dt(cc_mhd2d* h, double* dt);

/* One forward-Euler stage in place, with the stored g (advance_stage, M2:690-834). */
cc_status cc_mhd2d_stage(cc_mhd2d* h, double dt);

/* One Heun step in place with the stored g, including the inner Poisson solve on U1
 * (rk2_step, M2:838-869). */
cc_status cc_mhd2d_rk2_step(cc_mhd2d* h, double dt);

/* nsteps iterations of the driver loop M2:1111-1117 (gravity, dt, clamp to t_end, rk2_step),
 * with dt kept on the device and the loop captured in HIP graphs.  *t_inout holds t on entry
 * and the final t on return; dts (nullable, nsteps doubles) receives each step's dt.  Syncs. */
cc_status cc_mhd2d_run(cc_mhd2d* h, int nsteps, double* t_inout, double t_end, double* dts);

/* max |div B| over cells of the face field (divB_cc, M2:873-879, printed at M2:1119-1126).  Syncs. */
cc_status cc_mhd2d_divb_max(cc_mhd2d* h, double* out);

/* Wait for all queued work of the handle. */
cc_status cc_mhd2d_sync(cc_mhd2d* h);

/* Per-kernel timing with HIP events recorded on the handle's stream around each launch
 * (kernel names: "sor", "flux", "corner", "ct", "grad", "mean", "dt").  Enabling adds an
 * event pair per launch; totals accumulate until reset. */
cc_status cc_mhd2d_set_profiling(cc_mhd2d* h, int on);
cc_status cc_mhd2d_kernel_time(cc_mhd2d* h, const char* kernel, int* count, double* total_ms);
cc_status cc_mhd2d_reset_profiling(cc_mhd2d* h);

/* Snapshot frame of the current state, computed on the device as the store lambda of M2:1065-1106
 * does: CC_MHD2D_FRAME_FIELDS planes of Ny*Nx cells, in the order of the output datasets
 * {rho, P, Mach, u, v, vabs, Bx, By, Bz, divB} (M2:908-922), written to the host buffer `out`.  Syncs. */
#define CC_MHD2D_FRAME_FIELDS 10
cc_status cc_mhd2d_frame(cc_mhd2d* h, double* out);
/* In-situ totals of the current state, reduced on the device: out[0..5] = {mass, x-, y-, z-momentum,
 * total energy} as cell sums times dx*dy (the integrals of viz_h5.py:640-657 taken on the conserved
 * variables), out[5] = max|divB| (M2:1119-1126).  Syncs. */
cc_status cc_mhd2d_totals(cc_mhd2d* h, double* out6);
/* L1/Linf comparison of the frames of two handles on one grid (e.g. CC_MIXED against CC_FP64), on the
 * device.  For each frame field f: out[3f] = mean |a-b|, out[3f+1] = sum |a-b| / sum |b| (relative L1;
 * 0 when sum |b| is 0), out[3f+2] = max |a-b|.  Syncs both handles. */
cc_status cc_mhd2d_compare(cc_mhd2d* a, cc_mhd2d* b, double* out);

/* Batches of independent 2D MHD runs (no reference counterpart: the benchmark's batch of collapse
 * instances of uneven size).  The instances' Heun steps run as the parallel branches of one HIP
 * graph -- one graph launch per step for the whole batch -- each with its own dt on the device.
 * The handles stay owned by the caller and must outlive the batch; all must be at the same point of
 * their state rotation (e.g. freshly uploaded, or stepped only through the batch), have profiling
 * off, and must not be used by other calls while the batch runs.  At most 65536 steps per call. */
typedef struct cc_mhd2d_batch cc_mhd2d_batch;
cc_status cc_mhd2d_batch_create(cc_mhd2d* const* handles, int n, cc_mhd2d_batch** out);
cc_status cc_mhd2d_batch_destroy(cc_mhd2d_batch* b);
/* merged = 1 when the instances' SOR tile levels run as merged launches (every instance with
 * gravity, clamped boundaries and one precision), with the tile shape (S sweeps, WJ rows or, with
 * *sheared = 1, WJ values of i + j + 2 sweep) the batch uses -- the first instance's, or with
 * cc_sor_autotune the fastest merged shape, *fp64_kernels = 1 when a mixed batch runs the FP64 SOR kernels;
 * merged = 0 when each instance's step is a separate graph branch (or CC_BATCH_SEPARATE_SOR is set).
 * sheared and fp64_kernels may be null. */
cc_status cc_mhd2d_batch_info(cc_mhd2d_batch* b, int* merged, int* S, int* WJ, int* sheared, int* fp64_kernels);
/* nsteps steps of every instance (the loop of cc_mhd2d_run: gravity, dt, clamp to t_end[i], Heun).
 * t_inout[i] holds instance i's t on entry and return.  Syncs. */
cc_status cc_mhd2d_batch_run(cc_mhd2d_batch* b, int nsteps, double* t_inout, const double* t_end);

/* Batched HLLD fluxes on independent face states (hlld_x M2:372-530 for axis 0, hlld_y
 * M2:540-561 for axis 1), ideal EOS.  W* are n x 8 {rho,u,v,w,P,Bx,By,Bz}; out is n x 8:
 * axis 0 {frho,fmx,fmy,fmz,fE,fBy,fBz,Ez}, axis 1 {frho,fmx,fmy,fmz,fE,fBx,fBz,Ez}. */
cc_status cc_hlld2d_batch(int axis, int n, const double* WL, const double* WR, const double* Bn, double gamma,
                          double floor_rho, int precision, double* out);

/* Stand-alone Poisson gravity on a host density field (compute_poisson_gravity_periodic,
 * M2:178-237) with the given boundary flags, sweep count and omega.  phi (nullable) receives
 * the final iterate with its mean removed, as the reference holds it after its last sweep. */
cc_status cc_poisson2d_sor(int Nx, int Ny, int periodic_x, int periodic_y, double dx, double dy, double G,
                           int sor_iters, double omega, int precision, const double* rho, double* gx, double* gy,
                           double* phi);

/* ============================================================================================
 * 3D resistive MHD (M3), no gravity.
 * ========================================================================================== */

/* The hot-path subset of struct Params (M3:43-74) plus the EOS choice.  eos = 0 is the reference's
 * hybrid cold + thermal + bounce EOS (M3:114-212, parameters gamma1..Kbounce); eos = 1 an ideal
 * gas with `gamma` (the 3D benchmark problem; the reference has no such EOS). */
typedef struct cc_mhd3d_params {
    double gamma1, gamma2, K1, rho_nuc, gamma_th, rho_bounce, gamma_bounce, Kbounce;
    double cfl, x0, x1, y0, y1, z0, z1;
    int Nx, Ny, Nz;
    double floor_rho, floor_P;
    int periodic_x, periodic_y, periodic_z;
    double eta;
    int eos;
    double gamma;
    int precision;    /* cc_precision */
    double degen_tau; /* CC_MIXED only: faces that the FP32 HLLD cannot be trusted on are recomputed in FP64:
                         without the guard, faces within this relative margin of the Alfven degeneracy;
                         with hlld_guard > 0, faces whose |d| lies within this relative band of the guard
                         threshold (the one decision FP32 rounding can flip); <= 0 selects 0.02 without,
                         0.01 with the guard */
    /* As in cc_mhd2d_params (no reference counterpart; 0 keeps the reference's arithmetic): the HLLD
     * degeneracy guard on rho (S - u)(S - S_M) - Bn^2 (M3:463-484, HLL branch M3:491) and the compensated
     * face update of CT (M3:941-1015) and of the Heun average (M3:1042-1065). */
    double hlld_guard;
    int ct_compensated;
} cc_mhd3d_params;

typedef struct cc_mhd3d cc_mhd3d;

cc_status cc_mhd3d_create(const cc_mhd3d_params* p, cc_mhd3d** out);
cc_status cc_mhd3d_destroy(cc_mhd3d* h);
/* Reference layout (M3:91-99): cells (k*Ny+j)*Nx+i; Bx_f (Nx+1)*Ny*Nz at (k*Ny+j)*(Nx+1)+i_f;
 * By_f Nx*(Ny+1)*Nz at (k*(Ny+1)+j_f)*Nx+i; Bz_f Nx*Ny*(Nz+1) at (k_f*Ny+j)*Nx+i. */
cc_status cc_mhd3d_upload(cc_mhd3d* h, const double* rho, const double* mx, const double* my, const double* mz,
                          const double* E, const double* Bx_f, const double* By_f, const double* Bz_f);
cc_status cc_mhd3d_download(cc_mhd3d* h, double* rho, double* mx, double* my, double* mz, double* E, double* Bx_f,
                            double* By_f, double* Bz_f);
/* compute_dt, M3:626-648.  Syncs. */
cc_status cc_mhd3d_compute_dt(cc_mhd3d* h, double* dt);
/* advance_stage, M3:694-1030, in place. */
cc_status cc_mhd3d_stage(cc_mhd3d* h, double dt);
/* rk2_step, M3:1034-1066, in place. */
cc_status cc_mhd3d_rk2_step(cc_mhd3d* h, double dt);
/* nsteps of the loop M3:1285-1291 (compute_dt, clamp to t_end, rk2_step) with dt on the device,
 * captured in HIP graphs.  Syncs. */
cc_status cc_mhd3d_run(cc_mhd3d* h, int nsteps, double* t_inout, double t_end, double* dts);
/* max |divB_cc| (M3:1070-1078).  Syncs. */
cc_status cc_mhd3d_divb_max(cc_mhd3d* h, double* out);
cc_status cc_mhd3d_sync(cc_mhd3d* h);
/* Per-kernel HIP-event timing ("prim", "face" (x, y, z face kernels), "face_fp64" (the mixed path's FP64 pass over
 * faces queued near the Alfven degeneracy), "update", "edge", "ct", "joule", "dt"). */
cc_status cc_mhd3d_set_profiling(cc_mhd3d* h, int on);
cc_status cc_mhd3d_kernel_time(cc_mhd3d* h, const char* kernel, int* count, double* total_ms);
cc_status cc_mhd3d_reset_profiling(cc_mhd3d* h);
/* Snapshot frame (the store block of M3:1293-1325): CC_MHD3D_FRAME_FIELDS volumes of Nz*Ny*Nx cells in
 * dataset order {rho, P, Mach, u, v, w, vabs, Bx, By, Bz, divB} (M3:1112-1125).  Syncs. */
#define CC_MHD3D_FRAME_FIELDS 11
cc_status cc_mhd3d_frame(cc_mhd3d* h, double* out);
/* {mass, x-, y-, z-momentum, total energy} * dx*dy*dz and max|divB|, reduced on the device.  Syncs. */
cc_status cc_mhd3d_totals(cc_mhd3d* h, double* out6);
/* Per-field {mean|a-b|, sum|a-b|/sum|b|, max|a-b|} of two handles' frames (as cc_mhd2d_compare). */
cc_status cc_mhd3d_compare(cc_mhd3d* a, cc_mhd3d* b, double* out);
/* Batched hlld_x (M3:415-568) with the EOS and precision of p; W* n x 8, out n x 7
 * {frho,fmx,fmy,fmz,fE,fBy,fBz}. */
cc_status cc_hlld3d_batch(const cc_mhd3d_params* p, int n, const double* WL, const double* WR, const double* Bn,
                          double* out);

/* SOR tile autotuning (no reference counterpart).  With enable != 0, every 2D MHD handle created
 * afterwards with clamped boundaries times the compiled SOR tile shapes (sweeps per tile x rows) on
 * its grid once -- one full solve each, HIP events -- and keeps the fastest; the choice is cached per
 * (Nx, Ny, sor_iters, precision) for the process.  The environment variable CC_SOR_TILE overrides it.
 * For a mixed-precision grid the FP64 kernels are candidates too (full FP64 sweeps: exact, and faster where
 * the FP32 correction sweeps do not pay for their FP64 residual passes).  cc_sor_tile_choice reports the
 * shape a grid gets (the default shape if it was never tuned): S sweeps per tile, WJ rows (rectangular
 * tile) or WJ values of i + j + 2 sweep (sheared tile, *sheared = 1), and *fp64_kernels = 1 for a mixed
 * grid that runs the FP64 kernels (sheared and fp64_kernels may be null). */
cc_status cc_sor_autotune(int enable);
cc_status cc_sor_tile_choice(int Nx, int Ny, int sor_iters, int precision, int* S, int* WJ, int* sheared,
                             int* fp64_kernels);

/* The dataflow schedule of the tiled SOR (no reference counterpart; host only, touches no GPU): for an
 * Nx x Ny clamped grid, `iters` sweeps and tiles of S sweeps x W x WJ cells (W = 64 in the kernels;
 * sheared != 0: S sweeps x W values of i + sweep x WJ values of i + j + 2 sweep),
 * the tiles (TS, TJ, TI) that update at least one cell, in launch order (3 ints each), and per tile the
 * tiles it waits for before it starts, as CSR (dep_off: ntiles + 1 offsets into dep_idx).  Call with
 * tiles, dep_off or dep_idx null to get *ntiles and *ndeps; then with arrays of 3 *ntiles, *ntiles + 1
 * and *ndeps ints and *ntiles, *ndeps set to their capacities. */
cc_status cc_sor_flow_schedule(int Nx, int Ny, int iters, int W, int WJ, int S, int sheared, int* ntiles, int* ndeps,
                               int* tiles, int* dep_off, int* dep_idx);

/* ============================================================================================
 * 2D Euler HLLC + SOR self-gravity (EH).
 * ========================================================================================== */

/* The hot-path subset of struct Params (EH:29-56) plus the solver constants of EH:329-331
 * (max_iter 50000, tol 1e-10, omega 1.7 in the reference). */
typedef struct cc_euler2d_params {
    double gamma, cfl, x0, x1, y0, y1;
    int Nx, Ny;
    double floor_rho, floor_P;
    int periodic_x, periodic_y; /* hydro boundaries; the Poisson solve is always periodic (EH:300-304) */
    int gravity_on;
    double G;
    int max_iter;
    double tol, omega;
    int precision; /* cc_precision: CC_MIXED runs the HLLC in FP32 */
} cc_euler2d_params;

typedef struct cc_euler2d cc_euler2d;

cc_status cc_euler2d_create(const cc_euler2d_params* p, cc_euler2d** out);
cc_status cc_euler2d_destroy(cc_euler2d* h);
/* State in the reference layout (struct State, EH:58-60): SoA, q = j*Nx+i. */
cc_status cc_euler2d_upload(cc_euler2d* h, const double* rho, const double* mx, const double* my, const double* E);
cc_status cc_euler2d_download(cc_euler2d* h, double* rho, double* mx, double* my, double* E);
/* g (and phi, nullable) of the current state (compute_poisson_gravity_periodic, EH:288-386), and the
 * number of SOR sweeps it took.  Syncs. */
cc_status cc_euler2d_gravity(cc_euler2d* h, double* gx, double* gy, double* phi, int* sweeps);
/* compute_dt, EH:390-404.  Syncs. */
cc_status cc_euler2d_compute_dt(cc_euler2d* h, double* dt);
/* rhs(U) including its Poisson solve (EH:406-527) into caller buffers.  Syncs. */
cc_status cc_euler2d_rhs(cc_euler2d* h, double* drho, double* dmx, double* dmy, double* dE);
/* rk2_step, EH:529-550, in place. */
cc_status cc_euler2d_rk2_step(cc_euler2d* h, double dt);
/* nsteps of the loop EH:697-708 (compute_dt, "Bad dt" check -> CC_ERR_BAD_DT, clamp to t_end,
 * rk2_step).  *t_inout holds t on entry and the final t on return; dts (nullable) gets each dt. */
cc_status cc_euler2d_run(cc_euler2d* h, int nsteps, double* t_inout, double t_end, double* dts);
/* SOR sweep counts: of the last solve, summed over solves since the last reset, number of solves. */
cc_status cc_euler2d_sweep_stats(cc_euler2d* h, int* last, long long* total, int* solves);
cc_status cc_euler2d_sync(cc_euler2d* h);
/* Per-kernel HIP-event timing ("sor", "flux", "dt"). */
cc_status cc_euler2d_set_profiling(cc_euler2d* h, int on);
cc_status cc_euler2d_kernel_time(cc_euler2d* h, const char* kernel, int* count, double* total_ms);
cc_status cc_euler2d_reset_profiling(cc_euler2d* h);
/* Snapshot frame (the store lambda of EH:661-691: floors on a copy, then cons_to_prim_full):
 * CC_EULER2D_FRAME_FIELDS planes {rho, P, Mach} of Ny*Nx cells (EH:718-731).  Syncs. */
#define CC_EULER2D_FRAME_FIELDS 3
cc_status cc_euler2d_frame(cc_euler2d* h, double* out);
/* {mass, x-, y-momentum, total energy} * dx*dy, reduced on the device.  Syncs. */
cc_status cc_euler2d_totals(cc_euler2d* h, double* out4);
/* Per-field {mean|a-b|, sum|a-b|/sum|b|, max|a-b|} of two handles' frames (as cc_mhd2d_compare). */
cc_status cc_euler2d_compare(cc_euler2d* a, cc_euler2d* b, double* out);

/* Batched HLLC on independent face states.  variant 0: EH hllc_x / hllc_y (EH:125-219);
 * variant 1: NS hllc_x / hllc_y (NS:212-282).  W* are n x 4 {rho,u,v,P}; out n x 4
 * {f_rho, f_mx, f_my, f_E}; axis 0 = x faces, 1 = y faces. */
cc_status cc_hllc2d_batch(int variant, int axis, int n, const double* WL, const double* WR, double gamma,
                          double floor_rho, int precision, double* out);

/* ============================================================================================
 * 2D Navier-Stokes HLLC + FFT self-gravity (NS), periodic.
 * ========================================================================================== */

/* The hot-path subset of struct Params (NS:54-86). */
typedef struct cc_ns2d_params {
    double gamma, cfl, x0, x1, y0, y1;
    int Nx, Ny; /* powers of two, <= 4096 (NS:614-617 -> CC_ERR_NON_POW2) */
    double floor_rho, floor_P;
    int viscosity_on;
    double mu, cfl_visc;
    int gravity_on;
    double G;
    int precision; /* cc_precision: CC_MIXED runs the HLLC in FP32 */
} cc_ns2d_params;

typedef struct cc_ns2d cc_ns2d;

cc_status cc_ns2d_create(const cc_ns2d_params* p, cc_ns2d** out);
cc_status cc_ns2d_destroy(cc_ns2d* h);
/* AoS state as std::vector<U4> (NS:88-93): U[4*q + {0: rho, 1: mx, 2: my, 3: E}], q = i + Nx*j. */
cc_status cc_ns2d_upload(cc_ns2d* h, const double* U);
cc_status cc_ns2d_download(cc_ns2d* h, double* U);
/* phi of a host density field (solve_poisson_periodic_fft, NS:285-335).  Syncs. */
cc_status cc_ns2d_poisson(cc_ns2d* h, const double* rho, double* phi);
/* compute_rhs (NS:363-518) of the current state into dU (AoS).  Syncs. */
cc_status cc_ns2d_rhs(cc_ns2d* h, double* dU);
/* compute_dt (NS:521-546).  Syncs. */
cc_status cc_ns2d_compute_dt(cc_ns2d* h, double* dt);
/* The Heun update of the driver loop (NS:761-781) in place. */
cc_status cc_ns2d_heun_step(cc_ns2d* h, double dt);
/* nsteps of the loop NS:758-785 (compute_dt, clamp to t_end, Heun).  Syncs. */
cc_status cc_ns2d_run(cc_ns2d* h, int nsteps, double* t_inout, double t_end, double* dts);
cc_status cc_ns2d_sync(cc_ns2d* h);
/* Per-kernel HIP-event timing ("fft", "flux", "dt"). */
cc_status cc_ns2d_set_profiling(cc_ns2d* h, int on);
cc_status cc_ns2d_kernel_time(cc_ns2d* h, const char* kernel, int* count, double* total_ms);
cc_status cc_ns2d_reset_profiling(cc_ns2d* h);
/* Snapshot frame (the store lambda of NS:698-749, including the FFT solve for Phi of the frame's
 * density): CC_NS2D_FRAME_FIELDS planes {rho, P, Mach, u, v, momx, momy, Phi} of Ny*Nx cells
 * (NS:593-605).  Syncs. */
#define CC_NS2D_FRAME_FIELDS 8
cc_status cc_ns2d_frame(cc_ns2d* h, double* out);
/* {mass, x-, y-momentum, total energy} * dx*dy, reduced on the device.  Syncs. */
cc_status cc_ns2d_totals(cc_ns2d* h, double* out4);
/* Per-field {mean|a-b|, sum|a-b|/sum|b|, max|a-b|} of two handles' frames (as cc_mhd2d_compare). */
cc_status cc_ns2d_compare(cc_ns2d* a, cc_ns2d* b, double* out);

/* fft2d (NS:182-195) on the device: interleaved complex data[2*(i + Nx*j) + {0: re, 1: im}], rows
 * then columns, inverse scaled by 1/n per axis.  Powers of two <= 4096. */
cc_status cc_fft2d(int Nx, int Ny, int inverse, double* data);

/* cc_fft2d on device memory, allocating nothing (no counterpart in the reference: for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight serving loops).  `data` is a
 * device pointer to Nx*Ny interleaved complex values, transformed in place; `stream` is a hipStream_t
 * (null: the default stream) on which the call is asynchronous.  The twiddle tables go to a fixed
 * device array of 256 KiB that each GPU holds until the process ends: Nx and Ny must be powers of two
 * <= 4096 (else CC_ERR_NON_POW2 / CC_ERR_INVALID_ARG, before any device work).  Every call on one GPU
 * rebuilds the tables in that one array, so calls that overlap on a GPU (other streams, other threads)
 * race; serialise them.  Results are bit-identical to cc_fft2d. */
cc_status cc_fft2d_noalloc(int Nx, int Ny, int inverse, double* data, void* stream);

/* ============================================================================================
 * 3D Euler HLLC + FFT self-gravity, periodic (the uniform-sphere collapse benchmark).  No reference
 * program has this step: it is the EH step (EH:406-550: floored primitives, 4-point MUSCL-MC, the
 * EH HLLC, flux divergence + rho g source, Heun with floors) extended to three axes, with gravity
 * from the 7-point discrete Poisson equation solved by FFT and FP64 iterative refinement.  The
 * state layout is EH's SoA extended by mz: cell q = (k*Ny + j)*Nx + i.
 * ========================================================================================== */
typedef struct cc_euler3d_params {
    int Nx, Ny, Nz; /* powers of two in [8, 1024] (-> CC_ERR_NON_POW2) */
    double x0, x1, y0, y1, z0, z1;
    double gamma, cfl, floor_rho, floor_P;
    int gravity_on;
    double G;
    double poisson_tol;   /* stop refining when max|rhs - L_h phi| <= tol * max|rhs| */
    int poisson_max_iter; /* at most this many corrections per solve, in [0, 64] */
    int precision;        /* CC_FP64: everything FP64; CC_MIXED: FP32 primitives / fluxes / FFT */
} cc_euler3d_params;

typedef struct cc_euler3d cc_euler3d;

cc_status cc_euler3d_create(const cc_euler3d_params* p, cc_euler3d** out);
cc_status cc_euler3d_destroy(cc_euler3d* h);
cc_status cc_euler3d_upload(cc_euler3d* h, const double* rho, const double* mx, const double* my, const double* mz,
                            const double* E);
cc_status cc_euler3d_download(cc_euler3d* h, double* rho, double* mx, double* my, double* mz, double* E);
/* Gravity of the current state (any output may be null; phi has zero mean).  Syncs. */
cc_status cc_euler3d_gravity(cc_euler3d* h, double* gx, double* gy, double* gz, double* phi);
/* Refinement record over all solves so far: solves run, corrections of the last solve, the most any
 * needed and their total, the last and worst final relative residual, solves that stopped at max_iter. */
cc_status cc_euler3d_poisson_stats(cc_euler3d* h, int* solves, int* iters_last, int* iters_max, int* iters_total,
                                   double* resid_last, double* resid_worst, int* unconverged);
/* rhs (gravity included) of the current state.  Syncs. */
cc_status cc_euler3d_rhs(cc_euler3d* h, double* drho, double* dmx, double* dmy, double* dmz, double* dE);
cc_status cc_euler3d_compute_dt(cc_euler3d* h, double* dt);
cc_status cc_euler3d_rk2_step(cc_euler3d* h, double dt);
/* nsteps of the driver loop (dt on the device, clamp to t_end, Heun), graph-captured.  Syncs. */
cc_status cc_euler3d_run(cc_euler3d* h, int nsteps, double* t_inout, double t_end, double* dts);
/* {mass, x-, y-, z-momentum, total energy} * dx*dy*dz, reduced on the device.  Syncs. */
cc_status cc_euler3d_totals(cc_euler3d* h, double* out5);
cc_status cc_euler3d_sync(cc_euler3d* h);
/* Per-kernel HIP-event timing ("prim", "face", "update", "mean", "fft_x", "fft_y", "fft_z",
 * "check", "grad", "dt"). */
cc_status cc_euler3d_set_profiling(cc_euler3d* h, int on);
cc_status cc_euler3d_kernel_time(cc_euler3d* h, const char* kernel, int* count, double* total_ms);
cc_status cc_euler3d_reset_profiling(cc_euler3d* h);

/* ============================================================================================
 * Initial conditions (host only: no device is needed).  The reference builds its ICs inline in
 * main(); these functions restate them, with their knobs gathered in structs whose defaults
 * (cc_*_ic_defaults) are the reference values, and add the benchmark-problem ICs.  Output arrays
 * are caller-owned, in the layouts of the matching upload calls.
 * ========================================================================================== */

/* IC knobs of M2 (Params M2:74-85, gamma_eff of M2:1011). */
typedef struct cc_mhd2d_ic {
    double rho_bg, rho_core, r_core, r_edge, P0, vin, rot_omega, B0x, Bshear, gamma_eff;
} cc_mhd2d_ic;
void cc_mhd2d_ic_defaults(cc_mhd2d_ic* k);
/* The default IC of M2 (M2:954-1048): face B from a corner vector potential Az, tanh core with
 * infall and rotation, prim_to_cons + floors. */
cc_status cc_mhd2d_ic_reference(const cc_mhd2d_params* p, const cc_mhd2d_ic* k, double* rho, double* mx, double* my,
                                double* mz, double* E, double* Bz, double* Bx_f, double* By_f);

/* Gaussian collapse (no reference counterpart; the 2D benchmark problem): rho = 1 + (contrast-1)
 * exp(-r^2/width^2) about the box centre, p = P0, (u, v) = vamp * (z0, z1) with one Box-Muller pair
 * per cell in row-major order from std::mt19937(seed) (u1 = (r1+0.5)/2^32, u2 = (r2+0.5)/2^32,
 * z0 = sqrt(-2 ln u1) cos(2 pi u2), z1 = sqrt(-2 ln u1) sin(2 pi u2)), uniform Bx = B0. */
typedef struct cc_collapse2d_ic {
    unsigned seed;
    double contrast, width, P0, vamp, B0;
} cc_collapse2d_ic;
void cc_collapse2d_ic_defaults(cc_collapse2d_ic* k); /* seed 42, contrast 10, width 0.1, P0 0.1, vamp 1e-3, B0 0.05 */
cc_status cc_mhd2d_ic_collapse(const cc_mhd2d_params* p, const cc_collapse2d_ic* k, double* rho, double* mx, double* my,
                               double* mz, double* E, double* Bz, double* Bx_f, double* By_f);

/* The batch of collapse instances (no reference counterpart): for instance i, from std::mt19937(seed)
 * uniform draws u = (r + 0.5) / 2^32 in the order u_N, u_c, u_eta, u_B: grid N[i] = round(nmin *
 * (nmax / nmin)^u_N) (log-uniform), contrast 5 + 45 u_c, eta[i] = 1e-5 + (1e-3 - 1e-5) u_eta,
 * B0 = 0.01 + 0.09 u_B, IC seed 42 + i, width 0.1, P0 0.1, vamp 1e-3. */
cc_status cc_collapse2d_batch_sample(unsigned seed, int n, int nmin, int nmax, int* N, cc_collapse2d_ic* ic,
                                     double* eta);

/* IC knobs of M3 (Params M3:53, 76-88). */
typedef struct cc_mhd3d_ic {
    double rho_bg, rho_core, r_core, r_edge, P0, vin, rot_omega, B0x, B0z, Bshear, gamma_init;
} cc_mhd3d_ic;
void cc_mhd3d_ic_defaults(cc_mhd3d_ic* k);
/* The default IC of M3 (M3:1159-1272) through the EOS of p (hybrid for eos 0). */
cc_status cc_mhd3d_ic_reference(const cc_mhd3d_params* p, const cc_mhd3d_ic* k, double* rho, double* mx, double* my,
                                double* mz, double* E, double* Bx_f, double* By_f, double* Bz_f);

/* Solenoidal turbulence (no reference counterpart; the 3D benchmark problem, ideal EOS): rho = 1,
 * p = 1/gamma, uniform Bz = sqrt(2 p / beta); velocity from random-phase modes on integer wavevectors
 * kmin <= |k| <= kmax with |v(k)| = (0.5 + a) |k|^-2 and phase 2 pi b (a, b = mt19937 draws / 2^32,
 * component by component, modes in FFT order), projected solenoidal, inverse FFT, scaled to rms
 * Mach `mach`.  Nx, Ny, Nz powers of two. */
typedef struct cc_turbulence3d_ic {
    unsigned seed;
    int kmin, kmax;
    double mach, beta;
} cc_turbulence3d_ic;
void cc_turbulence3d_ic_defaults(cc_turbulence3d_ic* k); /* seed 42, k 1..8, Mach 1, beta 1 */
cc_status cc_mhd3d_ic_turbulence(const cc_mhd3d_params* p, const cc_turbulence3d_ic* k, double* rho, double* mx,
                                 double* my, double* mz, double* E, double* Bx_f, double* By_f, double* Bz_f);

/* init_random_blob of EH (EH:574-642) and its constants (EH:585-603). */
typedef struct cc_euler2d_ic {
    unsigned seed;
    double rho0, P0, amp_rho, amp_P, vmax, sigma, bump_amp;
} cc_euler2d_ic;
void cc_euler2d_ic_defaults(cc_euler2d_ic* k);
cc_status cc_euler2d_ic_reference(const cc_euler2d_params* p, const cc_euler2d_ic* k, double* rho, double* mx,
                                  double* my, double* E);

/* The collapse IC of NS (NS:638-688) and its constants (Params::seed NS:69, NS:642-660); AoS output. */
typedef struct cc_ns2d_ic {
    unsigned seed;
    double rho0, P0, A, sigma_frac, Arho, Mach0;
} cc_ns2d_ic;
void cc_ns2d_ic_defaults(cc_ns2d_ic* k);
cc_status cc_ns2d_ic_reference(const cc_ns2d_params* p, const cc_ns2d_ic* k, double* U);

/* The uniform sphere of the 3D Euler benchmark: rho = rho_in inside |x - centre| < radius, rho_out
 * outside, times (1 + noise z) with one standard normal z per cell (Box-Muller pairs on
 * std::mt19937(seed), u = (r + 0.5)/2^32, cells in storage order take z0, z1 of successive pairs);
 * p = P0, at rest. */
typedef struct cc_sphere3d_ic {
    unsigned seed;
    double rho_in, rho_out, radius, noise, P0;
} cc_sphere3d_ic;
void cc_sphere3d_ic_defaults(cc_sphere3d_ic* k);
cc_status cc_euler3d_ic_sphere(const cc_euler3d_params* p, const cc_sphere3d_ic* k, double* rho, double* mx,
                               double* my, double* mz, double* E);

/* ============================================================================================
 * Measurement helper (no reference counterpart): a plain device-to-device copy of nbytes,
 * launched reps times; avg_ms receives the mean kernel time.  Its rocprofv3 FETCH_SIZE /
 * WRITE_SIZE readings per dispatch calibrate those counters against a known byte count.
 * ========================================================================================== */
cc_status cc_copy_calibration(size_t nbytes, int reps, double* avg_ms);

#ifdef __cplusplus
}
#endif
#endif /* CCAMD_H */
