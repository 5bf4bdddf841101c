/*
 * oracle.h -- C API of the CPU parity oracle (test infrastructure only).
 *
 * The oracle restates the reference solver steps on the CPU in FP64; it is
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * only, as the checker.  The product library (libccamd.so) never uses it.
 *
 * The same API is implemented by oracle/_ref/libref_*.so (thin wrappers that
 * compile the reference programs' own sources, built in this container by
 * oracle/Makefile), so tests can compare the restatement with the reference.
 */
#ifndef CC_ORACLE_H
#define CC_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Compensated face update (ct_comp; no reference counterpart).  TwoSum (Knuth, TAOCP vol. 2, 4.2.2):
 * s = fl(a + b) and s + e = a + b exactly.  A face is hi + lo with lo the FP32 rounding error of hi; CT adds
 * its increment to the pair, the Heun average halves the exact sum of two pairs. */
static inline void or_two_sum(double a, double b, double* s, double* e) {
    const double ss = a + b, bb = ss - a;
    *s = ss;
    *e = (a - (ss - bb)) + (b - bb);
}
static inline void or_ct_add(double hi, float lo, double d, double* hi1, float* lo1) {
    double e;
    or_two_sum(hi, (double)lo + d, hi1, &e);
    *lo1 = (float)e;
}
static inline void or_ct_avg(double hn, float ln, double h2, float l2, double* hi, float* lo) {
    double s, e, h, l;
    or_two_sum(hn, h2, &s, &e);
    or_two_sum(s, e + ((double)ln + (double)l2), &h, &l);
    *hi = 0.5 * h;
    *lo = (float)(0.5 * l);
}

/* ---- 2D resistive MHD (mhd2d_resistive_hlld_ct_gravity.cpp, Params M2:44-86) ---- */
typedef struct {
    double gamma, cfl, x0, x1, y0, y1;
    int Nx, Ny;
    double floor_rho, floor_P;
    int periodic_x, periodic_y;
    double eta;
    int gravity_on;
    double G;
    int sor_iters;
    double sor_omega;
    /* No reference counterpart (0 keeps the reference's arithmetic; the reference harness ignores them):
     * hlld_guard > 0 sends an HLLD face whose d_L or d_R (M2:419-442) is below
     * hlld_guard * max(Bn^2, |rho s (S - S_M)|) to the HLL branch (M2:449); ct_comp carries each face's
     * rounding error in an FP32 low word through CT and the Heun average (state Bx_lo, By_lo). */
    double hlld_guard;
    int ct_comp;
} or_m2_params;

typedef struct {
    double *rho, *mx, *my, *mz, *E, *Bz; /* Nx*Ny, q = j*Nx+i        */
    double *Bx_f;                        /* (Nx+1)*Ny, j*(Nx+1)+i_f */
    double *By_f;                        /* Nx*(Ny+1), j_f*Nx+i     */
    float *Bx_lo, *By_lo;                /* low words of the faces (ct_comp only; else unused) */
} or_m2_state;

void or_m2_hlld_x(const or_m2_params* p, const double* WL, const double* WR, double Bn, double* out8);
void or_m2_hlld_y(const or_m2_params* p, const double* WL, const double* WR, double Bn, double* out8);
void or_m2_cons_to_prim(const or_m2_params* p, const double* c8, double* W8);
double or_m2_mc_limiter(double dm, double dp);
double or_m2_fast_speed_x(const or_m2_params* p, const double* W8);
double or_m2_fast_speed_y(const or_m2_params* p, const double* W8);
void or_m2_poisson(const or_m2_params* p, const double* rho, double dx, double dy, double* gx, double* gy);
void or_m2_sor_phi(const or_m2_params* p, const double* rho, double dx, double dy, double* phi);
double or_m2_compute_dt(const or_m2_params* p, const or_m2_state* U, double dx, double dy);
void or_m2_advance_stage(const or_m2_params* p, const or_m2_state* U0, or_m2_state* U1, const double* gx,
                         const double* gy, double dx, double dy, double dt);
void or_m2_rk2_step(const or_m2_params* p, or_m2_state* U, const double* gx, const double* gy, double dx, double dy,
                    double dt);
void or_m2_divb(const or_m2_params* p, const double* Bx_f, const double* By_f, double dx, double dy, double* out);
double or_m2_run(const or_m2_params* p, or_m2_state* U, double dx, double dy, int nsteps, double t0, double t_end,
                 double* dts_out);
int or_m2_last_branch(void); /* restatement only: branch of the last hlld call (see oracle_mhd2d.c) */
/* the snapshot frame of the store lambda (M2:1065-1106): 10 planes {rho,P,Mach,u,v,vabs,Bx,By,Bz,divB} */
void or_m2_frame(const or_m2_params* p, const or_m2_state* U, double dx, double dy, double* out);

/* ---- 3D resistive MHD (mhd3d_resistive_hlld_ct.cpp, Params M3:43-89) ---- */
typedef struct {
    double gamma1, gamma2, K1, rho_nuc, gamma_th, rho_bounce, gamma_bounce, Kbounce, gamma_init, cfl;
    double x0, x1, y0, y1, z0, z1;
    int Nx, Ny, Nz;
    double floor_rho, floor_P;
    int periodic_x, periodic_y, periodic_z;
    double eta;
    int ideal_eos;      /* 0: the reference hybrid EOS (M3:114-212); 1: ideal gas with gamma_ideal */
    double gamma_ideal; /* used when ideal_eos = 1 (no reference counterpart; see the EOS note in oracle_mhd3d.c) */
    double hlld_guard;  /* as in or_m2_params (M3:463-484, HLL branch M3:491) */
    int ct_comp;        /* as in or_m2_params (CT M3:941-1015, Heun M3:1042-1065; state B*_lo) */
} or_m3_params;

typedef struct {
    double *rho, *mx, *my, *mz, *E; /* Nx*Ny*Nz, (k*Ny+j)*Nx+i          */
    double *Bx_f;                   /* (Nx+1)*Ny*Nz, (k*Ny+j)*(Nx+1)+i_f */
    double *By_f;                   /* Nx*(Ny+1)*Nz, (k*(Ny+1)+j_f)*Nx+i */
    double *Bz_f;                   /* Nx*Ny*(Nz+1), (k_f*Ny+j)*Nx+i     */
    float *Bx_lo, *By_lo, *Bz_lo;   /* low words of the faces (ct_comp only; else unused) */
} or_m3_state;

void or_m3_hlld_x(const or_m3_params* p, const double* WL, const double* WR, double Bn, double* out7);
void or_m3_cons_to_prim(const or_m3_params* p, const double* c8, double* W8);
void or_m3_eos_cold(const or_m3_params* p, double rho, double* out3);
double or_m3_sound_speed2(const or_m3_params* p, const double* W8);
double or_m3_total_energy(const or_m3_params* p, const double* W8);
double or_m3_compute_dt(const or_m3_params* p, const or_m3_state* U, double dx, double dy, double dz);
void or_m3_advance_stage(const or_m3_params* p, const or_m3_state* U0, or_m3_state* U1, double dx, double dy,
                         double dz, double dt);
void or_m3_rk2_step(const or_m3_params* p, or_m3_state* U, double dx, double dy, double dz, double dt);
void or_m3_divb(const or_m3_params* p, const or_m3_state* U, double dx, double dy, double dz, double* out);
double or_m3_run(const or_m3_params* p, or_m3_state* U, double dx, double dy, double dz, int nsteps, double t0,
                 double t_end, double* dts_out);
/* the snapshot frame of the store block (M3:1293-1325): 11 volumes {rho,P,Mach,u,v,w,vabs,Bx,By,Bz,divB} */
void or_m3_frame(const or_m3_params* p, const or_m3_state* U, double dx, double dy, double dz, double* out);

/* ---- 2D Euler HLLC + SOR gravity (euler2d_gravity_hllc.cpp, Params EH:29-56) ---- */
typedef struct {
    double gamma, cfl, x0, x1, y0, y1;
    int Nx, Ny;
    double floor_rho, floor_P;
    int periodic_x, periodic_y;
    int gravity_on;
    double G;
    int max_iter;   /* EH:329 */
    double tol;     /* EH:330 */
    double omega;   /* EH:331 */
} or_eh_params;

typedef struct { double *rho, *mx, *my, *E; } or_eh_state;

void or_eh_hllc_x(const or_eh_params* p, const double* WL4, const double* WR4, double* out4);
void or_eh_hllc_y(const or_eh_params* p, const double* WL4, const double* WR4, double* out4);
int or_eh_poisson(const or_eh_params* p, const or_eh_state* U, double* gx, double* gy, double* phi_out);
double or_eh_compute_dt(const or_eh_params* p, const or_eh_state* U);
void or_eh_rhs(const or_eh_params* p, const or_eh_state* U, or_eh_state* dU);
void or_eh_rk2_step(const or_eh_params* p, or_eh_state* U, double dt);
double or_eh_run(const or_eh_params* p, or_eh_state* U, int nsteps, double t0, double t_end, double* dts_out);
/* the snapshot frame of the store lambda (EH:661-691): planes {rho, P, Mach} */
void or_eh_frame(const or_eh_params* p, const or_eh_state* U, double* out);
/* reference harness only: the program's default IC (init_random_blob, EH:574-642) */
void or_eh_init_random_blob(const or_eh_params* p, or_eh_state* U);

/* EH rhs (EH:406-527) with the gravity field given instead of solved (gx, gy of N cells) */
void or_eh_rhs_g(const or_eh_params* p, const or_eh_state* U, or_eh_state* dU, const double* gx, const double* gy);

/* ---- 3D Euler HLLC + FFT gravity (the periodic uniform-sphere benchmark; no reference program:
 *      the EH step extended to three axes, oracle_e3.c) ---- */
typedef struct {
    double gamma, cfl, x0, x1, y0, y1, z0, z1;
    int Nx, Ny, Nz;
    double floor_rho, floor_P;
    int gravity_on;
    double G;
} or_e3_params;

typedef struct { double *rho, *mx, *my, *mz, *E; } or_e3_state;

void or_e3_hllc(const or_e3_params* p, int axis, const double* WL5, const double* WR5, double* out5);
void or_e3_rhs_g(const or_e3_params* p, const or_e3_state* U, or_e3_state* dU, const double* gx, const double* gy,
                 const double* gz);
void or_e3_poisson(const or_e3_params* p, const double* rho, double* gx, double* gy, double* gz, double* phi_out);
void or_e3_rhs(const or_e3_params* p, const or_e3_state* U, or_e3_state* dU);
double or_e3_compute_dt(const or_e3_params* p, const or_e3_state* U);
void or_e3_rk2_step(const or_e3_params* p, or_e3_state* U, double dt);
double or_e3_run(const or_e3_params* p, or_e3_state* U, int nsteps, double t0, double t_end, double* dts_out);

/* ---- 2D Navier-Stokes HLLC + FFT gravity (navier_stokes_2d_hllc_poisson.cpp, Params NS:54-86) ---- */
typedef struct {
    double gamma, cfl, x0, x1, y0, y1;
    int Nx, Ny;
    double floor_rho, floor_P;
    int viscosity_on;
    double mu, cfl_visc;
    int gravity_on;
    double G;
} or_ns_params;

/* AoS state as in the reference (NS:88-93): U[q*4 + {0:d,1:mx,2:my,3:E}], q = i + Nx*j */
void or_ns_hllc_x(const or_ns_params* p, const double* WL4, const double* WR4, double* out4);
void or_ns_hllc_y(const or_ns_params* p, const double* WL4, const double* WR4, double* out4);
void or_ns_fft2d(int Nx, int Ny, int inverse, double* data); /* interleaved complex, q = i + Nx*j */
int or_ns_poisson_fft(const or_ns_params* p, const double* rho, double* phi);
void or_ns_gravity_accel(const or_ns_params* p, const double* phi, double* gx, double* gy);
void or_ns_compute_rhs(const or_ns_params* p, const double* U, double* dU);
double or_ns_compute_dt(const or_ns_params* p, const double* U);
void or_ns_heun_step(const or_ns_params* p, double* U, double dt);
double or_ns_run(const or_ns_params* p, double* U, int nsteps, double t0, double t_end, double* dts_out);
/* the snapshot frame of the store lambda (NS:698-749): planes {rho,P,Mach,u,v,momx,momy,Phi} */
int or_ns_frame(const or_ns_params* p, const double* U, double* out);
/* reference harness only: libstdc++'s std::normal_distribution / uniform_real_distribution over
 * std::mt19937(seed), n draws (the generators of the ICs at NS:638-639 and EH:585-587) */
void or_std_normal(unsigned seed, int n, double* out);
void or_std_uniform(unsigned seed, int n, double a, double b, double* out);

#ifdef __cplusplus
}
#endif
#endif
