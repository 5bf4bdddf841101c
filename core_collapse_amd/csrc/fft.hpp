// fft.hpp -- batched radix-2 FFT lines in LDS (gfx950), the on-device form of fft1d / fft2d of
// navier_stokes_2d_hllc_poisson.cpp (NS:152-195).
//
// The reference's fft1d is iterative radix-2 decimation in time after a bit-reversal permutation;
// its twiddles come from repeated multiplication w *= wlen (NS:162-171), so their rounding grows
// with the stage length.  The plan builds exactly that twiddle sequence on the host (same libm
// cos/sin, same complex product (ac - bd, ad + bc)) and the kernels run the same butterflies in the
// same stages, so a line transform is bit-identical to the reference's (kernels are compiled
// without FP contraction).  fft2d = all rows, then all columns; each launch transforms LPB lines
// of one axis, staged in LDS (line length <= 4096).
#pragma once
#include <vector>

#include "common.hpp"

namespace cc {

constexpr int kFftMaxLen = 4096;

struct FftAxis {
    int n = 0, logn = 0, lpb = 1;
    size_t tw_fwd = 0, tw_inv = 0;  // offsets (in complex entries) of the two twiddle tables
};

// Input / output forms of one lines pass.
enum FftIn { FFT_IN_COMPLEX = 0, FFT_IN_REAL_MINUS_MEAN = 1 };
enum FftOut { FFT_OUT_COMPLEX = 0, FFT_OUT_REAL = 1, FFT_OUT_GREEN = 2 };

struct FftGreen {  // A *= (-(4 pi G)) / k^2 with continuous k = 2 pi n / L, k = 0 -> 0 (NS:300-320)
    double neg4piG, two_pi, Lx, Ly;
};

class Fft2D {
   public:
    // Line lengths, launch shapes and twiddle-table offsets only (no allocation).
    cc_status plan(int Nx, int Ny);
    cc_status init(int Nx, int Ny);
    void destroy();
    ~Fft2D() { destroy(); }
    // One pass over all lines of one axis (cols = false: rows of length Nx).  `data` is the
    // interleaved complex array (Nx*Ny, q = i + Nx*j); in_real / mean / out_real as selected.
    cc_status pass(bool cols, bool inverse, FftIn in, FftOut out, double* data, const double* in_real,
                   const double* mean, double* out_real, const FftGreen* green, hipStream_t s);
    int Nx = 0, Ny = 0;
    FftAxis ax, ay;
    double* tw = nullptr;  // device twiddles, complex interleaved

    // Number of complex twiddle entries of the plan (all four tables).
    size_t twiddle_entries() const { return 2 * (size_t)(Nx - 1) + 2 * (size_t)(Ny - 1) + 1; }

   private:
    static void build_twiddles(int n, bool inverse, std::vector<double>& out);
};

}  // namespace cc
