// diag.hpp -- in-situ diagnostics shared by the solver handles: snapshot-frame buffers and
// per-field reductions (sums, L1 / Linf differences) on the device.
#pragma once
#include "common.hpp"

namespace cc {

constexpr int kDiagMaxFields = 12;

struct FieldPtrs {
    const double* p[kDiagMaxFields];
};

// Per-field statistics of n values, reduced on the device in a fixed order (so repeated calls on
// equal data give equal results):
//   without b: out[3f..3f+2] = {sum a, sum |a|, max |a|}
//   with b:    out[3f..3f+2] = {sum |a-b|, sum |b|, max |a-b|}
// The result is copied to host_out (3 * nf doubles) and the stream is synchronised.
class FieldStats {
   public:
    cc_status run(const FieldPtrs& a, const FieldPtrs* b, int nf, size_t n, hipStream_t s, double* host_out);
    void destroy();
    ~FieldStats() { destroy(); }

   private:
    double* part_ = nullptr;
    double* res_ = nullptr;
    int cap_fields_ = 0;
};

// A lazily allocated device buffer of nf planes of n cells, holding the last snapshot frame.
struct FrameBuffer {
    double* d = nullptr;
    size_t cap = 0;
    cc_status ensure(size_t count);
    void destroy();
    ~FrameBuffer() { destroy(); }
};

// Compare the frames of two handles: a's and b's frame buffers hold nf planes of n cells each.
// out[3f] = mean |a-b|, out[3f+1] = sum |a-b| / sum |b| (0 if sum |b| = 0), out[3f+2] = max |a-b|.
cc_status compare_frames(FieldStats& st, const double* fa, const double* fb, int nf, size_t n, hipStream_t s,
                         double* out);

}  // namespace cc
