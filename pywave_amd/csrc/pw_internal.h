// pw_internal.h -- shared definitions of the pywave_amd HIP library (gfx950 / MI355X).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/pywave_amd.h"

struct pw_ctx {
    int device = 0;
    hipStream_t stream = nullptr;   // library-owned stream (default priority)
    hipStream_t ext_stream = nullptr;  // caller stream (may be the null stream) when use_ext
    bool use_ext = false;
    std::string err;
    // autotuned 3D step choice per padded grid shape: (nz2, nx2, ny2) -> (variant, y chunk)
    std::map<std::tuple<int, int, int>, std::pair<int, int>> tuned;
    // autotuned 2D tile widths per workload key (kind, padded shapes, shots per launch and what
    // else decides which tiles run the general loop) -> the widths a plan's timing sweep chose,
    // then that sweep's log (kernel, width, ms per launch)
    std::map<std::vector<long long>, std::pair<std::vector<int>, std::vector<float>>> tuned2d;
    unsigned* d_count = nullptr;  // device counter of pw::count_nonfinite (allocated on first use)
    hipStream_t s() const { return use_ext ? ext_stream : stream; }
};

// Error helpers: every HIP call is checked; the message lands in ctx->err and the
// entry point returns a negative code.
#define PW_FAIL(ctx, code, msg)                                   \
    do {                                                          \
        if (ctx) (ctx)->err = (msg);                              \
        return (code);                                            \
    } while (0)

#define PW_HIP(ctx, call)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            if (ctx)                                                                       \
                (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);            \
            return PW_ERR_HIP;                                                             \
        }                                                                                  \
    } while (0)

namespace pw {

// kiss_fft_next_fast_size (pywave/src/wave_kissfft.c:412-424)
int next_fast_size(int n);

// padded extents of fdm/fdm2d (pywave/src/wave_fft2.c:45-51, wave_fft3.c:45-60)
void padded_sizes(int nz, int nx, int ny, int* nz2, int* nx2, int* ny2);

// abc_cal weights (pywave/src/wave_abc.c:141-143)
void abc_weights(int nb, float c, float* w);

// Zero-phase Ricker of pywave/src/wave_ricker.c:22-59 + wave_freqfilt.c:141-164 evaluated as
// the closed-form cosine sum of the filtered spike, in double precision.
int ricker(int nt, float dt, float f0, float t0, float A, float* out, std::string* err);

// Per-index damping factor pairs (lo, hi): the factor each face multiplies into index i of an
// axis of padded length n2 with nb boundary cells (1.0 where a face does not apply).
void damping_pairs(int n2, int nb_lo, int nb_hi, const float* w_lo, const float* w_hi,
                   std::vector<float>& out2);

template <typename T>
inline T ceil_div(T a, T b) { return (a + b - 1) / b; }

// Number of Inf/NaN values among n elements (elem_bytes 4 = f32, 8 = f64) of device array d,
// counted on the context's stream; synchronises the stream.  The result watchdog of every plan.
int count_nonfinite(pw_ctx* ctx, const void* d, size_t n, int elem_bytes, unsigned* count);

// PW_ERR_NONFINITE with a message naming what blew up, if count_nonfinite finds any
int check_finite(pw_ctx* ctx, const void* d, size_t n, int elem_bytes, const char* what);

}  // namespace pw

// Pseudo-spectral Laplacian of padded fields (pw_psp.hip): n1 (z, even, real transform), n2 (x),
// n3 (y; 1 in 2D) are the padded extents; fields use the plans' layout (z stride ldz, plane
// stride ldxy).  ps_apply enqueues wave = IFFT(-(2 pi |k|)^2 FFT(field)) on stream s.
struct PsLap;
PsLap* ps_create(int n1, int n2, int n3, int ldz, long long ldxy, float dz, float dx, float dy, std::string& err);
hipError_t ps_apply(PsLap* P, const float* field, float* wave, hipStream_t s);
void ps_destroy(PsLap* P);
