// pw_core.hip -- context, error reporting and host-side set-up (geometry, damping, wavelet).
#include <cmath>
#include <cstring>

#include "pw_internal.h"

namespace pw {

int next_fast_size(int n)
{
    if (n < 1) n = 1;
    for (;; ++n) {
        int m = n;
        while (m % 2 == 0) m /= 2;
        while (m % 3 == 0) m /= 3;
        while (m % 5 == 0) m /= 5;
        if (m <= 1) return n;
    }
}

void padded_sizes(int nz, int nx, int ny, int* nz2, int* nx2, int* ny2)
{
    // first axis uses the real-to-complex length 2*(nk-1), nk = nfs((n+1)/2)+1
    if (nz2) *nz2 = 2 * next_fast_size((nz + 1) / 2);
    if (nx2) *nx2 = next_fast_size(nx);
    if (ny2) *ny2 = ny <= 1 ? 1 : next_fast_size(ny);
}

void abc_weights(int nb, float c, float* w)
{
    for (int ib = 0; ib < nb; ++ib) {
        // f32 products, exp in double, stored to f32 (pywave/src/wave_abc.c:142)
        float e = -c * c * (nb - 1 - ib) * (nb - 1 - ib);
        w[ib] = (float)std::exp((double)e);
    }
}

int ricker(int nt, float dt, float f0, float t0, float A, float* out, std::string* err)
{
    if (nt <= 0 || !(dt > 0.f)) {
        if (err) *err = "ricker: nt and dt must be positive";
        return PW_ERR_ARG;
    }
    float q = t0 / dt;  // onset index (int)(t0/dt) in f32 (pywave/src/wave_fdm.c:508)
    int k0 = (int)q;
    if (!(q >= 0.f) || k0 >= nt) {
        if (err) *err = "ricker: onset index int(t/dt) must lie in [0, nt)";
        return PW_ERR_ARG;
    }
    const int nfft = 2 * next_fast_size((nt + 1) / 2);
    const int nw = nfft / 2 + 1;
    // spectral shape in f32 as the reference forms it (pywave/src/wave_ricker.c:33-51)
    float freq = f0 * dt;
    float dw = 1. / (nfft * freq);
    std::vector<double> S(nw);
    for (int iw = 0; iw < nw; ++iw) {
        float w = iw * dw;
        w *= w;
        S[iw] = (double)(w * expf(1 - w) / nfft);
    }
    // The filtered spike is A * g[(n - k0) mod nfft] with the real, even kernel
    //   g[m] = S0 + 2 sum_{k=1}^{nfft/2-1} S_k cos(2 pi k m / nfft) + S_{nfft/2} (-1)^m,
    // i.e. irfft(rfft(spike) * S) with the unnormalised inverse; evaluated in double.
    std::vector<double> cs(nfft);
    for (int k = 0; k < nfft; ++k) cs[k] = std::cos(2.0 * M_PI * (double)k / (double)nfft);
    for (int n = 0; n < nt; ++n) {
        long long m = ((long long)n - k0) % nfft;
        if (m < 0) m += nfft;
        double g = S[0] + S[nw - 1] * ((m & 1) ? -1.0 : 1.0);
        for (int k = 1; k < nw - 1; ++k) g += 2.0 * S[k] * cs[(k * m) % nfft];
        out[n] = (float)((double)A * g);
    }
    return PW_OK;
}

void damping_pairs(int n2, int nb_lo, int nb_hi, const float* w_lo, const float* w_hi,
                   std::vector<float>& out2)
{
    out2.assign(2 * (size_t)n2, 1.0f);
    for (int i = 0; i < n2; ++i) {
        if (i < nb_lo) out2[2 * i] = w_lo[i];
        int k = n2 - 1 - i;  // faces indexed from the padded extent (pywave/src/wave_abc.c:234,260,286)
        if (k < nb_hi) out2[2 * i + 1] = w_hi[k];
    }
}

}  // namespace pw

extern "C" {

pw_ctx* pw_create(int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return nullptr;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    pw_ctx* c = new pw_ctx;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return nullptr;
    }
    return c;
}

void pw_destroy(pw_ctx* ctx)
{
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) {
        hipStreamSynchronize(ctx->stream);
        hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

const char* pw_last_error(pw_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int pw_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int pw_set_stream(pw_ctx* ctx, void* stream)
{
    if (!ctx) return PW_ERR_ARG;
    ctx->ext_stream = (hipStream_t)stream;
    return PW_OK;
}

int pw_synchronize(pw_ctx* ctx)
{
    if (!ctx) return PW_ERR_ARG;
    PW_HIP(ctx, hipStreamSynchronize(ctx->s()));
    return PW_OK;
}

int pw_next_fast_size(int n) { return pw::next_fast_size(n); }

void pw_padded_sizes(int nz, int nx, int ny, int* nz2, int* nx2, int* ny2)
{
    pw::padded_sizes(nz, nx, ny, nz2, nx2, ny2);
}

void pw_abc_weights(int nb, float c, float* w) { pw::abc_weights(nb, c, w); }

int pw_ricker(int nt, float dt, float f0, float t0, float A, float* out)
{
    return pw::ricker(nt, dt, f0, t0, A, out, nullptr);
}

}  // extern "C"
