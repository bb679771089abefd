// pw_core.hip -- context, error reporting and host-side set-up (geometry, damping, wavelet).
#include <algorithm>
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

// ---------------------------------------------------------------------------------------------
// f32 mixed-radix FFT following the KISS FFT algorithm vendored by the reference
// (pywave/src/wave_kissfft.c:37-560): the same factorisation (radix 4 first, then 2, 3, 5,
// odd primes), the same recursive decimation, the same butterfly arithmetic in f32 and the same
// double-precision twiddle generation rounded to f32.  Reproducing its rounding makes the
// wavelet, and therefore every trace and wavefield, bit-identical to the reference.
namespace {

struct cpx {
    float r, i;
};

inline cpx cmul(cpx a, cpx b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
inline cpx cadd(cpx a, cpx b) { return {a.r + b.r, a.i + b.i}; }
inline cpx csub(cpx a, cpx b) { return {a.r - b.r, a.i - b.i}; }

struct CFft {
    int n = 0;
    bool inverse = false;
    std::vector<cpx> tw;
    std::vector<int> fac;  // p0, m0, p1, m1, ...

    CFft(int n_, bool inv) : n(n_), inverse(inv), tw(n_)
    {
        const double pi = 3.141592653589793238462643383279502884197169399375105820974944;
        for (int i = 0; i < n; ++i) {
            double phase = -2 * pi * i / n;
            if (inverse) phase *= -1;
            tw[i] = {(float)std::cos(phase), (float)std::sin(phase)};
        }
        int m = n, p = 4;
        double fsq = std::floor(std::sqrt((double)m));
        do {
            while (m % p) {
                p = p == 4 ? 2 : (p == 2 ? 3 : p + 2);
                if (p > fsq) p = m;
            }
            m /= p;
            fac.push_back(p);
            fac.push_back(m);
        } while (m > 1);
    }

    void bfly2(cpx* F, size_t fs, int m) const
    {
        for (int k = 0; k < m; ++k) {
            cpx t = cmul(F[m + k], tw[k * fs]);
            F[m + k] = csub(F[k], t);
            F[k] = cadd(F[k], t);
        }
    }

    void bfly3(cpx* F, size_t fs, size_t m) const
    {
        const cpx epi3 = tw[fs * m];
        for (size_t k = 0; k < m; ++k) {
            cpx s1 = cmul(F[k + m], tw[k * fs]);
            cpx s2 = cmul(F[k + 2 * m], tw[2 * k * fs]);
            cpx s3 = cadd(s1, s2);
            cpx s0 = csub(s1, s2);
            F[k + m].r = (float)((double)F[k].r - (double)s3.r * .5);  // HALF_OF promotes to double
            F[k + m].i = (float)((double)F[k].i - (double)s3.i * .5);
            s0.r *= epi3.i;
            s0.i *= epi3.i;
            F[k] = cadd(F[k], s3);
            F[k + 2 * m].r = F[k + m].r + s0.i;
            F[k + 2 * m].i = F[k + m].i - s0.r;
            F[k + m].r -= s0.i;
            F[k + m].i += s0.r;
        }
    }

    void bfly4(cpx* F, size_t fs, size_t m) const
    {
        for (size_t k = 0; k < m; ++k) {
            cpx s0 = cmul(F[k + m], tw[k * fs]);
            cpx s1 = cmul(F[k + 2 * m], tw[2 * k * fs]);
            cpx s2 = cmul(F[k + 3 * m], tw[3 * k * fs]);
            cpx s5 = csub(F[k], s1);
            F[k] = cadd(F[k], s1);
            cpx s3 = cadd(s0, s2);
            cpx s4 = csub(s0, s2);
            F[k + 2 * m] = csub(F[k], s3);
            F[k] = cadd(F[k], s3);
            if (inverse) {
                F[k + m] = {s5.r - s4.i, s5.i + s4.r};
                F[k + 3 * m] = {s5.r + s4.i, s5.i - s4.r};
            } else {
                F[k + m] = {s5.r + s4.i, s5.i - s4.r};
                F[k + 3 * m] = {s5.r - s4.i, s5.i + s4.r};
            }
        }
    }

    void bfly5(cpx* F, size_t fs, int m) const
    {
        const cpx ya = tw[fs * m], yb = tw[fs * 2 * m];
        cpx *F0 = F, *F1 = F + m, *F2 = F + 2 * m, *F3 = F + 3 * m, *F4 = F + 4 * m;
        for (int u = 0; u < m; ++u) {
            cpx s0 = F0[u];
            cpx s1 = cmul(F1[u], tw[u * fs]);
            cpx s2 = cmul(F2[u], tw[2 * u * fs]);
            cpx s3 = cmul(F3[u], tw[3 * u * fs]);
            cpx s4 = cmul(F4[u], tw[4 * u * fs]);
            cpx s7 = cadd(s1, s4), s10 = csub(s1, s4), s8 = cadd(s2, s3), s9 = csub(s2, s3);
            F0[u].r += s7.r + s8.r;
            F0[u].i += s7.i + s8.i;
            cpx s5 = {s0.r + s7.r * ya.r + s8.r * yb.r, s0.i + s7.i * ya.r + s8.i * yb.r};
            cpx s6 = {s10.i * ya.i + s9.i * yb.i, -(s10.r * ya.i) - s9.r * yb.i};
            F1[u] = csub(s5, s6);
            F4[u] = cadd(s5, s6);
            cpx s11 = {s0.r + s7.r * yb.r + s8.r * ya.r, s0.i + s7.i * yb.r + s8.i * ya.r};
            cpx s12 = {-(s10.i * yb.i) + s9.i * ya.i, s10.r * yb.i - s9.r * ya.i};
            F2[u] = cadd(s11, s12);
            F3[u] = csub(s11, s12);
        }
    }

    void bfly_generic(cpx* F, size_t fs, int m, int p) const
    {
        std::vector<cpx> s(p);
        for (int u = 0; u < m; ++u) {
            for (int q1 = 0, k = u; q1 < p; ++q1, k += m) s[q1] = F[k];
            for (int q1 = 0, k = u; q1 < p; ++q1, k += m) {
                size_t twidx = 0;
                F[k] = s[0];
                for (int q = 1; q < p; ++q) {
                    twidx += fs * k;
                    if (twidx >= (size_t)n) twidx -= n;
                    F[k] = cadd(F[k], cmul(s[q], tw[twidx]));
                }
            }
        }
    }

    void work(cpx* Fout, const cpx* f, size_t fs, const int* fp) const
    {
        const int p = fp[0], m = fp[1];
        if (m == 1) {
            for (int j = 0; j < p; ++j) Fout[j] = f[j * fs];
        } else {
            for (int j = 0; j < p; ++j) work(Fout + j * m, f + j * fs, fs * p, fp + 2);
        }
        switch (p) {
            case 2: bfly2(Fout, fs, m); break;
            case 3: bfly3(Fout, fs, m); break;
            case 4: bfly4(Fout, fs, m); break;
            case 5: bfly5(Fout, fs, m); break;
            default: bfly_generic(Fout, fs, m, p); break;
        }
    }

    void run(const cpx* in, cpx* out) const { work(out, in, 1, fac.data()); }
};

// real transforms of even length via a half-length complex transform and split twiddles
struct RFft {
    int half;
    CFft sub;
    std::vector<cpx> sup;

    RFft(int nfft, bool inv) : half(nfft / 2), sub(nfft / 2, inv), sup(nfft / 4)
    {
        for (int i = 0; i < half / 2; ++i) {
            double phase = -3.14159265358979323846264338327 * ((double)(i + 1) / half + .5);
            if (inv) phase *= -1;
            sup[i] = {(float)std::cos(phase), (float)std::sin(phase)};
        }
    }

    void forward(const float* x, cpx* X) const  // X has half+1 bins
    {
        std::vector<cpx> t(half);
        sub.run(reinterpret_cast<const cpx*>(x), t.data());
        X[0].r = t[0].r + t[0].i;
        X[half].r = t[0].r - t[0].i;
        X[0].i = X[half].i = 0;
        for (int k = 1; k <= half / 2; ++k) {
            cpx fpk = t[k], fpnk = {t[half - k].r, -t[half - k].i};
            cpx f1k = cadd(fpk, fpnk), f2k = csub(fpk, fpnk);
            cpx w = cmul(f2k, sup[k - 1]);
            X[k].r = (float)((double)(f1k.r + w.r) * .5);
            X[k].i = (float)((double)(f1k.i + w.i) * .5);
            X[half - k].r = (float)((double)(f1k.r - w.r) * .5);
            X[half - k].i = (float)((double)(w.i - f1k.i) * .5);
        }
    }

    void inverse(const cpx* X, float* x) const
    {
        std::vector<cpx> t(half);
        t[0].r = X[0].r + X[half].r;
        t[0].i = X[0].r - X[half].r;
        for (int k = 1; k <= half / 2; ++k) {
            cpx fk = X[k], fnkc = {X[half - k].r, -X[half - k].i};
            cpx fek = cadd(fk, fnkc), tmp = csub(fk, fnkc);
            cpx fok = cmul(tmp, sup[k - 1]);
            t[k] = cadd(fek, fok);
            t[half - k] = csub(fek, fok);
            t[half - k].i *= -1;
        }
        sub.run(t.data(), reinterpret_cast<cpx*>(x));
    }
};

}  // namespace

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
    // spike, then zero-phase filtering by w^2 e^(1-w^2)/nfft over nfft = 2*nfs((nt+1)/2) samples
    // (pywave/src/wave_ricker.c:22-59, wave_freqfilt.c:141-164, wave_fdm.c:505-513)
    const int nfft = 2 * next_fast_size((nt + 1) / 2);
    const int nw = nfft / 2 + 1;
    float freq = f0 * dt;
    float dw = 1. / (nfft * freq);
    std::vector<cpx> shape(nw);
    for (int iw = 0; iw < nw; ++iw) {
        float w = iw * dw;
        w *= w;
        shape[iw] = {w * expf(1 - w) / nfft, 0.f};
    }
    std::vector<float> buf(nfft, 0.f);
    buf[k0] = A;
    std::vector<cpx> spec(nw);
    RFft fw(nfft, false), iv(nfft, true);
    fw.forward(buf.data(), spec.data());
    for (int iw = 0; iw < nw; ++iw) spec[iw] = cmul(spec[iw], shape[iw]);
    iv.inverse(spec.data(), buf.data());
    for (int n = 0; n < nt; ++n) out[n] = buf[n];
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

namespace {

template <typename U>
__global__ void k_count_nonfinite(size_t n, const U* __restrict__ a, U expo, unsigned* __restrict__ count)
{
    unsigned c = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        c += (a[i] & expo) == expo;  // exponent all ones: Inf or NaN
    if (c) atomicAdd(count, c);
}

__global__ void k_zero_u32(unsigned* p) { *p = 0; }

// Streaming probe: o = a (+ b (+ c)), R input streams and one output stream, 16 B per lane per
// stream with the non-temporal cache policy (each byte is touched once), grid-stride: the plain
// streaming rate against which the step kernels' traffic is read
typedef float f32x4 __attribute__((ext_vector_type(4)));
// (CACHED 0: non-temporal loads and stores; 1: the default policy, under which buffers that fit
// the Infinity Cache stay there between launches)
template <int R, int CACHED>
__global__ __launch_bounds__(256) void k_stream(long long n4, const f32x4* __restrict__ a,
                                                const f32x4* __restrict__ b, const f32x4* __restrict__ c,
                                                f32x4* __restrict__ o)
{
    auto ld = [](const f32x4* p) { return CACHED ? *p : __builtin_nontemporal_load(p); };
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n4; i += (long long)gridDim.x * 256) {
        f32x4 v = ld(&a[i]);
        if constexpr (R > 1) v = v + ld(&b[i]);
        if constexpr (R > 2) v = v + ld(&c[i]);
        if constexpr (CACHED) o[i] = v;
        else __builtin_nontemporal_store(v, &o[i]);
    }
}

}  // namespace

int pw::count_nonfinite(pw_ctx* ctx, const void* d, size_t n, int elem_bytes, unsigned* count)
{
    *count = 0;
    if (!n) return PW_OK;
    hipStream_t s = ctx->s();
    if (!ctx->d_count) PW_HIP(ctx, hipMalloc((void**)&ctx->d_count, sizeof(unsigned)));
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, ctx->d_count);
    const unsigned blocks = (unsigned)std::min<size_t>(2048, (n + 255) / 256);
    if (elem_bytes == 8)
        hipLaunchKernelGGL(k_count_nonfinite<unsigned long long>, dim3(blocks), dim3(256), 0, s, n,
                           (const unsigned long long*)d, 0x7ff0000000000000ull, ctx->d_count);
    else
        hipLaunchKernelGGL(k_count_nonfinite<unsigned>, dim3(blocks), dim3(256), 0, s, n, (const unsigned*)d,
                           0x7f800000u, ctx->d_count);
    PW_HIP(ctx, hipGetLastError());
    PW_HIP(ctx, hipMemcpyAsync(count, ctx->d_count, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    PW_HIP(ctx, hipStreamSynchronize(s));
    return PW_OK;
}

int pw::check_finite(pw_ctx* ctx, const void* d, size_t n, int elem_bytes, const char* what)
{
    unsigned c = 0;
    if (int rc = count_nonfinite(ctx, d, n, elem_bytes, &c)) return rc;
    if (c)
        PW_FAIL(ctx, PW_ERR_NONFINITE,
                std::string(what) + ": " + std::to_string(c) + " of " + std::to_string(n) +
                    " values are Inf or NaN (the run blew up; is the time step within the CFL limit?)");
    return PW_OK;
}

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
    if (ctx->d_count) hipFree(ctx->d_count);
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
    ctx->use_ext = true;
    return PW_OK;
}

int pw_use_own_stream(pw_ctx* ctx)
{
    if (!ctx) return PW_ERR_ARG;
    ctx->ext_stream = nullptr;
    ctx->use_ext = false;
    return PW_OK;
}

int pw_synchronize(pw_ctx* ctx)
{
    if (!ctx) return PW_ERR_ARG;
    PW_HIP(ctx, hipStreamSynchronize(ctx->s()));
    return PW_OK;
}

// Streaming bandwidth of the device: `reads` input streams (1, 2 or 3) and one output stream of n
// floats each (n a positive multiple of 4), `reps` timed launches after two untimed ones;
// *gbs = bytes moved per second / 1e9.  Allocates its buffers for the call.
int pw_stream_bandwidth(pw_ctx* ctx, long long n, int reads, int reps, int cached, double* gbs)
{
    if (!ctx || !gbs || n <= 0 || n % 4 || reads < 1 || reads > 3 || reps < 1 || cached < 0 || cached > 1)
        PW_FAIL(ctx, PW_ERR_ARG,
                "pw_stream_bandwidth: n must be a positive multiple of 4, reads 1-3, reps >= 1, cached 0 or 1");
    PW_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->s();
    float* buf[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto release = [&] {
        for (float* b : buf)
            if (b) (void)hipFree(b);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    hipError_t r = hipSuccess;
    for (int k = 0; k <= reads && r == hipSuccess; ++k) {
        r = hipMalloc((void**)&buf[k], (size_t)n * 4);
        if (r == hipSuccess) r = hipMemsetAsync(buf[k], 0, (size_t)n * 4, s);
    }
    if (r == hipSuccess) r = hipEventCreate(&e0);
    if (r == hipSuccess) r = hipEventCreate(&e1);
    if (r != hipSuccess) {
        release();
        PW_FAIL(ctx, PW_ERR_HIP, std::string("pw_stream_bandwidth: ") + hipGetErrorString(r));
    }
    const long long n4 = n / 4;
    const unsigned blocks = (unsigned)std::min<long long>(32768, (n4 + 255) / 256);
    const f32x4 *a = (const f32x4*)buf[0], *b = (const f32x4*)buf[reads > 1 ? 1 : 0],
                *c = (const f32x4*)buf[reads > 2 ? 2 : 0];
    f32x4* o = (f32x4*)buf[reads];
    auto launch = [&] {
        const int k = 2 * (reads - 1) + cached;
        switch (k) {
            case 0: hipLaunchKernelGGL((k_stream<1, 0>), dim3(blocks), dim3(256), 0, s, n4, a, b, c, o); break;
            case 1: hipLaunchKernelGGL((k_stream<1, 1>), dim3(blocks), dim3(256), 0, s, n4, a, b, c, o); break;
            case 2: hipLaunchKernelGGL((k_stream<2, 0>), dim3(blocks), dim3(256), 0, s, n4, a, b, c, o); break;
            case 3: hipLaunchKernelGGL((k_stream<2, 1>), dim3(blocks), dim3(256), 0, s, n4, a, b, c, o); break;
            case 4: hipLaunchKernelGGL((k_stream<3, 0>), dim3(blocks), dim3(256), 0, s, n4, a, b, c, o); break;
            default: hipLaunchKernelGGL((k_stream<3, 1>), dim3(blocks), dim3(256), 0, s, n4, a, b, c, o); break;
        }
    };
    launch();
    launch();
    r = hipEventRecord(e0, s);
    for (int k = 0; k < reps && r == hipSuccess; ++k) launch();
    if (r == hipSuccess) r = hipGetLastError();
    if (r == hipSuccess) r = hipEventRecord(e1, s);
    if (r == hipSuccess) r = hipEventSynchronize(e1);
    float ms = 0.f;
    if (r == hipSuccess) r = hipEventElapsedTime(&ms, e0, e1);
    release();
    if (r != hipSuccess) PW_FAIL(ctx, PW_ERR_HIP, std::string("pw_stream_bandwidth: ") + hipGetErrorString(r));
    *gbs = (double)(reads + 1) * (double)n * 4.0 * reps / ((double)ms * 1e-3) / 1e9;
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
