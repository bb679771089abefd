// pw_psp.hip -- the pseudo-spectral Laplacian of pywave's aps2d/aps3d on MI355X.
//
// Reference path (pywave = github.com/chenyk1990/pywave):
//   entry        pywave/aps2d.py:4-90, pywave/aps3d.py:4-94 -> apscfun (pywave/src/aps.c), whose set-up
//                is afd's; only the Laplacian of the time loop differs
//   engine       pywave/src/wave_psp.c:367-635 (psm2d), :25-365 (psm): wave = IFFT(-(2 pi |k|)^2 FFT(curr))
//   transforms   pywave/src/wave_fft2.c:72-151, wave_fft3.c:96-200 over the KISS FFT the reference
//                vendors (pywave/src/wave_kissfft.c)
//
// Each transform runs the KISS algorithm's arithmetic exactly, laid out for the GPU: one workgroup
// per line (a z column, or a line of wavenumbers along x or y); the recursive decimation of KISS
// becomes a digit-reversed load followed by one pass per factorisation level, deepest first, in
// LDS.  All butterflies of a level are independent and each computes what KISS's radix-2/3/4/5
// butterfly computes for its index (same twiddles, same f32 operation order, the radix-3 half in
// double), so every wavenumber and every output sample is bit-identical to the reference's.  The
// real transforms along z pack pairs of samples into a half-length complex transform and split
// it with KISS's super-twiddles.  The last axis fuses forward transform, wavenumber multiply and
// inverse transform in one pass over each line.
#include <cmath>
#include <string>
#include <vector>

#include "pw_internal.h"

namespace {

struct cf {
    float r, i;
};

__device__ __forceinline__ cf cmul(cf a, cf b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ cf cadd(cf a, cf b) { return {a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cf csub(cf a, cf b) { return {a.r - b.r, a.i - b.i}; }

constexpr int kMaxLev = 24;

// one transform length and direction
struct FftDev {
    int n, nlev, inverse;
    int p[kMaxLev], m[kMaxLev], fs[kMaxLev];  // level 0 = the top of KISS's recursion
    const int* perm;                          // output position -> input element (digit reversal)
    const cf* tw;                             // n twiddles exp(-+2 pi i k/n)
};

__device__ __forceinline__ void bf2(cf* F, int k, int m, int fs, const cf* tw)
{
    const cf t = cmul(F[m + k], tw[k * fs]);
    F[m + k] = csub(F[k], t);
    F[k] = cadd(F[k], t);
}

__device__ __forceinline__ void bf3(cf* F, int k, int m, int fs, const cf* tw)
{
    const cf e = tw[fs * m];
    const cf s1 = cmul(F[k + m], tw[k * fs]);
    const cf s2 = cmul(F[k + 2 * m], tw[2 * k * fs]);
    const cf s3 = cadd(s1, s2);
    cf s0 = csub(s1, s2);
    const cf f0 = F[k];
    cf fm;
    fm.r = (float)((double)f0.r - (double)s3.r * .5);  // the half is taken in double
    fm.i = (float)((double)f0.i - (double)s3.i * .5);
    s0.r = s0.r * e.i;
    s0.i = s0.i * e.i;
    F[k] = cadd(f0, s3);
    F[k + 2 * m] = {fm.r + s0.i, fm.i - s0.r};
    F[k + m] = {fm.r - s0.i, fm.i + s0.r};
}

__device__ __forceinline__ void bf4(cf* F, int k, int m, int fs, const cf* tw, int inverse)
{
    const cf s0 = cmul(F[k + m], tw[k * fs]);
    const cf s1 = cmul(F[k + 2 * m], tw[2 * k * fs]);
    const cf s2 = cmul(F[k + 3 * m], tw[3 * k * fs]);
    const cf f = F[k];
    const cf s5 = csub(f, s1);
    const cf g = cadd(f, s1);
    const cf s3 = cadd(s0, s2), s4 = csub(s0, s2);
    F[k + 2 * m] = csub(g, s3);
    F[k] = cadd(g, s3);
    if (inverse) {
        F[k + m] = {s5.r - s4.i, s5.i + s4.r};
        F[k + 3 * m] = {s5.r + s4.i, s5.i - s4.r};
    } else {
        F[k + m] = {s5.r + s4.i, s5.i - s4.r};
        F[k + 3 * m] = {s5.r - s4.i, s5.i + s4.r};
    }
}

__device__ __forceinline__ void bf5(cf* F, int u, int m, int fs, const cf* tw)
{
    const cf ya = tw[fs * m], yb = tw[fs * 2 * m];
    const cf s0 = F[u];
    const cf s1 = cmul(F[m + u], tw[u * fs]);
    const cf s2 = cmul(F[2 * m + u], tw[2 * u * fs]);
    const cf s3 = cmul(F[3 * m + u], tw[3 * u * fs]);
    const cf s4 = cmul(F[4 * m + u], tw[4 * u * fs]);
    const cf s7 = cadd(s1, s4), s10 = csub(s1, s4), s8 = cadd(s2, s3), s9 = csub(s2, s3);
    F[u] = {s0.r + (s7.r + s8.r), s0.i + (s7.i + s8.i)};
    const cf s5 = {s0.r + s7.r * ya.r + s8.r * yb.r, s0.i + s7.i * ya.r + s8.i * yb.r};
    const cf s6 = {s10.i * ya.i + s9.i * yb.i, -(s10.r * ya.i) - s9.r * yb.i};
    F[m + u] = csub(s5, s6);
    F[4 * m + u] = cadd(s5, s6);
    const cf s11 = {s0.r + s7.r * yb.r + s8.r * ya.r, s0.i + s7.i * yb.r + s8.i * ya.r};
    const cf s12 = {-(s10.i * yb.i) + s9.i * ya.i, s10.r * yb.i - s9.r * ya.i};
    F[2 * m + u] = cadd(s11, s12);
    F[3 * m + u] = csub(s11, s12);
}

// the butterfly passes of one transform over LDS data already in digit-reversed order
__device__ void run_levels(cf* F, const FftDev& d, int tid, int T)
{
    for (int L = d.nlev - 1; L >= 0; --L) {
        const int p = d.p[L], m = d.m[L], fs = d.fs[L];
        const int nbf = d.n / p;
        for (int b = tid; b < nbf; b += T) {
            const int c = b / m, k = b - c * m;
            cf* G = F + c * p * m;
            if (p == 4) bf4(G, k, m, fs, d.tw, d.inverse);
            else if (p == 2) bf2(G, k, m, fs, d.tw);
            else if (p == 3) bf3(G, k, m, fs, d.tw);
            else bf5(G, k, m, fs, d.tw);
        }
        __syncthreads();
    }
}

struct PsGeo {
    int n1, n2, n3, nk, h;  // padded z (real), x, y; nk = n1/2 + 1; h = n1/2
    int ldz;
    long long ldxy;
    float wwt;              // 1/(n1 n2 n3)
};

// forward: centering sign, real transform of each z column -> tmp [n3][n2][nk]
__global__ void k_ps_rz(PsGeo g, const float* __restrict__ field, cf* __restrict__ tmp, FftDev d,
                        const cf* __restrict__ sup)
{
    extern __shared__ cf sh[];
    const int x = blockIdx.x, y = blockIdx.y, tid = threadIdx.x, T = blockDim.x;
    const float* col = field + (long long)y * g.ldxy + (long long)x * g.ldz;
    const bool neg = g.n3 > 1 ? ((y & 1) != (x & 1)) : (x & 1);
    for (int pos = tid; pos < g.h; pos += T) {
        const int j = d.perm[pos];
        float a = col[2 * j], b = col[2 * j + 1];
        if (neg) { a = -a; b = -b; }
        sh[pos] = {a, b};
    }
    __syncthreads();
    run_levels(sh, d, tid, T);
    cf* X = tmp + ((long long)y * g.n2 + x) * g.nk;
    const int h = g.h;
    for (int k = tid; k <= h / 2; k += T) {
        if (k == 0) {
            const cf t0 = sh[0];
            X[0] = {t0.r + t0.i, 0.f};
            X[h] = {t0.r - t0.i, 0.f};
            continue;
        }
        const cf fpk = sh[k], q = sh[h - k];
        const cf fpnk = {q.r, -q.i};
        const cf f1k = cadd(fpk, fpnk), f2k = csub(fpk, fpnk);
        const cf w = cmul(f2k, sup[k - 1]);
        const cf lo = {(float)((double)(f1k.r + w.r) * .5), (float)((double)(f1k.i + w.i) * .5)};
        const cf hi = {(float)((double)(f1k.r - w.r) * .5), (float)((double)(w.i - f1k.i) * .5)};
        if (k != h - k) X[k] = lo;  // at k == h-k the second store wins, as in KISS
        X[h - k] = hi;
    }
}

// complex transforms along x (stride nk, n2 points) or y (stride nk*n2, n3 points).
// mode 0: forward; 1: inverse; 2: forward, multiply by the wavenumber operator, inverse.
__global__ void k_ps_line(PsGeo g, cf* __restrict__ tmp, int axis, int mode, FftDev df, FftDev di,
                          const float* __restrict__ lapl)
{
    extern __shared__ cf sh[];
    const int l = blockIdx.x, tid = threadIdx.x, T = blockDim.x;
    const int i1 = l % g.nk, o = l / g.nk;
    long long base, stride;
    int n;
    if (axis == 0) { base = (long long)o * g.n2 * g.nk + i1; stride = g.nk; n = g.n2; }
    else { base = (long long)o * g.nk + i1; stride = (long long)g.nk * g.n2; n = g.n3; }
    cf* A = sh;
    cf* B = sh + n;
    const FftDev& d0 = mode == 1 ? di : df;
    for (int pos = tid; pos < n; pos += T) A[pos] = tmp[base + (long long)d0.perm[pos] * stride];
    __syncthreads();
    run_levels(A, d0, tid, T);
    if (mode == 2) {
        for (int i = tid; i < n; i += T) {
            const float w = lapl[base + (long long)i * stride];
            A[i] = {A[i].r * w, A[i].i * w};
        }
        __syncthreads();
        for (int pos = tid; pos < n; pos += T) B[pos] = A[di.perm[pos]];
        __syncthreads();
        run_levels(B, di, tid, T);
        A = B;
    }
    for (int i = tid; i < n; i += T) tmp[base + (long long)i * stride] = A[i];
}

// inverse: split of each z column's bins, half-length inverse transform, centering sign and
// 1/(n1 n2 n3) -> wave (field layout)
__global__ void k_ps_iz(PsGeo g, const cf* __restrict__ tmp, float* __restrict__ wave, FftDev d,
                        const cf* __restrict__ sup)
{
    extern __shared__ cf sh[];
    const int x = blockIdx.x, y = blockIdx.y, tid = threadIdx.x, T = blockDim.x;
    const int h = g.h;
    cf* A = sh;
    cf* B = sh + h;
    const cf* X = tmp + ((long long)y * g.n2 + x) * g.nk;
    for (int k = tid; k <= h / 2; k += T) {
        if (k == 0) {
            A[0] = {X[0].r + X[h].r, X[0].r - X[h].r};
            continue;
        }
        const cf fk = X[k], q = X[h - k];
        const cf fnkc = {q.r, -q.i};
        const cf fek = cadd(fk, fnkc), dd = csub(fk, fnkc);
        const cf fok = cmul(dd, sup[k - 1]);
        const cf lo = cadd(fek, fok);
        cf hi = csub(fek, fok);
        hi.i = hi.i * -1.f;
        if (k != h - k) A[k] = lo;
        A[h - k] = hi;
    }
    __syncthreads();
    for (int pos = tid; pos < h; pos += T) B[pos] = A[d.perm[pos]];
    __syncthreads();
    run_levels(B, d, tid, T);
    const bool neg = g.n3 > 1 ? ((y & 1) != (x & 1)) : (x & 1);
    const float w = neg ? -g.wwt : g.wwt;
    float* col = wave + (long long)y * g.ldxy + (long long)x * g.ldz;
    for (int j = tid; j < h; j += T) {
        col[2 * j] = w * B[j].r;
        col[2 * j + 1] = w * B[j].i;
    }
}

// ---------------------------------------------------------------- host side
struct HostFft {
    int n = 0, inverse = 0;
    std::vector<int> p, m, perm;
    std::vector<cf> tw;
};

// KISS's factorisation (4 first, then 2, 3, 5, ...), its twiddles and the input order its
// recursion reads; false if a radix other than 2, 3, 4, 5 appears
bool host_fft(int n, int inverse, HostFft& f)
{
    f.n = n;
    f.inverse = inverse;
    f.tw.resize(n);
    const double pi = 3.141592653589793238462643383279502884197169399375105820974944;
    for (int i = 0; i < n; ++i) {
        double ph = -2 * pi * i / n;
        if (inverse) ph *= -1;
        f.tw[i] = {(float)std::cos(ph), (float)std::sin(ph)};
    }
    int mm = n, pp = 4;
    const double fsq = std::floor(std::sqrt((double)mm));
    do {
        while (mm % pp) {
            pp = pp == 4 ? 2 : (pp == 2 ? 3 : pp + 2);
            if (pp > fsq) pp = mm;
        }
        mm /= pp;
        f.p.push_back(pp);
        f.m.push_back(mm);
    } while (mm > 1);
    for (int q : f.p)
        if (q != 2 && q != 3 && q != 4 && q != 5) return false;
    f.perm.assign(n, 0);
    // output position -> input index, following the recursion's strides
    std::vector<int> fsv(f.p.size());
    int fsacc = 1;
    for (size_t L = 0; L < f.p.size(); ++L) {
        fsv[L] = fsacc;
        fsacc *= f.p[L];
    }
    for (int pos = 0; pos < n; ++pos) {
        int rem = pos, in = 0;
        for (size_t L = 0; L < f.p.size(); ++L) {
            const int j = rem / f.m[L];
            rem -= j * f.m[L];
            in += j * fsv[L];
        }
        f.perm[pos] = in;
    }
    return true;
}

}  // namespace

struct PsLap {
    PsGeo g{};
    float* lapl = nullptr;
    cf* tmp = nullptr;
    std::vector<void*> bufs;
    FftDev rf{}, ri{}, xf{}, xi{}, yf{}, yi{};
    cf *sup_f = nullptr, *sup_i = nullptr;
    int thr_z = 64, thr_x = 64, thr_y = 64;
};

namespace {

int upload(PsLap* P, const void* host, size_t bytes, void** dev)
{
    if (hipMalloc(dev, bytes ? bytes : 8) != hipSuccess) return PW_ERR_NOMEM;
    P->bufs.push_back(*dev);
    if (host && bytes && hipMemcpy(*dev, host, bytes, hipMemcpyHostToDevice) != hipSuccess) return PW_ERR_HIP;
    return PW_OK;
}

int make_dev(PsLap* P, int n, int inverse, FftDev& d, std::string& err)
{
    HostFft h;
    if (!host_fft(n, inverse, h)) {
        err = "spectral Laplacian: transform length " + std::to_string(n) + " has a prime factor above 5";
        return PW_ERR_ARG;
    }
    if ((int)h.p.size() > kMaxLev) {
        err = "spectral Laplacian: too many factorisation levels";
        return PW_ERR_ARG;
    }
    d = FftDev{};
    d.n = n;
    d.nlev = (int)h.p.size();
    d.inverse = inverse;
    int fsacc = 1;
    for (int L = 0; L < d.nlev; ++L) {
        d.p[L] = h.p[L];
        d.m[L] = h.m[L];
        d.fs[L] = fsacc;
        fsacc *= h.p[L];
    }
    void* a = nullptr;
    void* b = nullptr;
    int rc = upload(P, h.perm.data(), h.perm.size() * sizeof(int), &a);
    if (!rc) rc = upload(P, h.tw.data(), h.tw.size() * sizeof(cf), &b);
    d.perm = (const int*)a;
    d.tw = (const cf*)b;
    return rc;
}

int threads_for(int n) { return std::min(256, std::max(64, (n / 2 + 63) / 64 * 64)); }

}  // namespace

void ps_destroy(PsLap* P)
{
    if (!P) return;
    for (void* b : P->bufs)
        if (b) hipFree(b);
    delete P;
}

// the operator of psm2d/psm (pywave/src/wave_psp.c:115-126, :471-480): lapl = -k*k with
// k = 2 pi hypot(kx, kz) (3D: hypot(ky, hypot(kx, kz))) in double, kx = kx0 + ix dkx in f32 with
// centred x (and y) wavenumbers and kz >= 0
PsLap* ps_create(int n1, int n2, int n3, int ldz, long long ldxy, float dz, float dx, float dy, std::string& err)
{
    if (n1 < 2 || n1 % 2 || n2 < 1 || n3 < 1 || ldz < n1) {
        err = "spectral Laplacian: bad padded extents";
        return nullptr;
    }
    PsLap* P = new PsLap;
    PsGeo& g = P->g;
    g.n1 = n1; g.n2 = n2; g.n3 = n3; g.h = n1 / 2; g.nk = n1 / 2 + 1;
    g.ldz = ldz; g.ldxy = ldxy;
    g.wwt = 1.0 / (n1 * n2 * n3);
    const int nk = g.nk;
    std::vector<float> lapl((size_t)nk * n2 * n3);
    const double PI = 3.14159265358979323846264338328;
    const float dkz = 1. / (n1 * dz), dkx = 1. / (n2 * dx), kx0 = -0.5 / dx;
    const float dky = n3 > 1 ? (float)(1. / (n3 * dy)) : 0.f, ky0 = n3 > 1 ? (float)(-0.5 / dy) : 0.f;
    for (int iy = 0; iy < n3; ++iy) {
        const float ky = ky0 + iy * dky;
        for (int ix = 0; ix < n2; ++ix) {
            const float kx = kx0 + ix * dkx;
            for (int iz = 0; iz < nk; ++iz) {
                const float kz = 0.f + iz * dkz;
                const float k = n3 > 1 ? (float)(2 * PI * std::hypot((double)ky, std::hypot((double)kx, (double)kz)))
                                       : (float)(2 * PI * std::hypot((double)kx, (double)kz));
                lapl[((size_t)iy * n2 + ix) * nk + iz] = -k * k;
            }
        }
    }
    // super-twiddles of the real transforms (kiss_fftr_alloc)
    std::vector<cf> sf(std::max(1, g.h / 2)), si(std::max(1, g.h / 2));
    for (int i = 0; i < g.h / 2; ++i) {
        const double ph = -3.14159265358979323846264338327 * ((double)(i + 1) / g.h + .5);
        sf[i] = {(float)std::cos(ph), (float)std::sin(ph)};
        si[i] = {(float)std::cos(-ph), (float)std::sin(-ph)};
    }
    void *a = nullptr, *b = nullptr, *c = nullptr, *t = nullptr;
    int rc = upload(P, lapl.data(), lapl.size() * 4, &a);
    if (!rc) rc = upload(P, sf.data(), sf.size() * sizeof(cf), &b);
    if (!rc) rc = upload(P, si.data(), si.size() * sizeof(cf), &c);
    if (!rc) rc = upload(P, nullptr, (size_t)nk * n2 * n3 * sizeof(cf), &t);
    P->lapl = (float*)a;
    P->sup_f = (cf*)b;
    P->sup_i = (cf*)c;
    P->tmp = (cf*)t;
    if (!rc) rc = make_dev(P, g.h, 0, P->rf, err);
    if (!rc) rc = make_dev(P, g.h, 1, P->ri, err);
    if (!rc) rc = make_dev(P, n2, 0, P->xf, err);
    if (!rc) rc = make_dev(P, n2, 1, P->xi, err);
    if (!rc && n3 > 1) rc = make_dev(P, n3, 0, P->yf, err);
    if (!rc && n3 > 1) rc = make_dev(P, n3, 1, P->yi, err);
    if (rc) {
        if (err.empty()) err = "spectral Laplacian: device allocation failed";
        ps_destroy(P);
        return nullptr;
    }
    P->thr_z = threads_for(g.h);
    P->thr_x = threads_for(n2);
    P->thr_y = threads_for(n3);
    return P;
}

// wave = spectral Laplacian of field (both in the plan's padded field layout)
hipError_t ps_apply(PsLap* P, const float* field, float* wave, hipStream_t s)
{
    const PsGeo& g = P->g;
    const size_t shz = 2 * (size_t)g.h * sizeof(cf);
    hipLaunchKernelGGL(k_ps_rz, dim3(g.n2, g.n3), dim3(P->thr_z), shz, s, g, field, P->tmp, P->rf, P->sup_f);
    if (g.n3 == 1) {
        hipLaunchKernelGGL(k_ps_line, dim3(g.nk), dim3(P->thr_x), 2 * (size_t)g.n2 * sizeof(cf), s, g, P->tmp, 0, 2,
                           P->xf, P->xi, P->lapl);
    } else {
        hipLaunchKernelGGL(k_ps_line, dim3(g.nk * g.n3), dim3(P->thr_x), 2 * (size_t)g.n2 * sizeof(cf), s, g, P->tmp,
                           0, 0, P->xf, P->xi, P->lapl);
        hipLaunchKernelGGL(k_ps_line, dim3(g.nk * g.n2), dim3(P->thr_y), 2 * (size_t)g.n3 * sizeof(cf), s, g, P->tmp,
                           1, 2, P->yf, P->yi, P->lapl);
        hipLaunchKernelGGL(k_ps_line, dim3(g.nk * g.n3), dim3(P->thr_x), 2 * (size_t)g.n2 * sizeof(cf), s, g, P->tmp,
                           0, 1, P->xf, P->xi, P->lapl);
    }
    hipLaunchKernelGGL(k_ps_iz, dim3(g.n2, g.n3), dim3(P->thr_z), shz, s, g, P->tmp, wave, P->ri, P->sup_i);
    return hipGetLastError();
}
