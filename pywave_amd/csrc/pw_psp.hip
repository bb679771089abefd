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

// a / b for 0 <= a < 2^22 and b >= 1 through the f32 reciprocal rb = 1/b (the product is within
// one of the quotient; one correction step each way)
__device__ __forceinline__ int qdiv(int a, int b, float rb)
{
    int q = (int)((float)a * rb);
    const int r = a - q * b;
    q += (r >= b) - (r < 0);
    return q;
}

// twiddles of a descriptor into LDS
__device__ __forceinline__ void stage_tw(cf* dst, const FftDev& d, int tid, int T)
{
    for (int i = tid; i < d.n; i += T) dst[i] = d.tw[i];
}

// the butterfly passes of nl transforms (line l at F + l*ld) over LDS data already in
// digit-reversed order; tw: the descriptor's twiddles staged in LDS
__device__ void run_levels(cf* F, const FftDev& d, const cf* tw, int tid, int T, int nl = 1, int ld = 0)
{
    for (int L = d.nlev - 1; L >= 0; --L) {
        const int p = d.p[L], m = d.m[L], fs = d.fs[L];
        const int nbf = d.n / p;
        const float rn = 1.f / (float)nbf, rm = 1.f / (float)m;
        for (int bb = tid; bb < nbf * nl; bb += T) {
            const int l = nl > 1 ? qdiv(bb, nbf, rn) : 0, b = bb - l * nbf;
            const int c = qdiv(b, m, rm), k = b - c * m;
            cf* G = F + l * ld + c * p * m;
            if (p == 4) bf4(G, k, m, fs, tw, d.inverse);
            else if (p == 2) bf2(G, k, m, fs, tw);
            else if (p == 3) bf3(G, k, m, fs, tw);
            else bf5(G, k, m, fs, tw);
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

// Global-to-LDS copy of 'total' elements: each thread issues its (up to KR per round) loads
// before the first LDS store, so the rounds' memory latencies overlap instead of queueing one
// behind the other.  ld(e) reads element e, st(e, v) stores it.
template <int KR, typename V, typename LD, typename ST>
__device__ __forceinline__ void gather(int total, int tid, int T, LD ld, ST st)
{
    for (int e0 = 0; e0 < total; e0 += KR * T) {
        V r[KR];
#pragma unroll
        for (int q = 0; q < KR; ++q) {
            const int e = e0 + tid + q * T;
            if (e < total) r[q] = ld(e);
        }
#pragma unroll
        for (int q = 0; q < KR; ++q) {
            const int e = e0 + tid + q * T;
            if (e < total) st(e, r[q]);
        }
    }
}

__device__ __forceinline__ void stage_perm(int* dst, const FftDev& d, int tid, int T)
{
    for (int i = tid; i < d.n; i += T) dst[i] = d.perm[i];
}

// forward: centering sign, real transform of each z column -> tmp [n3][n2][nk]; cpb neighbouring
// columns (x0 .. x0+cpb-1 of plane y) per block.  Columns are read whole and contiguously, the
// digit-reversed order is formed in LDS.
__global__ __launch_bounds__(256) void k_ps_rz(PsGeo g, const float* __restrict__ field, cf* __restrict__ tmp,
                                               FftDev d, const cf* __restrict__ sup, int cpb)
{
    extern __shared__ cf sh[];
    const int h = g.h, tid = threadIdx.x, T = blockDim.x;
    const int x0 = blockIdx.x * cpb, y = blockIdx.y;
    const int nc = min(cpb, g.n2 - x0);
    cf* A = sh;                             // [cpb][h] natural order
    cf* B = sh + cpb * h;                   // [cpb][h] digit-reversed
    cf* tw = sh + 2 * cpb * h;
    int* perm = (int*)(tw + h);
    stage_tw(tw, d, tid, T);
    stage_perm(perm, d, tid, T);
    float* Af = (float*)A;
    const float r2h = 1.f / (float)(2 * h);
    gather<8, float>(nc * 2 * h, tid, T,
        [&](int e) {
            const int c = qdiv(e, 2 * h, r2h), f = e - c * 2 * h;
            return field[(long long)y * g.ldxy + (long long)(x0 + c) * g.ldz + f];
        },
        [&](int e, float v) {
            const int c = qdiv(e, 2 * h, r2h);
            const int x = x0 + c;
            const bool neg = g.n3 > 1 ? ((y & 1) != (x & 1)) : (x & 1);
            Af[e] = neg ? -v : v;
        });
    __syncthreads();
    const float rh = 1.f / (float)h;
    for (int e = tid; e < nc * h; e += T) {
        const int c = qdiv(e, h, rh), pos = e - c * h;
        B[c * h + pos] = A[c * h + perm[pos]];
    }
    __syncthreads();
    run_levels(B, d, tw, tid, T, nc, h);
    const int nkk = h / 2 + 1;
    const float rk = 1.f / (float)nkk;
    for (int e = tid; e < nc * nkk; e += T) {
        const int c = qdiv(e, nkk, rk), k = e - c * nkk;
        const cf* S = B + c * h;
        cf* X = tmp + ((long long)y * g.n2 + x0 + c) * g.nk;
        if (k == 0) {
            const cf t0 = S[0];
            X[0] = {t0.r + t0.i, 0.f};
            X[h] = {t0.r - t0.i, 0.f};
            continue;
        }
        const cf fpk = S[k], q = S[h - k];
        const cf fpnk = {q.r, -q.i};
        const cf f1k = cadd(fpk, fpnk), f2k = csub(fpk, fpnk);
        const cf w = cmul(f2k, sup[k - 1]);
        const cf lo = {(float)((double)(f1k.r + w.r) * .5), (float)((double)(f1k.i + w.i) * .5)};
        const cf hi = {(float)((double)(f1k.r - w.r) * .5), (float)((double)(w.i - f1k.i) * .5)};
        if (k != h - k) X[k] = lo;  // at k == h-k the second store wins, as in KISS
        X[h - k] = hi;
    }
}

// complex transforms along x (stride nk, n2 points) or y (stride nk*n2, n3 points), lpb
// neighbouring wavenumber lines (consecutive i1, adjacent in memory) per block so that every
// load and store moves lpb contiguous elements; loads in natural order, digit reversal in LDS.
// MODE 0: forward; 1: inverse; 2: forward, multiply by the wavenumber operator, inverse.  (A
// compile-time mode keeps the descriptors in kernel arguments instead of private memory.)
template <int MODE>
__global__ __launch_bounds__(256) void k_ps_line(PsGeo g, cf* __restrict__ tmp, int axis, FftDev df, FftDev di,
                                                 const float* __restrict__ lapl, int lpb)
{
    extern __shared__ cf sh[];
    const int tid = threadIdx.x, T = blockDim.x;
    const int ngrp = (g.nk + lpb - 1) / lpb;
    const int i10 = (blockIdx.x % ngrp) * lpb, o = blockIdx.x / ngrp;
    const int nl = min(lpb, g.nk - i10);
    long long base, stride;
    int n;
    if (axis == 0) { base = (long long)o * g.n2 * g.nk + i10; stride = g.nk; n = g.n2; }
    else { base = (long long)o * g.nk + i10; stride = (long long)g.nk * g.n2; n = g.n3; }
    // lines in LDS at a stride of n + 1 elements: the lanes that touch element i of neighbouring
    // lines fall in different banks
    const int L = n + 1;
    cf* A = sh;                 // [lpb][L]
    cf* B = sh + lpb * L;       // [lpb][L]
    cf* tf = sh + 2 * lpb * L;  // twiddles of the first transform
    cf* ti = tf + n;            // and of the inverse (MODE 2)
    int* perm = (int*)(ti + n); // the same digit reversal for both directions
    const FftDev& d0 = MODE == 1 ? di : df;
    stage_tw(tf, d0, tid, T);
    if (MODE == 2) stage_tw(ti, di, tid, T);
    stage_perm(perm, d0, tid, T);
    const float rl = 1.f / (float)nl, rn = 1.f / (float)n;
    gather<8, cf>(n * nl, tid, T,
        [&](int e) { const int i = qdiv(e, nl, rl), l = e - i * nl; return tmp[base + l + (long long)i * stride]; },
        [&](int e, cf v) { const int i = qdiv(e, nl, rl), l = e - i * nl; A[l * L + i] = v; });
    __syncthreads();
    for (int e = tid; e < n * nl; e += T) {
        const int l = qdiv(e, n, rn), pos = e - l * n;
        B[l * L + pos] = A[l * L + perm[pos]];
    }
    __syncthreads();
    run_levels(B, d0, tf, tid, T, nl, L);
    cf* R = B;
    if (MODE == 2) {
        gather<8, float>(n * nl, tid, T,
            [&](int e) { const int i = qdiv(e, nl, rl), l = e - i * nl; return lapl[base + l + (long long)i * stride]; },
            [&](int e, float w) {
                const int i = qdiv(e, nl, rl), l = e - i * nl;
                const cf a = B[l * L + i];
                B[l * L + i] = {a.r * w, a.i * w};
            });
        __syncthreads();
        for (int e = tid; e < n * nl; e += T) {
            const int l = qdiv(e, n, rn), pos = e - l * n;
            A[l * L + pos] = B[l * L + perm[pos]];
        }
        __syncthreads();
        run_levels(A, di, ti, tid, T, nl, L);
        R = A;
    }
    for (int e = tid; e < n * nl; e += T) {
        const int i = qdiv(e, nl, rl), l = e - i * nl;
        tmp[base + l + (long long)i * stride] = R[l * L + i];
    }
}

// inverse: split of each z column's bins, half-length inverse transform, centering sign and
// 1/(n1 n2 n3) -> wave (field layout); cpb neighbouring columns per block
__global__ __launch_bounds__(256) void k_ps_iz(PsGeo g, const cf* __restrict__ tmp, float* __restrict__ wave,
                                               FftDev d, const cf* __restrict__ sup, int cpb)
{
    extern __shared__ cf sh[];
    const int h = g.h, tid = threadIdx.x, T = blockDim.x;
    const int x0 = blockIdx.x * cpb, y = blockIdx.y;
    const int nc = min(cpb, g.n2 - x0);
    cf* A = sh;
    cf* B = sh + cpb * h;
    cf* tw = sh + 2 * cpb * h;
    int* perm = (int*)(tw + h);
    stage_tw(tw, d, tid, T);
    stage_perm(perm, d, tid, T);
    const int nkk = h / 2 + 1;
    const float rk = 1.f / (float)nkk;
    struct Pair {
        cf a, b;
    };
    gather<4, Pair>(nc * nkk, tid, T,
        [&](int e) {
            const int c = qdiv(e, nkk, rk), k = e - c * nkk;
            const cf* X = tmp + ((long long)y * g.n2 + x0 + c) * g.nk;
            return Pair{X[k], X[h - k]};
        },
        [&](int e, Pair pr) {
        const int c = qdiv(e, nkk, rk), k = e - c * nkk;
        cf* S = A + c * h;
        if (k == 0) {  // pr = (X[0], X[h])
            S[0] = {pr.a.r + pr.b.r, pr.a.r - pr.b.r};
            return;
        }
        const cf fk = pr.a, q = pr.b;
        const cf fnkc = {q.r, -q.i};
        const cf fek = cadd(fk, fnkc), dd = csub(fk, fnkc);
        const cf fok = cmul(dd, sup[k - 1]);
        const cf lo = cadd(fek, fok);
        cf hi = csub(fek, fok);
        hi.i = hi.i * -1.f;
        if (k != h - k) S[k] = lo;
        S[h - k] = hi;
        });
    __syncthreads();
    const float rh = 1.f / (float)h;
    for (int e = tid; e < nc * h; e += T) {
        const int c = qdiv(e, h, rh), pos = e - c * h;
        B[c * h + pos] = A[c * h + perm[pos]];
    }
    __syncthreads();
    run_levels(B, d, tw, tid, T, nc, h);
    const float* Bf = (const float*)B;
    const float r2h = 1.f / (float)(2 * h);
    for (int e = tid; e < nc * 2 * h; e += T) {
        const int c = qdiv(e, 2 * h, r2h), f = e - c * 2 * h;
        const int x = x0 + c;
        const bool neg = g.n3 > 1 ? ((y & 1) != (x & 1)) : (x & 1);
        const float w = neg ? -g.wwt : g.wwt;
        wave[(long long)y * g.ldxy + (long long)x * g.ldz + f] = w * Bf[c * 2 * h + f];
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
    int cpb = 1;               // z columns per block of the real transforms
    int lpb_x = 1, lpb_y = 1;  // wavenumber lines per block of the x and y passes
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
    if (3 * (size_t)(n1 / 2) * sizeof(cf) + (size_t)n1 * 2 > 64 * 1024 ||
        (4 * (size_t)std::max(n2, n3) + 2) * sizeof(cf) + (size_t)std::max(n2, n3) * 4 > 64 * 1024) {
        err = "spectral Laplacian: a transform line exceeds the 64 KB of LDS per workgroup";
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
    // neighbouring z columns per block: about 256 samples' worth of butterflies per block
    P->cpb = std::max(1, std::min(8, 512 / g.h));
    while (P->cpb > 1 && ((size_t)(2 * P->cpb + 1) * g.h * sizeof(cf) + g.h * 4) > 48 * 1024) P->cpb /= 2;
    // 16 adjacent lines per block (128 B per load), within 48 KB of LDS for both line buffers
    auto lpb_for = [&](int n) {
        int l = 16;
        while (l > 1 && ((size_t)2 * l * (n + 1) + 2 * (size_t)n) * sizeof(cf) + (size_t)n * 4 > 48 * 1024) l /= 2;
        return std::min(l, nk);
    };
    P->lpb_x = lpb_for(n2);
    P->lpb_y = lpb_for(n3);
    return P;
}

// wave = spectral Laplacian of field (both in the plan's padded field layout)
hipError_t ps_apply(PsLap* P, const float* field, float* wave, hipStream_t s)
{
    const PsGeo& g = P->g;
    const size_t shz = (2 * (size_t)P->cpb + 1) * g.h * sizeof(cf) + (size_t)g.h * 4;
    const int gz = (g.n2 + P->cpb - 1) / P->cpb;
    hipLaunchKernelGGL(k_ps_rz, dim3(gz, g.n3), dim3(256), shz, s, g, field, P->tmp, P->rf, P->sup_f, P->cpb);
    const int gx = (g.nk + P->lpb_x - 1) / P->lpb_x, gy = (g.nk + P->lpb_y - 1) / P->lpb_y;
    const size_t shx = (2 * (size_t)P->lpb_x * (g.n2 + 1) + 2 * (size_t)g.n2) * sizeof(cf) + (size_t)g.n2 * 4,
                 shy = (2 * (size_t)P->lpb_y * (g.n3 + 1) + 2 * (size_t)g.n3) * sizeof(cf) + (size_t)g.n3 * 4;
    if (g.n3 == 1) {
        hipLaunchKernelGGL(k_ps_line<2>, dim3(gx), dim3(256), shx, s, g, P->tmp, 0, P->xf, P->xi, P->lapl, P->lpb_x);
    } else {
        hipLaunchKernelGGL(k_ps_line<0>, dim3(gx * g.n3), dim3(256), shx, s, g, P->tmp, 0, P->xf, P->xi, P->lapl,
                           P->lpb_x);
        hipLaunchKernelGGL(k_ps_line<2>, dim3(gy * g.n2), dim3(256), shy, s, g, P->tmp, 1, P->yf, P->yi, P->lapl,
                           P->lpb_y);
        hipLaunchKernelGGL(k_ps_line<1>, dim3(gx * g.n3), dim3(256), shx, s, g, P->tmp, 0, P->xf, P->xi, P->lapl,
                           P->lpb_x);
    }
    hipLaunchKernelGGL(k_ps_iz, dim3(gz, g.n3), dim3(256), shz, s, g, P->tmp, wave, P->ri, P->sup_i, P->cpb);
    return hipGetLastError();
}
