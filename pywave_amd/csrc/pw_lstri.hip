// pw_lstri.hip -- pfwi mode 4's passive modelling and least-squares time-reversal imaging
// (lstric) on MI355X.
//
// Reference path (pywave = github.com/chenyk1990/pywave):
//   entry        pywave/pfwi.py:120-165, :212-250 -> pfwicfun.lstric (pywave/src/pfwi.c:1031-1505)
//   driver       pywave/src/wave_fwi.c:121-210 (lstri), wave_fwigradient.c:454-525 (lstri_op)
//   operator     pywave/src/wave_triutil.c:14-180 (tri2d_make, timerev_init, timerev_lop), the
//                8th-order Lap/LapT of wave_triutil.h, decay boundary pfwiabc (wave_triutil.c:264-317)
//   solver       pywave/src/wave_bigsolver.c:1170-1450 (np_solver, unit model weight) with
//                wave_cgstep.c:17-67 (np_cgstep)
//
// The operator maps a space-time source field w[it][ix][iz] to data d[ix][it] recorded one cell
// below the top of the model: per step, record u1 at the receiver row, u2 = (Lx/dx^2 + Lz/dz^2)
// v^2dt^2 + 2 u1 - u0 on the padded interior, add w v^2dt^2 on the model, decay both levels at the
// edges.  One fused kernel per step and shot batch does all of it per cell: the stored level is
// already decayed, the previous one is decayed again on read (the reference decays both levels
// after every step).  The adjoint kernel runs the transposed recursion (LapT of u1 v^2dt^2, image
// = u1 v^2dt^2, data injected at the receiver row).  Per-cell f32 arithmetic follows the
// reference's order, so modelled data and images are bit-identical to the reference C.
//
// The solver is the reference's conjugate-gradient iteration on the device: model vectors
// (nt*nz*nx) and data vectors (nt*nx) stay in HBM; dot products are f64 reductions.  Its step
// vectors carry over from one shot to the next as np_cgstep's static state does within a call
// (restart only at the call's first shot), so shots of an inversion run one after the other.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pw_buffer.h"
#include "pw_internal.h"

namespace {

using pwbuf::bld;
using pwbuf::bldo;
using pwbuf::bsto;
using pwbuf::kOOB;
using pwbuf::win;

// stencil weights as float quotients (wave_triutil.h)
constexpr float W0 = -205.0f / 72.0f, W1 = 8.0f / 5.0f, W2 = -1.0f / 5.0f, W3 = 8.0f / 315.0f,
                W4 = -1.0f / 560.0f;

struct TriGeo {
    int pz, px, ldz;        // padded extents, z stride of a field
    int nz, nx, nb, depth;  // model extents, decay width, receiver row (padded index)
    float sx, sz;           // 1/dx^2, 1/dz^2
    long long fs;           // elements per field
};

struct TriArgs {
    const float* X;   // modelling: u1 (decayed); adjoint: the current level
    float* Y;         // in: the previous level (decayed once); out: the new level (decayed)
    const float* v2;  // v*v*dt^2 on the padded grid, edge-replicated
    const float* sfz; // decay factor pairs (top, bottom) per padded z
    const float* sfx; // per padded x (left, right)
    float* plane;     // modelling: source plane of this step (read); adjoint: image plane (written)
    long long pss;    // slot stride of the plane ([slot][nx][nz] within a step)
    float* row;       // modelling: recorded row of this step (written); adjoint: injected row (read)
    long long rss;    // slot stride of the row
};

__device__ __forceinline__ float decay(float v, float zl, float zh, float xl, float xh)
{
    v = zl * v; v = zh * v;  // top and bottom rows (pfwiabc_apply's first loop)
    v = xl * v; v = xh * v;  // left and right columns
    return v;
}

// One modelling step: one 64 x XCT tile of one shot per block (one wave), streamed along x with a
// 9-column register queue and an LDS z row; the loads of column x+1 and the centre of column x+5
// are issued before column x is computed.
template <int XCT>
__global__ __launch_bounds__(64) void k_trm(TriGeo g, TriArgs a, int ntz)
{
    const int slot = blockIdx.y;
    const int tz = blockIdx.x % ntz, xc = blockIdx.x / ntz;
    const int lane = threadIdx.x, z0 = tz * 64, z = z0 + lane;
    const int x0 = xc * XCT, x1 = min(x0 + XCT, g.px);
    __shared__ float row[2][72];
    const int L = g.ldz;
    const int bias = 4 * L + 4;  // windows start 4 columns + 4 cells before the field
    const unsigned wb = (unsigned)((g.fs + 2 * bias) * 4);
    const auto rX = win(a.X + slot * g.fs - bias, wb), rY = win(a.Y + slot * g.fs - bias, wb),
               rV = win(a.v2 - bias, wb);
    const bool zok = z < g.pz;
    const bool zin = z >= 4 && z < g.pz - 4;
    const float zl = zok ? a.sfz[2 * z] : 1.f, zh = zok ? a.sfz[2 * z + 1] : 1.f;
    const int zhz = lane < 4 ? z0 - 4 + lane : z0 + 60 + lane;  // z halo of lanes 0..7
    const int lzh = lane < 4 ? lane : 64 + lane;
    const bool zhv = lane < 8 && zhz >= 0 && zhz < g.pz;
    const unsigned oz = zok ? (unsigned)(bias + z) * 4 : kOOB, ozh = zhv ? (unsigned)(bias + zhz) * 4 : kOOB;
    const unsigned oc = (unsigned)L * 4;
    auto xv = [&](int x) { return x >= 0 && x < g.px; };
    auto ldc = [&](__amdgpu_buffer_rsrc_t r, unsigned voff, int x, bool ok) {
        return bldo<float>(r, ok ? voff : kOOB, ok ? (unsigned)x * oc : 0u);
    };
    // the source plane covers the model interior [nb, nb+nz) x [nb, nb+nx)
    const bool mz = z >= g.nb && z < g.nb + g.nz;
    const auto rS = win(a.plane + slot * a.pss, (unsigned)(g.nz * g.nx * 4));
    const unsigned ofz = mz ? (unsigned)(z - g.nb) * 4 : kOOB;
    auto mx = [&](int x) { return x >= g.nb && x < g.nb + g.nx; };
    auto lds = [&](int x) { return bldo<float>(rS, mx(x) ? ofz : kOOB, mx(x) ? (unsigned)((x - g.nb) * g.nz) * 4 : 0u); };
    const bool rlane = z == g.depth;
    float* rec = a.row + slot * a.rss;
    float q[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int xx = x0 - 4 + k;
        q[k] = ldc(rX, oz, xx, xv(xx));
    }
    float yv = ldc(rY, oz, x0, true), vv = ldc(rV, oz, x0, true), hz = ldc(rX, ozh, x0, true), sv = lds(x0);
    float xl = a.sfx[2 * x0], xh = a.sfx[2 * x0 + 1];
    int buf = 0;
    for (int x = x0; x < x1; ++x) {
        const int xn = x + 1;
        const bool nv = xn < x1;
        const float nq = ldc(rX, oz, x + 5, nv && x + 5 < g.px);
        const float nyv = ldc(rY, oz, xn, nv);
        const float nvv = ldc(rV, oz, xn, nv);
        const float nhz = ldc(rX, ozh, xn, nv);
        const float nsv = lds(xn);
        const int xs = xn < g.px ? xn : g.px - 1;
        const float nxl = a.sfx[2 * xs], nxh = a.sfx[2 * xs + 1];
        row[buf][lane + 4] = q[4];
        if (lane < 8) row[buf][lzh] = hz;
        __syncthreads();
        if (zok) {
            const float u1 = q[4];
            if (rlane && mx(x)) rec[x - g.nb] = 0.f + u1;  // record before the step
            float u2 = 0.f;                              // the 4-cell rim stays zero
            if (zin && x >= 4 && x < g.px - 4) {
                const float* R = &row[buf][lane + 4];
                const float lx = W4 * (q[8] + q[0]) + W3 * (q[7] + q[1]) + W2 * (q[6] + q[2]) + W1 * (q[5] + q[3]) +
                                 W0 * q[4];
                const float lz = W4 * (R[4] + R[-4]) + W3 * (R[3] + R[-3]) + W2 * (R[2] + R[-2]) +
                                 W1 * (R[1] + R[-1]) + W0 * R[0];
                u2 = (lx * g.sx + lz * g.sz) * vv + 2.0f * u1 - decay(yv, zl, zh, xl, xh);
            }
            if (mz && mx(x)) u2 = u2 + sv * vv;  // inject the source after the step
            bsto(rY, oz, (unsigned)x * oc, decay(u2, zl, zh, xl, xh));
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) q[k] = q[k + 1];
        q[8] = nq;
        yv = nyv; vv = nvv; hz = nhz; sv = nsv; xl = nxl; xh = nxh;
        buf ^= 1;
    }
}

// One adjoint step: image = A v^2dt^2 on the model, new level = decay(LapT(A v^2dt^2) + 2A -
// decay(B) + injected data at the receiver row).  Queues of the raw level and of its product with
// v^2dt^2 (the transposed stencil weights the neighbours' products).
template <int XCT>
__global__ __launch_bounds__(64) void k_tra(TriGeo g, TriArgs a, int ntz)
{
    const int slot = blockIdx.y;
    const int tz = blockIdx.x % ntz, xc = blockIdx.x / ntz;
    const int lane = threadIdx.x, z0 = tz * 64, z = z0 + lane;
    const int x0 = xc * XCT, x1 = min(x0 + XCT, g.px);
    __shared__ float row[2][72];
    const int L = g.ldz;
    const int bias = 4 * L + 4;
    const unsigned wb = (unsigned)((g.fs + 2 * bias) * 4);
    const auto rA = win(a.X + slot * g.fs - bias, wb), rB = win(a.Y + slot * g.fs - bias, wb),
               rV = win(a.v2 - bias, wb);
    const bool zok = z < g.pz;
    const bool zin = z >= 4 && z < g.pz - 4;
    const float zl = zok ? a.sfz[2 * z] : 1.f, zh = zok ? a.sfz[2 * z + 1] : 1.f;
    const int zhz = lane < 4 ? z0 - 4 + lane : z0 + 60 + lane;
    const int lzh = lane < 4 ? lane : 64 + lane;
    const bool zhv = lane < 8 && zhz >= 0 && zhz < g.pz;
    const unsigned oz = zok ? (unsigned)(bias + z) * 4 : kOOB, ozh = zhv ? (unsigned)(bias + zhz) * 4 : kOOB;
    const unsigned oc = (unsigned)L * 4;
    auto xv = [&](int x) { return x >= 0 && x < g.px; };
    auto ldc = [&](__amdgpu_buffer_rsrc_t r, unsigned voff, int x, bool ok) {
        return bldo<float>(r, ok ? voff : kOOB, ok ? (unsigned)x * oc : 0u);
    };
    const bool mz = z >= g.nb && z < g.nb + g.nz;
    const auto rI = win(a.plane + slot * a.pss, (unsigned)(g.nz * g.nx * 4));
    const unsigned ofz = mz ? (unsigned)(z - g.nb) * 4 : kOOB;
    auto mx = [&](int x) { return x >= g.nb && x < g.nb + g.nx; };
    const bool rlane = z == g.depth;
    const auto rD = win(a.row + slot * a.rss, (unsigned)(g.nx * 4));
    auto ldd = [&](int x) { return bld<float>(rD, (unsigned)(x - g.nb) * 4, rlane && mx(x)); };
    float qr[9], qp[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int xx = x0 - 4 + k;
        qr[k] = ldc(rA, oz, xx, xv(xx));
        qp[k] = qr[k] * ldc(rV, oz, xx, xv(xx));
    }
    float bv = ldc(rB, oz, x0, true);
    float hp = ldc(rA, ozh, x0, true) * ldc(rV, ozh, x0, true);
    float dv = ldd(x0);
    float xl = a.sfx[2 * x0], xh = a.sfx[2 * x0 + 1];
    int buf = 0;
    for (int x = x0; x < x1; ++x) {
        const int xn = x + 1, x5 = x + 5;
        const bool nv = xn < x1;
        const bool ok5 = nv && x5 < g.px;
        const float na = ldc(rA, oz, x5, ok5);
        const float nvp = ldc(rV, oz, x5, ok5);
        const float nbv = ldc(rB, oz, xn, nv);
        const float nhp = ldc(rA, ozh, xn, nv) * ldc(rV, ozh, xn, nv);
        const float ndv = ldd(xn);
        const int xs = xn < g.px ? xn : g.px - 1;
        const float nxl = a.sfx[2 * xs], nxh = a.sfx[2 * xs + 1];
        row[buf][lane + 4] = qp[4];
        if (lane < 8) row[buf][lzh] = hp;
        __syncthreads();
        if (zok) {
            const float u1 = qr[4];
            if (mz && mx(x))  // image of this step, u1 v^2dt^2 on the model
                bsto(rI, ofz, (unsigned)((x - g.nb) * g.nz) * 4, 0.f + qp[4]);
            float u2 = 0.f;
            if (zin && x >= 4 && x < g.px - 4) {
                const float* R = &row[buf][lane + 4];
                const float lx = W4 * (qp[8] + qp[0]) + W3 * (qp[7] + qp[1]) + W2 * (qp[6] + qp[2]) +
                                 W1 * (qp[5] + qp[3]) + W0 * qp[4];
                const float lz = W4 * (R[4] + R[-4]) + W3 * (R[3] + R[-3]) + W2 * (R[2] + R[-2]) +
                                 W1 * (R[1] + R[-1]) + W0 * R[0];
                u2 = (lx * g.sx + lz * g.sz) + 2.0f * u1 - decay(bv, zl, zh, xl, xh);
            }
            if (rlane && mx(x)) u2 = u2 + dv;  // inject the data of this step after it
            bsto(rB, oz, (unsigned)x * oc, decay(u2, zl, zh, xl, xh));
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) { qr[k] = qr[k + 1]; qp[k] = qp[k + 1]; }
        qr[8] = na;
        qp[8] = na * nvp;
        bv = nbv; hp = nhp; dv = ndv; xl = nxl; xh = nxh;
        buf ^= 1;
    }
}

// ---------------------------------------------------------------- two time steps per launch
// On the small grids of the passive workloads a step launch is latency-bound, so the sweeps can
// advance two steps per launch: a block owns an OZ2 x OX2 tile, stages the current level on the
// tile grown by 8 cells in LDS, forms the first step on the tile grown by 4 (LDS) and the second
// step on the tile, each cell with the single-step kernels' expressions, so every value equals
// that of two single steps bit for bit.  Both new levels go to two spare fields, since the
// neighbouring blocks still read the current and previous levels during the launch.
constexpr int OZ2 = 56, OX2 = 4;             // output tile
constexpr int Z1 = OZ2 + 8, X1 = OX2 + 8;    // first-step region
constexpr int Z0 = OZ2 + 16, X0 = OX2 + 16;  // staged current level
constexpr int T2 = 256;                      // threads per block

struct TriArgs2 {
    const float* X;  // current level (modelling: u1; adjoint: A)
    const float* Y;  // previous level (decayed once)
    float* C;        // out: the first step's level (the next launch's previous level)
    float* D;        // out: the second step's level (the next launch's current level)
    const float* v2;
    const float* sfz;
    const float* sfx;
    float* plane0;   // plane of the first step (modelling: source, read; adjoint: image, written)
    float* plane1;   // plane of the second step
    long long pss;
    float* row0;     // row of the first step (modelling: recorded; adjoint: injected)
    float* row1;     // row of the second step
    long long rss;
};

template <bool ADJ>
__global__ __launch_bounds__(T2) void k_tr2(TriGeo g, TriArgs2 a, int ntz2)
{
    __shared__ float s0[X0][Z0 + 1];            // current level
    __shared__ float p0[ADJ ? X0 : 1][Z0 + 1];  // adjoint: current level x v^2dt^2
    __shared__ float s1[X1][Z1 + 1];            // first step's level
    __shared__ float p1[ADJ ? X1 : 1][Z1 + 1];  // adjoint: first step's level x v^2dt^2
    __shared__ float yv[X1][Z1 + 1];            // previous level on the first-step region
    __shared__ float vr[X1][Z1 + 1];            // v^2dt^2 there
    __shared__ float sp[ADJ ? 1 : X1][Z1 + 1];  // modelling: first step's source plane there
    __shared__ float fz[Z0][2], fx[X0][2];      // decay factor pairs of the staged rows and columns
    __shared__ float rd0[X1];                   // adjoint: first step's injected row there
    const int slot = blockIdx.y;
    const int tz = blockIdx.x % ntz2, tx = blockIdx.x / ntz2;
    const int zo = tz * OZ2, xo = tx * OX2;
    const long long so = slot * g.fs;
    const float* __restrict__ X = a.X + so;
    const float* __restrict__ Y = a.Y + so;
    auto ing = [&](int z, int x) { return z >= 0 && z < g.pz && x >= 0 && x < g.px; };
    auto inr = [&](int z, int x) { return z >= 4 && z < g.pz - 4 && x >= 4 && x < g.px - 4; };
    auto mz = [&](int z) { return z >= g.nb && z < g.nb + g.nz; };
    auto mx = [&](int x) { return x >= g.nb && x < g.nb + g.nx; };
    auto mi = [&](int z, int x) { return slot * a.pss + (long long)(x - g.nb) * g.nz + (z - g.nb); };
    // every global operand of both steps is loaded into registers up front (all loads in flight
    // together), then written to LDS, so the launch waits on memory once
    constexpr int N0 = (Z0 * X0 + T2 - 1) / T2, N1 = (Z1 * X1 + T2 - 1) / T2, N2 = (OZ2 * OX2 + T2 - 1) / T2;
    float r0[N0], rv0[ADJ ? N0 : 1], ry[N1], rvv[N1], rsp[ADJ ? 1 : N1], pl1[N2];
#pragma unroll
    for (int j = 0; j < N0; ++j) {
        const int e = threadIdx.x + j * T2;
        const int zz = e % Z0, xx = e / Z0;
        const int z = zo - 8 + zz, x = xo - 8 + xx;
        const bool in = e < Z0 * X0 && ing(z, x);
        const long long k = (long long)x * g.ldz + z;
        r0[j] = in ? X[k] : 0.f;
        if constexpr (ADJ) rv0[j] = in ? a.v2[k] : 0.f;
    }
#pragma unroll
    for (int j = 0; j < N1; ++j) {
        const int e = threadIdx.x + j * T2;
        const int zz = e % Z1, xx = e / Z1;
        const int z = zo - 4 + zz, x = xo - 4 + xx;
        const bool in = e < Z1 * X1 && ing(z, x);
        const long long k = (long long)x * g.ldz + z;
        ry[j] = in ? Y[k] : 0.f;
        rvv[j] = in ? a.v2[k] : 0.f;
        if constexpr (!ADJ) rsp[j] = (in && mz(z) && mx(x)) ? a.plane0[mi(z, x)] : 0.f;
    }
    // the second step's own operands for this thread's output cells
#pragma unroll
    for (int j = 0; j < N2; ++j) {
        const int e = threadIdx.x + j * T2;
        const int zz = e % OZ2, xx = e / OZ2;
        const int z = zo + zz, x = xo + xx;
        pl1[j] = 0.f;
        if constexpr (!ADJ) {
            if (e < OZ2 * OX2 && ing(z, x) && mz(z) && mx(x)) pl1[j] = a.plane1[mi(z, x)];
        } else {
            if (e < OZ2 * OX2 && ing(z, x) && z == g.depth && mx(x)) pl1[j] = a.row1[slot * a.rss + (x - g.nb)];
        }
    }
    {
        // decay pairs (row zo-8+t, column xo-8+t) and the adjoint's injected row
        const int t = threadIdx.x;
        const int z = zo - 8 + t, x = xo - 8 + t;
        const bool zv = t < Z0 && z >= 0 && z < g.pz, xv = t < X0 && x >= 0 && x < g.px;
        const float a0 = zv ? a.sfz[2 * z] : 1.f, a1 = zv ? a.sfz[2 * z + 1] : 1.f;
        const float b0 = xv ? a.sfx[2 * x] : 1.f, b1 = xv ? a.sfx[2 * x + 1] : 1.f;
        const int xr = xo - 4 + t;
        float d0 = 0.f;
        if constexpr (ADJ)
            if (t < X1 && xr >= g.nb && xr < g.nb + g.nx) d0 = a.row0[slot * a.rss + (xr - g.nb)];
        if (t < Z0) { fz[t][0] = a0; fz[t][1] = a1; }
        if (t < X0) { fx[t][0] = b0; fx[t][1] = b1; }
        if constexpr (ADJ)
            if (t < X1) rd0[t] = d0;
    }
#pragma unroll
    for (int j = 0; j < N0; ++j) {
        const int e = threadIdx.x + j * T2;
        if (e < Z0 * X0) {
            const int zz = e % Z0, xx = e / Z0;
            s0[xx][zz] = r0[j];
            if constexpr (ADJ) p0[xx][zz] = r0[j] * rv0[j];
        }
    }
#pragma unroll
    for (int j = 0; j < N1; ++j) {
        const int e = threadIdx.x + j * T2;
        if (e < Z1 * X1) {
            const int zz = e % Z1, xx = e / Z1;
            yv[xx][zz] = ry[j];
            vr[xx][zz] = rvv[j];
            if constexpr (!ADJ) sp[xx][zz] = rsp[j];
        }
    }
    __syncthreads();
    // first step on the tile grown by 4
    for (int e = threadIdx.x; e < Z1 * X1; e += T2) {
        const int zz = e % Z1, xx = e / Z1;
        const int z = zo - 4 + zz, x = xo - 4 + xx;
        const int cz = zz + 4, cx = xx + 4;  // the cell in s0
        float v = 0.f;
        if (ing(z, x)) {
            const float zl = fz[cz][0], zh = fz[cz][1], xl = fx[cx][0], xh = fx[cx][1];
            const float u1 = s0[cx][cz];
            const float vv = vr[xx][zz];
            float u2 = 0.f;  // the 4-cell rim stays zero
            if constexpr (!ADJ) {
                if (inr(z, x)) {
                    const float lx = W4 * (s0[cx + 4][cz] + s0[cx - 4][cz]) + W3 * (s0[cx + 3][cz] + s0[cx - 3][cz]) +
                                     W2 * (s0[cx + 2][cz] + s0[cx - 2][cz]) + W1 * (s0[cx + 1][cz] + s0[cx - 1][cz]) +
                                     W0 * u1;
                    const float lz = W4 * (s0[cx][cz + 4] + s0[cx][cz - 4]) + W3 * (s0[cx][cz + 3] + s0[cx][cz - 3]) +
                                     W2 * (s0[cx][cz + 2] + s0[cx][cz - 2]) + W1 * (s0[cx][cz + 1] + s0[cx][cz - 1]) +
                                     W0 * u1;
                    u2 = (lx * g.sx + lz * g.sz) * vv + 2.0f * u1 - decay(yv[xx][zz], zl, zh, xl, xh);
                }
                if (mz(z) && mx(x)) u2 = u2 + sp[xx][zz] * vv;
            } else {
                if (inr(z, x)) {
                    const float c = p0[cx][cz];
                    const float lx = W4 * (p0[cx + 4][cz] + p0[cx - 4][cz]) + W3 * (p0[cx + 3][cz] + p0[cx - 3][cz]) +
                                     W2 * (p0[cx + 2][cz] + p0[cx - 2][cz]) + W1 * (p0[cx + 1][cz] + p0[cx - 1][cz]) +
                                     W0 * c;
                    const float lz = W4 * (p0[cx][cz + 4] + p0[cx][cz - 4]) + W3 * (p0[cx][cz + 3] + p0[cx][cz - 3]) +
                                     W2 * (p0[cx][cz + 2] + p0[cx][cz - 2]) + W1 * (p0[cx][cz + 1] + p0[cx][cz - 1]) +
                                     W0 * c;
                    u2 = (lx * g.sx + lz * g.sz) + 2.0f * u1 - decay(yv[xx][zz], zl, zh, xl, xh);
                }
                if (z == g.depth && mx(x)) u2 = u2 + rd0[xx];
            }
            v = decay(u2, zl, zh, xl, xh);
            if (zz >= 4 && zz < 4 + OZ2 && xx >= 4 && xx < 4 + OX2) a.C[so + (long long)x * g.ldz + z] = v;
            if constexpr (ADJ) p1[xx][zz] = v * vv;
        } else if constexpr (ADJ) {
            p1[xx][zz] = 0.f;
        }
        s1[xx][zz] = v;
    }
    __syncthreads();
    // second step on the tile
#pragma unroll
    for (int j = 0; j < (OZ2 * OX2 + T2 - 1) / T2; ++j) {
        const int e = threadIdx.x + j * T2;
        const int zz = e % OZ2, xx = e / OZ2;
        const int z = zo + zz, x = xo + xx;
        if (e >= OZ2 * OX2 || !ing(z, x)) continue;
        const int cz = zz + 4, cx = xx + 4;  // the cell in s1
        const float zl = fz[zz + 8][0], zh = fz[zz + 8][1], xl = fx[xx + 8][0], xh = fx[xx + 8][1];
        const float u1 = s1[cx][cz];
        const float prev = s0[xx + 8][zz + 8];  // the launch's current level is this step's previous one
        const bool rcell = z == g.depth && mx(x);
        float u2 = 0.f;
        if constexpr (!ADJ) {
            if (rcell) {  // record both steps' levels before they are stepped
                a.row0[slot * a.rss + (x - g.nb)] = 0.f + prev;
                a.row1[slot * a.rss + (x - g.nb)] = 0.f + u1;
            }
            const float vv = vr[cx][cz];
            if (inr(z, x)) {
                const float lx = W4 * (s1[cx + 4][cz] + s1[cx - 4][cz]) + W3 * (s1[cx + 3][cz] + s1[cx - 3][cz]) +
                                 W2 * (s1[cx + 2][cz] + s1[cx - 2][cz]) + W1 * (s1[cx + 1][cz] + s1[cx - 1][cz]) +
                                 W0 * u1;
                const float lz = W4 * (s1[cx][cz + 4] + s1[cx][cz - 4]) + W3 * (s1[cx][cz + 3] + s1[cx][cz - 3]) +
                                 W2 * (s1[cx][cz + 2] + s1[cx][cz - 2]) + W1 * (s1[cx][cz + 1] + s1[cx][cz - 1]) +
                                 W0 * u1;
                u2 = (lx * g.sx + lz * g.sz) * vv + 2.0f * u1 - decay(prev, zl, zh, xl, xh);
            }
            if (mz(z) && mx(x)) u2 = u2 + pl1[j] * vv;
        } else {
            if (mz(z) && mx(x)) {  // images of both steps, u1 v^2dt^2 on the model
                const long long m = mi(z, x);
                a.plane0[m] = 0.f + p0[xx + 8][zz + 8];
                a.plane1[m] = 0.f + p1[cx][cz];
            }
            if (inr(z, x)) {
                const float c = p1[cx][cz];
                const float lx = W4 * (p1[cx + 4][cz] + p1[cx - 4][cz]) + W3 * (p1[cx + 3][cz] + p1[cx - 3][cz]) +
                                 W2 * (p1[cx + 2][cz] + p1[cx - 2][cz]) + W1 * (p1[cx + 1][cz] + p1[cx - 1][cz]) +
                                 W0 * c;
                const float lz = W4 * (p1[cx][cz + 4] + p1[cx][cz - 4]) + W3 * (p1[cx][cz + 3] + p1[cx][cz - 3]) +
                                 W2 * (p1[cx][cz + 2] + p1[cx][cz - 2]) + W1 * (p1[cx][cz + 1] + p1[cx][cz - 1]) +
                                 W0 * c;
                u2 = (lx * g.sx + lz * g.sz) + 2.0f * u1 - decay(prev, zl, zh, xl, xh);
            }
            if (rcell) u2 = u2 + pl1[j];
        }
        a.D[so + (long long)x * g.ldz + z] = decay(u2, zl, zh, xl, xh);
    }
}

// ---------------------------------------------------------------- NS time steps per launch
// The two-step scheme carried to NS steps: a block owns an OZN x OX tile and stages the current
// level on the tile grown by 4*NS cells.  Step k (1..NS) forms its level on the tile grown by
// 4*(NS-k), reading the level before it through the stencil and the one before that pointwise, so
// two LDS level buffers suffice: A holds the even levels (the staged current level, then 2, ...),
// B the odd ones (the previous level as staged, then 1, 3, ...), and a step overwrites, cell by
// cell, the level it reads pointwise.  Every cell uses the single-step kernels' expressions, so each
// level equals NS single steps bit for bit.  The last two levels go to the spare fields C (step
// NS-1) and D (step NS).
constexpr int OZN = 56;  // output z per tile
constexpr int TN = 1024;  // threads per block (16 waves: the LDS reads of the steps need them in flight)

template <int NS>
struct TriArgsN {
    const float* X;       // current level (modelling: u1; adjoint: A)
    const float* Y;       // previous level (decayed once)
    float* C;             // out: level NS-1 (the next launch's previous level)
    float* D;             // out: level NS (the next launch's current level)
    const float* v2;
    const float* sfz;
    const float* sfx;
    float* plane[NS];     // per step: modelling source plane (read); adjoint image plane (written)
    long long pss;
    float* row[NS];       // per step: modelling recorded row (written); adjoint injected row (read)
    long long rss;
};

// cells of the region of step k (the tile grown by 4*(NS-k))
template <int NS, int OX>
__host__ __device__ constexpr int trn_zr(int k) { return OZN + 8 * (NS - k); }
template <int NS, int OX>
__host__ __device__ constexpr int trn_xr(int k) { return OX + 8 * (NS - k); }
template <int NS, int OX>
__host__ __device__ constexpr int trn_sp_off(int k)  // offset of step k's source plane in the staged planes
{
    return k <= 1 ? 0 : trn_sp_off<NS, OX>(k - 1) + trn_zr<NS, OX>(k - 1) * trn_xr<NS, OX>(k - 1);
}

template <bool ADJ, int NS, int OX>
__global__ __launch_bounds__(TN) void k_trn(TriGeo g, TriArgsN<NS> a, int ntz)
{
    static_assert(NS >= 2, "at least two steps");
    constexpr int G = 4 * NS;
    constexpr int ZA = OZN + 2 * G, XA = OX + 2 * G;              // staged region R_0
    constexpr int ZB = OZN + 2 * (G - 4), XB = OX + 2 * (G - 4);  // R_1
    constexpr int NSP = ADJ ? 1 : trn_sp_off<NS, OX>(NS + 1);     // staged source planes
    __shared__ float LA[XA][ZA + 1];                 // even levels, on R_0 coordinates
    __shared__ float LB[XB][ZB + 1];                 // the staged previous level, then the odd levels (R_1)
    __shared__ float PA[ADJ ? XA : 1][ZA + 1];       // adjoint: level x v^2dt^2
    __shared__ float PB[ADJ ? XB : 1][ZB + 1];
    __shared__ float VV[XA][ZA + 1];                 // v^2dt^2 on R_0
    __shared__ float SP[NSP];                        // modelling: each step's source plane on its region
    __shared__ float fz[ZA][2], fx[XA][2];           // decay factor pairs of R_0's rows and columns
    __shared__ float rd[ADJ ? NS : 1][XB];           // adjoint: each step's injected row over R_1's columns
    const int slot = blockIdx.y;
    const int tz = blockIdx.x % ntz, tx = blockIdx.x / ntz;
    const int zo = tz * OZN, xo = tx * OX;
    const long long so = slot * g.fs;
    const float* __restrict__ X = a.X + so;
    const float* __restrict__ Y = a.Y + so;
    auto ing = [&](int z, int x) { return z >= 0 && z < g.pz && x >= 0 && x < g.px; };
    auto inr = [&](int z, int x) { return z >= 4 && z < g.pz - 4 && x >= 4 && x < g.px - 4; };
    auto mz = [&](int z) { return z >= g.nb && z < g.nb + g.nz; };
    auto mx = [&](int x) { return x >= g.nb && x < g.nb + g.nx; };
    auto own = [&](int z, int x) { return z >= zo && z < zo + OZN && x >= xo && x < xo + OX; };
    auto mi = [&](int z, int x) { return slot * a.pss + (long long)(x - g.nb) * g.nz + (z - g.nb); };
    // ---- stage: current level, v^2dt^2 (and their product) on R_0, the previous level on R_1,
    // the decay pairs, and each step's source plane (modelling) or injected row (adjoint).  Every
    // global operand is loaded into registers first (all loads in flight together), then written
    // to LDS, so the launch waits on memory once.
    constexpr int N0 = (ZA * XA + TN - 1) / TN, N1 = (ZB * XB + TN - 1) / TN, NP = (NSP + TN - 1) / TN;
    float r0[N0], rv[N0], r1[N1], rp[ADJ ? 1 : NP];
#pragma unroll
    for (int j = 0; j < N0; ++j) {
        const int e = threadIdx.x + j * TN;
        const int zz = e % ZA, xx = e / ZA;
        const int z = zo - G + zz, x = xo - G + xx;
        const bool in = e < ZA * XA && ing(z, x);
        const long long k = (long long)x * g.ldz + z;
        r0[j] = in ? X[k] : 0.f;
        rv[j] = in ? a.v2[k] : 0.f;
    }
#pragma unroll
    for (int j = 0; j < N1; ++j) {
        const int e = threadIdx.x + j * TN;
        const int zz = e % ZB, xx = e / ZB;
        const int z = zo - G + 4 + zz, x = xo - G + 4 + xx;
        r1[j] = (e < ZB * XB && ing(z, x)) ? Y[(long long)x * g.ldz + z] : 0.f;
    }
    if constexpr (!ADJ) {
        // flat index over the steps' regions back to back: step k's region at trn_sp_off(k)
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            const int e = threadIdx.x + j * TN;
            int k = 1;
            while (k < NS && e >= trn_sp_off<NS, OX>(k + 1)) ++k;
            const int zr = trn_zr<NS, OX>(k), el = e - trn_sp_off<NS, OX>(k);
            const int zz = el % zr, xx = el / zr;
            const int z = zo - 4 * (NS - k) + zz, x = xo - 4 * (NS - k) + xx;
            rp[j] = (e < NSP && ing(z, x) && mz(z) && mx(x)) ? a.plane[k - 1][mi(z, x)] : 0.f;
        }
    } else {
        for (int e = threadIdx.x; e < NS * XB; e += TN) {
            const int k = e / XB, xx = e % XB;
            const int x = xo - G + 4 + xx;
            rd[k][xx] = (x >= 0 && x < g.px && mx(x)) ? a.row[k][slot * a.rss + (x - g.nb)] : 0.f;
        }
    }
#pragma unroll
    for (int j = 0; j < N0; ++j) {
        const int e = threadIdx.x + j * TN;
        if (e < ZA * XA) {
            const int zz = e % ZA, xx = e / ZA;
            LA[xx][zz] = r0[j];
            VV[xx][zz] = rv[j];
            if constexpr (ADJ) PA[xx][zz] = r0[j] * rv[j];
        }
    }
#pragma unroll
    for (int j = 0; j < N1; ++j) {
        const int e = threadIdx.x + j * TN;
        if (e < ZB * XB) LB[e / ZB][e % ZB] = r1[j];
    }
    if constexpr (!ADJ) {
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            const int e = threadIdx.x + j * TN;
            if (e < NSP) SP[e] = rp[j];
        }
    }
    for (int t = threadIdx.x; t < ZA + XA; t += TN) {
        if (t < ZA) {
            const int z = zo - G + t;
            const bool zv = z >= 0 && z < g.pz;
            fz[t][0] = zv ? a.sfz[2 * z] : 1.f;
            fz[t][1] = zv ? a.sfz[2 * z + 1] : 1.f;
        } else {
            const int u = t - ZA, x = xo - G + u;
            const bool xv = x >= 0 && x < g.px;
            fx[u][0] = xv ? a.sfx[2 * x] : 1.f;
            fx[u][1] = xv ? a.sfx[2 * x + 1] : 1.f;
        }
    }
    __syncthreads();
    // ---- steps 1..NS
#pragma unroll
    for (int k = 1; k <= NS; ++k) {
        const int zr = trn_zr<NS, OX>(k), xr = trn_xr<NS, OX>(k);
        const int da = 4 * k, db = 4 * (k - 1);  // R_k's offsets in R_0 and R_1 coordinates
        const bool odd = k & 1;                  // level k goes to B (odd) or A (even)
        for (int e = threadIdx.x; e < zr * xr; e += TN) {
            const int zz = e % zr, xx = e / zr;
            const int z = zo - 4 * (NS - k) + zz, x = xo - 4 * (NS - k) + xx;
            const int za = zz + da, xa = xx + da, zb = zz + db, xb = xx + db;
            float v = 0.f, pv = 0.f;
            if (ing(z, x)) {
                const float zl = fz[za][0], zh = fz[za][1], xl = fx[xa][0], xh = fx[xa][1];
                const float vv = VV[xa][za];
                // the level before this step (stencil) and the one before that (pointwise)
                const float u1 = odd ? LA[xa][za] : LB[xb][zb];
                const float prev = odd ? LB[xb][zb] : LA[xa][za];
                float u2 = 0.f;  // the 4-cell rim stays zero
                if constexpr (!ADJ) {
                    if (g.depth == z && mx(x) && own(z, x))  // record the level before the step
                        a.row[k - 1][slot * a.rss + (x - g.nb)] = 0.f + u1;
                    if (inr(z, x)) {
                        float lx, lz;
                        if (odd) {
                            lx = W4 * (LA[xa + 4][za] + LA[xa - 4][za]) + W3 * (LA[xa + 3][za] + LA[xa - 3][za]) +
                                 W2 * (LA[xa + 2][za] + LA[xa - 2][za]) + W1 * (LA[xa + 1][za] + LA[xa - 1][za]) +
                                 W0 * u1;
                            lz = W4 * (LA[xa][za + 4] + LA[xa][za - 4]) + W3 * (LA[xa][za + 3] + LA[xa][za - 3]) +
                                 W2 * (LA[xa][za + 2] + LA[xa][za - 2]) + W1 * (LA[xa][za + 1] + LA[xa][za - 1]) +
                                 W0 * u1;
                        } else {
                            lx = W4 * (LB[xb + 4][zb] + LB[xb - 4][zb]) + W3 * (LB[xb + 3][zb] + LB[xb - 3][zb]) +
                                 W2 * (LB[xb + 2][zb] + LB[xb - 2][zb]) + W1 * (LB[xb + 1][zb] + LB[xb - 1][zb]) +
                                 W0 * u1;
                            lz = W4 * (LB[xb][zb + 4] + LB[xb][zb - 4]) + W3 * (LB[xb][zb + 3] + LB[xb][zb - 3]) +
                                 W2 * (LB[xb][zb + 2] + LB[xb][zb - 2]) + W1 * (LB[xb][zb + 1] + LB[xb][zb - 1]) +
                                 W0 * u1;
                        }
                        u2 = (lx * g.sx + lz * g.sz) * vv + 2.0f * u1 - decay(prev, zl, zh, xl, xh);
                    }
                    if (mz(z) && mx(x)) u2 = u2 + SP[trn_sp_off<NS, OX>(k) + e] * vv;
                } else {
                    const float c = odd ? PA[xa][za] : PB[xb][zb];  // the level before, x v^2dt^2
                    if (mz(z) && mx(x) && own(z, x)) a.plane[k - 1][mi(z, x)] = 0.f + c;  // this step's image
                    if (inr(z, x)) {
                        float lx, lz;
                        if (odd) {
                            lx = W4 * (PA[xa + 4][za] + PA[xa - 4][za]) + W3 * (PA[xa + 3][za] + PA[xa - 3][za]) +
                                 W2 * (PA[xa + 2][za] + PA[xa - 2][za]) + W1 * (PA[xa + 1][za] + PA[xa - 1][za]) +
                                 W0 * c;
                            lz = W4 * (PA[xa][za + 4] + PA[xa][za - 4]) + W3 * (PA[xa][za + 3] + PA[xa][za - 3]) +
                                 W2 * (PA[xa][za + 2] + PA[xa][za - 2]) + W1 * (PA[xa][za + 1] + PA[xa][za - 1]) +
                                 W0 * c;
                        } else {
                            lx = W4 * (PB[xb + 4][zb] + PB[xb - 4][zb]) + W3 * (PB[xb + 3][zb] + PB[xb - 3][zb]) +
                                 W2 * (PB[xb + 2][zb] + PB[xb - 2][zb]) + W1 * (PB[xb + 1][zb] + PB[xb - 1][zb]) +
                                 W0 * c;
                            lz = W4 * (PB[xb][zb + 4] + PB[xb][zb - 4]) + W3 * (PB[xb][zb + 3] + PB[xb][zb - 3]) +
                                 W2 * (PB[xb][zb + 2] + PB[xb][zb - 2]) + W1 * (PB[xb][zb + 1] + PB[xb][zb - 1]) +
                                 W0 * c;
                        }
                        u2 = (lx * g.sx + lz * g.sz) + 2.0f * u1 - decay(prev, zl, zh, xl, xh);
                    }
                    if (z == g.depth && mx(x)) u2 = u2 + rd[k - 1][xb];
                }
                v = decay(u2, zl, zh, xl, xh);
                if constexpr (ADJ) pv = v * vv;
                if (k >= NS - 1 && own(z, x)) (k == NS ? a.D : a.C)[so + (long long)x * g.ldz + z] = v;
            }
            if (k < NS) {  // over the level this cell read pointwise (no other cell reads it now)
                if (odd) {
                    LB[xb][zb] = v;
                    if constexpr (ADJ) PB[xb][zb] = pv;
                } else {
                    LA[xa][za] = v;
                    if constexpr (ADJ) PA[xa][za] = pv;
                }
            }
        }
        if (k < NS) __syncthreads();
    }
}

// ---------------------------------------------------------------- small kernels
__global__ void k_tri_pad_v2(TriGeo g, const float* __restrict__ vel, float dt2, float* __restrict__ v2)
{
    const int iz = blockIdx.x * blockDim.x + threadIdx.x, ix = blockIdx.y;
    if (iz >= g.pz || ix >= g.px) return;
    const int jx = min(max(ix, g.nb), g.nb + g.nx - 1) - g.nb;
    const int jz = min(max(iz, g.nb), g.nb + g.nz - 1) - g.nb;
    const float v = vel[(long long)jx * g.nz + jz];
    v2[(long long)ix * g.ldz + iz] = v * v * dt2;
}

__global__ void k_zero(long long n, float* __restrict__ p)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0.f;
}

__global__ void k_negate(long long n, const float* __restrict__ a, float* __restrict__ out)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = -a[i];
}

// [R][C] -> [C][R]
__global__ void k_transpose32(const float* __restrict__ in, float* __restrict__ out, int R, int C)
{
    __shared__ float t[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    for (int k = threadIdx.y; k < 32; k += blockDim.y) {
        const int r = r0 + k, c = c0 + threadIdx.x;
        if (r < R && c < C) t[k][threadIdx.x] = in[(long long)r * C + c];
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += blockDim.y) {
        const int c = c0 + k, r = r0 + threadIdx.x;
        if (r < R && c < C) out[(long long)c * R + r] = t[threadIdx.x][k];
    }
}

constexpr int NRED = 512;  // blocks of the partial reductions

// partial sums over data vectors: gg.gg, Ss.Ss, gg.Ss, gg.rr, Ss.rr, rr.rr
__global__ __launch_bounds__(256) void k_dots_data(long long n, const float* __restrict__ gg,
                                                   const float* __restrict__ ss, const float* __restrict__ rr,
                                                   double* __restrict__ part)
{
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const double g = gg[i], t = ss[i], r = rr[i];
        s[0] += g * g; s[1] += t * t; s[2] += g * t; s[3] += g * r; s[4] += t * r; s[5] += r * r;
    }
    __shared__ double sh[6][256];
#pragma unroll
    for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 6) part[threadIdx.x * NRED + blockIdx.x] = sh[threadIdx.x][0];
}

// partial sum of squares over a model vector
__global__ __launch_bounds__(256) void k_sumsq(long long n, const float* __restrict__ a, double* __restrict__ part)
{
    double s = 0;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const double v = a[i];
        s += v * v;
    }
    __shared__ double sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// k sums of NRED partials each -> out[k]
__global__ __launch_bounds__(256) void k_reduce_parts(int k, const double* __restrict__ part, double* __restrict__ out)
{
    __shared__ double sh[256];
    for (int j = 0; j < k; ++j) {
        double s = 0;
        for (int i = threadIdx.x; i < NRED; i += 256) s += part[j * NRED + i];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[j] = sh[0];
        __syncthreads();
    }
}

// S = S*beta + alpha*g; x += S (model space) -- cblas_sscal, cblas_saxpy, then the model update
__global__ void k_cg_model(long long n, float alpha, float beta, const float* __restrict__ g, float* __restrict__ S,
                           float* __restrict__ x)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        float s = S[i] * beta;
        s = s + alpha * g[i];
        S[i] = s;
        x[i] = x[i] + s;
    }
}

// Ss = Ss*beta + alpha*gg; rr += Ss (data space)
__global__ void k_cg_data(long long n, float alpha, float beta, const float* __restrict__ gg, float* __restrict__ Ss,
                          float* __restrict__ rr)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        float s = Ss[i] * beta;
        s = s + alpha * gg[i];
        Ss[i] = s;
        rr[i] = rr[i] + s;
    }
}

inline int nblk(long long n, int bs = 256) { return (int)std::min<long long>((n + bs - 1) / bs, 1 << 16); }

}  // namespace

// ==================================================================================== plan
struct pw_lstri {
    pw_ctx* ctx = nullptr;
    int nz = 0, nx = 0, nt = 0, ns = 0, nb = 0;
    float dz = 0, dx = 0, dt = 0, coef = 0;
    TriGeo g{};
    int xct = 16, ntz = 0, nxc = 0;
    long long nzx = 0, nm = 0, nd = 0;  // model plane, model (per shot), data (per shot) sizes
    float* v2 = nullptr;
    float *sfz = nullptr, *sfx = nullptr;
    float* fields = nullptr;  // 4 * cap_slots fields (two levels, and two for the two-step launches)
    int cap_slots = 0;
    int spl = 1;              // time steps per sweep launch: 1, 2 or 4
    int oxn = 4;              // output columns of the four-step launches' tiles (4 or 8)
    // launch choice per shots-per-sweep count: {single-step tile width, steps per launch, four-step
    // tile columns}, from timing both sweeps' launches on the plan's own grid and buffers
    std::map<int, std::vector<int>> choice;
    int spl_forced = 0;       // 0: the timed choice; 1, 2 or 4 steps per launch
    int xct_forced = 0;       // 0: the timed choice; else this single-step tile width
    int oxn_forced = 0;       // 0: the timed choice; else the four-step tiles' output columns
    std::vector<float> tune_log;  // (kernel, width, ms per launch) of the last timing sweep
    int graph_xct = 0, graph_spl = -1, graph_oxn = 0;  // the launch choice the captured CG sweeps hold
    float *src = nullptr, *dat = nullptr, *stage = nullptr;  // batch buffers (modelling / imaging)
    int cap_src_slots = 0;
    float *x = nullptr, *gvec = nullptr, *S = nullptr, *rr = nullptr, *gg = nullptr, *Ss = nullptr;
    double* part = nullptr;
    double* sums = nullptr;
    bool have_model = false;
    double t_ms = 0;
    // the CG's two sweeps (rr -> g adjoint, g -> gg modelling) captured once and replayed
    bool use_graph = true;
    hipGraph_t g_adj = nullptr, g_fwd = nullptr;
    hipGraphExec_t x_adj = nullptr, x_fwd = nullptr;
};

namespace {

#define LS_HIP(call)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess) {                                                                         \
            p->ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);                            \
            return PW_ERR_HIP;                                                                          \
        }                                                                                               \
    } while (0)

int ls_alloc(pw_lstri* p, float** ptr, long long n)
{
    if (hipMalloc((void**)ptr, (size_t)std::max<long long>(n, 1) * sizeof(float)) != hipSuccess) {
        *ptr = nullptr;
        p->ctx->err = "pw_lstri: device allocation of " + std::to_string(n * 4) + " bytes failed";
        return PW_ERR_NOMEM;
    }
    return PW_OK;
}

int zero(pw_lstri* p, float* ptr, long long n)
{
    hipLaunchKernelGGL(k_zero, dim3(nblk(n)), dim3(256), 0, p->ctx->s(), n, ptr);
    LS_HIP(hipGetLastError());
    return PW_OK;
}

void drop_graphs(pw_lstri* p)
{
    if (p->x_adj) hipGraphExecDestroy(p->x_adj);
    if (p->x_fwd) hipGraphExecDestroy(p->x_fwd);
    if (p->g_adj) hipGraphDestroy(p->g_adj);
    if (p->g_fwd) hipGraphDestroy(p->g_fwd);
    p->x_adj = p->x_fwd = nullptr;
    p->g_adj = p->g_fwd = nullptr;
}

int ensure_fields(pw_lstri* p, int slots)
{
    if (slots <= p->cap_slots) return PW_OK;
    drop_graphs(p);  // captured sweeps refer to the old field buffers
    if (p->fields) hipFree(p->fields);
    p->fields = nullptr;
    p->cap_slots = 0;
    int rc = ls_alloc(p, &p->fields, 4LL * slots * p->g.fs);
    if (!rc) p->cap_slots = slots;
    return rc;
}

template <int XCT>
void launch_step(pw_lstri* p, bool adj, int nslot, const TriArgs& a)
{
    dim3 grid(p->ntz * ((p->g.px + XCT - 1) / XCT), nslot);
    if (adj) hipLaunchKernelGGL(k_tra<XCT>, grid, dim3(64), 0, p->ctx->s(), p->g, a, p->ntz);
    else hipLaunchKernelGGL(k_trm<XCT>, grid, dim3(64), 0, p->ctx->s(), p->g, a, p->ntz);
}

// two steps per launch (k_tr2): a.X, a.Y in; a.C, a.D out
void step2(pw_lstri* p, bool adj, int nslot, const TriArgs2& a)
{
    const int ntz2 = (p->g.pz + OZ2 - 1) / OZ2, ntx2 = (p->g.px + OX2 - 1) / OX2;
    dim3 grid(ntz2 * ntx2, nslot);
    if (adj) hipLaunchKernelGGL(k_tr2<true>, grid, dim3(T2), 0, p->ctx->s(), p->g, a, ntz2);
    else hipLaunchKernelGGL(k_tr2<false>, grid, dim3(T2), 0, p->ctx->s(), p->g, a, ntz2);
}

// four steps per launch (k_trn): a.X, a.Y in; a.C, a.D out
void step4(pw_lstri* p, bool adj, int nslot, const TriArgsN<4>& a)
{
    const int ntz = (p->g.pz + OZN - 1) / OZN, ntx = (p->g.px + p->oxn - 1) / p->oxn;
    dim3 grid(ntz * ntx, nslot);
    hipStream_t s = p->ctx->s();
    if (p->oxn == 8) {
        if (adj) hipLaunchKernelGGL((k_trn<true, 4, 8>), grid, dim3(TN), 0, s, p->g, a, ntz);
        else hipLaunchKernelGGL((k_trn<false, 4, 8>), grid, dim3(TN), 0, s, p->g, a, ntz);
    } else {
        if (adj) hipLaunchKernelGGL((k_trn<true, 4, 4>), grid, dim3(TN), 0, s, p->g, a, ntz);
        else hipLaunchKernelGGL((k_trn<false, 4, 4>), grid, dim3(TN), 0, s, p->g, a, ntz);
    }
}

void step(pw_lstri* p, bool adj, int nslot, const TriArgs& a)
{
    switch (p->xct) {
        case 1: launch_step<1>(p, adj, nslot, a); break;
        case 2: launch_step<2>(p, adj, nslot, a); break;
        case 4: launch_step<4>(p, adj, nslot, a); break;
        case 8: launch_step<8>(p, adj, nslot, a); break;
        default: launch_step<16>(p, adj, nslot, a); break;
    }
}

int ensure_batch(pw_lstri* p, int nslot);

// Milliseconds per launch of fn (two warm-up launches, then `reps` timed ones).
template <typename F>
int time_launches(pw_lstri* p, F fn, float* ms)
{
    hipStream_t s = p->ctx->s();
    const int reps = 10;
    hipEvent_t e0, e1;
    LS_HIP(hipEventCreate(&e0));
    LS_HIP(hipEventCreate(&e1));
    fn();
    fn();
    hipEventRecord(e0, s);
    for (int k = 0; k < reps; ++k) fn();
    hipEventRecord(e1, s);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    LS_HIP(e);
    *ms /= reps;
    return PW_OK;
}

// Launch choice for sweeps of nslot shots: each single-step tile width timed on a modelling and
// an adjoint step, then the two-step launches against two single steps of the fastest width.  The
// timing launches run on the sweep buffers before a sweep (or its graph capture): the sweep clears
// the fields and rewrites the data row and image plane they wrote.  Plans of one geometry on one
// context share the choice.
int select_launch(pw_lstri* p, int nslot)
{
    auto own = p->choice.find(nslot);
    if (own == p->choice.end()) {
        const std::vector<long long> key = {3, p->g.pz, p->g.px, p->g.nb, p->g.depth, p->nz, p->nx, p->nt, nslot};
        auto hit = p->ctx->tuned2d.find(key);
        if (hit != p->ctx->tuned2d.end()) {
            p->tune_log = hit->second.second;
            own = p->choice.emplace(nslot, hit->second.first).first;
        } else {
            int rc = ensure_fields(p, nslot);
            if (!rc) rc = ensure_batch(p, nslot);
            if (!rc) rc = zero(p, p->fields, 4LL * nslot * p->g.fs);
            if (rc) return rc;
            p->tune_log.clear();
            const long long fl = (long long)nslot * p->g.fs;
            TriArgs a{};
            a.X = p->fields; a.Y = p->fields + fl; a.v2 = p->v2; a.sfz = p->sfz; a.sfx = p->sfx;
            a.plane = p->src; a.pss = p->nm; a.row = p->dat; a.rss = p->nd;
            float best = 0.f;  // per time step
            int wbest = 1, sbest = 1, obest = 4;
            for (int W : {1, 2, 4, 8, 16}) {
                p->xct = W;
                float tm = 0.f, ta = 0.f;
                if ((rc = time_launches(p, [&] { step(p, false, nslot, a); }, &tm))) return rc;
                if ((rc = time_launches(p, [&] { step(p, true, nslot, a); }, &ta))) return rc;
                for (float v : {0.f, (float)W, tm, 1.f, (float)W, ta}) p->tune_log.push_back(v);
                if (W == 1 || tm + ta < best) { best = tm + ta; wbest = W; }
            }
            if (p->nt >= 2) {  // the two-step launch reads planes and rows it and it + 1
                TriArgs2 b{};
                b.X = p->fields; b.Y = p->fields + fl; b.C = p->fields + 2 * fl; b.D = p->fields + 3 * fl;
                b.v2 = p->v2; b.sfz = p->sfz; b.sfx = p->sfx;
                b.plane0 = p->src; b.plane1 = p->src + p->nzx; b.pss = p->nm;
                b.row0 = p->dat; b.row1 = p->dat + p->nx; b.rss = p->nd;
                float tm = 0.f, ta = 0.f;
                if ((rc = time_launches(p, [&] { step2(p, false, nslot, b); }, &tm))) return rc;
                if ((rc = time_launches(p, [&] { step2(p, true, nslot, b); }, &ta))) return rc;
                for (float v : {2.f, (float)OX2, tm, 3.f, (float)OX2, ta}) p->tune_log.push_back(v);
                if ((tm + ta) / 2.f < best) { best = (tm + ta) / 2.f; sbest = 2; }
            }
            if (p->nt >= 4) {  // the four-step launches read planes and rows it .. it + 3
                TriArgsN<4> b{};
                b.X = p->fields; b.Y = p->fields + fl; b.C = p->fields + 2 * fl; b.D = p->fields + 3 * fl;
                b.v2 = p->v2; b.sfz = p->sfz; b.sfx = p->sfx;
                b.pss = p->nm; b.rss = p->nd;
                for (int k = 0; k < 4; ++k) {
                    b.plane[k] = p->src + k * p->nzx;
                    b.row[k] = p->dat + k * p->nx;
                }
                for (int ox : {4, 8}) {
                    p->oxn = ox;
                    float tm = 0.f, ta = 0.f;
                    if ((rc = time_launches(p, [&] { step4(p, false, nslot, b); }, &tm))) return rc;
                    if ((rc = time_launches(p, [&] { step4(p, true, nslot, b); }, &ta))) return rc;
                    for (float v : {4.f, (float)ox, tm, 5.f, (float)ox, ta}) p->tune_log.push_back(v);
                    if ((tm + ta) / 4.f < best) { best = (tm + ta) / 4.f; sbest = 4; obest = ox; }
                }
            }
            own = p->choice.emplace(nslot, std::vector<int>{wbest, sbest, obest}).first;
            p->ctx->tuned2d[key] = {own->second, p->tune_log};
        }
    }
    p->xct = p->xct_forced ? p->xct_forced : own->second[0];
    p->spl = p->spl_forced ? p->spl_forced : own->second[1];
    p->oxn = p->oxn_forced ? p->oxn_forced : own->second[2];
    return PW_OK;
}

// modelling sweep: w [slot][nt][nx][nz] -> d [slot][nt][nx] (time-major), nslot shots at once
int sweep_model(pw_lstri* p, int nslot, const float* w, float* d)
{
    int rc = ensure_fields(p, nslot);
    if (!rc) rc = zero(p, p->fields, 2LL * nslot * p->g.fs);
    if (rc) return rc;
    const long long fl = (long long)nslot * p->g.fs;
    float* X = p->fields;
    float* Y = p->fields + fl;
    float* C = p->fields + 2 * fl;
    float* D = p->fields + 3 * fl;
    int it0 = 0;
    if (p->spl == 4) {
        for (; it0 + 3 < p->nt; it0 += 4) {
            TriArgsN<4> a{};
            a.X = X; a.Y = Y; a.C = C; a.D = D; a.v2 = p->v2; a.sfz = p->sfz; a.sfx = p->sfx;
            a.pss = p->nm;
            a.rss = p->nd;
            for (int k = 0; k < 4; ++k) {
                a.plane[k] = const_cast<float*>(w) + (long long)(it0 + k) * p->nzx;
                a.row[k] = d + (long long)(it0 + k) * p->nx;
            }
            step4(p, false, nslot, a);
            std::swap(X, D);  // the fourth step's level is current, the third step's previous
            std::swap(Y, C);
        }
    }
    if (p->spl >= 2) {
        for (; it0 + 1 < p->nt; it0 += 2) {
            TriArgs2 a{};
            a.X = X; a.Y = Y; a.C = C; a.D = D; a.v2 = p->v2; a.sfz = p->sfz; a.sfx = p->sfx;
            a.plane0 = const_cast<float*>(w) + (long long)it0 * p->nzx;
            a.plane1 = const_cast<float*>(w) + (long long)(it0 + 1) * p->nzx;
            a.pss = p->nm;
            a.row0 = d + (long long)it0 * p->nx;
            a.row1 = d + (long long)(it0 + 1) * p->nx;
            a.rss = p->nd;
            step2(p, false, nslot, a);
            std::swap(X, D);  // the second step's level is current, the first step's previous
            std::swap(Y, C);
        }
    }
    for (int it = it0; it < p->nt; ++it) {
        TriArgs a{};
        a.X = X; a.Y = Y; a.v2 = p->v2; a.sfz = p->sfz; a.sfx = p->sfx;
        a.plane = const_cast<float*>(w) + (long long)it * p->nzx;
        a.pss = p->nm;
        a.row = d + (long long)it * p->nx;
        a.rss = p->nd;
        step(p, false, nslot, a);
        std::swap(X, Y);
    }
    LS_HIP(hipGetLastError());
    return PW_OK;
}

// adjoint sweep: d [slot][nt][nx] -> image w [slot][nt][nx][nz] (every plane written)
int sweep_adjoint(pw_lstri* p, int nslot, const float* d, float* w)
{
    int rc = ensure_fields(p, nslot);
    if (!rc) rc = zero(p, p->fields, 2LL * nslot * p->g.fs);
    if (rc) return rc;
    const long long fl = (long long)nslot * p->g.fs;
    float* A = p->fields;
    float* B = p->fields + fl;
    float* C = p->fields + 2 * fl;
    float* D = p->fields + 3 * fl;
    int it0 = p->nt - 1;
    if (p->spl == 4) {
        for (; it0 >= 3; it0 -= 4) {
            TriArgsN<4> a{};
            a.X = A; a.Y = B; a.C = C; a.D = D; a.v2 = p->v2; a.sfz = p->sfz; a.sfx = p->sfx;
            a.pss = p->nm;
            a.rss = p->nd;
            for (int k = 0; k < 4; ++k) {
                a.plane[k] = w + (long long)(it0 - k) * p->nzx;
                a.row[k] = const_cast<float*>(d) + (long long)(it0 - k) * p->nx;
            }
            step4(p, true, nslot, a);
            std::swap(A, D);
            std::swap(B, C);
        }
    }
    if (p->spl >= 2) {
        for (; it0 >= 1; it0 -= 2) {
            TriArgs2 a{};
            a.X = A; a.Y = B; a.C = C; a.D = D; a.v2 = p->v2; a.sfz = p->sfz; a.sfx = p->sfx;
            a.plane0 = w + (long long)it0 * p->nzx;
            a.plane1 = w + (long long)(it0 - 1) * p->nzx;
            a.pss = p->nm;
            a.row0 = const_cast<float*>(d) + (long long)it0 * p->nx;
            a.row1 = const_cast<float*>(d) + (long long)(it0 - 1) * p->nx;
            a.rss = p->nd;
            step2(p, true, nslot, a);
            std::swap(A, D);
            std::swap(B, C);
        }
    }
    for (int it = it0; it >= 0; --it) {
        TriArgs a{};
        a.X = A; a.Y = B; a.v2 = p->v2; a.sfz = p->sfz; a.sfx = p->sfx;
        a.plane = w + (long long)it * p->nzx;
        a.pss = p->nm;
        a.row = const_cast<float*>(d) + (long long)it * p->nx;
        a.rss = p->nd;
        step(p, true, nslot, a);
        std::swap(A, B);
    }
    LS_HIP(hipGetLastError());
    return PW_OK;
}

// data (nt, nx, ns) Fortran = [is][ix][it] on the host <-> [is][it][ix] on the device
int upload_data(pw_lstri* p, const float* host, int s0, int nslot, float* dev)
{
    hipStream_t s = p->ctx->s();
    const long long nd = p->nd;
    LS_HIP(hipMemcpyAsync(p->stage, host + (size_t)s0 * nd, (size_t)nslot * nd * 4, hipMemcpyHostToDevice, s));
    for (int k = 0; k < nslot; ++k)
        hipLaunchKernelGGL(k_transpose32, dim3((p->nt + 31) / 32, (p->nx + 31) / 32), dim3(32, 8), 0, s,
                           p->stage + k * nd, dev + k * nd, p->nx, p->nt);
    LS_HIP(hipGetLastError());
    return PW_OK;
}

int download_data(pw_lstri* p, const float* dev, int s0, int nslot, float* host)
{
    hipStream_t s = p->ctx->s();
    const long long nd = p->nd;
    for (int k = 0; k < nslot; ++k)
        hipLaunchKernelGGL(k_transpose32, dim3((p->nx + 31) / 32, (p->nt + 31) / 32), dim3(32, 8), 0, s,
                           dev + k * nd, p->stage + k * nd, p->nt, p->nx);
    LS_HIP(hipGetLastError());
    LS_HIP(hipMemcpyAsync(host + (size_t)s0 * nd, p->stage, (size_t)nslot * nd * 4, hipMemcpyDeviceToHost, s));
    LS_HIP(hipStreamSynchronize(s));
    return PW_OK;
}

// batch buffers for nslot shots of modelling or imaging
int ensure_batch(pw_lstri* p, int nslot)
{
    if (nslot <= p->cap_src_slots) return PW_OK;
    float* bufs[] = {p->src, p->dat, p->stage};
    for (float* b : bufs)
        if (b) hipFree(b);
    p->src = p->dat = p->stage = nullptr;
    p->cap_src_slots = 0;
    int rc = ls_alloc(p, &p->src, (long long)nslot * p->nm);
    if (!rc) rc = ls_alloc(p, &p->dat, (long long)nslot * p->nd);
    if (!rc) rc = ls_alloc(p, &p->stage, (long long)nslot * p->nd);
    if (!rc) p->cap_src_slots = nslot;
    return rc;
}

// shots per batch so that the source/image batch stays within about a quarter of free HBM
int batch_slots(pw_lstri* p)
{
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = (size_t)8 << 30;
    const double per = (double)(p->nm + 2 * p->nd + 2 * p->g.fs) * 4;
    long long k = (long long)(fr / 4 / std::max(per, 1.0));
    return (int)std::max<long long>(1, std::min<long long>(k, p->ns));
}

int reduce(pw_lstri* p, int k, double* out_host)
{
    hipStream_t s = p->ctx->s();
    hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(256), 0, s, k, p->part, p->sums);
    LS_HIP(hipGetLastError());
    LS_HIP(hipMemcpyAsync(out_host, p->sums, k * sizeof(double), hipMemcpyDeviceToHost, s));
    LS_HIP(hipStreamSynchronize(s));
    return PW_OK;
}

void ls_free(pw_lstri* p)
{
    drop_graphs(p);
    float* bufs[] = {p->v2, p->sfz, p->sfx, p->fields, p->src, p->dat, p->stage, p->x, p->gvec, p->S, p->rr, p->gg, p->Ss};
    for (float* b : bufs)
        if (b) hipFree(b);
    if (p->part) hipFree(p->part);
    if (p->sums) hipFree(p->sums);
}

}  // namespace

// ==================================================================================== C ABI
extern "C" {

pw_lstri* pw_lstri_create(pw_ctx* ctx, const float* pararray)
{
    if (!ctx || !pararray) return nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess) {
        ctx->err = "hipSetDevice failed";
        return nullptr;
    }
    const float* q = pararray;
    pw_lstri* p = new pw_lstri;
    p->ctx = ctx;
    p->nz = (int)q[0]; p->nx = (int)q[1]; p->dz = q[2]; p->dx = q[3];
    p->nt = (int)q[6]; p->dt = q[7]; p->ns = (int)q[10]; p->nb = (int)q[14]; p->coef = q[15];
    if (p->nz < 1 || p->nx < 1 || p->nt < 1 || p->ns < 1 || !(p->dx > 0) || !(p->dz > 0) || !(p->dt > 0)) {
        ctx->err = "pw_lstri_create: invalid dimensions (nz, nx, nt, ns, dx, dz, dt must be positive)";
        delete p;
        return nullptr;
    }
    if (p->nb < 4) {
        ctx->err = "pw_lstri_create: nb must be >= 4 (sources and receivers must stay off the 4-cell rim)";
        delete p;
        return nullptr;
    }
    TriGeo& g = p->g;
    g.nz = p->nz; g.nx = p->nx; g.nb = p->nb;
    g.pz = p->nz + 2 * p->nb; g.px = p->nx + 2 * p->nb;
    g.ldz = (g.pz + 31) / 32 * 32;
    g.depth = 1 + p->nb;  // tri2d_make: depth (rz - nb = 1) + nb
    g.sx = 1.0f / (p->dx * p->dx);
    g.sz = 1.0f / (p->dz * p->dz);
    g.fs = (long long)g.ldz * g.px;
    p->nzx = (long long)p->nz * p->nx;
    p->nm = p->nzx * p->nt;
    p->nd = (long long)p->nx * p->nt;
    if ((double)(g.fs + 8 * g.ldz + 8) * 4 >= 2147483648.0 || (double)p->nzx * 4 >= 2147483648.0) {
        ctx->err = "pw_lstri_create: padded model too large for one field window (2 GiB)";
        delete p;
        return nullptr;
    }
    if (g.depth >= g.pz - 4) {
        ctx->err = "pw_lstri_create: receiver row reaches the 4-cell rim (nz too small)";
        delete p;
        return nullptr;
    }
    p->ntz = (g.pz + 63) / 64;  // tile width and steps per launch: select_launch, before a sweep
    int rc = ls_alloc(p, &p->v2, g.fs);
    if (!rc) rc = ls_alloc(p, &p->sfz, 2LL * g.pz);
    if (!rc) rc = ls_alloc(p, &p->sfx, 2LL * g.px);
    if (!rc && hipMalloc((void**)&p->part, 6 * NRED * sizeof(double)) != hipSuccess) rc = PW_ERR_NOMEM;
    if (!rc && hipMalloc((void**)&p->sums, 8 * sizeof(double)) != hipSuccess) rc = PW_ERR_NOMEM;
    // decay weights (pfwiabc_cal): w[i] = exp(-c*c*(nb-1-i)^2), f32 products, exp in double
    std::vector<float> dec(p->nb), fz(2 * (size_t)g.pz, 1.f), fx(2 * (size_t)g.px, 1.f);
    for (int i = 0; i < p->nb; ++i) {
        const int k = p->nb - 1 - i;
        dec[i] = (float)std::exp((double)(-p->coef * p->coef * (float)k * (float)k));
    }
    for (int i = 0; i < p->nb; ++i) {
        fz[2 * i] = dec[i];
        fz[2 * (size_t)(g.pz - 1 - i) + 1] = dec[i];
        fx[2 * i] = dec[i];
        fx[2 * (size_t)(g.px - 1 - i) + 1] = dec[i];
    }
    hipStream_t s = ctx->s();
    if (!rc && (hipMemcpyAsync(p->sfz, fz.data(), fz.size() * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
                hipMemcpyAsync(p->sfx, fx.data(), fx.size() * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess))
        rc = PW_ERR_HIP;
    if (!rc) rc = zero(p, p->v2, g.fs);
    if (rc) {
        if (ctx->err.empty()) ctx->err = "pw_lstri_create: set-up failed";
        ls_free(p);
        delete p;
        return nullptr;
    }
    return p;
}

void pw_lstri_destroy(pw_lstri* p)
{
    if (!p) return;
    hipSetDevice(p->ctx->device);
    hipStreamSynchronize(p->ctx->s());
    ls_free(p);
    delete p;
}

// vel (nz, nx) Fortran order, km/s or any unit consistent with dx, dz, dt
int pw_lstri_set_model(pw_lstri* p, const float* vel)
{
    if (!p || !vel) return PW_ERR_ARG;
    hipStream_t s = p->ctx->s();
    float* dv = nullptr;
    int rc = ls_alloc(p, &dv, p->nzx);
    if (rc) return rc;
    const float dt2 = p->dt * p->dt;
    if (hipMemcpyAsync(dv, vel, p->nzx * 4, hipMemcpyHostToDevice, s) != hipSuccess) rc = PW_ERR_HIP;
    if (!rc) {
        hipLaunchKernelGGL(k_tri_pad_v2, dim3((p->g.pz + 255) / 256, p->g.px), dim3(256), 0, s, p->g, dv, dt2, p->v2);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) rc = PW_ERR_HIP;
    }
    hipFree(dv);
    if (rc) PW_FAIL(p->ctx, rc, "pw_lstri_set_model: upload failed");
    p->have_model = true;
    return PW_OK;
}

// timerev_lop modelling for every shot: src (nz, nx, nt, ns) Fortran -> data (nt, nx, ns) Fortran
int pw_lstri_model(pw_lstri* p, const float* src, float* data)
{
    if (!p || !src || !data) return PW_ERR_ARG;
    if (!p->have_model) PW_FAIL(p->ctx, PW_ERR_STATE, "pw_lstri_model: model not set");
    hipStream_t s = p->ctx->s();
    const int bs = batch_slots(p);
    int rc = ensure_batch(p, bs);
    if (rc) return rc;
    hipEvent_t e0, e1;
    LS_HIP(hipEventCreate(&e0)); LS_HIP(hipEventCreate(&e1));
    float ms_tot = 0;
    for (int s0 = 0; s0 < p->ns; s0 += bs) {
        const int k = std::min(bs, p->ns - s0);
        if ((rc = select_launch(p, k))) break;
        LS_HIP(hipMemcpyAsync(p->src, src + (size_t)s0 * p->nm, (size_t)k * p->nm * 4, hipMemcpyHostToDevice, s));
        LS_HIP(hipEventRecord(e0, s));
        if ((rc = sweep_model(p, k, p->src, p->dat))) break;
        LS_HIP(hipEventRecord(e1, s));
        if ((rc = download_data(p, p->dat, s0, k, data))) break;
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        ms_tot += ms;
    }
    hipEventDestroy(e0); hipEventDestroy(e1);
    p->t_ms = ms_tot;
    return rc;
}

// timerev_lop adjoint for every shot: data (nt, nx, ns) Fortran -> image field (nz, nx, nt, ns)
int pw_lstri_image(pw_lstri* p, const float* data, float* img)
{
    if (!p || !data || !img) return PW_ERR_ARG;
    if (!p->have_model) PW_FAIL(p->ctx, PW_ERR_STATE, "pw_lstri_image: model not set");
    hipStream_t s = p->ctx->s();
    const int bs = batch_slots(p);
    int rc = ensure_batch(p, bs);
    if (rc) return rc;
    for (int s0 = 0; s0 < p->ns && !rc; s0 += bs) {
        const int k = std::min(bs, p->ns - s0);
        if ((rc = select_launch(p, k))) break;
        if ((rc = upload_data(p, data, s0, k, p->dat))) break;
        if ((rc = sweep_adjoint(p, k, p->dat, p->src))) break;
        LS_HIP(hipMemcpyAsync(img + (size_t)s0 * p->nm, p->src, (size_t)k * p->nm * 4, hipMemcpyDeviceToHost, s));
        LS_HIP(hipStreamSynchronize(s));
    }
    return rc;
}

// lstri with inversion: for each shot, niter conjugate-gradient iterations of timerev(w) ~= d
// from w = 0 (np_solver + np_cgstep).  The step vectors carry over from one shot to the next;
// the first shot of the call restarts.  src out: (nz, nx, nt, ns) Fortran.  iters_out (ns ints,
// may be NULL): iterations run per shot (fewer than niter after convergence).
int pw_lstri_invert(pw_lstri* p, const float* data, int niter, float* src, int* iters_out)
{
    if (!p || !data || !src || niter < 0) return PW_ERR_ARG;
    if (!p->have_model) PW_FAIL(p->ctx, PW_ERR_STATE, "pw_lstri_invert: model not set");
    hipStream_t s = p->ctx->s();
    int rc = 0;
    if (!p->x) {
        rc = ls_alloc(p, &p->x, p->nm);
        if (!rc) rc = ls_alloc(p, &p->gvec, p->nm);
        if (!rc) rc = ls_alloc(p, &p->S, p->nm);
        if (!rc) rc = ls_alloc(p, &p->rr, p->nd);
        if (!rc) rc = ls_alloc(p, &p->gg, p->nd);
        if (!rc) rc = ls_alloc(p, &p->Ss, p->nd);
        if (!rc) rc = zero(p, p->S, p->nm);
        if (!rc) rc = zero(p, p->Ss, p->nd);
        if (rc) return rc;
    }
    if ((rc = ensure_batch(p, 1)) || (rc = ensure_fields(p, 1)) || (rc = select_launch(p, 1))) return rc;
    if (p->graph_xct != p->xct || p->graph_spl != p->spl || p->graph_oxn != p->oxn) drop_graphs(p);
    const long long nm = p->nm, nd = p->nd;
    const bool graph = p->use_graph && s != nullptr;
    if (graph && !p->x_adj) {
        p->graph_xct = p->xct;
        p->graph_spl = p->spl;
        p->graph_oxn = p->oxn;
        LS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        rc = sweep_adjoint(p, 1, p->rr, p->gvec);
        hipError_t e = hipStreamEndCapture(s, &p->g_adj);
        if (rc) return rc;
        LS_HIP(e);
        LS_HIP(hipGraphInstantiate(&p->x_adj, p->g_adj, nullptr, nullptr, 0));
        LS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        rc = sweep_model(p, 1, p->gvec, p->gg);
        e = hipStreamEndCapture(s, &p->g_fwd);
        if (rc) return rc;
        LS_HIP(e);
        LS_HIP(hipGraphInstantiate(&p->x_fwd, p->g_fwd, nullptr, nullptr, 0));
    }
    bool started = false;
    hipEvent_t e0, e1;
    LS_HIP(hipEventCreate(&e0)); LS_HIP(hipEventCreate(&e1));
    LS_HIP(hipEventRecord(e0, s));
    for (int is = 0; is < p->ns && !rc; ++is) {
        // rr = -d, x = 0
        if ((rc = upload_data(p, data, is, 1, p->dat))) break;
        hipLaunchKernelGGL(k_negate, dim3(nblk(nd)), dim3(256), 0, s, nd, p->dat, p->rr);
        if ((rc = zero(p, p->x, nm))) break;
        double sums[7];
        hipLaunchKernelGGL(k_sumsq, dim3(NRED), dim3(256), 0, s, nd, p->rr, p->part);
        if ((rc = reduce(p, 1, sums))) break;
        const double r0 = sums[0];
        double g0 = 1.0;
        int iter = 0;
        for (; iter < niter; ++iter) {
            if (graph) {
                LS_HIP(hipGraphLaunch(p->x_adj, s));
                LS_HIP(hipGraphLaunch(p->x_fwd, s));
            } else {
                if ((rc = sweep_adjoint(p, 1, p->rr, p->gvec))) break;
                if ((rc = sweep_model(p, 1, p->gvec, p->gg))) break;
            }
            // data-space dots (gg.gg, Ss.Ss, gg.Ss, gg.rr, Ss.rr, rr.rr) and |g|^2
            hipLaunchKernelGGL(k_dots_data, dim3(NRED), dim3(256), 0, s, nd, p->gg, p->Ss, p->rr, p->part);
            hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(256), 0, s, 6, p->part, p->sums);
            hipLaunchKernelGGL(k_sumsq, dim3(NRED), dim3(256), 0, s, nm, p->gvec, p->part);
            hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(256), 0, s, 1, p->part, p->sums + 6);
            LS_HIP(hipGetLastError());
            LS_HIP(hipMemcpyAsync(sums, p->sums, 7 * sizeof(double), hipMemcpyDeviceToHost, s));
            LS_HIP(hipStreamSynchronize(s));
            // convergence test of np_solver (squared norms relative to the first iteration)
            double dpr = 1.0, dpg = 1.0;
            if (iter == 0) g0 = sums[6];
            else { dpr = sums[5] / r0; dpg = sums[6] / g0; }
            if (dpr < 1.e-12 || dpg < 1.e-12) break;
            // np_cgstep
            double alfa, beta;
            bool skip = false;
            if (!started) {
                started = true;
                if ((rc = zero(p, p->S, nm)) || (rc = zero(p, p->Ss, nd))) break;
                beta = 0.0;
                if (sums[0] <= 0.) { skip = true; alfa = 0; }
                else alfa = -sums[3] / sums[0];
            } else {
                const double gdg = sums[0], sds = sums[1], gds = sums[2];
                if (gdg == 0. || sds == 0.) {
                    skip = true; alfa = beta = 0;
                } else {
                    double det = 1.0 - (gds / gdg) * (gds / sds);
                    if (det > (double)1.e-12f) det *= gdg * sds;
                    else det = gdg * sds * (double)1.e-12f;
                    const double gdr = -sums[3], sdr = -sums[4];
                    alfa = (sds * gdr - gds * sdr) / det;
                    beta = (-gds * gdr + gdg * sdr) / det;
                }
            }
            if (skip) continue;
            hipLaunchKernelGGL(k_cg_model, dim3(nblk(nm)), dim3(256), 0, s, nm, (float)alfa, (float)beta, p->gvec, p->S,
                               p->x);
            hipLaunchKernelGGL(k_cg_data, dim3(nblk(nd)), dim3(256), 0, s, nd, (float)alfa, (float)beta, p->gg, p->Ss,
                               p->rr);
            LS_HIP(hipGetLastError());
        }
        if (rc) break;
        if (iters_out) iters_out[is] = iter;
        LS_HIP(hipMemcpyAsync(src + (size_t)is * nm, p->x, (size_t)nm * 4, hipMemcpyDeviceToHost, s));
        LS_HIP(hipStreamSynchronize(s));
    }
    LS_HIP(hipEventRecord(e1, s));
    LS_HIP(hipEventSynchronize(e1));
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    p->t_ms = ms;
    hipEventDestroy(e0); hipEventDestroy(e1);
    return rc;
}

// replay the CG's sweeps from captured hipGraphs (default) or launch them step by step
int pw_lstri_use_graph(pw_lstri* p, int on)
{
    if (!p) return PW_ERR_ARG;
    p->use_graph = on != 0;
    return PW_OK;
}

// time steps per launch of the sweeps: 1, 2, 4, or 0 for the plan's timed choice; the results
// are the same bit for bit
int pw_lstri_steps_per_launch(pw_lstri* p, int n)
{
    if (!p || !(n == 0 || n == 1 || n == 2 || n == 4)) return PW_ERR_ARG;
    p->spl_forced = n;
    return PW_OK;
}

// output columns of the four-step launches' tiles: 4, 8, or 0 for the plan's timed choice
int pw_lstri_set_four_step_columns(pw_lstri* p, int c)
{
    if (!p || !(c == 0 || c == 4 || c == 8)) return PW_ERR_ARG;
    p->oxn_forced = c;
    return PW_OK;
}

// single-step tile width: 1, 2, 4, 8 or 16 columns, or 0 for the plan's timed choice
int pw_lstri_set_tile_width(pw_lstri* p, int w)
{
    if (!p || !(w == 0 || w == 1 || w == 2 || w == 4 || w == 8 || w == 16)) return PW_ERR_ARG;
    p->xct_forced = w;
    return PW_OK;
}

// the launch choice of the last sweep: w[0] single-step tile width, w[1] steps per launch (1, 2, 4),
// w[2] the four-step tiles' output columns; the timing sweep's (kernel, width, ms per launch)
// triples into log[3*cap] (kernel 0 modelling step, 1 adjoint step, 2 two-step modelling, 3
// two-step adjoint, 4 four-step modelling, 5 four-step adjoint); returns their number
int pw_lstri_launch_choice(pw_lstri* p, int* w, float* log, int cap)
{
    if (!p) return PW_ERR_ARG;
    if (w) { w[0] = p->xct; w[1] = p->spl; w[2] = p->oxn; }
    const int n = (int)p->tune_log.size() / 3;
    for (int i = 0; i < std::min(n, cap) * 3 && log; ++i) log[i] = p->tune_log[i];
    return n;
}

// elapsed milliseconds of the last model / invert call (device work; invert includes its transfers)
int pw_lstri_timing(pw_lstri* p, double* ms)
{
    if (!p || !ms) return PW_ERR_ARG;
    *ms = p->t_ms;
    return PW_OK;
}

// pfwicfun.lstric (pywave/src/pfwi.c:1031-1505): pararray[9] (inv) != 0 -> in = data (nt*nx*ns),
// out = source field (nz*nx*nt*ns) after niter = pararray[19] CG iterations per shot; else in =
// source field, out = data.  q and wav are accepted and unused, as in the reference.
int pw_lstric(pw_ctx* ctx, const float* vel, const float* q, const float* wav, const float* datasrc,
              const float* pararray, float* out, long long out_len, long long in_len)
{
    (void)q; (void)wav;
    if (!ctx || !vel || !datasrc || !pararray || !out) return PW_ERR_ARG;
    const long long nz = (long long)pararray[0], nx = (long long)pararray[1], nt = (long long)pararray[6],
                    ns = (long long)pararray[10];
    const bool inv = pararray[9] != 0.f;
    const long long nd = nt * nx * ns, nm = nz * nx * nt * ns;
    if (in_len != (inv ? nd : nm)) PW_FAIL(ctx, PW_ERR_ARG, "lstric: input length");
    if (out_len != (inv ? nm : nd)) PW_FAIL(ctx, PW_ERR_ARG, "lstric: output length");
    pw_lstri* p = pw_lstri_create(ctx, pararray);
    if (!p) return PW_ERR_ARG;
    int rc = pw_lstri_set_model(p, vel);
    if (!rc) rc = inv ? pw_lstri_invert(p, datasrc, (int)pararray[19], out, nullptr) : pw_lstri_model(p, datasrc, out);
    pw_lstri_destroy(p);
    return rc;
}

}  // extern "C"
