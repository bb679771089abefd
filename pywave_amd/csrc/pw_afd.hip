// pw_afd.hip -- afd2d/afd3d on MI355X: one fused kernel per time step.
//
// Reference path: pywave/src/wave_fdm.c:397-686 (fdm2d) and :28-381 (fdm), set up by
// pywave/src/afd.c:23-578.  Per step the reference does, in separate passes,
//   (i)   wave = 4th-order Laplacian of curr on [2, n2-2) per axis (0 on the rim)
//   (ii)  c = curr; c += c - prev; prev = c_old; c += wave*vv      (updated domain [0,n))
//   (iii) source 3x3(x3) injection  curr += (vv*rick)/(|dz|+|dx|+|dy|+1)   or TRI plane
//   (iv)  damp curr and damp prev (faces applied top, bottom, lx, rx, ly, ry)
//   (v)   record the receiver row/plane, (vi) snapshot every jsnap steps.
// Here one kernel does all of it per cell.  The second damping of prev is applied when prev
// is read in the next step: the kernel reads C_n (already damped) and the raw C_{n-1},
// forms D(C_{n-1}) on the fly and writes D(C_n + (C_n - D(C_{n-1})) + lap*vv [+src]) over
// C_{n-1}.  Every cell uses the reference's f32 operation order; the library is built with
// -ffp-contract=off, so results are bit-identical to the reference C.
//
// HBM layout: z fastest, padded extents nz2 x nx2 (x ny2) of the reference with the z stride
// rounded up to 32 floats (128 B rows), slack cells zero.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>

#include <hip/hip_ext.h>

#include "pw_internal.h"

namespace {

constexpr float C0 = -30. / 12., C1 = +16. / 12., C2 = -1. / 12.;  // pywave/src/wave_fdm.c:15
// z stride of the fields: nz2 rounded up to 32 floats (rows on 128 B boundaries), or to 16 floats
// (64 B) where that lets a step's three fields fit the 256 MB Infinity Cache with room to spare.
// Measured on the plain 3D step: the 261^3 grid (nz2 = 270) with three 79 MB fields at a stride of
// 272 took 48.4 us per step against 49.1 us with 84 MB fields at 288; the 321^3 grid, whose fields
// exceed the cache at either stride, took 121 us at 336 against 108 us at 352.
constexpr double kCacheFitBytes = 240e6;
int z_stride(int nz2, int nx2, int ny2)
{
    const int a32 = pw::ceil_div(nz2, 32) * 32, a16 = pw::ceil_div(nz2, 16) * 16;
    const double fields16 = 3.0 * 4.0 * a16 * (double)nx2 * (double)ny2;
    return (a16 < a32 && fields16 <= kCacheFitBytes) ? a16 : a32;
}

struct Geo {
    int nz, nx, ny;     // updated domain (model + 2*nb)
    int nz2, nx2, ny2;  // padded extents
    int ldz;            // z stride
    long long ldxy;     // plane stride (ldz * nx2)
    int nb;             // absorbing width, all faces
    int nz1, nx1, ny1;  // model
    float idz, idx, idy, cc;
};

struct Step {
    const float* __restrict__ cur;  // C_n (damped)
    float* __restrict__ pv;         // in: C_{n-1} undamped-once; out: C_{n+1}
    const float* __restrict__ vv;   // v^2 dt^2, same layout
    const float2* __restrict__ fz;  // damping (lo, hi) per index
    const float2* __restrict__ fx;
    const float2* __restrict__ fy;
    const int* __restrict__ src;    // 3*ns: gz, gx, gy in padded coordinates
    const float* __restrict__ wav;  // ns*nt
    const float* __restrict__ tri;  // this step's receiver plane [ny1][nx1] or null
    float* __restrict__ rec;        // this step's receiver plane or null
    float* __restrict__ snap;       // this snapshot [ny1][nx1][nz1] or null
    const float* __restrict__ fz1;  // one damping factor per index (lean kernel; faces do not overlap)
    const float* __restrict__ fx1;
    const float* __restrict__ fy1;
    int ns, nt, it;
};

__device__ __forceinline__ float damp3(float v, float2 a, float2 b, float2 c)
{
    v = v * a.x; v = v * a.y;  // top, bottom
    v = v * b.x; v = v * b.y;  // left x, right x
    v = v * c.x; v = v * c.y;  // left y, right y
    return v;
}

__device__ __forceinline__ float damp2(float v, float2 a, float2 b)
{
    v = v * a.x; v = v * a.y;
    v = v * b.x; v = v * b.y;
    return v;
}

// source stencil contributions at one cell, in source order
__device__ __forceinline__ float inject(float c, float vvc, const Step& a, int z, int x, int y)
{
    for (int s = 0; s < a.ns; ++s) {
        int dz = z - a.src[3 * s], dx = x - a.src[3 * s + 1], dy = y - a.src[3 * s + 2];
        int adz = dz < 0 ? -dz : dz, adx = dx < 0 ? -dx : dx, ady = dy < 0 ? -dy : dy;
        if (adz <= 1 && adx <= 1 && ady <= 1) {
            float k = (float)(adz + adx + ady + 1);
            c = c + (vvc * a.wav[(size_t)s * a.nt + a.it]) / k;
        }
    }
    return c;
}

// ---------------------------------------------------------------- 3D: y-streaming 2.5D tiles
template <int TX, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_afd3d(Geo g, Step a, int ntz, int ntx, int nyc,
                                                      int nblk)
{
    constexpr int TZ = 64, RPW = TX / WAVES, LW = TZ + 4;
    static_assert(TX % WAVES == 0, "rows must split evenly over waves");
    __shared__ float lds[2][TX + 4][LW];

    // XCD-aware order: hardware deals blocks round-robin over the 8 XCDs; give each XCD a
    // contiguous range of logical tiles so neighbouring tiles share an L2.
    const int per = gridDim.x >> 3;
    const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (lb >= nblk) return;
    const int tz = lb % ntz;
    const int rest = lb / ntz;
    const int tx = rest % ntx;
    const int ch = rest / ntx;
    const int z0 = tz * TZ, x0 = tx * TX;
    const int y0 = ch * nyc, y1 = min(y0 + nyc, g.ny);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int z = z0 + lane;
    const bool zok = z < g.nz2;

    bool blk_src = false;
    for (int s = 0; s < a.ns; ++s) {
        int gz = a.src[3 * s], gx = a.src[3 * s + 1], gy = a.src[3 * s + 2];
        blk_src |= gz >= z0 - 1 && gz <= z0 + TZ && gx >= x0 - 1 && gx <= x0 + TX && gy >= y0 - 1 &&
                   gy <= y1;
    }

    int xoff[RPW];
    bool xok[RPW], act[RPW];
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        int x = x0 + w * RPW + r;
        xok[r] = x < g.nx2;
        act[r] = (z < g.nz) && (x < g.nx);
        xoff[r] = x * g.ldz;
    }

    float q[RPW][5];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int yy = y0 - 2 + k;
        bool yok = yy >= 0 && yy < g.ny2;
        const float* P = a.cur + (long long)(yok ? yy : 0) * g.ldxy;
#pragma unroll
        for (int r = 0; r < RPW; ++r) q[r][k] = (yok && zok && xok[r]) ? P[xoff[r] + z] : 0.f;
    }

    int buf = 0;
    for (int y = y0; y < y1; ++y) {
        const long long poff = (long long)y * g.ldxy;
        const float* Pc = a.cur + poff;
        {
            int yy = y + 2;
            bool yok = yy < g.ny2;
            const float* P = a.cur + (long long)(yok ? yy : 0) * g.ldxy;
#pragma unroll
            for (int r = 0; r < RPW; ++r) q[r][4] = (yok && zok && xok[r]) ? P[xoff[r] + z] : 0.f;
        }
        float pr[RPW], vc[RPW];
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            pr[r] = act[r] ? a.pv[poff + xoff[r] + z] : 0.f;
            vc[r] = act[r] ? a.vv[poff + xoff[r] + z] : 0.f;
        }
        float(*L)[LW] = lds[buf];
#pragma unroll
        for (int r = 0; r < RPW; ++r) L[w * RPW + r + 2][lane + 2] = q[r][2];
        if (lane < 4) {
            int zz = lane < 2 ? z0 - 2 + lane : z0 + TZ + lane - 2;
            int lz = lane < 2 ? lane : TZ + lane;
            bool zzok = zz >= 0 && zz < g.nz2;
#pragma unroll
            for (int r = 0; r < RPW; ++r) L[w * RPW + r + 2][lz] = (zzok && xok[r]) ? Pc[xoff[r] + zz] : 0.f;
        }
        for (int h = w; h < 4; h += WAVES) {
            int xx = h < 2 ? x0 - 2 + h : x0 + TX + h - 2;
            int lrow = h < 2 ? h : TX + h;
            L[lrow][lane + 2] = (zok && xx >= 0 && xx < g.nx2) ? Pc[xx * g.ldz + z] : 0.f;
        }
        __syncthreads();

        const float2 fyy = a.fy[y];
        const bool yin = y >= 2 && y < g.ny2 - 2;
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            if (!act[r]) continue;
            const int x = x0 + w * RPW + r;
            const int lr = w * RPW + r + 2;
            const float u = q[r][2];
            float lap = 0.f;
            if (yin && z >= 2 && z < g.nz2 - 2 && x >= 2 && x < g.nx2 - 2) {
                lap = C0 * u * g.cc
                    + C1 * (q[r][1] + q[r][3]) * g.idy
                    + C2 * (q[r][0] + q[r][4]) * g.idy
                    + C1 * (L[lr - 1][lane + 2] + L[lr + 1][lane + 2]) * g.idx
                    + C2 * (L[lr - 2][lane + 2] + L[lr + 2][lane + 2]) * g.idx
                    + C1 * (L[lr][lane + 1] + L[lr][lane + 3]) * g.idz
                    + C2 * (L[lr][lane] + L[lr][lane + 4]) * g.idz;
            }
            const float2 fzz = a.fz[z], fxx = a.fx[x];
            const float p = damp3(pr[r], fzz, fxx, fyy);
            float c = u + (u - p);
            c = c + lap * vc[r];
            if (blk_src) c = inject(c, vc[r], a, z, x, y);
            const bool onrec = z == g.nb && x >= g.nb && x < g.nb + g.nx1 && y >= g.nb && y < g.nb + g.ny1;
            if (a.tri && onrec) c = c + vc[r] * a.tri[(y - g.nb) * g.nx1 + (x - g.nb)];
            c = damp3(c, fzz, fxx, fyy);
            a.pv[poff + xoff[r] + z] = c;
            if (a.rec && onrec) a.rec[(y - g.nb) * g.nx1 + (x - g.nb)] = c;
            if (a.snap && z >= g.nb && z < g.nb + g.nz1 && x >= g.nb && x < g.nb + g.nx1 && y >= g.nb &&
                y < g.nb + g.ny1)
                a.snap[((size_t)(y - g.nb) * g.nx1 + (x - g.nb)) * g.nz1 + (z - g.nb)] = c;
        }
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            q[r][0] = q[r][1]; q[r][1] = q[r][2]; q[r][2] = q[r][3]; q[r][3] = q[r][4];
        }
        buf ^= 1;
    }
}

// ---------------------------------------------------------------- 3D lean variant
// Same arithmetic as k_afd3d with the per-cell overhead removed: the fields carry zero guard
// planes around the padded grid, so every load is unconditional and in bounds (lanes past the
// end of a row read the next row or the guard; they are outside the updated domain and their
// values only reach other such lanes); the rim rule is a select; damping uses one factor per axis (exact when the two faces of
// an axis do not overlap, checked on the host); offsets are 32-bit.
template <int RPW, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_afd3d_lean(Geo g, Step a, int ntz, int ntx, int nyc, int nblk)
{
    constexpr int TZ = 64, TX = RPW * WAVES, LW = TZ + 4;
    __shared__ float lds[2][TX + 4][LW];
    const int per = gridDim.x >> 3;
    const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (lb >= nblk) return;
    const int tz = lb % ntz;
    const int rest = lb / ntz;
    const int tx = rest % ntx;
    const int ch = rest / ntx;
    const int z0 = tz * TZ, x0 = tx * TX;
    const int y0 = ch * nyc, y1 = min(y0 + nyc, g.ny);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int z = z0 + lane;
    const int ldz = g.ldz, ldxy = (int)g.ldxy;
    bool blk_src = false;
    for (int s = 0; s < a.ns; ++s) {
        int gz = a.src[3 * s], gx = a.src[3 * s + 1], gy = a.src[3 * s + 2];
        blk_src |= gz >= z0 - 1 && gz <= z0 + TZ && gx >= x0 - 1 && gx <= x0 + TX && gy >= y0 - 1 && gy <= y1;
    }
    const float fz = a.fz1[z];
    const bool zin = z >= 2 && z < g.nz2 - 2;
    const bool zact = z < g.nz;
    int xo[RPW];
    float fx[RPW];
    bool xin[RPW], act[RPW];
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        const int x = x0 + w * RPW + r;
        xo[r] = x * ldz + z;
        fx[r] = a.fx1[x];
        xin[r] = x >= 2 && x < g.nx2 - 2;
        act[r] = zact && x < g.nx;
    }
    const float* __restrict__ cur = a.cur;
    float q[RPW][5];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int r = 0; r < RPW; ++r) q[r][k] = cur[(y0 - 2 + k) * ldxy + xo[r]];
    int buf = 0;
    for (int y = y0; y < y1; ++y) {
        const int po = y * ldxy;
#pragma unroll
        for (int r = 0; r < RPW; ++r) q[r][4] = cur[po + 2 * ldxy + xo[r]];
        float pr[RPW], vc[RPW];
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            pr[r] = a.pv[po + xo[r]];
            vc[r] = a.vv[po + xo[r]];
        }
        float(*L)[LW] = lds[buf];
#pragma unroll
        for (int r = 0; r < RPW; ++r) L[w * RPW + r + 2][lane + 2] = q[r][2];
        if (lane < 4) {
            const int zz = lane < 2 ? z0 - 2 + lane : z0 + TZ - 2 + lane;
            const int lz = lane < 2 ? lane : TZ + lane;
#pragma unroll
            for (int r = 0; r < RPW; ++r) L[w * RPW + r + 2][lz] = cur[po + xo[r] - z + zz];
        }
        for (int h = w; h < 4; h += WAVES) {
            const int xx = h < 2 ? x0 - 2 + h : x0 + TX - 2 + h;
            const int lrow = h < 2 ? h : TX + h;
            L[lrow][lane + 2] = cur[po + xx * ldz + z];
        }
        __syncthreads();
        const float fyv = a.fy1[y];
        const bool yin = y >= 2 && y < g.ny2 - 2;
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const int lr = w * RPW + r + 2;
            const float u = q[r][2];
            float lap = C0 * u * g.cc
                      + C1 * (q[r][1] + q[r][3]) * g.idy
                      + C2 * (q[r][0] + q[r][4]) * g.idy
                      + C1 * (L[lr - 1][lane + 2] + L[lr + 1][lane + 2]) * g.idx
                      + C2 * (L[lr - 2][lane + 2] + L[lr + 2][lane + 2]) * g.idx
                      + C1 * (L[lr][lane + 1] + L[lr][lane + 3]) * g.idz
                      + C2 * (L[lr][lane] + L[lr][lane + 4]) * g.idz;
            lap = (yin && zin && xin[r]) ? lap : 0.f;
            const float p = pr[r] * fz * fx[r] * fyv;
            float c = u + (u - p);
            c = c + lap * vc[r];
            const int x = x0 + w * RPW + r;
            if (blk_src) c = inject(c, vc[r], a, z, x, y);
            const bool onrec = z == g.nb && x >= g.nb && x < g.nb + g.nx1 && y >= g.nb && y < g.nb + g.ny1;
            if (a.tri && onrec) c = c + vc[r] * a.tri[(y - g.nb) * g.nx1 + (x - g.nb)];
            c = c * fz * fx[r] * fyv;
            if (act[r]) a.pv[po + xo[r]] = c;
            if (a.rec && onrec) a.rec[(y - g.nb) * g.nx1 + (x - g.nb)] = c;
            if (a.snap && act[r] && z >= g.nb && z < g.nb + g.nz1 && x >= g.nb && x < g.nb + g.nx1 && y >= g.nb &&
                y < g.nb + g.ny1)
                a.snap[((size_t)(y - g.nb) * g.nx1 + (x - g.nb)) * g.nz1 + (z - g.nb)] = c;
        }
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            q[r][0] = q[r][1]; q[r][1] = q[r][2]; q[r][2] = q[r][3]; q[r][3] = q[r][4];
        }
        buf ^= 1;
    }
}

// ---------------------------------------------------------------- 3D pipelined variant
// The lean kernel's arithmetic with a software-pipelined y loop: the loads for plane y+1 (prev,
// vv, both halos) and for the centre column of plane y+3 are issued before plane y is computed,
// so a wave only ever waits on memory it requested one iteration earlier.  Fields are read and
// written through buffer resources (scalar base per plane, 32-bit lane offsets).  A lane whose
// cell lies outside the updated domain -- where every field value is zero -- gets an offset
// beyond the resource's range: the hardware answers such a load with 0 without touching
// memory and drops such a store, so the z slack of the 64-aligned rows, rows x >= nx and planes
// outside [0, ny) cost no HBM traffic and need no branches.  NTP: non-temporal cache policy on
// the streamed-once fields (prev/next, vv), leaving L2 to the halo-shared current field.
constexpr unsigned kOOB = 0x80000000u;  // out-of-range lane offset (resource windows are < 2 GiB)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const float* p, unsigned bytes)
{
    // raw buffer: stride 0, num_records = window bytes, 32-bit data format
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)bytes, 0x00020000);
}

// load through a plane window; ok == false loads 0 without a memory access
template <int AUX>
__device__ __forceinline__ float bld(__amdgpu_buffer_rsrc_t rs, unsigned off, bool ok)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, ok ? off : kOOB, 0, AUX));
}

template <int AUX>
__device__ __forceinline__ void bst(float v, __amdgpu_buffer_rsrc_t rs, unsigned off)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rs, off, 0, AUX);
}

template <int RPW, int WAVES, int NTP, int ORD>
__global__ __launch_bounds__(64 * WAVES) void k_afd3d_pipe(Geo g, Step a, int ntz, int ntx, int nyc, int nblk)
{
    static_assert(WAVES >= 4, "one x-halo row per wave for waves 0..3");
    constexpr int TZ = 64, TX = RPW * WAVES, LW = TZ + 4;
    constexpr int AUX = NTP ? 2 : 0;  // slc: non-temporal
    __shared__ float lds[2][TX + 4][LW];
    const int per = gridDim.x >> 3;
    const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (lb >= nblk) return;
    // logical tile order inside an XCD's range: ORD 0 z-tile fastest, then x-tile, then y chunk;
    // ORD 1 x-tile fastest (x neighbours, which read each other's halo rows, are dispatched
    // together), then z-tile, then y chunk
    int tz, tx, ch;
    if (ORD == 0) {
        tz = lb % ntz;
        const int rest = lb / ntz;
        tx = rest % ntx;
        ch = rest / ntx;
    } else {
        tx = lb % ntx;
        const int rest = lb / ntx;
        tz = rest % ntz;
        ch = rest / ntz;
    }
    const int z0 = tz * TZ, x0 = tx * TX;
    const int y0 = ch * nyc, y1 = min(y0 + nyc, g.ny);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int z = z0 + lane;
    const int ldz = g.ldz;
    // one resource per field over the planes this block touches (yb .. ye-1: its chunk and two
    // planes either side, within [0, ny)), from 2 rows + 2 cells before plane yb; lane offsets
    // (within a plane, constant over the loop) are vector, the plane offset from yb is the scalar
    // soffset; a plane outside [0, ny) or outside this block's needs uses a resource of size 0, so
    // none of its loads touches memory.  The window of a chunk stays below 2 GiB (the launch
    // bounds the chunk length), whatever the size of the field.
    const int bias = 2 * ldz + 2;
    const int yb = y0 >= 2 ? y0 - 2 : 0, ye = min(y1 + 2, g.ny);
    const long long base = (long long)yb * g.ldxy - bias;
    const unsigned fbytes = (unsigned)(((long long)(ye - yb) * g.ldxy + 2 * bias) * 4);
    const unsigned pstep = (unsigned)(g.ldxy * 4);  // bytes per plane
    const auto rC = plane_rsrc(a.cur + base, fbytes), rP = plane_rsrc(a.pv + base, fbytes),
               rV = plane_rsrc(a.vv + base, fbytes), rN = plane_rsrc(a.cur - bias, 0u);
    bool blk_src = false;
    for (int s = 0; s < a.ns; ++s) {
        int gz = a.src[3 * s], gx = a.src[3 * s + 1], gy = a.src[3 * s + 2];
        blk_src |= gz >= z0 - 1 && gz <= z0 + TZ && gx >= x0 - 1 && gx <= x0 + TX && gy >= y0 - 1 && gy <= y1;
    }
    const bool zact = z < g.nz;
    const float fz = a.fz1[z];
    const bool zin = z >= 2 && z < g.nz2 - 2;
    // z halo: lanes 0..3 stage cells z0-2, z0-1, z0+64, z0+65; the other lanes rewrite their own
    // centre value, so the halo store needs no branch
    const int zh = lane < 2 ? z0 - 2 + lane : z0 + TZ - 2 + lane;
    const int lzh = lane < 4 ? (lane < 2 ? lane : TZ + lane) : lane + 2;
    unsigned oc[RPW], ohz[RPW];
    float fx[RPW];
    bool xin[RPW];
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        const int x = x0 + w * RPW + r;
        const bool xact = x < g.nx;
        oc[r] = (zact && xact) ? (unsigned)(bias + x * ldz + z) * 4u : kOOB;
        ohz[r] = (lane < 4 && zh >= 0 && zh < g.nz && xact) ? (unsigned)(bias + x * ldz + zh) * 4u : kOOB;
        fx[r] = a.fx1[x];
        xin[r] = x >= 2 && x < g.nx2 - 2;
    }
    const int xh = w < 2 ? x0 - 2 + w : x0 + TX - 2 + w;  // x halo row of waves 0..3
    const int lrh = w < 2 ? w : TX + w;
    const unsigned ohx = (w < 4 && xh >= 0 && xh < g.nx && zact) ? (unsigned)(bias + xh * ldz + z) * 4u : kOOB;
    auto ld = [](__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    };
    auto ldn = [](__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, AUX));
    };
    auto yok = [&](int yy) { return yy >= 0 && yy < g.ny; };

    // prologue: centre planes y0-2 .. y0+2, and plane y0's prev, vv and halos
    float q[RPW][6];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const int yy = y0 - 2 + k;
        const auto rs = yok(yy) ? rC : rN;
        const unsigned so = yok(yy) ? (unsigned)(yy - yb) * pstep : 0u;
#pragma unroll
        for (int r = 0; r < RPW; ++r) q[r][k] = ld(rs, oc[r], so);
    }
    // The operands of one plane are double-buffered by plane parity, and the centre planes sit in
    // a ring of six slots indexed by plane mod 6: five in use and one free for the plane being
    // prefetched.  The y loop is unrolled by 6 so that every ring slot and buffer is a fixed
    // register: no value is copied from plane to plane, and a prefetch never targets a register
    // whose previous load is still to be read (which would make the wave wait for all its loads in
    // flight before issuing the next plane's).
    float pr[2][RPW], vc[2][RPW], hz[2][RPW], hx[2], fyb[2];
    fyb[0] = a.fy1[y0];
    {
        const unsigned so = (unsigned)(y0 - yb) * pstep;
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            pr[0][r] = ldn(rP, oc[r], so);
            vc[0][r] = ldn(rV, oc[r], so);
            hz[0][r] = ld(rC, ohz[r], so);
        }
        hx[0] = ld(rC, ohx, so);
    }
    // block-uniform: does this z tile hold the receiver plane (z == nb)?
    const bool rtile = g.nb >= z0 && g.nb < z0 + TZ;
    // plane y at phase K = (y - y0) mod 6: its operands in buffers [K & 1] (LDS buffer too), the
    // centre planes y-2 .. y+3 in ring slots (K + j) mod 6, j = 0 .. 5 (plane y+3 being loaded)
    auto plane = [&](auto Kc, int y) {
        constexpr int K = decltype(Kc)::value;
        constexpr int cu = K & 1, nu = cu ^ 1;
        auto Q = [&](int r, int j) -> float& { return q[r][(K + j) % 6]; };
        // prefetch for the next plane: plane y+1 (prev, vv, halos) and the centre of y+3
        const unsigned s0 = (unsigned)(y - yb) * pstep;
        const bool nv = y + 1 < y1;
        const bool n3 = nv && y + 3 < g.ny;
        const unsigned s1 = s0 + pstep, s3 = s0 + 3 * pstep;
        const auto rc1 = nv ? rC : rN, rp1 = nv ? rP : rN, rv1 = nv ? rV : rN, rc3 = n3 ? rC : rN;
        // vector load (the field stores in this kernel keep the compiler off the scalar cache),
        // so it is prefetched like the rest
        fyb[nu] = a.fy1[y + 1];
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            Q(r, 5) = ld(rc3, oc[r], s3);
            pr[nu][r] = ldn(rp1, oc[r], s1);
            vc[nu][r] = ldn(rv1, oc[r], s1);
            hz[nu][r] = ld(rc1, ohz[r], s1);
        }
        hx[nu] = ld(rc1, ohx, s1);
        float(*L)[LW] = lds[cu];
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            L[w * RPW + r + 2][lane + 2] = Q(r, 2);
            L[w * RPW + r + 2][lzh] = lane < 4 ? hz[cu][r] : Q(r, 2);
        }
        if (w < 4) L[lrh][lane + 2] = hx[cu];
        __syncthreads();
        const bool yin = y >= 2 && y < g.ny2 - 2;
        const float fyv = fyb[cu];
        float c[RPW];
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const int lr = w * RPW + r + 2;
            const float u = Q(r, 2);
            float lap = C0 * u * g.cc
                      + C1 * (Q(r, 1) + Q(r, 3)) * g.idy
                      + C2 * (Q(r, 0) + Q(r, 4)) * g.idy
                      + C1 * (L[lr - 1][lane + 2] + L[lr + 1][lane + 2]) * g.idx
                      + C2 * (L[lr - 2][lane + 2] + L[lr + 2][lane + 2]) * g.idx
                      + C1 * (L[lr][lane + 1] + L[lr][lane + 3]) * g.idz
                      + C2 * (L[lr][lane] + L[lr][lane + 4]) * g.idz;
            lap = (yin && zin && xin[r]) ? lap : 0.f;
            const float p = pr[cu][r] * fz * fx[r] * fyv;
            c[r] = u + (u - p);
            c[r] = c[r] + lap * vc[cu][r];
        }
        if (blk_src) {
#pragma unroll
            for (int r = 0; r < RPW; ++r) c[r] = inject(c[r], vc[cu][r], a, z, x0 + w * RPW + r, y);
        }
        const bool yint = y >= g.nb && y < g.nb + g.ny1;
        if (a.tri && rtile && yint) {  // time-reversal injection on the receiver plane
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                const int x = x0 + w * RPW + r;
                if (z == g.nb && x >= g.nb && x < g.nb + g.nx1)
                    c[r] = c[r] + vc[cu][r] * a.tri[(y - g.nb) * g.nx1 + (x - g.nb)];
            }
        }
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            c[r] = c[r] * fz * fx[r] * fyv;
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, c[r]), rP, oc[r], s0, AUX);
        }
        if (a.rec && rtile && yint) {
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                const int x = x0 + w * RPW + r;
                if (z == g.nb && x >= g.nb && x < g.nb + g.nx1) a.rec[(y - g.nb) * g.nx1 + (x - g.nb)] = c[r];
            }
        }
        if (a.snap && yint) {
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                const int x = x0 + w * RPW + r;
                if (zact && x < g.nx && z >= g.nb && z < g.nb + g.nz1 && x >= g.nb && x < g.nb + g.nx1)
                    a.snap[((size_t)(y - g.nb) * g.nx1 + (x - g.nb)) * g.nz1 + (z - g.nb)] = c[r];
            }
        }
    };
    for (int yb = y0; yb < y1; yb += 6) {
        plane(std::integral_constant<int, 0>{}, yb);
        if (yb + 1 < y1) plane(std::integral_constant<int, 1>{}, yb + 1);
        if (yb + 2 < y1) plane(std::integral_constant<int, 2>{}, yb + 2);
        if (yb + 3 < y1) plane(std::integral_constant<int, 3>{}, yb + 3);
        if (yb + 4 < y1) plane(std::integral_constant<int, 4>{}, yb + 4);
        if (yb + 5 < y1) plane(std::integral_constant<int, 5>{}, yb + 5);
    }
}

// ---------------------------------------------------------------- 3D, two time steps per launch
// Temporal blocking of the pipelined kernel: one launch advances C_n, C_{n-1} to C_{n+2}, C_{n+1}.
// A block owns an output tile of 60 z x (TX1 - 4) x rows over a y chunk.  Streaming along y, at
// plane j it computes step 1 (C_{n+1}) on the tile grown by 2 cells in z and x (64 lanes x TX1
// rows, from C_n staged in LDS with a further 2-cell halo) and, two planes behind, step 2
// (C_{n+2}) at plane j-2 from the step-1 values kept in an LDS ring of five planes: the y
// neighbours are a lane's own ring entries, the z and x neighbours those of plane j-2, so one
// barrier per plane suffices and no step-1 value occupies a register beyond its plane.  The grown ring of step 1 is
// recomputed by neighbouring blocks instead of exchanged; every cell of both steps uses the
// single-step arithmetic, so results are bit-identical.  C_{n+1} and C_{n+2} go to two other
// buffers (neighbours still read C_n and C_{n-1} around the tile).  HBM traffic per two steps:
// C_n and C_{n-1} and vv read once (plus halos), C_{n+1} and C_{n+2} written once.
struct Step2 {
    const float* cur;  // C_n
    const float* pv;   // C_{n-1} as stored
    const float* vv;
    float* out1;       // C_{n+1}
    float* out2;       // C_{n+2}
    const float* __restrict__ fz1;
    const float* __restrict__ fx1;
    const float* __restrict__ fy1;
    const int* __restrict__ src;
    const float* __restrict__ wav;
    const float* __restrict__ tri1;  // receiver planes of the two steps (time reversal) or null
    const float* __restrict__ tri2;
    float* __restrict__ rec1;  // receiver planes of the two steps or null
    float* __restrict__ rec2;
    float* __restrict__ snap1;  // snapshots of the two steps or null
    float* __restrict__ snap2;
    int ns, nt, it1, it2;
};

__device__ __forceinline__ float inject_at(float c, float vvc, const int* __restrict__ src,
                                           const float* __restrict__ wav, int ns, int nt, int it, int z, int x,
                                           int y)
{
    for (int s = 0; s < ns; ++s) {
        int dz = z - src[3 * s], dx = x - src[3 * s + 1], dy = y - src[3 * s + 2];
        int adz = dz < 0 ? -dz : dz, adx = dx < 0 ? -dx : dx, ady = dy < 0 ? -dy : dy;
        if (adz <= 1 && adx <= 1 && ady <= 1) {
            float k = (float)(adz + adx + ady + 1);
            c = c + (vvc * wav[(size_t)s * nt + it]) / k;
        }
    }
    return c;
}

template <int RPW, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_afd3d_2step(Geo g, Step2 a, int ntz, int ntx, int nyc, int nblk)
{
    static_assert(WAVES >= 4, "one x-halo row per wave for waves 0..3");
    constexpr int TX1 = RPW * WAVES;  // step-1 rows
    constexpr int TXO = TX1 - 4;      // output rows
    constexpr int TZO = 60;           // output z per tile
    static_assert(TXO >= 4, "tile too narrow");
    __shared__ float lc[2][TX1 + 4][68];  // C_n of plane j: z from zs-4, x from xs-4 (double buffer)
    __shared__ float l1[5][TX1][64];      // C_{n+1} of planes j-4 .. j (ring, slot rb holds plane j)
    const int per = gridDim.x >> 3;
    const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (lb >= nblk) return;
    const int tz = lb % ntz;
    const int rest = lb / ntz;
    const int tx = rest % ntx;
    const int ch = rest / ntx;
    const int zs = tz * TZO, xs = tx * TXO;
    const int y0 = ch * nyc, y1 = min(y0 + nyc, g.ny);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int z = zs - 2 + lane;
    const int ldz = g.ldz;
    // Whole-field resources from plane -2 (minus 4 rows + 4 cells) to plane ny2 (plus as much):
    // plane y is at scalar offset (y + 2) * pstep, lane offsets are vector.  The planes before the
    // field are its zero guard and every field is zero outside the updated domain, so planes
    // -2 .. ny2 - 1 load without a mask; beyond the window a load returns 0.
    const int bias = 4 * ldz + 4;
    const long long gb = 2 * g.ldxy + bias;
    const unsigned fbytes = (unsigned)((g.ldxy * (g.ny2 + 2) + 2 * bias) * 4);
    const unsigned pstep = (unsigned)(g.ldxy * 4);
    const auto rC = plane_rsrc(a.cur - gb, fbytes), rP = plane_rsrc(a.pv - gb, fbytes),
               rV = plane_rsrc(a.vv - gb, fbytes), rO1 = plane_rsrc(a.out1 - gb, fbytes),
               rO2 = plane_rsrc(a.out2 - gb, fbytes);
    auto pof = [&](int yy) { return (unsigned)(yy + 2) * pstep; };
    const bool zact = z >= 0 && z < g.nz;
    const bool zout = lane >= 2 && lane < 62 && z < g.nz;
    const float fz = a.fz1[z < 0 ? 0 : z];
    const bool zin = z >= 2 && z < g.nz2 - 2;
    // C_n z halo: lanes 0..3 stage zs-4, zs-3, zs+62, zs+63; the others rewrite their centre
    const int zh = lane < 2 ? zs - 4 + lane : zs + 60 + lane;
    const int lzh = lane < 4 ? (lane < 2 ? lane : 64 + lane) : lane + 2;
    const int cl = lane < 2 ? 2 : (lane > 61 ? 61 : lane);  // step-2 z neighbours stay in the ring row
    unsigned oc[RPW], ohz[RPW], oown[RPW];
    float fx[RPW];
    bool xin[RPW];
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int r = w * RPW + k;
        const int x = xs - 2 + r;
        const bool xact = x >= 0 && x < g.nx;
        oc[k] = (zact && xact) ? (unsigned)(bias + x * ldz + z) * 4u : kOOB;
        ohz[k] = (lane < 4 && zh >= 0 && zh < g.nz && xact) ? (unsigned)(bias + x * ldz + zh) * 4u : kOOB;
        oown[k] = (zout && r >= 2 && r < TX1 - 2 && x < g.nx) ? oc[k] : kOOB;  // cells this block writes
        fx[k] = a.fx1[x < 0 ? 0 : x];
        xin[k] = x >= 2 && x < g.nx2 - 2;
    }
    auto xrow = [&](int k) { return xs - 2 + w * RPW + k; };
    const int xh = w < 2 ? xs - 4 + w : xs - 4 + TX1 + w;  // C_n x halo row of waves 0..3
    const int lrh = w < 2 ? w : TX1 + w;
    const unsigned ohx = (w < 4 && xh >= 0 && xh < g.nx && zact) ? (unsigned)(bias + xh * ldz + z) * 4u : kOOB;
    bool blk_src = false;
    for (int s = 0; s < a.ns; ++s) {
        int gz = a.src[3 * s], gx = a.src[3 * s + 1], gy = a.src[3 * s + 2];
        blk_src |= gz >= zs - 3 && gz <= zs + 62 && gx >= xs - 3 && gx <= xs + TX1 - 2 && gy >= y0 - 3 && gy <= y1 + 2;
    }
    const bool rtile = g.nb >= zs - 2 && g.nb < zs + 62;  // the receiver plane z == nb is in the tile
    auto ld = [](__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    };
    auto st = [](float v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
    };
    auto fyat = [&](int yy) { return a.fy1[yy < 0 ? 0 : (yy > g.ny2 ? g.ny2 : yy)]; };

    // step-1 plane j runs over [y0-2, y1+1]; step 2 trails it by two planes
    const int j0 = y0 - 2, j1 = y1 + 1;
    float q[RPW][5], vq[RPW][3];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const int yy = j0 - 2 + k;  // planes -4 and -3 lie before the guard: not loaded
        const bool ok = yy >= -2;
#pragma unroll
        for (int r = 0; r < RPW; ++r) q[r][k] = ld(rC, ok ? oc[r] : kOOB, ok ? pof(yy) : 0u);
    }
#pragma unroll
    for (int r = 0; r < RPW; ++r) vq[r][0] = vq[r][1] = vq[r][2] = 0.f;
    float pr[RPW], vc[RPW], hz[RPW], hx;
    float fyq0 = 1.f, fyq1 = 1.f, fyv = fyat(j0);
    {
        const unsigned so = pof(j0);
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            pr[r] = ld(rP, oc[r], so);
            vc[r] = ld(rV, oc[r], so);
            hz[r] = ld(rC, ohz[r], so);
        }
        hx = ld(rC, ohx, so);
    }
    int cb = 0, rb = 0;  // C_n double buffer; ring slot of step-1 plane j
    for (int j = j0; j <= j1; ++j) {
        // prefetch plane j+1 (and the centre of j+3): planes >= -1, unmasked (a plane past the
        // last one needed, on the last iteration, is loaded and not used)
        const int jn = j + 1;
        const unsigned s1 = pof(jn), s3 = pof(j + 3);
        float npr[RPW], nvc[RPW], nhz[RPW], nq[RPW], nhx;
        const float nfy = fyat(jn);
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            nq[r] = ld(rC, oc[r], s3);
            npr[r] = ld(rP, oc[r], s1);
            nvc[r] = ld(rV, oc[r], s1);
            nhz[r] = ld(rC, ohz[r], s1);
        }
        nhx = ld(rC, ohx, s1);
        float(*C)[68] = lc[cb];
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            C[w * RPW + r + 2][lane + 2] = q[r][2];
            C[w * RPW + r + 2][lzh] = lane < 4 ? hz[r] : q[r][2];
        }
        if (w < 4) C[lrh][lane + 2] = hx;
        __syncthreads();
        // ---- step 1 at plane j (the tile grown by 2 rows and 2 cells), into the ring slot rb
        const bool jown = j >= y0 && j < y1;
        const bool jint = j >= g.nb && j < g.nb + g.ny1;
        const bool yin = j >= 2 && j < g.ny2 - 2;
        const unsigned sj = pof(j);
        float c1[RPW];
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const int lr = w * RPW + r + 2;
            const float u = q[r][2];
            float lap = C0 * u * g.cc
                      + C1 * (q[r][1] + q[r][3]) * g.idy
                      + C2 * (q[r][0] + q[r][4]) * g.idy
                      + C1 * (C[lr - 1][lane + 2] + C[lr + 1][lane + 2]) * g.idx
                      + C2 * (C[lr - 2][lane + 2] + C[lr + 2][lane + 2]) * g.idx
                      + C1 * (C[lr][lane + 1] + C[lr][lane + 3]) * g.idz
                      + C2 * (C[lr][lane] + C[lr][lane + 4]) * g.idz;
            lap = (yin && zin && xin[r]) ? lap : 0.f;
            const float p = pr[r] * fz * fx[r] * fyv;
            c1[r] = u + (u - p);
            c1[r] = c1[r] + lap * vc[r];
        }
        if (blk_src) {
#pragma unroll
            for (int r = 0; r < RPW; ++r)
                c1[r] = inject_at(c1[r], vc[r], a.src, a.wav, a.ns, a.nt, a.it1, z, xrow(r), j);
        }
        if (a.tri1 && rtile && jint) {
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                const int x = xrow(r);
                if (z == g.nb && x >= g.nb && x < g.nb + g.nx1)
                    c1[r] = c1[r] + vc[r] * a.tri1[(j - g.nb) * g.nx1 + (x - g.nb)];
            }
        }
        float(*R0)[64] = l1[rb];
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            c1[r] = c1[r] * fz * fx[r] * fyv;
            R0[w * RPW + r][lane] = c1[r];
            if (jown) st(c1[r], rO1, oown[r], sj);
            vq[r][0] = vq[r][1]; vq[r][1] = vq[r][2]; vq[r][2] = vc[r];
        }
        if (a.rec1 && rtile && jown && jint) {
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                const int x = xrow(r);
                if (oown[r] != kOOB && z == g.nb && x >= g.nb && x < g.nb + g.nx1)
                    a.rec1[(j - g.nb) * g.nx1 + (x - g.nb)] = c1[r];
            }
        }
        if (a.snap1 && jown && jint) {
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                const int x = xrow(r);
                if (oown[r] != kOOB && z >= g.nb && z < g.nb + g.nz1 && x >= g.nb && x < g.nb + g.nx1)
                    a.snap1[((size_t)(j - g.nb) * g.nx1 + (x - g.nb)) * g.nz1 + (z - g.nb)] = c1[r];
            }
        }
        // ---- step 2 at plane m = j-2: its step-1 planes m-2 .. m+2 are ring slots rb+1 .. rb+5 (mod
        // 5).  The y neighbours are this lane's own ring entries (written by this lane, in order);
        // the z and x neighbours come from plane m, written two iterations ago by other waves and
        // published by the last barrier.
        const int m = j - 2;
        if (m >= y0) {
            const float fym = fyq0;
            const bool yin2 = m >= 2 && m < g.ny2 - 2;
            const bool mint = m >= g.nb && m < g.nb + g.ny1;
            const unsigned sm = pof(m);
            const float(*Sm2)[64] = l1[rb >= 4 ? rb - 4 : rb + 1];  // plane m-2 = j-4
            const float(*Sm1)[64] = l1[rb >= 3 ? rb - 3 : rb + 2];  // plane m-1
            const float(*S)[64] = l1[rb >= 2 ? rb - 2 : rb + 3];    // plane m
            const float(*Sp1)[64] = l1[rb >= 1 ? rb - 1 : rb + 4];  // plane m+1
            float c2[RPW];
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                const int lr = w * RPW + r;
                if (lr < 2 || lr >= TX1 - 2) { c2[r] = 0.f; continue; }  // uniform per wave
                const float u = S[lr][lane];
                float lap = C0 * u * g.cc
                          + C1 * (Sm1[lr][lane] + Sp1[lr][lane]) * g.idy
                          + C2 * (Sm2[lr][lane] + R0[lr][lane]) * g.idy
                          + C1 * (S[lr - 1][cl] + S[lr + 1][cl]) * g.idx
                          + C2 * (S[lr - 2][cl] + S[lr + 2][cl]) * g.idx
                          + C1 * (S[lr][cl - 1] + S[lr][cl + 1]) * g.idz
                          + C2 * (S[lr][cl - 2] + S[lr][cl + 2]) * g.idz;
                lap = (yin2 && zin && xin[r]) ? lap : 0.f;
                const float p = q[r][0] * fz * fx[r] * fym;  // C_n at plane m
                c2[r] = u + (u - p);
                c2[r] = c2[r] + lap * vq[r][0];
            }
            if (blk_src) {
#pragma unroll
                for (int r = 0; r < RPW; ++r)
                    c2[r] = inject_at(c2[r], vq[r][0], a.src, a.wav, a.ns, a.nt, a.it2, z, xrow(r), m);
            }
            if (a.tri2 && rtile && mint) {
#pragma unroll
                for (int r = 0; r < RPW; ++r) {
                    const int x = xrow(r);
                    if (z == g.nb && x >= g.nb && x < g.nb + g.nx1)
                        c2[r] = c2[r] + vq[r][0] * a.tri2[(m - g.nb) * g.nx1 + (x - g.nb)];
                }
            }
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                c2[r] = c2[r] * fz * fx[r] * fym;
                st(c2[r], rO2, oown[r], sm);
            }
            if (a.rec2 && rtile && mint) {
#pragma unroll
                for (int r = 0; r < RPW; ++r) {
                    const int x = xrow(r);
                    if (oown[r] != kOOB && z == g.nb && x >= g.nb && x < g.nb + g.nx1)
                        a.rec2[(m - g.nb) * g.nx1 + (x - g.nb)] = c2[r];
                }
            }
            if (a.snap2 && mint) {
#pragma unroll
                for (int r = 0; r < RPW; ++r) {
                    const int x = xrow(r);
                    if (oown[r] != kOOB && z >= g.nb && z < g.nb + g.nz1 && x >= g.nb && x < g.nb + g.nx1)
                        a.snap2[((size_t)(m - g.nb) * g.nx1 + (x - g.nb)) * g.nz1 + (z - g.nb)] = c2[r];
                }
            }
        }
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            q[r][0] = q[r][1]; q[r][1] = q[r][2]; q[r][2] = q[r][3]; q[r][3] = q[r][4]; q[r][4] = nq[r];
            pr[r] = npr[r]; vc[r] = nvc[r]; hz[r] = nhz[r];
        }
        hx = nhx;
        fyq0 = fyq1; fyq1 = fyv; fyv = nfy;
        cb ^= 1;
        rb = rb == 4 ? 0 : rb + 1;
    }
}

// ---------------------------------------------------------------- 2D: one cell per lane
template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_afd2d(Geo g, Step a)
{
    const int z = blockIdx.x * 64 + (threadIdx.x & 63);
    const int x = blockIdx.y * WAVES + (threadIdx.x >> 6);
    if (z >= g.nz || x >= g.nx) return;
    const float* __restrict__ cur = a.cur;
    const int L = g.ldz;
    const int j = x * L + z;
    const float u = cur[j];
    float lap = 0.f;
    if (z >= 2 && z < g.nz2 - 2 && x >= 2 && x < g.nx2 - 2) {
        lap = C0 * u * g.cc
            + C1 * (cur[j - L] + cur[j + L]) * g.idx
            + C2 * (cur[j - 2 * L] + cur[j + 2 * L]) * g.idx
            + C1 * (cur[j - 1] + cur[j + 1]) * g.idz
            + C2 * (cur[j - 2] + cur[j + 2]) * g.idz;
    }
    const float vc = a.vv[j];
    const float2 fzz = a.fz[z], fxx = a.fx[x];
    const float p = damp2(a.pv[j], fzz, fxx);
    float c = u + (u - p);
    c = c + lap * vc;
    c = inject(c, vc, a, z, x, 0);
    const bool onrec = z == g.nb && x >= g.nb && x < g.nb + g.nx1;
    if (a.tri && onrec) c = c + vc * a.tri[x - g.nb];
    c = damp2(c, fzz, fxx);
    a.pv[j] = c;
    if (a.rec && onrec) a.rec[x - g.nb] = c;
    if (a.snap && z >= g.nb && z < g.nb + g.nz1 && x >= g.nb && x < g.nb + g.nx1)
        a.snap[(size_t)(x - g.nb) * g.nz1 + (z - g.nb)] = c;
}

// ---------------------------------------------------------------- pseudo-spectral step
// psm2d/psm's update (pywave/src/wave_psp.c:518-581, :193-256) with the spectral Laplacian 'wave'
// computed beforehand over the whole padded grid: the same leapfrog, injection, damping of both
// levels, receivers and snapshots as the finite-difference step, on every cell of the updated
// domain.  One thread per cell; D3 selects the y faces.
template <bool D3>
__global__ __launch_bounds__(256) void k_ps_update(Geo g, Step a, const float* __restrict__ wave)
{
    const int z = blockIdx.x * 64 + (threadIdx.x & 63);
    const int x = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int y = blockIdx.z;
    if (z >= g.nz || x >= g.nx) return;
    const long long j = (long long)y * g.ldxy + (long long)x * g.ldz + z;
    const float u = a.cur[j];
    const float vc = a.vv[j];
    const float2 fzz = a.fz[z], fxx = a.fx[x];
    const float2 fyy = D3 ? a.fy[y] : make_float2(1.f, 1.f);
    const float p = D3 ? damp3(a.pv[j], fzz, fxx, fyy) : damp2(a.pv[j], fzz, fxx);
    float c = u + (u - p);
    c = c + wave[j] * vc;
    c = inject(c, vc, a, z, x, y);
    const int nby = D3 ? g.nb : 0;
    const bool onrec = z == g.nb && x >= g.nb && x < g.nb + g.nx1 && y >= nby && y < nby + g.ny1;
    const long long r = (long long)(y - nby) * g.nx1 + (x - g.nb);
    if (a.tri && onrec) c = c + vc * a.tri[r];
    c = D3 ? damp3(c, fzz, fxx, fyy) : damp2(c, fzz, fxx);
    a.pv[j] = c;
    if (a.rec && onrec) a.rec[r] = c;
    if (a.snap && z >= g.nb && z < g.nb + g.nz1 && x >= g.nb && x < g.nb + g.nx1 && y >= nby && y < nby + g.ny1)
        a.snap[((long long)(y - nby) * g.nx1 + (x - g.nb)) * g.nz1 + (z - g.nb)] = c;
}

// ---------------------------------------------------------------- helpers
// vv = ((v*v)*dt)*dt on the padded layout from the clamp-expanded model velocity
// (vel_expand, pywave/src/wave_abc.c:14-125, replicates the model edges into the boundary)
__global__ void k_expand_vv(Geo g, const float* __restrict__ vel, float* __restrict__ vv, float dt)
{
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int x = blockIdx.y;
    const int y = blockIdx.z;
    if (z >= g.nz) return;
    const int sz = min(max(z - g.nb, 0), g.nz1 - 1);
    const int sx = min(max(x - g.nb, 0), g.nx1 - 1);
    const int nby = g.ny2 > 1 ? g.nb : 0;
    const int sy = min(max(y - nby, 0), g.ny1 - 1);
    const float v = vel[((size_t)sy * g.nx1 + sx) * g.nz1 + sz];
    vv[(long long)y * g.ldxy + (long long)x * g.ldz + z] = v * v * dt * dt;
}

// out[c][r] = in[r][c] for an R x C matrix (receiver data time-major <-> time-fastest)
__global__ void k_transpose(const float* __restrict__ in, float* __restrict__ out, int R, int C)
{
    __shared__ float t[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    for (int k = threadIdx.y; k < 32; k += blockDim.y) {
        int r = r0 + k, c = c0 + threadIdx.x;
        if (r < R && c < C) t[k][threadIdx.x] = in[(size_t)r * C + c];
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += blockDim.y) {
        int c = c0 + k, r = r0 + threadIdx.x;
        if (r < R && c < C) out[(size_t)c * R + r] = t[threadIdx.x][k];
    }
}

// copy the interior of a padded field (img of the reference, wave_fdm.c:356-367,666-672)
__global__ void k_interior(Geo g, const float* __restrict__ f, float* __restrict__ out)
{
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int x = blockIdx.y;
    const int y = blockIdx.z;
    if (z >= g.nz1) return;
    const int nby = g.ny2 > 1 ? g.nb : 0;
    out[((size_t)y * g.nx1 + x) * g.nz1 + z] =
        f[(long long)(y + nby) * g.ldxy + (long long)(x + g.nb) * g.ldz + z + g.nb];
}

// ---------------------------------------------------------------- 3D time reversal of G groups
// G receiver groups back-propagated together (the caller's loop over groups in
// demos/test_third_tri.py:57-77 with afd3d(tri=True) per group): one launch per step reads vv
// once for all G fields; at snapshot steps (it % jsnap == 0) the source-location image
// sum_snap prod_g u_g^2 is accumulated in f64 (product formed in f32 as ((u0*u0)*u1)*u1 ...).
template <int G>
struct TriStep {
    const float* cur[G];
    float* pv[G];
    const float* __restrict__ vv;
    const float2* __restrict__ fz;
    const float2* __restrict__ fx;
    const float2* __restrict__ fy;
    const float* __restrict__ dat;          // this step's receiver plane [ny1][nx1]
    const unsigned char* __restrict__ mask;  // [G][ny1][nx1]: receiver belongs to group
    double* __restrict__ img;                // [ny1][nx1][nz1] or null (no image this step)
    const unsigned char* __restrict__ mbits; // [ny1][nx1]: bit f set when the receiver is in group f
    const float* __restrict__ fz1;           // one damping factor per index (faces do not overlap)
    const float* __restrict__ fx1;
    const float* __restrict__ fy1;
    const float* pool;                       // the 2G fields back to back (pipelined kernel)
    unsigned pool_bytes;
    int off_cur[G], off_pv[G];               // element offsets of cur[f], pv[f] in the pool
    int img_mul;                             // 0: img += prod; 1: img *= prod (a later pass of groups)
};

template <int G>
__global__ __launch_bounds__(256) void k_tri3d(Geo g, TriStep<G> a, int ntz, int ntx, int nyc, int nblk)
{
    constexpr int TX = 16, WAVES = 4, TZ = 64, RPW = TX / WAVES, LW = TZ + 4;
    __shared__ float lds[G][2][TX + 4][LW];
    const int per = gridDim.x >> 3;
    const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (lb >= nblk) return;
    const int tz = lb % ntz;
    const int rest = lb / ntz;
    const int tx = rest % ntx;
    const int ch = rest / ntx;
    const int z0 = tz * TZ, x0 = tx * TX;
    const int y0 = ch * nyc, y1 = min(y0 + nyc, g.ny);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int z = z0 + lane;
    const bool zok = z < g.nz2;
    int xoff[RPW];
    bool xok[RPW], act[RPW];
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        int x = x0 + w * RPW + r;
        xok[r] = x < g.nx2;
        act[r] = (z < g.nz) && (x < g.nx);
        xoff[r] = x * g.ldz;
    }
    float q[G][RPW][5];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int yy = y0 - 2 + k;
        bool yok = yy >= 0 && yy < g.ny2;
        long long po = (long long)(yok ? yy : 0) * g.ldxy;
#pragma unroll
        for (int f = 0; f < G; ++f)
#pragma unroll
            for (int r = 0; r < RPW; ++r) q[f][r][k] = (yok && zok && xok[r]) ? a.cur[f][po + xoff[r] + z] : 0.f;
    }
    int buf = 0;
    for (int y = y0; y < y1; ++y) {
        const long long poff = (long long)y * g.ldxy;
        {
            int yy = y + 2;
            bool yok = yy < g.ny2;
            long long po = (long long)(yok ? yy : 0) * g.ldxy;
#pragma unroll
            for (int f = 0; f < G; ++f)
#pragma unroll
                for (int r = 0; r < RPW; ++r) q[f][r][4] = (yok && zok && xok[r]) ? a.cur[f][po + xoff[r] + z] : 0.f;
        }
        float vc[RPW];
#pragma unroll
        for (int r = 0; r < RPW; ++r) vc[r] = act[r] ? a.vv[poff + xoff[r] + z] : 0.f;
#pragma unroll
        for (int f = 0; f < G; ++f) {
            float(*L)[LW] = lds[f][buf];
            const float* Pc = a.cur[f] + poff;
#pragma unroll
            for (int r = 0; r < RPW; ++r) L[w * RPW + r + 2][lane + 2] = q[f][r][2];
            if (lane < 4) {
                int zz = lane < 2 ? z0 - 2 + lane : z0 + TZ + lane - 2;
                int lz = lane < 2 ? lane : TZ + lane;
                bool zzok = zz >= 0 && zz < g.nz2;
#pragma unroll
                for (int r = 0; r < RPW; ++r) L[w * RPW + r + 2][lz] = (zzok && xok[r]) ? Pc[xoff[r] + zz] : 0.f;
            }
            {
                const int h = w;  // WAVES == 4: one halo row per wave
                int xx = h < 2 ? x0 - 2 + h : x0 + TX + h - 2;
                int lrow = h < 2 ? h : TX + h;
                L[lrow][lane + 2] = (zok && xx >= 0 && xx < g.nx2) ? Pc[xx * g.ldz + z] : 0.f;
            }
        }
        __syncthreads();
        const float2 fyy = a.fy[y];
        const bool yin = y >= 2 && y < g.ny2 - 2;
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            if (!act[r]) continue;
            const int x = x0 + w * RPW + r;
            const int lr = w * RPW + r + 2;
            const float2 fzz = a.fz[z], fxx = a.fx[x];
            const bool inner = yin && z >= 2 && z < g.nz2 - 2 && x >= 2 && x < g.nx2 - 2;
            const bool onrec = z == g.nb && x >= g.nb && x < g.nb + g.nx1 && y >= g.nb && y < g.nb + g.ny1;
            const bool inside = z >= g.nb && z < g.nb + g.nz1 && x >= g.nb && x < g.nb + g.nx1 && y >= g.nb &&
                                y < g.nb + g.ny1;
            const int rk = onrec ? (y - g.nb) * g.nx1 + (x - g.nb) : 0;
            const float dv = onrec ? a.dat[rk] : 0.f;
            float prod = 1.f;
#pragma unroll
            for (int f = 0; f < G; ++f) {
                const float(*L)[LW] = lds[f][buf];
                const float u = q[f][r][2];
                float lap = 0.f;
                if (inner) {
                    lap = C0 * u * g.cc
                        + C1 * (q[f][r][1] + q[f][r][3]) * g.idy
                        + C2 * (q[f][r][0] + q[f][r][4]) * g.idy
                        + C1 * (L[lr - 1][lane + 2] + L[lr + 1][lane + 2]) * g.idx
                        + C2 * (L[lr - 2][lane + 2] + L[lr + 2][lane + 2]) * g.idx
                        + C1 * (L[lr][lane + 1] + L[lr][lane + 3]) * g.idz
                        + C2 * (L[lr][lane] + L[lr][lane + 4]) * g.idz;
                }
                const float p = damp3(a.pv[f][poff + xoff[r] + z], fzz, fxx, fyy);
                float c = u + (u - p);
                c = c + lap * vc[r];
                if (onrec && a.mask[(size_t)f * g.nx1 * g.ny1 + rk]) c = c + vc[r] * dv;
                c = damp3(c, fzz, fxx, fyy);
                a.pv[f][poff + xoff[r] + z] = c;
                prod = f == 0 ? c * c : prod * c * c;
            }
            if (a.img && inside) {
                double& im = a.img[((size_t)(y - g.nb) * g.nx1 + (x - g.nb)) * g.nz1 + (z - g.nb)];
                im = a.img_mul ? im * (double)prod : im + (double)prod;
            }
        }
#pragma unroll
        for (int f = 0; f < G; ++f)
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                q[f][r][0] = q[f][r][1]; q[f][r][1] = q[f][r][2]; q[f][r][2] = q[f][r][3]; q[f][r][3] = q[f][r][4];
            }
        buf ^= 1;
    }
}

// Pipelined form of k_tri3d (the layout and arithmetic of k_afd3d_pipe, G fields per cell): the
// loads of plane y+1 (prev of every field, vv, halos, the receiver sample and group bits, the image
// cell) and the centre columns of plane y+3 are in flight while plane y is computed; cells outside
// the updated domain, receivers off the row and the image outside the model are masked through the
// buffer range check.  TX = WAVES rows per block, one per wave.
template <int G, int WAVES, int NT>
__global__ __launch_bounds__(64 * WAVES) void k_tri3d_pipe(Geo g, TriStep<G> a, int ntz, int ntx, int nyc,
                                                           int nblk)
{
    constexpr int AUX = NT ? 2 : 0;  // non-temporal policy on the streamed-once operands (previous
                                     // levels, vv, the new levels' stores) when NT
    static_assert(WAVES >= 4, "one x-halo row per wave for waves 0..3");
    constexpr int TZ = 64, TX = WAVES, LW = TZ + 4;
    __shared__ float lds[G][2][TX + 4][LW];
    const int per = gridDim.x >> 3;
    const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (lb >= nblk) return;
    const int tz = lb % ntz;
    const int rest = lb / ntz;
    const int tx = rest % ntx;
    const int ch = rest / ntx;
    const int z0 = tz * TZ, x0 = tx * TX;
    const int y0 = ch * nyc, y1 = min(y0 + nyc, g.ny);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int z = z0 + lane, x = x0 + w;
    const int ldz = g.ldz;
    // one resource over the pool of all fields (field and plane go in the scalar offset), one
    // over vv, and a size-0 resource for planes outside the grid
    const int bias = 2 * ldz + 2;
    const unsigned pstep = (unsigned)(g.ldxy * 4);
    const auto rF = plane_rsrc(a.pool - bias, a.pool_bytes),
               rV = plane_rsrc(a.vv - bias, (unsigned)((g.ldxy * g.ny2 + 2 * bias) * 4)),
               rN = plane_rsrc(a.pool - bias, 0u);
    const bool zact = z < g.nz, xact = x < g.nx;
    const unsigned oc = (zact && xact) ? (unsigned)(bias + x * ldz + z) * 4u : kOOB;
    const int zh = lane < 2 ? z0 - 2 + lane : z0 + TZ - 2 + lane;
    const int lzh = lane < 4 ? (lane < 2 ? lane : TZ + lane) : lane + 2;  // others rewrite their centre
    const unsigned ohz = (lane < 4 && zh >= 0 && zh < g.nz && xact) ? (unsigned)(bias + x * ldz + zh) * 4u : kOOB;
    const int xh = w < 2 ? x0 - 2 + w : x0 + TX - 2 + w;
    const int lrh = w < 2 ? w : TX + w;
    const unsigned ohx = (w < 4 && xh >= 0 && xh < g.nx && zact) ? (unsigned)(bias + xh * ldz + z) * 4u : kOOB;
    const float fz = a.fz1[z], fx = a.fx1[x];
    const bool zxin = z >= 2 && z < g.nz2 - 2 && x >= 2 && x < g.nx2 - 2;
    // receiver row (z == nb) and image interior, per (z, x); the y range is checked per plane
    const bool rtile = g.nb >= z0 && g.nb < z0 + TZ;  // block-uniform
    const bool rx = x >= g.nb && x < g.nb + g.nx1;
    const unsigned orc = (z == g.nb && rx) ? (unsigned)(x - g.nb) : kOOB;  // element index in the row
    const bool ins = zact && rx && z >= g.nb && z < g.nb + g.nz1;
    const unsigned oim = ins ? (unsigned)((x - g.nb) * g.nz1 + (z - g.nb)) * 8u : kOOB;
    const unsigned rbytes = (unsigned)(g.nx1 * 4), mbytes = (unsigned)g.nx1, ibytes = (unsigned)(g.nx1 * g.nz1 * 8);
    auto ld = [](__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    };
    auto yok = [&](int yy) { return yy >= 0 && yy < g.ny; };
    auto yrec = [&](int yy) { return yy >= g.nb && yy < g.nb + g.ny1; };
    // resource bases depend on the (uniform) plane only; the lane's receiver goes in the offset
    auto ld_dat = [&](int yy) {
        const bool yr = rtile && yrec(yy);
        const auto r = plane_rsrc(a.dat + (long long)(yr ? yy - g.nb : 0) * g.nx1, yr ? rbytes : 0u);
        return ld(r, orc != kOOB ? orc * 4u : kOOB, 0u);
    };
    auto ld_bits = [&](int yy) -> unsigned {
        const bool yr = rtile && yrec(yy);
        const auto r = __builtin_amdgcn_make_buffer_rsrc((void*)(a.mbits + (long long)(yr ? yy - g.nb : 0) * g.nx1),
                                                         (short)0, (int)(yr ? mbytes : 0u), 0x00020000);
        return __builtin_amdgcn_raw_buffer_load_b8(r, orc, 0, 0);
    };
    auto img_rsrc = [&](int yy) {
        const bool ok = a.img && yrec(yy);
        return __builtin_amdgcn_make_buffer_rsrc((void*)(a.img + (long long)(ok ? yy - g.nb : 0) * g.nx1 * g.nz1),
                                                 (short)0, (int)(ok ? ibytes : 0u), 0x00020000);
    };
    auto ld_img = [&](int yy) -> double {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(img_rsrc(yy), oim, 0, 0));
    };

    float q[G][5], pr[G], hz[G], hx[G];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const int yy = y0 - 2 + k;
        const auto rs = yok(yy) ? rF : rN;
        const unsigned so = yok(yy) ? (unsigned)yy * pstep : 0u;
#pragma unroll
        for (int f = 0; f < G; ++f) q[f][k] = ld(rs, oc, so + (unsigned)a.off_cur[f] * 4u);
    }
    {
        const unsigned so = (unsigned)y0 * pstep;
#pragma unroll
        for (int f = 0; f < G; ++f) {
            pr[f] = ld(rF, oc, so + (unsigned)a.off_pv[f] * 4u);
            hz[f] = ld(rF, ohz, so + (unsigned)a.off_cur[f] * 4u);
            hx[f] = ld(rF, ohx, so + (unsigned)a.off_cur[f] * 4u);
        }
    }
    float vc = ld(rV, oc, (unsigned)y0 * pstep);
    float dv = ld_dat(y0);
    unsigned bits = ld_bits(y0);
    double im = ld_img(y0);
    float fyv = a.fy1[y0];
    unsigned s0 = (unsigned)y0 * pstep;
    int buf = 0;
    for (int y = y0; y < y1; ++y) {
        const int yn = y + 1;
        const bool nv = yn < y1;
        const bool n3 = nv && y + 3 < g.ny;
        const unsigned s1 = s0 + pstep, s3 = s0 + 3 * pstep;
        const auto r1 = nv ? rF : rN, r3 = n3 ? rF : rN;
        float nq[G], npr[G], nhz[G], nhx[G];
#pragma unroll
        for (int f = 0; f < G; ++f) {
            const unsigned oc_f = (unsigned)a.off_cur[f] * 4u, op_f = (unsigned)a.off_pv[f] * 4u;
            nq[f] = ld(r3, oc, s3 + oc_f);
            npr[f] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r1, oc, s1 + op_f, AUX));
            nhz[f] = ld(r1, ohz, s1 + oc_f);
            nhx[f] = ld(r1, ohx, s1 + oc_f);
        }
        const float nvc = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(nv ? rV : rN, oc, s1, AUX));
        const float ndv = ld_dat(yn);
        const unsigned nbits = ld_bits(yn);
        const double nim = ld_img(yn);
        const float nfy = a.fy1[yn];
#pragma unroll
        for (int f = 0; f < G; ++f) {
            float(*L)[LW] = lds[f][buf];
            L[w + 2][lane + 2] = q[f][2];
            L[w + 2][lzh] = lane < 4 ? hz[f] : q[f][2];
            if (w < 4) L[lrh][lane + 2] = hx[f];
        }
        __syncthreads();
        const bool inner = zxin && y >= 2 && y < g.ny2 - 2;
        const bool rplane = rtile && yrec(y);  // uniform
        float c[G];
#pragma unroll
        for (int f = 0; f < G; ++f) {
            const float(*L)[LW] = lds[f][buf];
            const int lr = w + 2;
            const float u = q[f][2];
            float lap = C0 * u * g.cc
                      + C1 * (q[f][1] + q[f][3]) * g.idy
                      + C2 * (q[f][0] + q[f][4]) * g.idy
                      + C1 * (L[lr - 1][lane + 2] + L[lr + 1][lane + 2]) * g.idx
                      + C2 * (L[lr - 2][lane + 2] + L[lr + 2][lane + 2]) * g.idx
                      + C1 * (L[lr][lane + 1] + L[lr][lane + 3]) * g.idz
                      + C2 * (L[lr][lane] + L[lr][lane + 4]) * g.idz;
            lap = inner ? lap : 0.f;
            const float p = pr[f] * fz * fx * fyv;
            c[f] = u + (u - p);
            c[f] = c[f] + lap * vc;
        }
        if (rplane) {
#pragma unroll
            for (int f = 0; f < G; ++f)
                if (orc != kOOB && ((bits >> f) & 1u)) c[f] = c[f] + vc * dv;
        }
        float prod = 1.f;
#pragma unroll
        for (int f = 0; f < G; ++f) {
            c[f] = c[f] * fz * fx * fyv;
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, c[f]), rF, oc,
                                                  s0 + (unsigned)a.off_pv[f] * 4u, AUX);
            prod = f == 0 ? c[f] * c[f] : prod * c[f] * c[f];
        }
        if (a.img) {
            const double v = a.img_mul ? im * (double)prod : im + (double)prod;
            using u32x2 = unsigned __attribute__((ext_vector_type(2)));
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), img_rsrc(y), oim, 0, 0);
        }
#pragma unroll
        for (int f = 0; f < G; ++f) {
            q[f][0] = q[f][1]; q[f][1] = q[f][2]; q[f][2] = q[f][3]; q[f][3] = q[f][4]; q[f][4] = nq[f];
            pr[f] = npr[f]; hz[f] = nhz[f]; hx[f] = nhx[f];
        }
        vc = nvc; dv = ndv; bits = nbits; im = nim; fyv = nfy;
        s0 = s1;
        buf ^= 1;
    }
}

// 3D step kernels: 0-5 general (bounds-checked), 6-11 lean (guarded layout, single damping
// factor per axis), 12-19 pipelined (lean arithmetic, prefetching y loop, buffer loads; 16-17
// with non-temporal prev/vv traffic), 20-23 pipelined with x tiles fastest in the block order,
// 24-27 two time steps per launch (tx = step-1 rows; tx - 4 output rows).
// tx = rows per block, waves per block; rows per thread of the lean and pipelined kernels =
// tx / waves.
struct Variant3 {
    int tx, waves, lean;  // lean: 0 general, 1 lean, 2 pipelined, 3 pipelined non-temporal,
                          // 4 pipelined x-major block order, 5 two steps per launch
};
const Variant3 kVariants3[] = {{16, 4, 0}, {8, 4, 0},  {16, 8, 0},  {32, 8, 0},  {8, 8, 0},  {4, 4, 0},
                               {16, 8, 1}, {8, 8, 1},  {16, 4, 1},  {32, 8, 1},  {16, 16, 1}, {32, 16, 1},
                               {16, 8, 2}, {16, 4, 2}, {32, 8, 2},  {8, 4, 2},   {16, 8, 3}, {16, 4, 3},
                               {32, 16, 2}, {8, 8, 2}, {16, 4, 4}, {8, 4, 4}, {16, 8, 4}, {32, 8, 4},
                               {16, 4, 5}, {16, 8, 5}, {24, 8, 5}, {32, 8, 5}};
constexpr int kNumVariants3 = sizeof(kVariants3) / sizeof(kVariants3[0]);
constexpr int kLeanDefault = 12;
constexpr int kLeanOnlyDefault = 6;  // lean kernel when fields exceed the pipelined kernels' 2 GiB windows
constexpr int kGeneralDefault = 2;

}  // namespace

struct pw_afd {
    pw_ctx* ctx = nullptr;
    int ndim = 2;
    Geo g{};
    int nt = 0, jsnap = 0, ntsnap = 0;
    float dt = 0.f, ct = 0.f;
    float* d_a = nullptr;  // field buffers (ping-pong)
    float* d_b = nullptr;
    float* d_vv = nullptr;
    float2* d_fz = nullptr;
    float2* d_fx = nullptr;
    float2* d_fy = nullptr;
    int ns = 0;
    int* d_src = nullptr;
    float* d_wav = nullptr;
    float* d_tri = nullptr;   // [nt][ny1][nx1]
    float* d_rec = nullptr;   // [nt][ny1][nx1]
    float* d_snap = nullptr;  // [ntsnap][ny1][nx1][nz1]
    float* d_tmp = nullptr;   // transpose scratch (nrec*nt)
    float* raw[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // behind d_a, d_b, d_vv, d_c, d_d (guarded)
    float* d_c = nullptr;  // the two extra fields of the two-step kernels (allocated on first use)
    float* d_d = nullptr;
    size_t raw_elems = 0, guard_lo = 0;
    float* d_fz1 = nullptr;  // single damping factor per index (lean kernels)
    float* d_fx1 = nullptr;
    float* d_fy1 = nullptr;
    bool lean_ok = false;  // faces do not overlap, 32-bit offsets
    bool pipe_ok = false;  // lean_ok and every field fits one 2 GiB buffer window (two-step kernels)
    bool chunk_ok = false;  // lean_ok and a y chunk of one plane fits a 2 GiB window (single-step
                            // pipelined kernels, whose windows cover one chunk)
    std::vector<float> tri_host;  // echoed back as the trace part in TRI mode
    bool have_vel = false, have_src = false, have_tri = false;
    int last_mode = -1;  // 0 modelling, 1 tri
    const float* final_field = nullptr;
    int variant = -1, ny_chunk = 0;
    size_t field_elems = 0;
    long long resident[kNumVariants3] = {};  // co-resident blocks per 3D variant (0: not queried)
    float dz = 0.f, dx = 0.f, dy = 0.f;
    PsLap* ps = nullptr;      // pseudo-spectral Laplacian (aps2d/aps3d) instead of the stencil
    float* d_wave = nullptr;  // its output, field layout
};

namespace {

// longest y chunk whose planes (the chunk and two either side, plus the 2-row bias at each end)
// fit the pipelined kernels' 2 GiB buffer windows
int max_chunk_planes(const Geo& g)
{
    const double plane = 4.0 * g.ldxy, bias = 4.0 * (2.0 * g.ldz + 2.0);
    return (int)std::min(1.0e6, std::floor((2147483648.0 - 2.0 * bias) / plane)) - 5;
}

// the 3D kernel used when none was chosen
int default_variant(const pw_afd* p)
{
    return p->chunk_ok ? kLeanDefault : (p->lean_ok ? kLeanOnlyDefault : kGeneralDefault);
}

// whether 3D variant vi can run on this plan's grid
bool variant_ok(const pw_afd* p, int vi)
{
    const int k = kVariants3[vi].lean;
    return k == 0 || (k == 1 && p->lean_ok) || (k >= 2 && k <= 4 && p->chunk_ok) || (k == 5 && p->pipe_ok);
}

// blocks of 3D variant vi that the device holds at once (CUs x blocks per CU), cached per plan
long long resident_blocks(pw_afd* p, int vi)
{
    if (p->resident[vi] > 0) return p->resident[vi];
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->ctx->device) != hipSuccess || cus < 1)
        cus = 256;
    const Variant3 v = kVariants3[vi];
    hipError_t e = hipErrorInvalidValue;
    switch (vi) {
#define PW_OCC(I, K) \
    case I: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, K, 64 * v.waves, 0); break;
        PW_OCC(0, (k_afd3d<16, 4>)) PW_OCC(1, (k_afd3d<8, 4>)) PW_OCC(2, (k_afd3d<16, 8>))
        PW_OCC(3, (k_afd3d<32, 8>)) PW_OCC(4, (k_afd3d<8, 8>)) PW_OCC(5, (k_afd3d<4, 4>))
        PW_OCC(6, (k_afd3d_lean<2, 8>)) PW_OCC(7, (k_afd3d_lean<1, 8>)) PW_OCC(8, (k_afd3d_lean<4, 4>))
        PW_OCC(9, (k_afd3d_lean<4, 8>)) PW_OCC(10, (k_afd3d_lean<1, 16>)) PW_OCC(11, (k_afd3d_lean<2, 16>))
        PW_OCC(12, (k_afd3d_pipe<2, 8, 0, 0>)) PW_OCC(13, (k_afd3d_pipe<4, 4, 0, 0>))
        PW_OCC(14, (k_afd3d_pipe<4, 8, 0, 0>)) PW_OCC(15, (k_afd3d_pipe<2, 4, 0, 0>))
        PW_OCC(16, (k_afd3d_pipe<2, 8, 1, 0>)) PW_OCC(17, (k_afd3d_pipe<4, 4, 1, 0>))
        PW_OCC(18, (k_afd3d_pipe<2, 16, 0, 0>)) PW_OCC(19, (k_afd3d_pipe<1, 8, 0, 0>))
        PW_OCC(20, (k_afd3d_pipe<4, 4, 0, 1>)) PW_OCC(21, (k_afd3d_pipe<2, 4, 0, 1>))
        PW_OCC(22, (k_afd3d_pipe<2, 8, 0, 1>)) PW_OCC(23, (k_afd3d_pipe<4, 8, 0, 1>))
        PW_OCC(24, (k_afd3d_2step<4, 4>)) PW_OCC(25, (k_afd3d_2step<2, 8>))
        PW_OCC(26, (k_afd3d_2step<3, 8>)) PW_OCC(27, (k_afd3d_2step<4, 8>))
#undef PW_OCC
        default: break;
    }
    if (e != hipSuccess || per_cu < 1) per_cu = 2;
    p->resident[vi] = (long long)cus * per_cu;
    return p->resident[vi];
}

// e0/e1 (optional): events stamped by the dispatch itself at kernel start and end
int afd_launch_step(pw_afd* p, const Step& a, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr)
{
    pw_ctx* ctx = p->ctx;
    hipStream_t s = ctx->s();
    const Geo& g = p->g;
    if (p->ndim == 2) {
        constexpr int W = 4;
        dim3 grid(pw::ceil_div(g.nz, 64), pw::ceil_div(g.nx, W));
        hipExtLaunchKernelGGL(k_afd2d<W>, grid, dim3(64 * W), 0, s, e0, e1, 0, g, a);
    } else {
        const int vi = p->variant >= 0 ? p->variant : default_variant(p);
        const Variant3 v = kVariants3[vi];
        if (!variant_ok(p, vi)) PW_FAIL(ctx, PW_ERR_ARG, "3D kernel variant not applicable to this grid");
        const int ntz = pw::ceil_div(g.nz, 64), ntx = pw::ceil_div(g.nx, v.tx);
        int nyc = p->ny_chunk;
        if (nyc <= 0) {
            // as many y chunks per tile column as fit one round of co-resident blocks: every
            // block starts at once (no tail round) and neighbours stream in step through L2
            const long long tiles = (long long)ntz * ntx;
            const long long cap = resident_blocks(p, vi);
            const int nch = (int)std::max(1LL, cap / tiles);
            nyc = std::max(8, pw::ceil_div(g.ny, nch));
        }
        if (v.lean >= 2 && v.lean <= 4) nyc = std::min(nyc, max_chunk_planes(g));  // chunk window < 2 GiB
        const int nch = pw::ceil_div(g.ny, nyc);
        const int nblk = ntz * ntx * nch;
        const int grid = pw::ceil_div(nblk, 8) * 8;
        switch (vi) {
#define PW_V3(I, TX_, W_)                                                                          \
    case I:                                                                                        \
        hipExtLaunchKernelGGL((k_afd3d<TX_, W_>), dim3(grid), dim3(64 * W_), 0, s, e0, e1, 0, g, a, ntz, ntx, \
                              nyc, nblk);                                                                  \
        break;
            PW_V3(0, 16, 4)
            PW_V3(1, 8, 4)
            PW_V3(2, 16, 8)
            PW_V3(3, 32, 8)
            PW_V3(4, 8, 8)
            PW_V3(5, 4, 4)
#undef PW_V3
#define PW_VL(I, TX_, W_)                                                                          \
    case I:                                                                                        \
        hipExtLaunchKernelGGL((k_afd3d_lean<TX_ / W_, W_>), dim3(grid), dim3(64 * W_), 0, s, e0, e1, 0, g, a, \
                              ntz, ntx, nyc, nblk);                                                             \
        break;
            PW_VL(6, 16, 8)
            PW_VL(7, 8, 8)
            PW_VL(8, 16, 4)
            PW_VL(9, 32, 8)
            PW_VL(10, 16, 16)
            PW_VL(11, 32, 16)
#undef PW_VL
#define PW_VP(I, TX_, W_, NT_, ORD_)                                                                \
    case I:                                                                                         \
        hipExtLaunchKernelGGL((k_afd3d_pipe<TX_ / W_, W_, NT_, ORD_>), dim3(grid), dim3(64 * W_), 0, s, e0, e1, \
                              0, g, a, ntz, ntx, nyc, nblk);                                                         \
        break;
            PW_VP(12, 16, 8, 0, 0)
            PW_VP(13, 16, 4, 0, 0)
            PW_VP(14, 32, 8, 0, 0)
            PW_VP(15, 8, 4, 0, 0)
            PW_VP(16, 16, 8, 1, 0)
            PW_VP(17, 16, 4, 1, 0)
            PW_VP(18, 32, 16, 0, 0)
            PW_VP(19, 8, 8, 0, 0)
            PW_VP(20, 16, 4, 0, 1)
            PW_VP(21, 8, 4, 0, 1)
            PW_VP(22, 16, 8, 0, 1)
            PW_VP(23, 32, 8, 0, 1)
#undef PW_VP
            default:
                PW_FAIL(ctx, PW_ERR_ARG, "unknown 3D kernel variant");
        }
    }
    PW_HIP(ctx, hipGetLastError());
    return PW_OK;
}

int afd_launch_2step(pw_afd* p, int vi, const Step2& a, hipEvent_t e0, hipEvent_t e1)
{
    pw_ctx* ctx = p->ctx;
    hipStream_t s = ctx->s();
    const Geo& g = p->g;
    const Variant3 v = kVariants3[vi];
    const int txo = v.tx - 4;
    const int ntz = pw::ceil_div(g.nz, 60), ntx = pw::ceil_div(g.nx, txo);
    int nyc = p->ny_chunk;
    if (nyc <= 0) {
        const long long tiles = (long long)ntz * ntx;
        const int nch = (int)std::max(1LL, resident_blocks(p, vi) / tiles);
        nyc = std::max(8, pw::ceil_div(g.ny, nch));
    }
    const int nch = pw::ceil_div(g.ny, nyc);
    const int nblk = ntz * ntx * nch;
    const int grid = pw::ceil_div(nblk, 8) * 8;
    switch (vi) {
#define PW_V2(I, R_, W_)                                                                                  \
    case I:                                                                                               \
        hipExtLaunchKernelGGL((k_afd3d_2step<R_, W_>), dim3(grid), dim3(64 * W_), 0, s, e0, e1, 0, g, a, ntz, ntx, \
                              nyc, nblk);                                                                 \
        break;
        PW_V2(24, 4, 4)
        PW_V2(25, 2, 8)
        PW_V2(26, 3, 8)
        PW_V2(27, 4, 8)
#undef PW_V2
        default:
            PW_FAIL(ctx, PW_ERR_ARG, "not a two-step 3D kernel variant");
    }
    PW_HIP(ctx, hipGetLastError());
    return PW_OK;
}

void afd_free(pw_afd* p)
{
    float* bufs[] = {p->raw[0], p->raw[1], p->raw[2], p->raw[3], p->raw[4], p->d_wav, p->d_tri, p->d_rec,
                     p->d_snap, p->d_tmp, p->d_fz1, p->d_fx1, p->d_fy1};
    for (float* b : bufs)
        if (b) hipFree(b);
    if (p->d_fz) hipFree(p->d_fz);
    if (p->d_fx) hipFree(p->d_fx);
    if (p->d_fy) hipFree(p->d_fy);
    if (p->d_src) hipFree(p->d_src);
    if (p->d_wave) hipFree(p->d_wave);
    ps_destroy(p->ps);
}

int afd_check(pw_afd* p) { return p && p->ctx ? PW_OK : PW_ERR_ARG; }

}  // namespace

extern "C" {

long long pw_afd_output_len(int nt, int nz, int nx, int ny, int jsnap)
{
    if (ny < 1) ny = 1;
    long long ntsnap = 0;
    if (jsnap > 0) ntsnap = (nt - 1) / jsnap + 1;
    return (long long)nt * nx * ny + (jsnap > 0 ? (long long)nz * nx * ny * ntsnap : 0);
}

int pw_afd_variant_count(void) { return kNumVariants3; }

int pw_afd_variant_info(int variant, int* info)
{
    if (variant < 0 || variant >= kNumVariants3 || !info) return PW_ERR_ARG;
    const Variant3 v = kVariants3[variant];
    info[0] = v.tx;
    info[1] = v.waves;
    info[2] = v.lean;
    info[3] = v.lean == 5 ? 2 : 1;  // time steps per launch
    return PW_OK;
}

pw_afd* pw_afd_create(pw_ctx* ctx, int ndim, int nz, int nx, int ny, float dz, float dx, float dy,
                      int nt, float dt, int nbt, float ct, int jsnap)
{
    if (!ctx) return nullptr;
    if ((ndim != 2 && ndim != 3) || nz < 1 || nx < 1 || (ndim == 3 && ny < 1) || nt < 1 || nbt < 0 ||
        jsnap < 0 || !(dz > 0.f) || !(dx > 0.f) || (ndim == 3 && !(dy > 0.f)) || !(dt > 0.f)) {
        ctx->err = "pw_afd_create: invalid dimensions or spacings";
        return nullptr;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) {
        ctx->err = "hipSetDevice failed";
        return nullptr;
    }
    pw_afd* p = new pw_afd;
    p->ctx = ctx;
    p->ndim = ndim;
    p->nt = nt;
    p->dt = dt;
    p->ct = ct;
    p->jsnap = jsnap;
    p->ntsnap = jsnap > 0 ? (nt - 1) / jsnap + 1 : 0;
    p->dz = dz; p->dx = dx; p->dy = ndim == 3 ? dy : 1.f;
    Geo& g = p->g;
    g.nz1 = nz; g.nx1 = nx; g.ny1 = ndim == 3 ? ny : 1;
    g.nb = nbt;
    g.nz = nz + 2 * nbt; g.nx = nx + 2 * nbt; g.ny = ndim == 3 ? ny + 2 * nbt : 1;
    pw::padded_sizes(g.nz, g.nx, g.ny, &g.nz2, &g.nx2, &g.ny2);
    if (ndim == 2) g.ny2 = 1;
    // z stride (z_stride): rows on 128 B or 64 B boundaries.  A 64-lane z tile may run past the
    // end of a row (into the next row or the guard); such lanes lie outside the updated domain and
    // are masked or only feed other masked lanes.
    g.ldz = z_stride(g.nz2, g.nx2, g.ny2);
    g.ldxy = (long long)g.ldz * g.nx2;
    g.idz = 1 / (dz * dz);
    g.idx = 1 / (dx * dx);
    g.idy = ndim == 3 ? 1 / (dy * dy) : 0.f;
    g.cc = ndim == 3 ? g.idy + g.idx + g.idz : g.idx + g.idz;
    p->field_elems = (size_t)g.ldxy * g.ny2;
    // 3D fields carry zero guard planes (2 plus 4 rows before, 5 plus a tile of rows after) so the
    // lean and two-step kernels can load the neighbourhood of any tile cell unconditionally (the
    // two-step kernel reads planes -2 .. ny + 4)
    constexpr int kMaxTx = 32;
    p->guard_lo = ndim == 3 ? (size_t)(2 * g.ldxy + 4 * g.ldz + 128) : 0;
    const size_t guard_hi = ndim == 3 ? (size_t)(5 * g.ldxy + (kMaxTx + 4) * g.ldz + 128) : 0;
    p->raw_elems = p->field_elems + p->guard_lo + guard_hi;

    // damping pairs per axis; all six faces use nbt and ct (pywave/src/afd.c:101-116,392-402)
    std::vector<float> w(std::max(nbt, 1));
    pw::abc_weights(nbt, ct, w.data());
    std::vector<float> fz, fx, fy;
    pw::damping_pairs(g.nz2, nbt, nbt, w.data(), w.data(), fz);
    pw::damping_pairs(g.nx2, nbt, nbt, w.data(), w.data(), fx);
    if (ndim == 3) pw::damping_pairs(g.ny2, nbt, nbt, w.data(), w.data(), fy);
    else fy.assign(2, 1.0f);
    // single factor per index when the lo and hi faces of an axis never both apply
    auto single = [](const std::vector<float>& pr, size_t len, std::vector<float>& out) -> bool {
        out.assign(len, 1.0f);
        for (size_t i = 0; i < pr.size() / 2; ++i) {
            float lo = pr[2 * i], hi = pr[2 * i + 1];
            if (lo != 1.0f && hi != 1.0f) return false;
            out[i] = lo != 1.0f ? lo : hi;
        }
        return true;
    };
    std::vector<float> fz1, fx1, fy1;
    bool lean = ndim == 3;
    lean = lean && single(fz, (size_t)pw::ceil_div(g.nz2, 64) * 64 + 64, fz1);  // every lane of every z tile
    lean = lean && single(fx, (size_t)g.nx2 + kMaxTx + 8, fx1);
    lean = lean && single(fy, (size_t)g.ny2 + 8, fy1);
    lean = lean && (double)p->raw_elems < 2.0e9;  // 32-bit offsets
    lean = lean && (double)(g.ldxy + 8 * g.ldz + 8) * 4.0 < 2147483648.0;  // pipelined plane windows
    p->lean_ok = lean;
    p->pipe_ok = lean && (double)p->raw_elems * 4.0 < 2147483648.0;
    p->chunk_ok = lean && max_chunk_planes(g) >= 1;

    const size_t nrec = (size_t)g.nx1 * g.ny1;
    bool ok = hipMalloc(&p->raw[0], p->raw_elems * 4) == hipSuccess &&
              hipMalloc(&p->raw[1], p->raw_elems * 4) == hipSuccess &&
              hipMalloc(&p->raw[2], p->raw_elems * 4) == hipSuccess &&
              hipMalloc(&p->d_fz, fz.size() * 4) == hipSuccess &&
              hipMalloc(&p->d_fx, fx.size() * 4) == hipSuccess &&
              hipMalloc(&p->d_fy, fy.size() * 4) == hipSuccess &&
              hipMalloc(&p->d_rec, nrec * nt * 4) == hipSuccess &&
              hipMalloc(&p->d_tmp, nrec * nt * 4) == hipSuccess;
    if (ok && lean)
        ok = hipMalloc(&p->d_fz1, fz1.size() * 4) == hipSuccess && hipMalloc(&p->d_fx1, fx1.size() * 4) == hipSuccess &&
             hipMalloc(&p->d_fy1, fy1.size() * 4) == hipSuccess;
    if (ok) {
        p->d_a = p->raw[0] + p->guard_lo;
        p->d_b = p->raw[1] + p->guard_lo;
        p->d_vv = p->raw[2] + p->guard_lo;
    }
    if (ok && p->ntsnap > 0)
        ok = hipMalloc(&p->d_snap, (size_t)p->ntsnap * g.nz1 * g.nx1 * g.ny1 * 4) == hipSuccess;
    if (!ok) {
        ctx->err = "pw_afd_create: device allocation failed";
        afd_free(p);
        delete p;
        return nullptr;
    }
    hipStream_t s = ctx->s();
    ok = hipMemcpyAsync(p->d_fz, fz.data(), fz.size() * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(p->d_fx, fx.data(), fx.size() * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(p->d_fy, fy.data(), fy.size() * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemsetAsync(p->raw[2], 0, p->raw_elems * 4, s) == hipSuccess &&
         (!lean || (hipMemcpyAsync(p->d_fz1, fz1.data(), fz1.size() * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
                    hipMemcpyAsync(p->d_fx1, fx1.data(), fx1.size() * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
                    hipMemcpyAsync(p->d_fy1, fy1.data(), fy1.size() * 4, hipMemcpyHostToDevice, s) == hipSuccess)) &&
         hipStreamSynchronize(s) == hipSuccess;
    if (!ok) {
        ctx->err = "pw_afd_create: upload failed";
        afd_free(p);
        delete p;
        return nullptr;
    }
    return p;
}

void pw_afd_destroy(pw_afd* p)
{
    if (!p) return;
    hipSetDevice(p->ctx->device);
    hipStreamSynchronize(p->ctx->s());
    afd_free(p);
    delete p;
}

int pw_afd_set_velocity(pw_afd* p, const float* vel)
{
    if (afd_check(p) || !vel) return PW_ERR_ARG;
    pw_ctx* ctx = p->ctx;
    const Geo& g = p->g;
    size_t n = (size_t)g.nz1 * g.nx1 * g.ny1;
    float* d_vel = nullptr;
    PW_HIP(ctx, hipMalloc(&d_vel, n * 4));
    hipStream_t s = ctx->s();
    hipError_t e = hipMemcpyAsync(d_vel, vel, n * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        dim3 grid(pw::ceil_div(g.nz, 256), g.nx, g.ny);
        hipLaunchKernelGGL(k_expand_vv, grid, dim3(256), 0, s, g, d_vel, p->d_vv, p->dt);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(d_vel);
    if (e != hipSuccess) PW_FAIL(ctx, PW_ERR_HIP, std::string("set_velocity: ") + hipGetErrorString(e));
    p->have_vel = true;
    return PW_OK;
}

int pw_afd_set_sources(pw_afd* p, int ns, const float* src_pack, const float* wavelet)
{
    if (afd_check(p)) return PW_ERR_ARG;
    pw_ctx* ctx = p->ctx;
    const Geo& g = p->g;
    if (ns < 0 || (ns > 0 && !src_pack)) PW_FAIL(ctx, PW_ERR_ARG, "set_sources: bad source pack");
    const int per = p->ndim == 3 ? 6 : 5;
    std::vector<int> cen(3 * std::max(ns, 1), 0);
    std::vector<float> wav((size_t)std::max(ns, 1) * p->nt, 0.f);
    for (int i = 0; i < ns; ++i) {
        // float -> int truncation of the grid positions (pywave/src/afd.c:136-147, 416-425)
        int sx = (int)src_pack[i];
        int sy = p->ndim == 3 ? (int)src_pack[ns + i] : 0;
        int sz = (int)src_pack[(p->ndim == 3 ? 2 : 1) * ns + i];
        float f0 = src_pack[(per - 3) * ns + i], t0 = src_pack[(per - 2) * ns + i];
        float A = src_pack[(per - 1) * ns + i];
        int gz = sz + g.nb, gx = sx + g.nb, gy = p->ndim == 3 ? sy + g.nb : 0;
        bool inside = gz - 1 >= 0 && gz + 1 < g.nz && gx - 1 >= 0 && gx + 1 < g.nx &&
                      (p->ndim == 2 || (gy - 1 >= 0 && gy + 1 < g.ny));
        if (!inside)
            PW_FAIL(ctx, PW_ERR_ARG,
                    "set_sources: source " + std::to_string(i) +
                        " stencil leaves the padded domain (the reference would write out of bounds)");
        cen[3 * i] = gz; cen[3 * i + 1] = gx; cen[3 * i + 2] = gy;
        if (wavelet) {
            std::memcpy(&wav[(size_t)i * p->nt], wavelet + (size_t)i * p->nt, sizeof(float) * p->nt);
        } else {
            std::string e;
            int rc = pw::ricker(p->nt, p->dt, f0, t0, A, &wav[(size_t)i * p->nt], &e);
            if (rc) PW_FAIL(ctx, rc, "set_sources: source " + std::to_string(i) + ": " + e);
        }
    }
    if (p->d_src) { hipFree(p->d_src); p->d_src = nullptr; }
    if (p->d_wav) { hipFree(p->d_wav); p->d_wav = nullptr; }
    PW_HIP(ctx, hipMalloc(&p->d_src, cen.size() * 4));
    PW_HIP(ctx, hipMalloc(&p->d_wav, wav.size() * 4));
    hipStream_t s = ctx->s();
    PW_HIP(ctx, hipMemcpyAsync(p->d_src, cen.data(), cen.size() * 4, hipMemcpyHostToDevice, s));
    PW_HIP(ctx, hipMemcpyAsync(p->d_wav, wav.data(), wav.size() * 4, hipMemcpyHostToDevice, s));
    PW_HIP(ctx, hipStreamSynchronize(s));
    p->ns = ns;
    p->have_src = true;
    return PW_OK;
}

int pw_afd_set_tri_data(pw_afd* p, const float* dat)
{
    if (afd_check(p) || !dat) return PW_ERR_ARG;
    pw_ctx* ctx = p->ctx;
    const Geo& g = p->g;
    const size_t nrec = (size_t)g.nx1 * g.ny1, n = nrec * p->nt;
    p->tri_host.assign(dat, dat + n);
    if (!p->d_tri) PW_HIP(ctx, hipMalloc(&p->d_tri, n * 4));
    hipStream_t s = ctx->s();
    // host layout [iy][ix][it] (time fastest) -> device [it][iy][ix]
    PW_HIP(ctx, hipMemcpyAsync(p->d_tmp, dat, n * 4, hipMemcpyHostToDevice, s));
    dim3 grid(pw::ceil_div(p->nt, 32), pw::ceil_div((int)nrec, 32));
    hipLaunchKernelGGL(k_transpose, grid, dim3(32, 8), 0, s, p->d_tmp, p->d_tri, (int)nrec, p->nt);
    PW_HIP(ctx, hipGetLastError());
    PW_HIP(ctx, hipStreamSynchronize(s));
    p->have_tri = true;
    return PW_OK;
}

// Enqueue nsteps time steps (all nt when nsteps <= 0) from zeroed fields (zero == false keeps the
// fields, for timing).  evs: per launch a (start, end) pair stamped by the dispatch; returns the
// number of launches in *nlaunch.
static int afd_enqueue(pw_afd* p, int tri, hipEvent_t* evs, int nsteps = 0, bool zero = true, int* nlaunch = nullptr)
{
    pw_ctx* ctx = p->ctx;
    const Geo& g = p->g;
    if (!p->have_vel) PW_FAIL(ctx, PW_ERR_STATE, "run: velocity not set");
    if (tri && !p->have_tri) PW_FAIL(ctx, PW_ERR_STATE, "run: time-reversal data not set");
    if (!tri && !p->have_src) PW_FAIL(ctx, PW_ERR_STATE, "run: sources not set");
    hipStream_t s = ctx->s();
    const int vi = p->ndim == 3 && !p->ps ? (p->variant >= 0 ? p->variant : default_variant(p)) : -1;
    const bool two = vi >= 0 && kVariants3[vi].lean == 5;
    if (two && !variant_ok(p, vi)) PW_FAIL(ctx, PW_ERR_ARG, "two-step 3D kernel not applicable to this grid");
    if (two && !p->d_c) {
        PW_HIP(ctx, hipMalloc(&p->raw[3], p->raw_elems * 4));
        PW_HIP(ctx, hipMalloc(&p->raw[4], p->raw_elems * 4));
        PW_HIP(ctx, hipMemsetAsync(p->raw[3], 0, p->raw_elems * 4, s));
        PW_HIP(ctx, hipMemsetAsync(p->raw[4], 0, p->raw_elems * 4, s));
        p->d_c = p->raw[3] + p->guard_lo;
        p->d_d = p->raw[4] + p->guard_lo;
    }
    if (zero) {
        PW_HIP(ctx, hipMemsetAsync(p->raw[0], 0, p->raw_elems * 4, s));
        PW_HIP(ctx, hipMemsetAsync(p->raw[1], 0, p->raw_elems * 4, s));
    }
    const size_t nrec = (size_t)g.nx1 * g.ny1;
    const size_t snap_elems = (size_t)g.nz1 * g.nx1 * g.ny1;
    const int total = nsteps > 0 ? std::min(nsteps, p->nt) : p->nt;
    float* A = p->d_a;  // C_n
    float* B = p->d_b;  // C_{n-1}
    float* Cb = p->d_c;
    float* Db = p->d_d;
    const int its = tri ? -1 : 1;
    int it = tri ? p->nt - 1 : 0;
    auto rec_of = [&](int t) { return tri ? nullptr : p->d_rec + (size_t)t * nrec; };
    auto tri_of = [&](int t) { return tri ? p->d_tri + (size_t)t * nrec : nullptr; };
    auto snap_of = [&](int t) {
        return (p->jsnap > 0 && t % p->jsnap == 0) ? p->d_snap + (size_t)(t / p->jsnap) * snap_elems : nullptr;
    };
    int k = 0, done = 0;
    while (done < total) {
        if (two && total - done >= 2) {
            Step2 a{};
            a.cur = A; a.pv = B; a.vv = p->d_vv; a.out1 = Cb; a.out2 = Db;
            a.fz1 = p->d_fz1; a.fx1 = p->d_fx1; a.fy1 = p->d_fy1;
            a.ns = tri ? 0 : p->ns;
            a.src = p->d_src; a.wav = p->d_wav; a.nt = p->nt;
            a.it1 = it; a.it2 = it + its;
            a.tri1 = tri_of(a.it1); a.tri2 = tri_of(a.it2);
            a.rec1 = rec_of(a.it1); a.rec2 = rec_of(a.it2);
            a.snap1 = snap_of(a.it1); a.snap2 = snap_of(a.it2);
            int rc = afd_launch_2step(p, vi, a, evs ? evs[2 * k] : nullptr, evs ? evs[2 * k + 1] : nullptr);
            if (rc) return rc;
            // (C_{n+2}, C_{n+1}) become (C_n, C_{n-1}); the old pair is free for the next outputs
            float* oa = A;
            float* ob = B;
            A = Db; B = Cb; Cb = oa; Db = ob;
            it += 2 * its;
            done += 2;
        } else {
            Step a{};
            a.cur = A; a.pv = B; a.vv = p->d_vv;
            a.fz = p->d_fz; a.fx = p->d_fx; a.fy = p->d_fy;
            a.fz1 = p->d_fz1; a.fx1 = p->d_fx1; a.fy1 = p->d_fy1;
            a.ns = tri ? 0 : p->ns;
            a.src = p->d_src; a.wav = p->d_wav; a.nt = p->nt; a.it = it;
            a.tri = tri_of(it);
            a.rec = rec_of(it);
            a.snap = snap_of(it);
            int rc = PW_OK;
            if (p->ps) {
                // spectral Laplacian of C_n, then the update (the pair of launches is one step)
                PW_HIP(ctx, ps_apply(p->ps, A, p->d_wave, s));
                dim3 grid(pw::ceil_div(g.nz, 64), pw::ceil_div(g.nx, 4), g.ny);
                if (p->ndim == 3)
                    hipExtLaunchKernelGGL(k_ps_update<true>, grid, dim3(256), 0, s, evs ? evs[2 * k] : nullptr,
                                          evs ? evs[2 * k + 1] : nullptr, 0, g, a, (const float*)p->d_wave);
                else
                    hipExtLaunchKernelGGL(k_ps_update<false>, grid, dim3(256), 0, s, evs ? evs[2 * k] : nullptr,
                                          evs ? evs[2 * k + 1] : nullptr, 0, g, a, (const float*)p->d_wave);
                PW_HIP(ctx, hipGetLastError());
            } else {
                const int saved = p->variant;
                if (two) p->variant = kLeanDefault;  // odd remainder: one single step
                rc = afd_launch_step(p, a, evs ? evs[2 * k] : nullptr, evs ? evs[2 * k + 1] : nullptr);
                p->variant = saved;
            }
            if (rc) return rc;
            std::swap(A, B);
            it += its;
            done += 1;
        }
        ++k;
    }
    // keep the buffer roles for the next run: d_a..d_d are only names of the four fields
    p->final_field = A;
    p->last_mode = tri ? 1 : 0;
    if (nlaunch) *nlaunch = k;
    return PW_OK;
}

int pw_afd_run(pw_afd* p, int tri)
{
    if (afd_check(p)) return PW_ERR_ARG;
    return afd_enqueue(p, tri, nullptr);
}

int pw_afd_run_timed(pw_afd* p, int tri, int per_step, double* total_ms, double* step_ms)
{
    if (afd_check(p)) return PW_ERR_ARG;
    pw_ctx* ctx = p->ctx;
    hipStream_t s = ctx->s();
    hipEvent_t e0, e1;
    PW_HIP(ctx, hipEventCreate(&e0));
    PW_HIP(ctx, hipEventCreate(&e1));
    std::vector<hipEvent_t> evs;
    if (per_step) {
        evs.resize(2 * (size_t)p->nt);
        for (auto& e : evs) PW_HIP(ctx, hipEventCreate(&e));
    }
    PW_HIP(ctx, hipEventRecord(e0, s));
    int nl = 0;
    int rc = afd_enqueue(p, tri, per_step ? evs.data() : nullptr, 0, true, &nl);
    if (rc) return rc;
    PW_HIP(ctx, hipEventRecord(e1, s));
    PW_HIP(ctx, hipEventSynchronize(e1));
    float ms = 0.f;
    PW_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
    if (total_ms) *total_ms = ms;
    double acc = 0.0;
    if (per_step) {
        for (int k = 0; k < nl; ++k) {
            float m = 0.f;
            PW_HIP(ctx, hipEventElapsedTime(&m, evs[2 * k], evs[2 * k + 1]));
            acc += m;
        }
        for (auto& e : evs) hipEventDestroy(e);
    } else {
        acc = ms;
    }
    if (step_ms) *step_ms = acc / p->nt;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return PW_OK;
}

int pw_afd_output(pw_afd* p, float* out, long long out_len)
{
    if (afd_check(p) || !out) return PW_ERR_ARG;
    pw_ctx* ctx = p->ctx;
    const Geo& g = p->g;
    const size_t nrec = (size_t)g.nx1 * g.ny1, ntr = nrec * p->nt;
    const long long need = pw_afd_output_len(p->nt, g.nz1, g.nx1, g.ny1, p->jsnap);
    if (out_len != need) PW_FAIL(ctx, PW_ERR_ARG, "output: length mismatch");
    if (p->last_mode < 0) PW_FAIL(ctx, PW_ERR_STATE, "output: nothing has run");
    hipStream_t s = ctx->s();
    // result watchdog: a run that blew up (CFL) leaves Inf/NaN in the last wavefield and traces
    if (int rc = pw::check_finite(ctx, p->final_field, p->field_elems, 4, "afd: final wavefield")) return rc;
    if (p->last_mode == 0)
        if (int rc = pw::check_finite(ctx, p->d_rec, ntr, 4, "afd: receiver traces")) return rc;
    if (p->last_mode == 1) {
        std::memcpy(out, p->tri_host.data(), ntr * 4);  // TRI echoes its input data
    } else {
        dim3 grid(pw::ceil_div((int)nrec, 32), pw::ceil_div(p->nt, 32));
        hipLaunchKernelGGL(k_transpose, grid, dim3(32, 8), 0, s, p->d_rec, p->d_tmp, p->nt, (int)nrec);
        PW_HIP(ctx, hipGetLastError());
        PW_HIP(ctx, hipMemcpyAsync(out, p->d_tmp, ntr * 4, hipMemcpyDeviceToHost, s));
    }
    if (p->jsnap > 0)
        PW_HIP(ctx, hipMemcpyAsync(out + ntr, p->d_snap, (size_t)p->ntsnap * g.nz1 * nrec * 4,
                                   hipMemcpyDeviceToHost, s));
    PW_HIP(ctx, hipStreamSynchronize(s));
    return PW_OK;
}

int pw_afd_final_field(pw_afd* p, float* out, long long out_len)
{
    if (afd_check(p) || !out) return PW_ERR_ARG;
    pw_ctx* ctx = p->ctx;
    const Geo& g = p->g;
    const size_t n = (size_t)g.nz1 * g.nx1 * g.ny1;
    if ((size_t)out_len != n) PW_FAIL(ctx, PW_ERR_ARG, "final_field: length mismatch");
    if (!p->final_field) PW_FAIL(ctx, PW_ERR_STATE, "final_field: nothing has run");
    hipStream_t s = ctx->s();
    if (int rc = pw::check_finite(ctx, p->final_field, p->field_elems, 4, "afd: final wavefield")) return rc;
    dim3 grid(pw::ceil_div(g.nz1, 256), g.nx1, g.ny1);
    hipLaunchKernelGGL(k_interior, grid, dim3(256), 0, s, g, p->final_field, p->d_tmp);
    PW_HIP(ctx, hipGetLastError());
    PW_HIP(ctx, hipMemcpyAsync(out, p->d_tmp, n * 4, hipMemcpyDeviceToHost, s));
    PW_HIP(ctx, hipStreamSynchronize(s));
    return PW_OK;
}

// Time every applicable 3D kernel variant and y-chunk length on `steps` real steps of this plan
// (fields are reset by the next run) and keep the fastest; the choice is cached in the context
// per padded grid shape.  Returns the chosen variant in *variant_out and its chunk in *chunk_out.
int pw_afd_autotune(pw_afd* p, int steps, int* variant_out, int* chunk_out)
{
    if (afd_check(p) || steps < 1) return PW_ERR_ARG;
    pw_ctx* ctx = p->ctx;
    const Geo& g = p->g;
    if (p->ndim != 3) {
        if (variant_out) *variant_out = -1;
        if (chunk_out) *chunk_out = 0;
        return PW_OK;
    }
    const auto key = std::make_tuple(g.nz2, g.nx2, g.ny2);
    auto it = ctx->tuned.find(key);
    if (it == ctx->tuned.end()) {
        if (!p->have_vel) PW_FAIL(ctx, PW_ERR_STATE, "autotune: velocity not set");
        hipStream_t s = ctx->s();
        hipEvent_t e0, e1;
        PW_HIP(ctx, hipEventCreate(&e0));
        PW_HIP(ctx, hipEventCreate(&e1));
        const int chunks[] = {0, 4, 6, 8, 12, 16, 24, 32, 48};
        float best = 1e30f;
        std::pair<int, int> choice(default_variant(p), 0);
        const int saved_v = p->variant, saved_c = p->ny_chunk;
        const int nst = std::max(2, steps + (steps & 1));  // even: two-step kernels run whole launches
        const int mode = p->have_src ? 0 : 1;                 // time the plan's own kind of run
        if (mode == 1 && !p->have_tri) PW_FAIL(ctx, PW_ERR_STATE, "autotune: set sources or time-reversal data first");
        std::vector<std::tuple<int, int, float>> timed;
        for (int v = 0; v < kNumVariants3; ++v) {
            if (!variant_ok(p, v)) continue;
            for (int c : chunks) {
                p->variant = v;
                p->ny_chunk = c;
                int rc = afd_enqueue(p, mode, nullptr, 2, false);  // warm-up (and first-use allocation)
                if (rc) return rc;
                PW_HIP(ctx, hipEventRecord(e0, s));
                if ((rc = afd_enqueue(p, mode, nullptr, nst, false))) return rc;
                PW_HIP(ctx, hipEventRecord(e1, s));
                PW_HIP(ctx, hipEventSynchronize(e1));
                float ms = 0.f;
                PW_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
                timed.emplace_back(v, c, ms);
                best = std::min(best, ms);
            }
        }
        // the first configuration (table order) within 1.5% of the fastest: timing noise between
        // near-equal kernels does not flip the choice from run to run
        for (const auto& t : timed)
            if (std::get<2>(t) <= best * 1.015f) {
                choice = std::make_pair(std::get<0>(t), std::get<1>(t));
                break;
            }
        p->variant = saved_v;
        p->ny_chunk = saved_c;
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        it = ctx->tuned.emplace(key, choice).first;
    }
    p->variant = it->second.first;
    p->ny_chunk = it->second.second;
    if (variant_out) *variant_out = p->variant;
    if (chunk_out) *chunk_out = p->ny_chunk;
    return PW_OK;
}

// the pseudo-spectral Laplacian of aps2d/aps3d (on != 0) instead of the finite-difference stencil
int pw_afd_set_spectral(pw_afd* p, int on)
{
    if (afd_check(p)) return PW_ERR_ARG;
    pw_ctx* ctx = p->ctx;
    if (!on) {
        ps_destroy(p->ps);
        p->ps = nullptr;
        return PW_OK;
    }
    if (p->ps) return PW_OK;
    const Geo& g = p->g;
    std::string err;
    p->ps = ps_create(g.nz2, g.nx2, g.ny2, g.ldz, g.ldxy, p->dz, p->dx, p->dy, err);
    if (!p->ps) PW_FAIL(ctx, PW_ERR_ARG, "set_spectral: " + err);
    if (!p->d_wave) {
        PW_HIP(ctx, hipMalloc(&p->d_wave, p->field_elems * 4));
        PW_HIP(ctx, hipMemsetAsync(p->d_wave, 0, p->field_elems * 4, ctx->s()));
    }
    return PW_OK;
}

int pw_afd_set_variant(pw_afd* p, int variant, int ny_chunk)
{
    if (afd_check(p)) return PW_ERR_ARG;
    if (variant < 0 || variant >= kNumVariants3 || ny_chunk < 0)
        PW_FAIL(p->ctx, PW_ERR_ARG, "set_variant: unknown variant");
    p->variant = variant;
    p->ny_chunk = ny_chunk;
    return PW_OK;
}

static int afd_oneshot(pw_ctx* ctx, int ndim, const float* vel, const float* src, int tri, int nt,
                       int nx, int ny, int nz, int ns, int jsnap, int nbt, float ct, float dt,
                       float dx, float dy, float dz, float* out, long long out_len, int spectral = 0)
{
    if (!ctx) return PW_ERR_ARG;
    if (!vel || !out || (!src && (tri || ns > 0))) PW_FAIL(ctx, PW_ERR_ARG, "afd: null array");
    // ns = 0 models no source: zero traces and snapshots, as the reference's loops over no sources
    if (!tri && ns < 0) PW_FAIL(ctx, PW_ERR_ARG, "afd: ns must be >= 0 for modelling");
    if (jsnap < 0) jsnap = 0;
    if (out_len != pw_afd_output_len(nt, nz, nx, ny, jsnap))
        PW_FAIL(ctx, PW_ERR_ARG, "afd: output length mismatch");
    pw_afd* p = pw_afd_create(ctx, ndim, nz, nx, ny, dz, dx, dy, nt, dt, nbt, ct, jsnap);
    if (!p) return PW_ERR_ARG;
    int rc = spectral ? pw_afd_set_spectral(p, 1) : PW_OK;
    if (!rc) rc = pw_afd_set_velocity(p, vel);
    if (!rc) rc = tri ? pw_afd_set_tri_data(p, src) : pw_afd_set_sources(p, ns, src, nullptr);
    if (!rc) rc = pw_afd_run(p, tri);
    if (!rc) rc = pw_afd_output(p, out, out_len);
    pw_afd_destroy(p);
    return rc;
}

int pw_afd2d(pw_ctx* ctx, const float* vel, const float* src, int tri, int nt, int nx, int nz,
             int ns, int verb, int jsnap, int ifsnaps, int abc, int nbt, float ct, float dt,
             float ox, float dx, float oz, float dz, float* out, long long out_len)
{
    (void)verb; (void)ifsnaps; (void)abc; (void)ox; (void)oz;
    return afd_oneshot(ctx, 2, vel, src, tri, nt, nx, 1, nz, ns, jsnap, nbt, ct, dt, dx, 1.f, dz, out,
                       out_len);
}

}  // extern "C"

// ---------------------------------------------------------------- allocation-free 2D modelling
// pw_afd2d_dev_noalloc keeps every field of its run in one file-scope device array instead of
// memory it allocates: callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops, harnesses that forbid it) pass device inputs and outputs only.  The array holds
// kNoAllocBytes per GPU until the process ends; each GPU has its own copy at its own address,
// found once per device by hipGetSymbolAddress.  In host code the name 'scratch' is a host-side
// stand-in, never a device address: only the looked-up pointer is passed to kernels.
static constexpr size_t kNoAllocBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch_ptr;  // device -> the array's device address

int scratch_on_device(pw_ctx* ctx, unsigned char** out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(ctx->device);
    if (it == g_scratch_ptr.end()) {
        PW_HIP(ctx, hipSetDevice(ctx->device));
        void* ptr = nullptr;
        PW_HIP(ctx, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) PW_FAIL(ctx, PW_ERR_HIP, "noalloc: no device address for the scratch array");
        it = g_scratch_ptr.emplace(ctx->device, (unsigned char*)ptr).first;
    }
    *out = it->second;
    return PW_OK;
}

__global__ void k_fill_f32(float* __restrict__ p, size_t n, float v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

}  // namespace

extern "C" {

size_t pw_afd2d_noalloc_capacity(void) { return kNoAllocBytes; }

int pw_afd2d_dev_noalloc(pw_ctx* ctx, const float* d_vel, int nz, int nx, float dz, float dx, int nt, float dt,
                         int nbt, float ct, int ns, const int* src_xz, const float* d_wav, float* d_traces)
{
    if (!ctx) return PW_ERR_ARG;
    if (!d_vel || !d_traces || nz < 1 || nx < 1 || nt < 1 || nbt < 0 || ns < 0 || !(dz > 0.f) || !(dx > 0.f) ||
        !(dt > 0.f) || (ns > 0 && (!src_xz || !d_wav)))
        PW_FAIL(ctx, PW_ERR_ARG, "afd2d_dev_noalloc: invalid dimensions, spacings or pointers");
    unsigned char* base = nullptr;
    if (int rc = scratch_on_device(ctx, &base)) return rc;
    Geo g{};
    g.nz1 = nz; g.nx1 = nx; g.ny1 = 1;
    g.nb = nbt;
    g.nz = nz + 2 * nbt; g.nx = nx + 2 * nbt; g.ny = 1;
    pw::padded_sizes(g.nz, g.nx, 1, &g.nz2, &g.nx2, &g.ny2);
    g.ny2 = 1;
    g.ldz = z_stride(g.nz2, g.nx2, g.ny2);
    g.ldxy = (long long)g.ldz * g.nx2;
    g.idz = 1 / (dz * dz);
    g.idx = 1 / (dx * dx);
    g.idy = 0.f;
    g.cc = g.idx + g.idz;
    // carve the array: three fields, damping pairs, source cells, time-major receiver rows
    const size_t field = (size_t)g.ldxy;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) / 256 * 256; return o; };
    const size_t oA = take(field * 4), oB = take(field * 4), oV = take(field * 4);
    const size_t oFz = take((size_t)g.nz2 * 8), oFx = take((size_t)g.nx2 * 8), oFy = take(8);
    const size_t oS = take((size_t)3 * std::max(ns, 1) * 4), oR = take((size_t)nt * nx * 4);
    if (off > kNoAllocBytes)
        PW_FAIL(ctx, PW_ERR_ARG,
                "afd2d_dev_noalloc: the run needs " + std::to_string(off) + " bytes of device scratch, more than the " +
                    std::to_string(kNoAllocBytes) + " the allocation-free entry point holds");
    float* A = (float*)(base + oA);
    float* B = (float*)(base + oB);
    float* V = (float*)(base + oV);
    float2* fz = (float2*)(base + oFz);
    float2* fx = (float2*)(base + oFx);
    float2* fy = (float2*)(base + oFy);
    int* src = (int*)(base + oS);
    float* rec = (float*)(base + oR);
    // damping pairs (pywave/src/afd.c:392-402: every face nbt cells with ct) and source cells,
    // checked as pw_afd_set_sources checks them
    std::vector<float> w(std::max(nbt, 1));
    pw::abc_weights(nbt, ct, w.data());
    std::vector<float> pz, px, py(2, 1.0f);
    pw::damping_pairs(g.nz2, nbt, nbt, w.data(), w.data(), pz);
    pw::damping_pairs(g.nx2, nbt, nbt, w.data(), w.data(), px);
    std::vector<int> cen(3 * std::max(ns, 1), 0);
    for (int i = 0; i < ns; ++i) {
        const int gz = src_xz[ns + i] + nbt, gx = src_xz[i] + nbt;
        if (!(gz - 1 >= 0 && gz + 1 < g.nz && gx - 1 >= 0 && gx + 1 < g.nx))
            PW_FAIL(ctx, PW_ERR_ARG, "afd2d_dev_noalloc: source " + std::to_string(i) +
                                         " stencil leaves the padded domain");
        cen[3 * i] = gz; cen[3 * i + 1] = gx; cen[3 * i + 2] = 0;
    }
    hipStream_t s = ctx->s();
    PW_HIP(ctx, hipSetDevice(ctx->device));
    // the array keeps the previous call's contents: clear the fields (a fill kernel, so that the
    // sequence also records into a graph)
    const unsigned fb = (unsigned)std::min<size_t>(4096, pw::ceil_div<size_t>(3 * field, 256));
    hipLaunchKernelGGL(k_fill_f32, dim3(fb), dim3(256), 0, s, A, (size_t)(oV + field * 4 - oA) / 4, 0.f);
    PW_HIP(ctx, hipGetLastError());
    PW_HIP(ctx, hipMemcpyAsync(fz, pz.data(), pz.size() * 4, hipMemcpyHostToDevice, s));
    PW_HIP(ctx, hipMemcpyAsync(fx, px.data(), px.size() * 4, hipMemcpyHostToDevice, s));
    PW_HIP(ctx, hipMemcpyAsync(fy, py.data(), py.size() * 4, hipMemcpyHostToDevice, s));
    PW_HIP(ctx, hipMemcpyAsync(src, cen.data(), cen.size() * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_expand_vv, dim3(pw::ceil_div(g.nz, 256), g.nx, 1), dim3(256), 0, s, g, d_vel, V, dt);
    PW_HIP(ctx, hipGetLastError());
    Step a{};
    a.vv = V; a.fz = fz; a.fx = fx; a.fy = fy;
    a.src = src; a.wav = d_wav; a.ns = ns; a.nt = nt;
    constexpr int W = 4;
    const dim3 grid(pw::ceil_div(g.nz, 64), pw::ceil_div(g.nx, W));
    float* cur = A;
    float* pv = B;
    for (int it = 0; it < nt; ++it) {
        a.cur = cur; a.pv = pv; a.it = it;
        a.rec = rec + (size_t)it * nx;
        hipLaunchKernelGGL(k_afd2d<W>, grid, dim3(64 * W), 0, s, g, a);
        std::swap(cur, pv);
    }
    PW_HIP(ctx, hipGetLastError());
    // time-major rows -> the reference's time-fastest traces [ix][it]
    hipLaunchKernelGGL(k_transpose, dim3(pw::ceil_div(nx, 32), pw::ceil_div(nt, 32)), dim3(32, 8), 0, s, rec, d_traces,
                       nt, nx);
    PW_HIP(ctx, hipGetLastError());
    return PW_OK;
}

int pw_afd3d(pw_ctx* ctx, const float* vel, const float* src, int tri, int nt, int nx, int ny,
             int nz, int ns, int verb, int jsnap, int ifsnaps, int abc, int nbt, float ct,
             float dt, float ox, float dx, float oy, float dy, float oz, float dz, float* out,
             long long out_len)
{
    (void)verb; (void)ifsnaps; (void)abc; (void)ox; (void)oy; (void)oz;
    return afd_oneshot(ctx, 3, vel, src, tri, nt, nx, ny, nz, ns, jsnap, nbt, ct, dt, dx, dy, dz, out,
                       out_len);
}

// apscfun.aps2dc / aps3dc (pywave/src/aps.c): the afd entry points with the pseudo-spectral engine
int pw_aps2d(pw_ctx* ctx, const float* vel, const float* src, int tri, int nt, int nx, int nz,
             int ns, int verb, int jsnap, int ifsnaps, int abc, int nbt, float ct, float dt,
             float ox, float dx, float oz, float dz, float* out, long long out_len)
{
    (void)verb; (void)ifsnaps; (void)abc; (void)ox; (void)oz;
    return afd_oneshot(ctx, 2, vel, src, tri, nt, nx, 1, nz, ns, jsnap, nbt, ct, dt, dx, 1.f, dz, out,
                       out_len, 1);
}

int pw_aps3d(pw_ctx* ctx, const float* vel, const float* src, int tri, int nt, int nx, int ny,
             int nz, int ns, int verb, int jsnap, int ifsnaps, int abc, int nbt, float ct,
             float dt, float ox, float dx, float oy, float dy, float oz, float dz, float* out,
             long long out_len)
{
    (void)verb; (void)ifsnaps; (void)abc; (void)ox; (void)oy; (void)oz;
    return afd_oneshot(ctx, 3, vel, src, tri, nt, nx, ny, nz, ns, jsnap, nbt, ct, dt, dx, dy, dz, out,
                       out_len, 1);
}

}  // extern "C"

// ==================================================================================== TRI plans
// Receiver groups are back-propagated up to kTriPassGroups at a time.  With more groups the
// image needs every group's field at every snapshot step, so the passes keep the running product
// of the groups done so far per snapshot (f64, [snapshot][interior cell]): the first pass stores
// its f32 group product, later passes multiply theirs in, and the image is the sum over
// snapshots at the end.
constexpr int kTriPassGroups = 4;
constexpr int kTriMaxGroups = 64;

struct pw_tri {
    pw_ctx* ctx = nullptr;
    Geo g{};
    int nt = 0, G = 0, jsnap = 0;
    int npass = 1;            // passes of at most kTriPassGroups groups
    int nsnap = 0;
    double* d_part = nullptr;  // [nsnap][ny1][nx1][nz1] running products (npass > 1)
    float dt = 0.f;
    size_t field_elems = 0;
    float* d_f[8] = {nullptr};  // 2*G fields (inside d_pool when the pipelined kernel runs)
    float* d_pool = nullptr;
    unsigned pool_bytes = 0;  // resource size over the pool (pipelined kernel)
    float* d_vv = nullptr;
    float2 *d_fz = nullptr, *d_fx = nullptr, *d_fy = nullptr;
    float* d_dat = nullptr;  // [nt][ny1][nx1]
    float* d_tmp = nullptr;
    unsigned char* d_mask = nullptr;
    unsigned char* d_mbits = nullptr;  // [ny1][nx1] group bits (pipelined kernel)
    double* d_img = nullptr;
    float *d_fz1 = nullptr, *d_fx1 = nullptr, *d_fy1 = nullptr;
    bool pipe_ok = false;  // single damping factor per axis and 32-bit plane windows
    bool have_vel = false, have_dat = false, ran = false;
    // pipelined kernel: rows per block (one per wave) and y-chunk length, timed before the first run
    int tri_w = 8, tri_nyc = 0, tri_nt = 0;  // tri_nt: non-temporal policy on the streamed operands
    bool tuned = false;
    std::vector<float> tune_log;  // (rows per block, y chunk, non-temporal, ms per launch) of the timing sweep
};

namespace {

void tri_free(pw_tri* p)
{
    if (p->d_pool) {
        hipFree(p->d_pool);
    } else {
        for (int i = 0; i < 8; ++i)
            if (p->d_f[i]) hipFree(p->d_f[i]);
    }
    void* b[] = {p->d_vv, p->d_fz, p->d_fx, p->d_fy, p->d_dat, p->d_tmp, p->d_mask, p->d_img,
                 p->d_mbits, p->d_fz1, p->d_fx1, p->d_fy1, p->d_part};
    for (void* x : b)
        if (x) hipFree(x);
}

template <int G, int W, int NT>
void tri_launch_pipe(pw_tri* p, const TriStep<G>& a, hipStream_t s)
{
    const Geo& g = p->g;
    const int ntz = pw::ceil_div(g.nz, 64), ntx = pw::ceil_div(g.nx, W);
    const long long tiles = (long long)ntz * ntx;
    int nyc = p->tri_nyc;
    if (nyc <= 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->ctx->device) != hipSuccess || cus < 1)
            cus = 256;
        // y chunks of about 8 planes, at least two blocks per CU
        const int nch = std::max(pw::ceil_div(g.ny, 8), (int)std::max(1LL, (2LL * cus + tiles - 1) / tiles));
        nyc = std::max(1, pw::ceil_div(g.ny, nch));
    }
    const int nch = pw::ceil_div(g.ny, nyc);
    const int nblk = (int)(tiles * nch);
    const int grid = pw::ceil_div(nblk, 8) * 8;
    hipLaunchKernelGGL((k_tri3d_pipe<G, W, NT>), dim3(grid), dim3(64 * W), 0, s, g, a, ntz, ntx, nyc, nblk);
}

template <int G>
void tri_launch(pw_tri* p, const TriStep<G>& a, hipStream_t s)
{
    const Geo& g = p->g;
    if (p->pipe_ok) {
        if (p->tri_nt) {
            switch (p->tri_w) {
                case 4: tri_launch_pipe<G, 4, 1>(p, a, s); break;
                case 16: tri_launch_pipe<G, 16, 1>(p, a, s); break;
                default: tri_launch_pipe<G, 8, 1>(p, a, s); break;
            }
        } else {
            switch (p->tri_w) {
                case 4: tri_launch_pipe<G, 4, 0>(p, a, s); break;
                case 16: tri_launch_pipe<G, 16, 0>(p, a, s); break;
                default: tri_launch_pipe<G, 8, 0>(p, a, s); break;
            }
        }
        return;
    }
    const int ntz = pw::ceil_div(g.nz, 64), ntx = pw::ceil_div(g.nx, 16);
    long long tiles = (long long)ntz * ntx;
    int nch = (int)std::max(1LL, (6 * 256 + tiles - 1) / tiles);
    int nyc = std::max(16, pw::ceil_div(g.ny, nch));
    nch = pw::ceil_div(g.ny, nyc);
    const int nblk = ntz * ntx * nch;
    const int grid = pw::ceil_div(nblk, 8) * 8;
    hipLaunchKernelGGL(k_tri3d<G>, dim3(grid), dim3(256), 0, s, g, a, ntz, ntx, nyc, nblk);
}

// image = sum over snapshots of the running products (latest snapshot first, the order in which a
// single pass adds them)
__global__ void k_tri_sum_snaps(size_t ncell, int nsnap, const double* __restrict__ part, double* __restrict__ img)
{
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += (size_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int k = nsnap - 1; k >= 0; --k) v += part[(size_t)k * ncell + c];
        img[c] = v;
    }
}

// the launch arguments of step it of pass `pass` (fields A = current, B = previous), no image
template <int G>
TriStep<G> tri_step(const pw_tri* p, int pass, int it, float* const* A, float* const* B)
{
    const size_t nrec = (size_t)p->g.nx1 * p->g.ny1;
    TriStep<G> a{};
    for (int f = 0; f < G; ++f) { a.cur[f] = A[f]; a.pv[f] = B[f]; }
    a.vv = p->d_vv; a.fz = p->d_fz; a.fx = p->d_fx; a.fy = p->d_fy;
    a.dat = p->d_dat + (size_t)it * nrec;
    a.mask = p->d_mask + (size_t)pass * kTriPassGroups * nrec;
    a.img = nullptr;
    a.img_mul = pass > 0;
    a.mbits = p->d_mbits ? p->d_mbits + (size_t)pass * nrec : nullptr;
    a.fz1 = p->d_fz1; a.fx1 = p->d_fx1; a.fy1 = p->d_fy1;
    a.pool = p->d_pool;
    a.pool_bytes = p->pool_bytes;
    if (p->d_pool)
        for (int f = 0; f < G; ++f) {
            a.off_cur[f] = (int)(A[f] - p->d_pool);
            a.off_pv[f] = (int)(B[f] - p->d_pool);
        }
    return a;
}

// Launch-shape sweep of the pipelined kernel: rows per block 4, 8, 16 (one per wave), y chunks of
// 4 to 24 planes and the cache policy of the streamed operands (default or non-temporal), each
// timed on the last time step of pass 0 (no image) on the plan's own fields; the fastest is kept.  The run that follows clears the fields the timing launches wrote.
template <int G>
int tri_tune(pw_tri* p)
{
    pw_ctx* ctx = p->ctx;
    hipStream_t s = ctx->s();
    float* A[G];
    float* B[G];
    for (int f = 0; f < G; ++f) { A[f] = p->d_f[2 * f]; B[f] = p->d_f[2 * f + 1]; }
    const TriStep<G> a = tri_step<G>(p, 0, p->nt - 1, A, B);
    hipEvent_t e0, e1;
    PW_HIP(ctx, hipEventCreate(&e0));
    PW_HIP(ctx, hipEventCreate(&e1));
    p->tune_log.clear();
    float best = 0.f;
    int bw = 8, bc = 0, bn = 0, rc = PW_OK;
    constexpr int kRows[] = {4, 8, 16}, kChunks[] = {4, 6, 8, 12, 16, 24};
    for (int combo = 0; combo < 36 && !rc; ++combo) {
        const int W = kRows[combo / 12], nyc = kChunks[combo % 6];
        const int nt = (combo / 6) & 1;
        {
            p->tri_w = W;
            p->tri_nyc = std::min(nyc, p->g.ny);
            p->tri_nt = nt;
            const int reps = 6;
            for (int r = 0; r < 2; ++r) tri_launch<G>(p, a, s);
            hipEventRecord(e0, s);
            for (int r = 0; r < reps; ++r) tri_launch<G>(p, a, s);
            hipEventRecord(e1, s);
            float ms = 0.f;
            if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                ctx->err = "tri: launch-shape timing failed";
                rc = PW_ERR_HIP;
                break;
            }
            ms /= reps;
            for (float v : {(float)W, (float)p->tri_nyc, (float)nt, ms}) p->tune_log.push_back(v);
            if (best == 0.f || ms < best) { best = ms; bw = W; bc = p->tri_nyc; bn = nt; }
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    p->tri_w = bw;
    p->tri_nyc = bc;
    p->tri_nt = bn;
    p->tuned = true;
    return PW_OK;
}

int tri_ensure_tuned(pw_tri* p)
{
    if (p->tuned || !p->pipe_ok) return PW_OK;
    switch (std::min(kTriPassGroups, p->G)) {
        case 1: return tri_tune<1>(p);
        case 2: return tri_tune<2>(p);
        case 3: return tri_tune<3>(p);
        default: return tri_tune<4>(p);
    }
}

// one back-propagation of G groups, the pass-th group of passes (groups pass*kTriPassGroups ...)
template <int G>
int tri_enqueue(pw_tri* p, int pass)
{
    pw_ctx* ctx = p->ctx;
    hipStream_t s = ctx->s();
    const Geo& g = p->g;
    const size_t nrec = (size_t)g.nx1 * g.ny1;
    const size_t ncell = (size_t)g.nz1 * nrec;
    for (int i = 0; i < 2 * G; ++i) PW_HIP(ctx, hipMemsetAsync(p->d_f[i], 0, p->field_elems * 4, s));
    if (p->npass == 1) PW_HIP(ctx, hipMemsetAsync(p->d_img, 0, ncell * 8, s));
    else if (pass == 0) PW_HIP(ctx, hipMemsetAsync(p->d_part, 0, (size_t)p->nsnap * ncell * 8, s));
    float* A[G];
    float* B[G];
    for (int f = 0; f < G; ++f) { A[f] = p->d_f[2 * f]; B[f] = p->d_f[2 * f + 1]; }
    for (int it = p->nt - 1; it >= 0; --it) {
        TriStep<G> a = tri_step<G>(p, pass, it, A, B);
        const bool snap = p->jsnap > 0 && it % p->jsnap == 0;
        if (p->npass == 1) a.img = snap ? p->d_img : nullptr;
        else a.img = snap ? p->d_part + (size_t)(it / p->jsnap) * ncell : nullptr;
        tri_launch<G>(p, a, s);
        PW_HIP(ctx, hipGetLastError());
        for (int f = 0; f < G; ++f) std::swap(A[f], B[f]);
    }
    p->ran = true;
    return PW_OK;
}

}  // namespace

extern "C" {

pw_tri* pw_tri_create(pw_ctx* ctx, int nz, int nx, int ny, float dz, float dx, float dy, int nt, float dt,
                      int nbt, float ct, int ngroups, int jsnap)
{
    if (!ctx) return nullptr;
    if (nz < 1 || nx < 1 || ny < 1 || nt < 1 || nbt < 0 || jsnap < 1 || ngroups < 1 || ngroups > kTriMaxGroups ||
        !(dz > 0.f) || !(dx > 0.f) || !(dy > 0.f) || !(dt > 0.f)) {
        ctx->err = "pw_tri_create: invalid arguments (1 <= ngroups <= " + std::to_string(kTriMaxGroups) +
                   ", jsnap >= 1)";
        return nullptr;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) { ctx->err = "hipSetDevice failed"; return nullptr; }
    pw_tri* p = new pw_tri;
    p->ctx = ctx;
    p->nt = nt; p->dt = dt; p->G = ngroups; p->jsnap = jsnap;
    p->npass = pw::ceil_div(ngroups, kTriPassGroups);
    p->nsnap = (nt - 1) / jsnap + 1;
    const int gp = std::min(ngroups, kTriPassGroups);  // fields held at once: two per group of a pass
    Geo& g = p->g;
    g.nz1 = nz; g.nx1 = nx; g.ny1 = ny; g.nb = nbt;
    g.nz = nz + 2 * nbt; g.nx = nx + 2 * nbt; g.ny = ny + 2 * nbt;
    pw::padded_sizes(g.nz, g.nx, g.ny, &g.nz2, &g.nx2, &g.ny2);
    g.ldz = z_stride(g.nz2, g.nx2, g.ny2);
    g.ldxy = (long long)g.ldz * g.nx2;
    g.idz = 1 / (dz * dz); g.idx = 1 / (dx * dx); g.idy = 1 / (dy * dy);
    g.cc = g.idy + g.idx + g.idz;
    p->field_elems = (size_t)g.ldxy * g.ny2;
    std::vector<float> w(std::max(nbt, 1));
    pw::abc_weights(nbt, ct, w.data());
    std::vector<float> fz, fx, fy;
    pw::damping_pairs(g.nz2, nbt, nbt, w.data(), w.data(), fz);
    pw::damping_pairs(g.nx2, nbt, nbt, w.data(), w.data(), fx);
    pw::damping_pairs(g.ny2, nbt, nbt, w.data(), w.data(), fy);
    const size_t nrec = (size_t)nx * ny;
    // single damping factor per index when the two faces of an axis never overlap
    auto single = [](const std::vector<float>& pr, size_t len, std::vector<float>& out) -> bool {
        out.assign(len, 1.0f);
        for (size_t i = 0; i < pr.size() / 2; ++i) {
            float lo = pr[2 * i], hi = pr[2 * i + 1];
            if (lo != 1.0f && hi != 1.0f) return false;
            out[i] = lo != 1.0f ? lo : hi;
        }
        return true;
    };
    std::vector<float> fz1, fx1, fy1;
    bool pipe = single(fz, (size_t)pw::ceil_div(g.nz2, 64) * 64 + 64, fz1) && single(fx, (size_t)g.nx2 + 40, fx1) &&
                single(fy, (size_t)g.ny2 + 8, fy1);
    pipe = pipe && (double)(g.ldxy + 4 * g.ldz + 4) * 4.0 < 2147483648.0 && (double)nx * nz * 8.0 < 2147483648.0;
    p->pipe_ok = pipe;
    bool ok = true;
    // the pipelined kernel addresses all 2G fields through one buffer resource (< 2 GiB)
    const double pool_b = ((double)2 * gp * p->field_elems + 4.0 * g.ldz + 4) * 4.0;
    pipe = pipe && pool_b < 2147483648.0;
    p->pipe_ok = pipe;
    if (pipe) {
        ok = hipMalloc(&p->d_pool, (size_t)2 * gp * p->field_elems * 4) == hipSuccess;
        for (int i = 0; i < 2 * gp && ok; ++i) p->d_f[i] = p->d_pool + (size_t)i * p->field_elems;
        p->pool_bytes = (unsigned)((2LL * gp * (long long)p->field_elems + 2LL * (2 * g.ldz + 2)) * 4);
    } else {
        for (int i = 0; i < 2 * gp && ok; ++i) ok = hipMalloc(&p->d_f[i], p->field_elems * 4) == hipSuccess;
    }
    if (ok && p->npass > 1)
        ok = hipMalloc(&p->d_part, (size_t)p->nsnap * nz * nrec * 8) == hipSuccess;
    if (ok && pipe)
        ok = hipMalloc(&p->d_fz1, fz1.size() * 4) == hipSuccess && hipMalloc(&p->d_fx1, fx1.size() * 4) == hipSuccess &&
             hipMalloc(&p->d_fy1, fy1.size() * 4) == hipSuccess &&
             hipMalloc(&p->d_mbits, (size_t)p->npass * nrec) == hipSuccess &&
             hipMemcpy(p->d_fz1, fz1.data(), fz1.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(p->d_fx1, fx1.data(), fx1.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(p->d_fy1, fy1.data(), fy1.size() * 4, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMalloc(&p->d_vv, p->field_elems * 4) == hipSuccess &&
         hipMalloc(&p->d_fz, fz.size() * 4) == hipSuccess && hipMalloc(&p->d_fx, fx.size() * 4) == hipSuccess &&
         hipMalloc(&p->d_fy, fy.size() * 4) == hipSuccess &&
         hipMalloc(&p->d_dat, nrec * nt * 4) == hipSuccess && hipMalloc(&p->d_tmp, nrec * nt * 4) == hipSuccess &&
         hipMalloc(&p->d_mask, nrec * ngroups) == hipSuccess &&
         hipMalloc(&p->d_img, (size_t)nz * nrec * 8) == hipSuccess;
    hipStream_t s = ctx->s();
    ok = ok && hipMemcpyAsync(p->d_fz, fz.data(), fz.size() * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(p->d_fx, fx.data(), fx.size() * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(p->d_fy, fy.data(), fy.size() * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemsetAsync(p->d_vv, 0, p->field_elems * 4, s) == hipSuccess && hipStreamSynchronize(s) == hipSuccess;
    if (!ok) {
        ctx->err = "pw_tri_create: device allocation/upload failed";
        tri_free(p);
        delete p;
        return nullptr;
    }
    return p;
}

void pw_tri_destroy(pw_tri* p)
{
    if (!p) return;
    hipSetDevice(p->ctx->device);
    hipStreamSynchronize(p->ctx->s());
    tri_free(p);
    delete p;
}

int pw_tri_set_velocity(pw_tri* p, const float* vel)
{
    if (!p || !vel) return PW_ERR_ARG;
    pw_ctx* ctx = p->ctx;
    const Geo& g = p->g;
    size_t n = (size_t)g.nz1 * g.nx1 * g.ny1;
    float* d_vel = nullptr;
    PW_HIP(ctx, hipMalloc(&d_vel, n * 4));
    hipStream_t s = ctx->s();
    hipError_t e = hipMemcpyAsync(d_vel, vel, n * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_expand_vv, dim3(pw::ceil_div(g.nz, 256), g.nx, g.ny), dim3(256), 0, s, g, d_vel,
                           p->d_vv, p->dt);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(d_vel);
    if (e != hipSuccess) PW_FAIL(ctx, PW_ERR_HIP, std::string("tri set_velocity: ") + hipGetErrorString(e));
    p->have_vel = true;
    return PW_OK;
}

// dat: nt*nx*ny receiver data, time fastest ((nt,nx,ny) Fortran order); masks: ngroups*nx*ny
// bytes, [g][iy][ix], non-zero where the receiver belongs to group g
int pw_tri_set_data(pw_tri* p, const float* dat, const unsigned char* masks)
{
    if (!p || !dat || !masks) return PW_ERR_ARG;
    pw_ctx* ctx = p->ctx;
    const Geo& g = p->g;
    const size_t nrec = (size_t)g.nx1 * g.ny1;
    hipStream_t s = ctx->s();
    PW_HIP(ctx, hipMemcpyAsync(p->d_tmp, dat, nrec * p->nt * 4, hipMemcpyHostToDevice, s));
    dim3 grid(pw::ceil_div(p->nt, 32), pw::ceil_div((int)nrec, 32));
    hipLaunchKernelGGL(k_transpose, grid, dim3(32, 8), 0, s, p->d_tmp, p->d_dat, (int)nrec, p->nt);
    PW_HIP(ctx, hipGetLastError());
    PW_HIP(ctx, hipMemcpyAsync(p->d_mask, masks, nrec * p->G, hipMemcpyHostToDevice, s));
    // per pass, bit f of a receiver's byte: it belongs to the pass's f-th group
    std::vector<unsigned char> bits((size_t)p->npass * nrec, 0);
    for (int f = 0; f < p->G; ++f)
        for (size_t k = 0; k < nrec; ++k)
            if (masks[(size_t)f * nrec + k])
                bits[(size_t)(f / kTriPassGroups) * nrec + k] |= (unsigned char)(1u << (f % kTriPassGroups));
    if (p->d_mbits) PW_HIP(ctx, hipMemcpyAsync(p->d_mbits, bits.data(), bits.size(), hipMemcpyHostToDevice, s));
    PW_HIP(ctx, hipStreamSynchronize(s));
    p->have_dat = true;
    return PW_OK;
}

int pw_tri_run(pw_tri* p)
{
    if (!p) return PW_ERR_ARG;
    if (!p->have_vel || !p->have_dat) PW_FAIL(p->ctx, PW_ERR_STATE, "tri run: velocity or data not set");
    if (int rc = tri_ensure_tuned(p)) return rc;
    for (int pass = 0; pass < p->npass; ++pass) {
        const int gk = std::min(kTriPassGroups, p->G - pass * kTriPassGroups);
        int rc = PW_ERR_ARG;
        switch (gk) {
            case 1: rc = tri_enqueue<1>(p, pass); break;
            case 2: rc = tri_enqueue<2>(p, pass); break;
            case 3: rc = tri_enqueue<3>(p, pass); break;
            case 4: rc = tri_enqueue<4>(p, pass); break;
        }
        if (rc) return rc;
    }
    if (p->npass > 1) {
        const Geo& g = p->g;
        const size_t ncell = (size_t)g.nz1 * g.nx1 * g.ny1;
        const unsigned blocks = (unsigned)std::min<size_t>(4096, pw::ceil_div<size_t>(ncell, 256));
        hipLaunchKernelGGL(k_tri_sum_snaps, dim3(blocks), dim3(256), 0, p->ctx->s(), ncell, p->nsnap,
                           (const double*)p->d_part, p->d_img);
        PW_HIP(p->ctx, hipGetLastError());
    }
    return PW_OK;
}

int pw_tri_run_timed(pw_tri* p, double* total_ms)
{
    if (!p) return PW_ERR_ARG;
    pw_ctx* ctx = p->ctx;
    hipStream_t s = ctx->s();
    if (!p->have_vel || !p->have_dat) PW_FAIL(ctx, PW_ERR_STATE, "tri run: velocity or data not set");
    if (int rc = tri_ensure_tuned(p)) return rc;  // the launch-shape sweep is not part of the timed run
    hipEvent_t e0, e1;
    PW_HIP(ctx, hipEventCreate(&e0));
    PW_HIP(ctx, hipEventCreate(&e1));
    PW_HIP(ctx, hipEventRecord(e0, s));
    int rc = pw_tri_run(p);
    if (rc) return rc;
    PW_HIP(ctx, hipEventRecord(e1, s));
    PW_HIP(ctx, hipEventSynchronize(e1));
    float ms = 0.f;
    PW_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
    if (total_ms) *total_ms = ms;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return PW_OK;
}

// launch shape of the pipelined kernel: rows per block 4, 8 or 16, a y chunk of at least one
// plane and the cache policy (non_temporal 0 or 1); rows 0 returns to the timed choice
int pw_tri_set_launch(pw_tri* p, int rows, int y_chunk, int non_temporal)
{
    if (!p) return PW_ERR_ARG;
    if (rows == 0) {
        p->tuned = false;
        return PW_OK;
    }
    if (!(rows == 4 || rows == 8 || rows == 16) || y_chunk < 1)
        PW_FAIL(p->ctx, PW_ERR_ARG, "tri set_launch: rows must be 4, 8 or 16 and the y chunk positive");
    p->tri_w = rows;
    p->tri_nyc = y_chunk;
    p->tri_nt = non_temporal ? 1 : 0;
    p->tune_log.clear();
    p->tuned = true;
    return PW_OK;
}

// the pipelined kernel's launch shape: w[0] rows per block, w[1] y-chunk planes (0: not yet
// chosen or the general kernel runs), w[2] 1 for the non-temporal policy; the timing sweep's
// (rows, chunk, non-temporal, ms per launch) quadruples into log[4*cap]; returns their number
int pw_tri_launch_choice(pw_tri* p, int* w, float* log, int cap)
{
    if (!p) return PW_ERR_ARG;
    if (w) {
        w[0] = p->pipe_ok ? p->tri_w : 0;
        w[1] = p->pipe_ok ? p->tri_nyc : 0;
        w[2] = p->pipe_ok ? p->tri_nt : 0;
    }
    const int n = (int)p->tune_log.size() / 4;
    for (int i = 0; i < std::min(n, cap) * 4 && log; ++i) log[i] = p->tune_log[i];
    return n;
}

// image (nz*nx*ny, z fastest) in f64, unnormalised
int pw_tri_image(pw_tri* p, double* img, long long len)
{
    if (!p || !img) return PW_ERR_ARG;
    const Geo& g = p->g;
    if (len != (long long)g.nz1 * g.nx1 * g.ny1) PW_FAIL(p->ctx, PW_ERR_ARG, "tri image: length mismatch");
    if (!p->ran) PW_FAIL(p->ctx, PW_ERR_STATE, "tri image: nothing has run");
    hipStream_t s = p->ctx->s();
    if (int rc = pw::check_finite(p->ctx, p->d_img, (size_t)len, 8, "tri: image")) return rc;
    PW_HIP(p->ctx, hipMemcpyAsync(img, p->d_img, len * 8, hipMemcpyDeviceToHost, s));
    PW_HIP(p->ctx, hipStreamSynchronize(s));
    return PW_OK;
}

}  // extern "C"
