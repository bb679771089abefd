// pw_fwi.hip -- pfwi's acoustic FWI discretisation on MI355X.
//
// Reference path (pywave = github.com/chenyk1990/pywave):
//   geometry/sponge   pywave/src/wave_fwiutil.c:19-121 (preparation), :201-223 (apply_sponge)
//   8th-order stencil pywave/src/wave_fwiutil.c:17,179-199 (laplace)
//   source map        pywave/src/wave_fwiutil.c:149-177, pywave/src/wave_triangle.c:177-238
//   modelling         pywave/src/wave_fwimodeling.c:21-114 (forward_modeling_a)
//   gradient          pywave/src/wave_fwigradient.c:20-262 (gradient_init, gradient_av)
//   smoothing         pywave/src/wave_fwiutil.c:298-335 (gradient_smooth2)
//   optimiser         pywave/src/wave_fwi.c:24-119, pywave/src/wave_fwilbfgs.c:9-165
//
// One fused kernel per time step and shot batch.  As in the afd engine the reference's separate
// passes (record, snapshot, Laplacian, source add, leapfrog, rotate, sponge both levels) become
// one pass per cell: the kernel reads p1 (sponged) and the previous p1 (sponged once), forms the
// second sponge of p0 on read, writes S(2 p1 - S(p0raw) + (v v dt^2)(lap p1 + rr w)) over p0raw.
// Per-cell f32 arithmetic follows the reference's order (built without FP contraction), so traces
// and the imaging-condition gradient reproduce the reference C bit for bit for one shot per
// launch; batched launches differ only in the order per-shot gradients are summed.
//
// Two gradients are provided:
//   mode 0  the reference's imaging condition: -2/v^3 * (second time difference of forward
//           snapshots) * adjoint field, residual windowed, first-call 0.1/max scaling, triangle
//           smoothing (gradient_av);
//   mode 1  the exact discrete adjoint of the modelling map (the gradient of 0.5||F(v)-d||^2,
//           also used by the torch.autograd.Function), checked against finite differences.
// Both use forward checkpoints every K steps and recompute each segment during the backward
// sweep; when all segment buffers fit the memory budget the whole run is one segment.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "pw_buffer.h"
#include "pw_internal.h"

namespace {

constexpr double PI_D = 3.14159265358979323846264338328;
constexpr int PMAX = 7;  // largest source-map patch edge (rect <= 4)
// Tile widths (x columns streamed per wave) the step kernels are built for; the plan picks one per
// kernel by timing them on its own grid and batch (tune_widths), 8 until then.
constexpr int kWidths[] = {4, 8, 16};
constexpr int kDefaultWidth = 8;

// Tiles of the step kernels: 64 z cells x W columns of one shot slot, one wave each.
struct TileMap {
    int ntz, nxc;  // z tiles and column groups per slot
    int ns;        // slots (shots) of the launch
    int nblk;      // tiles over all slots of the launch
};

TileMap tile_map(int padnz, int padnx, int nslots, int width)
{
    TileMap t;
    t.ntz = (padnz + 63) / 64;
    t.nxc = (padnx + width - 1) / width;
    t.ns = nslots;
    t.nblk = t.ntz * t.nxc * nslots;
    return t;
}

// blocks to launch for a tile map: a multiple of the 8 XCDs
inline dim3 tile_grid(const TileMap& t) { return dim3((unsigned)((t.nblk + 7) / 8 * 8)); }

// XCD-aware tile order: the hardware deals blocks round-robin over the 8 XCDs; each XCD takes a
// contiguous range of tiles, slot fastest, then z tile, then column group: the shots' tiles of one
// grid tile run together on one XCD and read the fields the shots share (v^2dt^2, the sponge)
// through one L2, and column neighbours, which read each other's halo columns, share it too.
// False for the padding blocks.
__device__ __forceinline__ bool tile_of(const TileMap& t, int& slot, int& tz, int& xc)
{
    const int per = gridDim.x >> 3;
    const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (lb >= t.nblk) return false;
    const int r = lb / t.ns;
    slot = lb - r * t.ns;
    tz = r % t.ntz;
    xc = r / t.ntz;
    return true;
}

template <typename T>
struct Coef {
    // pywave/src/wave_fwiutil.c:17 (f32 constants from double literals; f64 for T = double)
    static constexpr T K0 = (T)(-205. / 72), K1 = (T)(8. / 5), K2 = (T)(-1. / 5), K3 = (T)(8. / 315),
                       K4 = (T)(-1. / 560);
};

template <typename T>
struct FGeo {
    int pnz, pnx, ldz;  // padded extents, z stride
    int nz, nx, nb;
    int nt, nr, rz, dr_v;
    T dx2, dz2;
    double wdt2;  // imaging denominator, f32 value promoted (reference promotes to double)
    long long fs;  // field stride
};

struct ShotInfo {
    int sx0, sz0, pw, ph;  // source-map patch origin (padded coords) and extent
    int r0_v, r02, nr2, shot;
};

template <typename T>
struct StepArgs {
    const T* X;     // p1 (sponged)
    T* Y;           // in: previous p1; out: new p1
    const T* m;     // (v*v)*dt^2 on the padded grid
    const T* sfz;   // sponge factor pairs per z (top, bottom)
    const T* sfx;   // per x (left, right)
    const ShotInfo* shots;
    const T* patch;     // per slot PMAX*PMAX, x-major
    const T* wav;       // nt
    T* rec;             // forward: recorded p1 at receivers [slot][nt][nr]
    const T* inj;       // backward: injected receiver signal [slot][nt][nr]
    long long trs;      // trace stride per slot (nt*nr)
    T* snap;            // interior snapshot of p1 for this step: [slot][nx][nz] or null
    long long snaps;    // snapshot stride per slot
    T* full;            // the step's term (masked L p1 + sources) on the full field [slot][fs], or
    long long fulls;    // null: what the exact adjoint's dJ/dm needs of the forward step
    const T* fsrc;      // passive source field of this step on the interior, [slot][nx][nz], or null
    long long fss;      // its slot stride (nt*nz*nx)
    // visco-acoustic modelling: tau, taus on the padded grid, memory variable r (damped once per
    // step) read from r1 and written to r2, per slot [slot][fs]; idt = 1/dt in f32
    const float* tau;
    const float* taus;
    const float* r1;
    float* r2;
    float idt;
    int it;
};

// v where k holds, +0 elsewhere, as a bit mask (a select the compiler keeps branch-free)
__device__ __forceinline__ float keep_if(float v, bool k)
{
    return __builtin_bit_cast(float, __builtin_bit_cast(unsigned, v) & (k ? ~0u : 0u));
}
__device__ __forceinline__ double keep_if(double v, bool k)
{
    using u64 = unsigned long long;
    return __builtin_bit_cast(double, __builtin_bit_cast(u64, v) & (k ? ~(u64)0 : (u64)0));
}

template <typename T>
__device__ __forceinline__ T sponge(T v, T zl, T zh, T xl, T xh)
{
    v = zl * v; v = zh * v;  // top, bottom faces (apply_sponge z loop)
    v = xl * v; v = xh * v;  // left, right faces (x loop)
    return v;
}

// ---------------------------------------------------------------- buffer-resource access
// (pw_buffer.h): a lane whose cell is outside the padded grid gets an out-of-range offset,
// answered with 0 without a memory access; its stores are dropped.
using pwbuf::kOOB;
using pwbuf::win;
using pwbuf::bld;
using pwbuf::bldo;
using pwbuf::bsto;
using pwbuf::bldo_nt;
using pwbuf::bsto_nt;


// ---------------------------------------------------------------- forward step (+ modelling)
// One 64 x W tile of one shot per block (one wave), streamed along x with a 9-column register
// queue (radius 4) and an LDS z row; the operands of column x+1 (previous level, model, z halo,
// sponge pair, and the queue entry x+5) are in flight while column x is computed.
template <typename T, bool VISCO, int XC>
__global__ __launch_bounds__(64) void k_fwd(FGeo<T> g, StepArgs<T> a, TileMap tm)
{
    using C = Coef<T>;
    int slot, tz, xc;
    if (!tile_of(tm, slot, tz, xc)) return;
    const int lane = threadIdx.x, z0 = tz * 64, z = z0 + lane;
    const int x0 = xc * XC, x1 = min(x0 + XC, g.pnx);
    __shared__ T row[2][128];  // z row (4..67) and halo (0..3, 68..71); lanes 8..63 may write 72..127
    const ShotInfo sh = a.shots[slot];
    const int L = g.ldz;
    const int bias = 4 * L + 4;  // windows start 4 columns + 4 cells before the field
    const unsigned wb = (unsigned)((g.fs + 2 * bias) * (long long)sizeof(T));
    const auto rX = win(a.X + slot * g.fs - bias, wb), rY = win(a.Y + slot * g.fs - bias, wb),
               rM = win(a.m - bias, wb);
    const bool zok = z < g.pnz;
    const bool zin = z >= 4 && z < g.pnz - 4;
    const T zl = zok ? a.sfz[2 * z] : T(1), zh = zok ? a.sfz[2 * z + 1] : T(1);
    const int zhz = lane < 4 ? z0 - 4 + lane : z0 + 60 + lane;  // z halo of lanes 0..7
    const int lzh = lane < 4 ? lane : 64 + lane;
    const bool zhv = lane < 8 && zhz >= 0 && zhz < g.pnz;
    constexpr unsigned E = sizeof(T);
    // lane offsets are fixed (out of range for lanes outside the grid); the column is the scalar
    // offset; a column outside the grid gives every lane the out-of-range offset
    const unsigned oz = zok ? (unsigned)(bias + z) * E : kOOB, ozh = zhv ? (unsigned)(bias + zhz) * E : kOOB;
    const unsigned oc = (unsigned)L * E;
    auto xv = [&](int x) { return x >= 0 && x < g.pnx; };
    auto ldc = [&](__amdgpu_buffer_rsrc_t r, unsigned voff, int x, bool ok) {  // column x of a field
        return bldo<T>(r, ok ? voff : kOOB, ok ? (unsigned)x * oc : 0u);
    };
    const T wv = a.wav[a.it];
    // passive source field (gradient_pas_av's ww3): added to the Laplacian on the model interior
    const bool fzin = a.fsrc && z >= g.nb && z < g.nb + g.nz;
    const auto rS = win(a.fsrc ? a.fsrc + slot * a.fss : nullptr, a.fsrc ? (unsigned)(g.nz * g.nx * sizeof(T)) : 0u);
    const unsigned ofz = fzin ? (unsigned)(z - g.nb) * E : kOOB;
    auto fxin = [&](int x) { return x >= g.nb && x < g.nb + g.nx; };
    auto ldf = [&](int x) { return bldo<T>(rS, fxin(x) ? ofz : kOOB, fxin(x) ? (unsigned)((x - g.nb) * g.nz) * E : 0u); };
    // visco-acoustic fields (f32 only)
    const auto rTa = win(VISCO ? a.tau - bias : nullptr, VISCO ? wb / sizeof(T) * 4 : 0u);
    const auto rTs = win(VISCO ? a.taus - bias : nullptr, VISCO ? wb / sizeof(T) * 4 : 0u);
    const auto rR1 = win(VISCO ? a.r1 + slot * g.fs - bias : nullptr, VISCO ? wb / sizeof(T) * 4 : 0u);
    const auto rR2 = win(VISCO ? a.r2 + slot * g.fs - bias : nullptr, VISCO ? wb / sizeof(T) * 4 : 0u);
    const unsigned oz4 = zok ? (unsigned)(bias + z) * 4 : kOOB, oc4 = (unsigned)L * 4;
    auto ldv = [&](__amdgpu_buffer_rsrc_t r, int x, bool ok) {
        return bldo<float>(r, ok ? oz4 : kOOB, ok ? (unsigned)x * oc4 : 0u);
    };
    // operands of one column: its previous level, model, z halo, passive source and visco
    // fields, sponge pair, and the queue entry four columns ahead; fetched one column early
    struct Col {
        T q, y, m, h, f, xl, xh;
        float ta, ts, r1;
    };
    auto fetch = [&](int xx) -> Col {
        const bool v = xx < x1;
        Col c;
        c.q = ldc(rX, oz, xx + 4, v && xx + 4 < g.pnx);
        c.y = ldc(rY, oz, xx, v);
        c.m = ldc(rM, oz, xx, v);
        c.h = ldc(rX, ozh, xx, v);
        c.f = (v && a.fsrc) ? ldf(xx) : T(0);
        c.ta = 0.f; c.ts = 1.f; c.r1 = 0.f;
        if constexpr (VISCO) {
            c.ta = ldv(rTa, xx, v);
            c.ts = ldv(rTs, xx, v);
            c.r1 = ldv(rR1, xx, v);
        }
        const int xs = xx < g.pnx ? xx : g.pnx - 1;
        c.xl = a.sfx[2 * xs];
        c.xh = a.sfx[2 * xs + 1];
        return c;
    };
    // tile-uniform flags: the per-lane work below runs only in tiles that hold the receiver row,
    // snapshot rows or source rows, behind scalar branches
    const bool rec_rows = a.rec && g.rz >= z0 && g.rz < z0 + 64;
    const bool snap_rows = a.snap && z0 < g.nb + g.nz && z0 + 64 > g.nb;
    const bool src_rows = sh.ph > 0 && sh.sz0 < z0 + 64 && sh.sz0 + sh.ph > z0;
    const bool fsrc_rows = a.fsrc && z0 < g.nb + g.nz && z0 + 64 > g.nb;
    // Lean tiles: full-width columns inside [4, pnx-4), no source-map cell, no passive-source
    // rows, and receivers (if any) at every grid point.  Every fetch is in range (the
    // last column fetches nothing), the Laplacian applies on every column, the tile's sponge pairs
    // are staged in LDS and the loop is unrolled, so the per-column work carries no scalar
    // branches.  The snapshot and receiver stores, the optional parts, are buffer stores whose
    // out-of-range offset drops them off the model interior and off the receiver line.
    const bool src_tile = src_rows && sh.sx0 < x0 + XC && sh.sx0 + sh.pw > x0;
    if (!src_tile && !fsrc_rows && !(a.full && snap_rows) && (!rec_rows || g.dr_v == 1) && x0 >= 4 &&
        x0 + XC + 4 <= g.pnx) {
        __shared__ T sxp[2 * XC];
        if (lane < 2 * XC) sxp[lane] = a.sfx[2 * x0 + lane];
        auto lean = [&](auto snap_c, auto rec_c, auto term_c) {
            constexpr bool SNAP = decltype(snap_c)::value, REC = decltype(rec_c)::value;
            constexpr bool TERM = !VISCO && decltype(term_c)::value;
            const auto rTo = win(TERM ? a.full + slot * a.fulls - bias : nullptr, TERM ? wb : 0u);
            // receiver k = r02 + (x - r0_v) records lane z == rz (wave_fwigradient.c:136-139)
            const auto rRc = win(REC ? a.rec + slot * a.trs + (long long)a.it * g.nr : nullptr,
                                 REC ? (unsigned)(g.nr * E) : 0u);
            const unsigned orz = z == g.rz ? 0u : kOOB;
            const bool szin = z >= g.nb && z < g.nb + g.nz;
            const auto rSn = win(SNAP ? a.snap + slot * a.snaps : nullptr,
                                 SNAP ? (unsigned)((long long)g.nz * g.nx * E) : 0u);
            const unsigned osn = szin ? (unsigned)(z - g.nb) * E : kOOB;
            auto get = [&](int xx) -> Col {
                Col c;
                c.q = bldo<T>(rX, oz, (unsigned)(xx + 4) * oc);
                c.y = bldo<T>(rY, oz, (unsigned)xx * oc);
                c.m = bldo<T>(rM, oz, (unsigned)xx * oc);
                c.h = bldo<T>(rX, ozh, (unsigned)xx * oc);
                c.f = T(0);
                c.ta = 0.f; c.ts = 1.f; c.r1 = 0.f;
                if constexpr (VISCO) {
                    c.ta = bldo<float>(rTa, oz4, (unsigned)xx * oc4);
                    c.ts = bldo<float>(rTs, oz4, (unsigned)xx * oc4);
                    c.r1 = bldo<float>(rR1, oz4, (unsigned)xx * oc4);
                }
                c.xl = sxp[2 * (xx - x0)];
                c.xh = sxp[2 * (xx - x0) + 1];
                return c;
            };
            T q[9];
#pragma unroll
            for (int k = 0; k < 8; ++k) q[k] = bldo<T>(rX, oz, (unsigned)(x0 - 4 + k) * oc);
            __syncthreads();
            Col pend = get(x0);
#pragma unroll
            for (int i = 0; i < XC; ++i) {
                const int x = x0 + i;
                const Col cur = pend;
                if (i + 1 < XC) pend = get(x + 1);
                q[8] = cur.q;
                row[i & 1][lane + 4] = q[4];
                row[i & 1][lzh] = cur.h;
                __syncthreads();
                const T p1 = q[4];
                if constexpr (SNAP) {
                    const bool sx = x >= g.nb && x < g.nb + g.nx;
                    bsto(rSn, sx ? osn : kOOB, sx ? (unsigned)((x - g.nb) * g.nz) * E : 0u, p1);
                }
                if constexpr (REC) {
                    const int d = x - sh.r0_v, k = sh.r02 + d;
                    const bool rx = d >= 0 && d < sh.nr2 && k >= 0 && k < g.nr;
                    bsto(rRc, rx ? orz : kOOB, rx ? (unsigned)k * E : 0u, p1);
                }
                const T* R = &row[i & 1][lane + 4];
                const T lap = (C::K0 * p1 + C::K1 * (q[5] + q[3]) + C::K2 * (q[6] + q[2]) + C::K3 * (q[7] + q[1]) +
                               C::K4 * (q[8] + q[0])) / g.dx2 +
                              (C::K0 * p1 + C::K1 * (R[1] + R[-1]) + C::K2 * (R[2] + R[-2]) +
                               C::K3 * (R[3] + R[-3]) + C::K4 * (R[4] + R[-4])) / g.dz2;
                T term = keep_if(lap, zin);
                if constexpr (TERM) bsto_nt(rTo, oz, (unsigned)x * oc, term);  // read once, by the adjoint
                const T xl = cur.xl, xh = cur.xh;
                if constexpr (VISCO) {
                    const float tf = (float)term;
                    const float tav = cur.ta, tsv = cur.ts, r1v = cur.r1;
                    const double num = (double)(tav / tsv * tf) + ((double)a.idt - 0.5 / (double)tsv) * (double)r1v;
                    const float r2v = (float)(num / ((double)a.idt + 0.5 / (double)tsv));
                    const float sv = 0.f * (float)wv;
                    const double tt = (double)tf * (1. + (double)tav) - (double)(r2v + r1v) * 0.5;
                    term = (T)(float)(tt + (double)sv);
                    bsto(rR2, oz4, (unsigned)x * oc4, (float)sponge((T)r2v, zl, zh, xl, xh));
                }
                const T p0 = sponge(cur.y, zl, zh, xl, xh);
                T p2 = 2 * p1 - p0 + cur.m * term;
                bsto(rY, oz, (unsigned)x * oc, sponge(p2, zl, zh, xl, xh));
#pragma unroll
                for (int k = 0; k < 8; ++k) q[k] = q[k + 1];
            }
        };
        using Y_ = std::true_type;
        using N_ = std::false_type;
        if (a.full) {  // terms for the exact adjoint (never with snapshots: the general loop has both)
            if (rec_rows) lean(N_{}, Y_{}, Y_{});
            else lean(N_{}, N_{}, Y_{});
        } else if (rec_rows) {
            if (snap_rows) lean(Y_{}, Y_{}, N_{});
            else lean(N_{}, Y_{}, N_{});
        } else {
            if (snap_rows) lean(Y_{}, N_{}, N_{});
            else lean(N_{}, N_{}, N_{});
        }
        return;
    }
    T q[9];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int xx = x0 - 4 + k;
        q[k] = ldc(rX, oz, xx, xv(xx));
    }
    Col pend = fetch(x0);
    int buf = 0;
    for (int x = x0; x < x1; ++x) {
        const Col cur = pend;
        pend = fetch(x + 1);
        q[8] = cur.q;
        const T yv = cur.y, mv = cur.m, hz = cur.h, xl = cur.xl, xh = cur.xh;
        row[buf][lane + 4] = q[4];
        if (lane < 8) row[buf][lzh] = hz;
        __syncthreads();
        const T p1 = q[4];
        if (rec_rows && z == g.rz) {  // record before the update (wave_fwigradient.c:136-139)
            T* r = a.rec + slot * a.trs + (long long)a.it * g.nr;
            if (g.dr_v > 0) {
                int d = x - sh.r0_v;
                if (d >= 0 && d % g.dr_v == 0 && d / g.dr_v < sh.nr2) {
                    int k = sh.r02 + d / g.dr_v;
                    if (k >= 0 && k < g.nr) r[k] = p1;
                }
            } else if (x == sh.r0_v) {
                for (int ir = 0; ir < sh.nr2; ++ir)
                    if (sh.r02 + ir >= 0 && sh.r02 + ir < g.nr) r[sh.r02 + ir] = p1;
            }
        }
        if (snap_rows && x >= g.nb && x < g.nb + g.nx && z >= g.nb && z < g.nb + g.nz)
            a.snap[slot * a.snaps + (long long)(x - g.nb) * g.nz + (z - g.nb)] = p1;
        // the Laplacian on [4, pad-4); formed on every lane (its operands are in range) and
        // selected
        const T* R = &row[buf][lane + 4];
        const T lap = (C::K0 * p1 + C::K1 * (q[5] + q[3]) + C::K2 * (q[6] + q[2]) + C::K3 * (q[7] + q[1]) +
                       C::K4 * (q[8] + q[0])) / g.dx2 +
                      (C::K0 * p1 + C::K1 * (R[1] + R[-1]) + C::K2 * (R[2] + R[-2]) + C::K3 * (R[3] + R[-3]) +
                       C::K4 * (R[4] + R[-4])) / g.dz2;
        T term = (zin && x >= 4 && x < g.pnx - 4) ? lap : T(0);
        bool inp = false;
        int px = 0, pz = 0;
        if (src_rows && x >= sh.sx0 && x < sh.sx0 + sh.pw) {
            px = x - sh.sx0;
            pz = z - sh.sz0;
            inp = pz >= 0 && pz < sh.ph;
        }
        if constexpr (VISCO) {
            // forward_modeling (wave_fwimodeling.c:186-192): r2 and the source term formed in
            // double where the reference's expressions promote, stored in f32
            const float tf = (float)term;
            const float tav = cur.ta, tsv = cur.ts, r1v = cur.r1;
            const double num = (double)(tav / tsv * tf) + ((double)a.idt - 0.5 / (double)tsv) * (double)r1v;
            const float r2v = (float)(num / ((double)a.idt + 0.5 / (double)tsv));
            const float sv = (inp ? (float)a.patch[slot * PMAX * PMAX + px * PMAX + pz] : 0.f) * (float)wv;
            const double tt = (double)tf * (1. + (double)tav) - (double)(r2v + r1v) * 0.5;
            term = (T)(float)(tt + (double)sv);
            bsto(rR2, oz4, (unsigned)x * oc4, (float)sponge((T)r2v, zl, zh, xl, xh));
        } else {
            if (inp) term = term + a.patch[slot * PMAX * PMAX + px * PMAX + pz] * wv;
        }
        if (fsrc_rows && fzin && fxin(x)) term = term + cur.f;
        if constexpr (!VISCO)
            if (a.full && zok) __builtin_nontemporal_store(term, &a.full[slot * a.fulls + (long long)x * g.ldz + z]);
        const T p0 = sponge(yv, zl, zh, xl, xh);
        T p2 = 2 * p1 - p0 + mv * term;
        bsto(rY, oz, (unsigned)x * oc, sponge(p2, zl, zh, xl, xh));
#pragma unroll
        for (int k = 0; k < 8; ++k) q[k] = q[k + 1];
        buf ^= 1;
    }
}

// ---------------------------------------------------------------- two forward steps per launch
// Temporal blocking of k_fwd for batched sweeps: one launch advances (X_t, X_{t-1}) to
// (X_{t+2}, X_{t+1}), written to two other buffers (neighbouring tiles still read X_t and X_{t-1}).
// A tile owns 56 z rows x W columns.  Its wave's 64 lanes hold the owned rows and 4 rows on either
// side; streaming along x, it forms step 1 (X_{t+1}) on all 64 lanes at column c, four columns
// ahead of step 2 (X_{t+2}) at column c - 4, which takes its x neighbours from a 9-column register
// queue of step-1 values and its z neighbours from LDS.  Step 1 therefore covers the tile grown by
// 4 cells on every side (recomputed by the neighbouring tiles rather than exchanged).  Every cell
// of both steps uses k_fwd's arithmetic in the same order (records before the update, the second
// sponge of the previous level on read), so the results equal two single-step launches bit for
// bit.  HBM traffic per two steps: X_t, X_{t-1} and v^2 dt^2 read once (plus halos), X_{t+1} and
// X_{t+2} written once: 20 B per cell for two updates against 32 B for two single steps.
constexpr int kOwnZ = 56;  // owned z rows of a two-step tile

template <typename T>
struct Fwd2Args {
    const T* X;        // X_t (sponged)
    const T* Y;        // X_{t-1} as stored
    T* C;              // out: X_{t+1}
    T* D;              // out: X_{t+2}
    const T* m;        // (v*v)*dt^2 on the padded grid
    const T* sfz;
    const T* sfx;
    const ShotInfo* shots;
    const T* patch;
    const T* wav;
    T* rec;            // traces [slot][nt][nr] or null
    long long trs;
    T* snap1;          // interior snapshots of X_t and X_{t+1} [slot][nx][nz], or null
    T* snap2;
    long long snaps;
    T* term1;          // the terms of steps t and t+1 on the full field [slot][fs], or null
    T* term2;
    long long terms;
    int it;            // step t; the launch makes steps t and t+1
};

template <typename T, int W>
__global__ __launch_bounds__(64) void k_fwd2(FGeo<T> g, Fwd2Args<T> a, TileMap tm)
{
    using C = Coef<T>;
    int slot, tz, xc;
    if (!tile_of(tm, slot, tz, xc)) return;
    const int lane = threadIdx.x;
    const int zo = tz * kOwnZ;   // owned rows [zo, zo + 56)
    const int z0 = zo - 4;       // row of lane 0
    const int z = z0 + lane;
    const int x0 = xc * W;       // owned columns [x0, x0 + W) within the grid
    __shared__ T rowA[2][128];   // X_t column with its z halo (lanes 8..63 may write 72..127)
    __shared__ T rowB[2][72];    // X_{t+1} column, lane l at l + 4
    const ShotInfo sh = a.shots[slot];
    const int L = g.ldz;
    const int bias = 4 * L + 4;
    const unsigned wb = (unsigned)((g.fs + 2 * bias) * (long long)sizeof(T));
    constexpr unsigned E = sizeof(T);
    const auto rX = win(a.X + slot * g.fs - bias, wb), rY = win(a.Y + slot * g.fs - bias, wb),
               rM = win(a.m - bias, wb), rC = win(a.C + slot * g.fs - bias, wb),
               rD = win(a.D + slot * g.fs - bias, wb);
    const auto rT1 = win(a.term1 ? a.term1 + slot * a.terms - bias : nullptr, a.term1 ? wb : 0u);
    const auto rT2 = win(a.term2 ? a.term2 + slot * a.terms - bias : nullptr, a.term2 ? wb : 0u);
    const bool zok = z >= 0 && z < g.pnz;
    const bool zin = z >= 4 && z < g.pnz - 4;
    const bool own = zok && lane >= 4 && lane < 4 + kOwnZ;
    const T zl = zok ? a.sfz[2 * z] : T(1), zh = zok ? a.sfz[2 * z + 1] : T(1);
    const int zhz = lane < 4 ? z0 - 4 + lane : z0 + 60 + lane;  // z halo of lanes 0..7
    const int lzh = lane < 4 ? lane : 64 + lane;
    const bool zhv = lane < 8 && zhz >= 0 && zhz < g.pnz;
    const unsigned oz = zok ? (unsigned)(bias + z) * E : kOOB, ozh = zhv ? (unsigned)(bias + zhz) * E : kOOB;
    const unsigned oo = own ? oz : kOOB;  // stores: owned rows only
    const unsigned oc = (unsigned)L * E;
    auto xv = [&](int x) { return x >= 0 && x < g.pnx; };
    auto ld = [&](__amdgpu_buffer_rsrc_t r, unsigned voff, int x) {  // column x of a field (0 off the grid)
        return bldo<T>(r, xv(x) ? voff : kOOB, xv(x) ? (unsigned)x * oc : 0u);
    };
    const T wv1 = a.wav[a.it], wv2 = a.wav[a.it + 1];
    const bool rec_rows = a.rec && g.rz >= zo && g.rz < zo + kOwnZ;
    const bool src_rows = sh.ph > 0 && sh.sz0 < z0 + 64 && sh.sz0 + sh.ph > z0;
    const bool snap_rows = (a.snap1 || a.snap2) && zo < g.nb + g.nz && zo + kOwnZ > g.nb;
    auto source = [&](int x, T wv) -> T {  // source-map term of cell (z, x)
        const int px = x - sh.sx0, pz = z - sh.sz0;
        return (px >= 0 && px < sh.pw && pz >= 0 && pz < sh.ph) ? a.patch[slot * PMAX * PMAX + px * PMAX + pz] * wv
                                                                : T(0);
    };
    // records before the update (wave_fwigradient.c:136-139), snapshots of the interior: owned cells
    auto record = [&](int x, int it, T p1) {
        if (!(rec_rows && own && z == g.rz)) return;
        T* r = a.rec + slot * a.trs + (long long)it * g.nr;
        if (g.dr_v > 0) {
            const int d = x - sh.r0_v;
            if (d >= 0 && d % g.dr_v == 0 && d / g.dr_v < sh.nr2) {
                const int k = sh.r02 + d / g.dr_v;
                if (k >= 0 && k < g.nr) r[k] = p1;
            }
        } else if (x == sh.r0_v) {
            for (int ir = 0; ir < sh.nr2; ++ir)
                if (sh.r02 + ir >= 0 && sh.r02 + ir < g.nr) r[sh.r02 + ir] = p1;
        }
    };
    auto snapshot = [&](T* snap, int x, T p1) {
        if (snap && snap_rows && own && x >= g.nb && x < g.nb + g.nx && z >= g.nb && z < g.nb + g.nz)
            snap[slot * a.snaps + (long long)(x - g.nb) * g.nz + (z - g.nb)] = p1;
    };
    // X_t queue: columns c-4 .. c+4 while step 1 runs at column c
    T qx[9], q1[9], mq[5];
#pragma unroll
    for (int k = 0; k < 8; ++k) qx[k] = ld(rX, oz, x0 - 8 + k);
#pragma unroll
    for (int k = 0; k < 9; ++k) q1[k] = T(0);
#pragma unroll
    for (int k = 0; k < 5; ++k) mq[k] = T(0);
    struct Col {
        T q, h, y, m;
    };
    auto fetch = [&](int c) -> Col {  // the operands step 1 needs at column c
        Col o;
        o.q = ld(rX, oz, c + 4);
        o.h = ld(rX, ozh, c);
        o.y = ld(rY, oz, c);
        o.m = ld(rM, oz, c);
        return o;
    };
    Col pend = fetch(x0 - 4);
#pragma unroll 1
    for (int i = 0; i < W + 8; ++i) {
        const int c = x0 - 4 + i;  // step-1 column
        const Col cur = pend;
        if (i + 1 < W + 8) pend = fetch(c + 1);
        qx[8] = cur.q;
        mq[4] = cur.m;
        // ---- step 1 at column c, every lane
        rowA[i & 1][lane + 4] = qx[4];
        rowA[i & 1][lzh] = cur.h;
        __syncthreads();
        const T p1 = qx[4];
        const bool cv = xv(c);
        const bool cown = c >= x0 && c < x0 + W && cv;
        const T* R = &rowA[i & 1][lane + 4];
        const T lap1 = (C::K0 * p1 + C::K1 * (qx[5] + qx[3]) + C::K2 * (qx[6] + qx[2]) + C::K3 * (qx[7] + qx[1]) +
                        C::K4 * (qx[8] + qx[0])) / g.dx2 +
                       (C::K0 * p1 + C::K1 * (R[1] + R[-1]) + C::K2 * (R[2] + R[-2]) + C::K3 * (R[3] + R[-3]) +
                        C::K4 * (R[4] + R[-4])) / g.dz2;
        T term1 = (zin && c >= 4 && c < g.pnx - 4) ? lap1 : T(0);
        if (src_rows) term1 = term1 + source(c, wv1);
        const T cx = cv ? a.sfx[2 * c] : T(1), cxh = cv ? a.sfx[2 * c + 1] : T(1);
        if (cown) {
            record(c, a.it, p1);
            snapshot(a.snap1, c, p1);
            if (a.term1) bsto(rT1, oo, (unsigned)c * oc, term1);
        }
        const T p0 = sponge(cur.y, zl, zh, cx, cxh);
        const T p2 = 2 * p1 - p0 + cur.m * term1;
        const T x1 = (zok && cv) ? sponge(p2, zl, zh, cx, cxh) : T(0);
        if (cown) bsto(rC, oo, (unsigned)c * oc, x1);
        q1[8] = x1;
        // ---- step 2 at column s = c - 4, once step 1 has covered s - 4 .. s + 4
        if (i >= 8) {
            const int s = c - 4;
            rowB[i & 1][lane + 4] = q1[4];
            __syncthreads();
            const T u = q1[4];
            const bool sv = xv(s);
            const bool sown = s < x0 + W && sv;
            const T* R2 = &rowB[i & 1][lane + 4];
            const T lap2 = (C::K0 * u + C::K1 * (q1[5] + q1[3]) + C::K2 * (q1[6] + q1[2]) + C::K3 * (q1[7] + q1[1]) +
                            C::K4 * (q1[8] + q1[0])) / g.dx2 +
                           (C::K0 * u + C::K1 * (R2[1] + R2[-1]) + C::K2 * (R2[2] + R2[-2]) +
                            C::K3 * (R2[3] + R2[-3]) + C::K4 * (R2[4] + R2[-4])) / g.dz2;
            T term2 = (zin && s >= 4 && s < g.pnx - 4) ? lap2 : T(0);
            if (src_rows) term2 = term2 + source(s, wv2);
            if (sown) {
                record(s, a.it + 1, u);
                snapshot(a.snap2, s, u);
                if (a.term2) bsto(rT2, oo, (unsigned)s * oc, term2);
            }
            const T sx = sv ? a.sfx[2 * s] : T(1), sxh = sv ? a.sfx[2 * s + 1] : T(1);
            const T p0b = sponge(qx[0], zl, zh, sx, sxh);  // X_t at s, its second sponge
            const T p2b = 2 * u - p0b + mq[0] * term2;
            if (sown) bsto(rD, oo, (unsigned)s * oc, sponge(p2b, zl, zh, sx, sxh));
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) { qx[k] = qx[k + 1]; q1[k] = q1[k + 1]; }
#pragma unroll
        for (int k = 0; k < 4; ++k) mq[k] = mq[k + 1];
        __builtin_amdgcn_sched_barrier(0);  // keep each column's loads one column ahead, no further
    }
}

// ---------------------------------------------------------------- reference imaging-condition adjoint
template <typename T>
struct ImgArgs {
    const T* w_prev;  // snapshots at it-I, it, it+I: [slot][nx][nz] (null when no imaging this step)
    const T* w_cur;
    const T* w_next;
    long long snaps;
    const float* tmpv;  // -2/v^3 per interior cell (f32 as the reference)
    T* grad;            // accumulator [slot][nx][nz] (slot stride gstride; 0 = shared)
    long long gstride;
};

// ---------------------------------------------------------------- visco-acoustic imaging adjoint
// gradient_v's backward step (pywave/src/wave_fwigradient.c:372-420): on [4, pad-4) the memory
// variable r2 = ((-tau)/taus p1 + (-idt + 0.5/taus) r1) / (-idt - 0.5/taus) and the Laplacian's
// argument tmp = p1 (1 + tau) - 0.5 (r2 + r1), both in double where C promotes and stored in f32;
// zero outside.  r2 is damped once, the pressure levels as in gradient_av.
struct ViscoBack {
    const float* tau;  // padded [fs]
    const float* taus;
    const float* r1;   // backward memory variable [slot][fs], read
    float* r2;         // written (damped)
    float idt;
};

__device__ __forceinline__ float visco_back(float p1, float r1, float ta, float ts, float idt, float& r2)
{
    const double num = (double)(-ta / ts * p1) + (-(double)idt + 0.5 / (double)ts) * (double)r1;
    r2 = (float)(num / (-(double)idt - 0.5 / (double)ts));
    return (float)((double)p1 * (1. + (double)ta) - 0.5 * (double)(r2 + r1));
}

// The reference imaging-condition adjoint step, pipelined as k_fwd.  gradient_av
// (pywave/src/wave_fwigradient.c:201-241): term = L p1 + residual at the receivers,
// p2 = 2 p1 - p0 + v^2 dt^2 term, and at imaging steps grad += (u[w+1] - 2u[w] + u[w-1])/wdt2 * p1
// * (-2/v^3) in double.  VISCO: gradient_v (:372-420), the Laplacian of tmp instead of p1 and the
// memory variable r2 written.  The column entering the 9-column queue (its p1, and for VISCO r1,
// tau, taus, forming tmp and r2 once), the z halo, Y, m and the imaging operands of the next
// column are loaded one iteration ahead through buffer resources.
template <bool VISCO, int W>
__global__ __launch_bounds__(64) void k_adj_ref(FGeo<float> g, StepArgs<float> a, ImgArgs<float> im, ViscoBack vb,
                                                TileMap tm)
{
    using C = Coef<float>;
    int slot, tz, xc;
    if (!tile_of(tm, slot, tz, xc)) return;
    const int lane = threadIdx.x, z0 = tz * 64, z = z0 + lane;
    const int x0 = xc * W, x1 = min(x0 + W, g.pnx);
    __shared__ float row[2][128];  // z row, halo, spare slots for lanes 8..63
    const ShotInfo sh = a.shots[slot];
    const int L = g.ldz;
    const int bias = 4 * L + 4;  // windows start 4 columns + 4 cells before the field
    const unsigned wb = (unsigned)((g.fs + 2 * bias) * 4LL);
    const auto rX = win(a.X + slot * g.fs - bias, wb), rY = win(a.Y + slot * g.fs - bias, wb),
               rM = win(a.m - bias, wb);
    const auto rR1 = win(VISCO ? vb.r1 + slot * g.fs - bias : nullptr, VISCO ? wb : 0u);
    const auto rR2 = win(VISCO ? vb.r2 + slot * g.fs - bias : nullptr, VISCO ? wb : 0u);
    const auto rTa = win(VISCO ? vb.tau - bias : nullptr, VISCO ? wb : 0u);
    const auto rTs = win(VISCO ? vb.taus - bias : nullptr, VISCO ? wb : 0u);
    const bool zok = z < g.pnz;
    const bool zin = z >= 4 && z < g.pnz - 4;
    const float zl = zok ? a.sfz[2 * z] : 1.f, zh = zok ? a.sfz[2 * z + 1] : 1.f;
    const int zhz = lane < 4 ? z0 - 4 + lane : z0 + 60 + lane;  // z halo of lanes 0..7
    const int lzh = lane < 4 ? lane : 64 + lane;
    const bool zhv = lane < 8 && zhz >= 0 && zhz < g.pnz;
    const bool zhin = zhv && zhz >= 4 && zhz < g.pnz - 4;
    // lane offsets (out-of-range offsets read 0 and drop stores): every lane of the grid, lanes
    // where tmp is formed (z in [4, pad-4)), halo lanes (VISCO: where tmp is formed)
    const unsigned oz = zok ? (unsigned)(bias + z) * 4u : kOOB;
    const unsigned ozi = (zok && zin) ? oz : kOOB;
    const unsigned ozh = (VISCO ? zhin : zhv) ? (unsigned)(bias + zhz) * 4u : kOOB;
    const unsigned oc = (unsigned)L * 4u;
    auto xv = [&](int x) { return x >= 0 && x < g.pnx; };
    auto xin = [&](int x) { return x >= 4 && x < g.pnx - 4; };
    auto ld = [&](__amdgpu_buffer_rsrc_t r, unsigned voff, int x, bool ok) {
        return bldo<float>(r, ok ? voff : kOOB, ok ? (unsigned)x * oc : 0u);
    };
    const bool lin = zok && zin;
    // tmp and r2 of a cell from its raw values; zero where the reference leaves them unwritten
    auto form = [&](bool ok, float p1, float r1, float ta, float ts, float& r2) -> float {
        r2 = 0.f;
        return ok ? visco_back(p1, r1, ta, ts, vb.idt, r2) : 0.f;
    };
    // imaging operands on the model interior [slot][nx][nz] (windows empty when this step does
    // not image)
    const bool img = im.w_cur != nullptr;
    const unsigned nzx4 = img ? (unsigned)(g.nz * g.nx * 4) : 0u;
    const auto rWp = win(img ? im.w_prev + slot * im.snaps : nullptr, nzx4);
    const auto rWc = win(img ? im.w_cur + slot * im.snaps : nullptr, nzx4);
    const auto rWn = win(img ? im.w_next + slot * im.snaps : nullptr, nzx4);
    const auto rTv = win(img ? im.tmpv : nullptr, nzx4);
    const auto rG = win(img ? im.grad + slot * im.gstride : nullptr, nzx4);
    const bool gz = zok && z >= g.nb && z < g.nb + g.nz;
    const unsigned og = gz ? (unsigned)(z - g.nb) * 4u : kOOB;
    auto gxin = [&](int x) { return x >= g.nb && x < g.nb + g.nx; };
    auto ldi = [&](__amdgpu_buffer_rsrc_t r, int x) {
        const bool ok = img && gxin(x);
        return bldo<float>(r, ok ? og : kOOB, ok ? (unsigned)((x - g.nb) * g.nz) * 4u : 0u);
    };
    const float* inj = a.inj + slot * a.trs + (long long)a.it * g.nr;
    // Lean tiles (acoustic): full-width columns inside [4, pnx-4) and receivers (if the tile holds
    // the receiver row) at every grid point, as in k_fwd: in-range fetches, LDS sponge pairs, an
    // unrolled loop without scalar branches; the imaging update is a buffer store dropped outside
    // the model interior, the receiver sample a load answered with 0 off the receiver line.
    if constexpr (!VISCO) {
        const bool rec_rows = g.rz >= z0 && g.rz < z0 + 64;
        if ((!rec_rows || g.dr_v == 1) && x0 >= 4 && x0 + W + 4 <= g.pnx) {
            __shared__ float sxp[2 * W];
            if (lane < 2 * W) sxp[lane] = a.sfx[2 * x0 + lane];
            auto lean = [&](auto img_c, auto rec_c) {
                constexpr bool IMG = decltype(img_c)::value, REC = decltype(rec_c)::value;
                const auto rIj = win(REC ? inj : nullptr, REC ? (unsigned)(g.nr * 4) : 0u);
                const bool rl = z == g.rz;
                struct ACol {
                    float q, h, y, m, wp, wc, wn, tv, gv, xl, xh, ij;
                };
                // every operand of a column, the receiver sample included, is fetched one column
                // ahead: a load issued after them would make each column wait for all of them
                auto get = [&](int xx) -> ACol {
                    ACol c;
                    c.q = bldo<float>(rX, oz, (unsigned)(xx + 4) * oc);
                    c.h = bldo<float>(rX, ozh, (unsigned)xx * oc);
                    c.y = bldo<float>(rY, oz, (unsigned)xx * oc);
                    c.m = bldo<float>(rM, oz, (unsigned)xx * oc);
                    c.wp = c.wc = c.wn = c.tv = c.gv = c.ij = 0.f;
                    if constexpr (REC) {
                        const int d = xx - sh.r0_v, kk = sh.r02 + d;
                        const bool rx = d >= 0 && d < sh.nr2 && kk >= 0 && kk < g.nr;
                        c.ij = bldo<float>(rIj, (rx && rl) ? 0u : kOOB, rx ? (unsigned)kk * 4u : 0u);
                    }
                    if constexpr (IMG) {
                        const bool ok = gxin(xx);
                        const unsigned vo = ok ? og : kOOB, so = ok ? (unsigned)((xx - g.nb) * g.nz) * 4u : 0u;
                        c.wp = bldo<float>(rWp, vo, so);
                        c.wc = bldo<float>(rWc, vo, so);
                        c.wn = bldo<float>(rWn, vo, so);
                        c.tv = bldo<float>(rTv, vo, so);
                        c.gv = bldo<float>(rG, vo, so);
                    }
                    c.xl = sxp[2 * (xx - x0)];
                    c.xh = sxp[2 * (xx - x0) + 1];
                    return c;
                };
                float q[9];
#pragma unroll
                for (int k = 0; k < 8; ++k) q[k] = bldo<float>(rX, oz, (unsigned)(x0 - 4 + k) * oc);
                __syncthreads();
                ACol pend = get(x0);
#pragma unroll
                for (int i = 0; i < W; ++i) {
                    const int x = x0 + i;
                    const ACol cur = pend;
                    if (i + 1 < W) pend = get(x + 1);
                    q[8] = cur.q;
                    row[i & 1][lane + 4] = q[4];
                    row[i & 1][lzh] = cur.h;
                    __syncthreads();
                    const float p1 = q[4];
                    const float* R = &row[i & 1][lane + 4];
                    const float lap = (C::K0 * p1 + C::K1 * (q[5] + q[3]) + C::K2 * (q[6] + q[2]) +
                                       C::K3 * (q[7] + q[1]) + C::K4 * (q[8] + q[0])) / g.dx2 +
                                      (C::K0 * p1 + C::K1 * (R[1] + R[-1]) + C::K2 * (R[2] + R[-2]) +
                                       C::K3 * (R[3] + R[-3]) + C::K4 * (R[4] + R[-4])) / g.dz2;
                    float term = keep_if(lap, zin);
                    if constexpr (REC) {  // the windowed residual at the receivers (:208-211)
                        const int d = x - sh.r0_v, kk = sh.r02 + d;
                        const bool rx = d >= 0 && d < sh.nr2 && kk >= 0 && kk < g.nr;
                        term = (rx && rl) ? term + cur.ij : term;
                    }
                    const float p0 = sponge(cur.y, zl, zh, cur.xl, cur.xh);
                    const float p2 = 2 * p1 - p0 + cur.m * term;
                    bsto(rY, oz, (unsigned)x * oc, sponge(p2, zl, zh, cur.xl, cur.xh));
                    if constexpr (IMG) {
                        const bool ok = gxin(x);
                        const double e = ((double)cur.wn - 2. * (double)cur.wc + (double)cur.wp) / g.wdt2 *
                                         (double)p1 * (double)cur.tv;
                        bsto(rG, ok ? og : kOOB, ok ? (unsigned)((x - g.nb) * g.nz) * 4u : 0u,
                             (float)((double)cur.gv + e));
                    }
#pragma unroll
                    for (int k = 0; k < 8; ++k) q[k] = q[k + 1];
                }
            };
            if (rec_rows) {
                if (img) lean(std::true_type{}, std::true_type{});
                else lean(std::false_type{}, std::true_type{});
            } else {
                if (img) lean(std::true_type{}, std::false_type{});
                else lean(std::false_type{}, std::false_type{});
            }
            return;
        }
    }
    // Lean tiles (visco-acoustic): every column of the stencil reach inside [4, pnx-4) (x0 >= 8,
    // x0 + W + 8 <= pnx), so tmp and r2 are formed on every lane of the Laplacian rows without
    // column masks; receivers and imaging as in the acoustic lean loop.
    if constexpr (VISCO) {
        const bool rec_rows = g.rz >= z0 && g.rz < z0 + 64;
        if ((!rec_rows || g.dr_v == 1) && x0 >= 8 && x0 + W + 8 <= g.pnx) {
            __shared__ float sxv[2 * W];
            if (lane < 2 * W) sxv[lane] = a.sfx[2 * x0 + lane];
            auto lean = [&](auto img_c, auto rec_c) {
                constexpr bool IMG = decltype(img_c)::value, REC = decltype(rec_c)::value;
                const auto rIj = win(REC ? inj : nullptr, REC ? (unsigned)(g.nr * 4) : 0u);
                const bool rl = z == g.rz;
                struct VCol {
                    float p5, r5, ta5, ts5, hp, hr, hta, hts, y, m, wp, wc, wn, tv, gv, xl, xh, ij;
                };
                // operands of column xx (and the queue entry xx + 4), the receiver sample included
                auto get = [&](int xx) -> VCol {
                    VCol c;
                    const unsigned s4 = (unsigned)(xx + 4) * oc, s0 = (unsigned)xx * oc;
                    c.p5 = bldo<float>(rX, oz, s4);
                    c.r5 = bldo<float>(rR1, ozi, s4);
                    c.ta5 = bldo<float>(rTa, ozi, s4);
                    c.ts5 = bldo<float>(rTs, ozi, s4);
                    c.hp = bldo<float>(rX, ozh, s0);
                    c.hr = bldo<float>(rR1, ozh, s0);
                    c.hta = bldo<float>(rTa, ozh, s0);
                    c.hts = bldo<float>(rTs, ozh, s0);
                    c.y = bldo<float>(rY, oz, s0);
                    c.m = bldo<float>(rM, oz, s0);
                    c.wp = c.wc = c.wn = c.tv = c.gv = c.ij = 0.f;
                    if constexpr (REC) {
                        const int d = xx - sh.r0_v, kk = sh.r02 + d;
                        const bool rx = d >= 0 && d < sh.nr2 && kk >= 0 && kk < g.nr;
                        c.ij = bldo<float>(rIj, (rx && rl) ? 0u : kOOB, rx ? (unsigned)kk * 4u : 0u);
                    }
                    if constexpr (IMG) {
                        const bool ok = gxin(xx);
                        const unsigned vo = ok ? og : kOOB, so = ok ? (unsigned)((xx - g.nb) * g.nz) * 4u : 0u;
                        c.wp = bldo<float>(rWp, vo, so);
                        c.wc = bldo<float>(rWc, vo, so);
                        c.wn = bldo<float>(rWn, vo, so);
                        c.tv = bldo<float>(rTv, vo, so);
                        c.gv = bldo<float>(rG, vo, so);
                    }
                    c.xl = sxv[2 * (xx - x0)];
                    c.xh = sxv[2 * (xx - x0) + 1];
                    return c;
                };
                float tq[9], pq[5], rq[5];
                // the queue's first eight columns: every load issued before the first use
                float p0[8], r0[8], ta0[8], ts0[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const unsigned so = (unsigned)(x0 - 4 + k) * oc;
                    p0[k] = bldo<float>(rX, oz, so);
                    r0[k] = bldo<float>(rR1, ozi, so);
                    ta0[k] = bldo<float>(rTa, ozi, so);
                    ts0[k] = bldo<float>(rTs, ozi, so);
                }
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    float r2;
                    tq[k] = form(lin, p0[k], r0[k], ta0[k], ts0[k], r2);
                    if (k >= 4) {
                        pq[k - 4] = p0[k];
                        rq[k - 4] = r2;
                    }
                }
                __syncthreads();
                VCol pend = get(x0);
#pragma unroll
                for (int i = 0; i < W; ++i) {
                    const int x = x0 + i;
                    const VCol cur = pend;
                    if (i + 1 < W) pend = get(x + 1);
                    float r2n;
                    tq[8] = form(lin, cur.p5, cur.r5, cur.ta5, cur.ts5, r2n);
                    pq[4] = cur.p5;
                    rq[4] = r2n;
                    float hr2;
                    const float ht = form(zhin, cur.hp, cur.hr, cur.hta, cur.hts, hr2);
                    row[i & 1][lane + 4] = tq[4];
                    row[i & 1][lzh] = ht;
                    __syncthreads();
                    const float p1 = pq[0];
                    const float* R = &row[i & 1][lane + 4];
                    const float lap = (C::K0 * tq[4] + C::K1 * (tq[5] + tq[3]) + C::K2 * (tq[6] + tq[2]) +
                                       C::K3 * (tq[7] + tq[1]) + C::K4 * (tq[8] + tq[0])) / g.dx2 +
                                      (C::K0 * tq[4] + C::K1 * (R[1] + R[-1]) + C::K2 * (R[2] + R[-2]) +
                                       C::K3 * (R[3] + R[-3]) + C::K4 * (R[4] + R[-4])) / g.dz2;
                    float term = keep_if(lap, zin);
                    if constexpr (REC) {
                        const int d = x - sh.r0_v, kk = sh.r02 + d;
                        const bool rx = d >= 0 && d < sh.nr2 && kk >= 0 && kk < g.nr;
                        term = (rx && rl) ? term + cur.ij : term;
                    }
                    const float p0 = sponge(cur.y, zl, zh, cur.xl, cur.xh);
                    const float p2 = 2 * p1 - p0 + cur.m * term;
                    bsto(rY, oz, (unsigned)x * oc, sponge(p2, zl, zh, cur.xl, cur.xh));
                    bsto(rR2, oz, (unsigned)x * oc, sponge(rq[0], zl, zh, cur.xl, cur.xh));
                    if constexpr (IMG) {
                        const bool ok = gxin(x);
                        const double e = ((double)cur.wn - 2. * (double)cur.wc + (double)cur.wp) / g.wdt2 *
                                         (double)p1 * (double)cur.tv;
                        bsto(rG, ok ? og : kOOB, ok ? (unsigned)((x - g.nb) * g.nz) * 4u : 0u,
                             (float)((double)cur.gv + e));
                    }
#pragma unroll
                    for (int k = 0; k < 8; ++k) tq[k] = tq[k + 1];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        pq[k] = pq[k + 1];
                        rq[k] = rq[k + 1];
                    }
                }
            };
            if (rec_rows) {
                if (img) lean(std::true_type{}, std::true_type{});
                else lean(std::false_type{}, std::true_type{});
            } else {
                if (img) lean(std::true_type{}, std::false_type{});
                else lean(std::false_type{}, std::false_type{});
            }
            return;
        }
    }
    float tq[9], pq[5], rq[5];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int xx = x0 - 4 + k;
        const float p1 = ld(rX, oz, xx, xv(xx));
        float r2 = 0.f;
        if constexpr (VISCO) {
            const bool ci = xin(xx);
            const float r1 = ld(rR1, ozi, xx, ci), ta = ld(rTa, ozi, xx, ci), ts = ld(rTs, ozi, xx, ci);
            tq[k] = form(lin && ci, p1, r1, ta, ts, r2);
        } else {
            tq[k] = p1;
        }
        if (k >= 4) {
            pq[k - 4] = p1;
            rq[k - 4] = r2;
        }
    }
    float hp, hr = 0.f, hta = 0.f, hts = 0.f;  // halo raw values of the centre column
    if constexpr (VISCO) {
        const bool ci = xin(x0);
        hp = ld(rX, ozh, x0, ci); hr = ld(rR1, ozh, x0, ci);
        hta = ld(rTa, ozh, x0, ci); hts = ld(rTs, ozh, x0, ci);
    } else {
        hp = ld(rX, ozh, x0, true);
    }
    float yv = ld(rY, oz, x0, true), mv = ld(rM, oz, x0, true);
    float wp = ldi(rWp, x0), wc = ldi(rWc, x0), wn = ldi(rWn, x0), tv = ldi(rTv, x0), gv = ldi(rG, x0);
    int buf = 0;
    for (int x = x0; x < x1; ++x) {
        const int xn = x + 1, xq = x + 5;
        const bool nv = xn < x1;
        // next iteration's loads
        const bool qv = nv && xq < g.pnx, qi = nv && xin(xq);
        const float np1 = ld(rX, oz, xq, qv);
        float nr1 = 0.f, nta = 0.f, nts = 0.f, nhr = 0.f, nhta = 0.f, nhts = 0.f, nhp;
        if constexpr (VISCO) {
            nr1 = ld(rR1, ozi, xq, qi); nta = ld(rTa, ozi, xq, qi); nts = ld(rTs, ozi, xq, qi);
            const bool hi = nv && xin(xn);
            nhp = ld(rX, ozh, xn, hi); nhr = ld(rR1, ozh, xn, hi);
            nhta = ld(rTa, ozh, xn, hi); nhts = ld(rTs, ozh, xn, hi);
        } else {
            nhp = ld(rX, ozh, xn, nv);
        }
        const float nyv = ld(rY, oz, xn, nv), nmv = ld(rM, oz, xn, nv);
        float nwp = 0.f, nwc = 0.f, nwn = 0.f, ntv = 0.f, ngv = 0.f;
        if (nv) {
            nwp = ldi(rWp, xn); nwc = ldi(rWc, xn); nwn = ldi(rWn, xn); ntv = ldi(rTv, xn); ngv = ldi(rG, xn);
        }
        float ht;
        if constexpr (VISCO) {
            float hr2;
            ht = form(zhin && xin(x), hp, hr, hta, hts, hr2);
        } else {
            ht = hp;
        }
        row[buf][lane + 4] = tq[4];
        if (lane < 8) row[buf][lzh] = ht;
        __syncthreads();
        if (zok) {
            const float p1 = pq[0];
            float term = 0.f;
            if (zin && xin(x)) {
                const float* R = &row[buf][lane + 4];
                term = (C::K0 * tq[4] + C::K1 * (tq[5] + tq[3]) + C::K2 * (tq[6] + tq[2]) + C::K3 * (tq[7] + tq[1]) +
                        C::K4 * (tq[8] + tq[0])) / g.dx2 +
                       (C::K0 * tq[4] + C::K1 * (R[1] + R[-1]) + C::K2 * (R[2] + R[-2]) + C::K3 * (R[3] + R[-3]) +
                        C::K4 * (R[4] + R[-4])) / g.dz2;
            }
            if (z == g.rz) {  // the windowed residual at the receivers (:208-211, :390-394)
                if (g.dr_v > 0) {
                    int d = x - sh.r0_v;
                    if (d >= 0 && d % g.dr_v == 0 && d / g.dr_v < sh.nr2) {
                        int kk = sh.r02 + d / g.dr_v;
                        if (kk >= 0 && kk < g.nr) term = term + inj[kk];
                    }
                } else if (x == sh.r0_v) {
                    for (int ir = 0; ir < sh.nr2; ++ir)
                        if (sh.r02 + ir >= 0 && sh.r02 + ir < g.nr) term = term + inj[sh.r02 + ir];
                }
            }
            const float xl = a.sfx[2 * x], xh = a.sfx[2 * x + 1];
            const float p0 = sponge(yv, zl, zh, xl, xh);
            const float p2 = 2 * p1 - p0 + mv * term;
            bsto(rY, oz, (unsigned)x * oc, sponge(p2, zl, zh, xl, xh));
            if constexpr (VISCO) bsto(rR2, oz, (unsigned)x * oc, sponge(rq[0], zl, zh, xl, xh));
            if (img && gz && gxin(x)) {
                const double e = ((double)wn - 2. * (double)wc + (double)wp) / g.wdt2 * (double)p1 * (double)tv;
                bsto(rG, og, (unsigned)((x - g.nb) * g.nz) * 4u, (float)((double)gv + e));
            }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) tq[k] = tq[k + 1];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            pq[k] = pq[k + 1];
            rq[k] = rq[k + 1];
        }
        if constexpr (VISCO) {
            float r2;
            tq[8] = form(lin && qi, np1, nr1, nta, nts, r2);
            rq[4] = r2;
        } else {
            tq[8] = np1;
        }
        pq[4] = np1;
        hp = nhp; hr = nhr; hta = nhta; hts = nhts;
        yv = nyv; mv = nmv;
        wp = nwp; wc = nwc; wn = nwn; tv = ntv; gv = ngv;
        buf ^= 1;
    }
}

// ---------------------------------------------------------------- exact discrete adjoint
// q_t = S(2 q_{t+1} - S q_{t+2} + L^T(m q_{t+1}) + R^T g_t);  dJ/dm += (L X_t + r w_t) q_{t+1}
template <typename T>
struct ExArgs {
    const T* F;      // the forward step's term (masked L X_t + sources) on the full field [slot][fs]
    long long fsl;   // slot stride of F
    T* gm;           // dJ/dm accumulator [slot][fs]
};

// Exact adjoint step (pipelined like k_fwd): q_t = S(2 q_{t+1} - S q_{t+2} + L^T(m q_{t+1}) + R^T g_t)
// with dJ/dm += term_t q_{t+1}, term_t = L X_t + source (what the forward step multiplied by m,
// stored by the forward sweep or its recomputation).  A = q_{t+1}, B = q_{t+2} (overwritten by
// q_t); the m*A product is zero outside the Laplacian's rows [4, n-4).  The adjoint has no
// counterpart in the reference whose rounding it must follow, so it contracts to FMAs and divides
// by the squared spacings through reciprocals.
template <typename T, int XC>
__global__ __launch_bounds__(64) void k_adj_exact(FGeo<T> g, StepArgs<T> a, ExArgs<T> ex, TileMap tm)
{
#pragma clang fp contract(fast)
    using C = Coef<T>;
    int slot, tz, xc;
    if (!tile_of(tm, slot, tz, xc)) return;
    const int lane = threadIdx.x, z0 = tz * 64, z = z0 + lane;
    const int x0 = xc * XC, x1 = min(x0 + XC, g.pnx);
    __shared__ T rowa[2][128];  // z rows, halo, spare slots for lanes 8..63
    const ShotInfo sh = a.shots[slot];
    const int L = g.ldz;
    const int bias = 4 * L + 4;
    const unsigned wb = (unsigned)((g.fs + 2 * bias) * (long long)sizeof(T));
    const auto rA = win(a.X + slot * g.fs - bias, wb), rB = win(a.Y + slot * g.fs - bias, wb),
               rM = win(a.m - bias, wb), rF = win(ex.F + slot * ex.fsl - bias, wb),
               rG = win(ex.gm + slot * g.fs - bias, wb);
    const T rdx = T(1) / g.dx2, rdz = T(1) / g.dz2;
    const bool zok = z < g.pnz;
    const bool zin = z >= 4 && z < g.pnz - 4;
    const T zl = zok ? a.sfz[2 * z] : T(1), zh = zok ? a.sfz[2 * z + 1] : T(1);
    const int zhz = lane < 4 ? z0 - 4 + lane : z0 + 60 + lane;  // z halo of lanes 0..7
    const int lzh = lane < 4 ? lane : 64 + lane;
    const bool zhv = lane < 8 && zhz >= 0 && zhz < g.pnz;
    const bool zhin = zhz >= 4 && zhz < g.pnz - 4;
    constexpr unsigned E = sizeof(T);
    // lane offsets fixed (out of range outside the grid / the Laplacian rows), column = scalar offset
    const unsigned oz = zok ? (unsigned)(bias + z) * E : kOOB, ozh = zhv ? (unsigned)(bias + zhz) * E : kOOB;
    const unsigned ozi = (zok && zin) ? oz : kOOB, ozhi = (zhv && zhin) ? ozh : kOOB;
    const unsigned oc = (unsigned)L * E;
    auto xv = [&](int x) { return x >= 0 && x < g.pnx; };
    auto xin = [&](int x) { return x >= 4 && x < g.pnx - 4; };
    auto ldc = [&](__amdgpu_buffer_rsrc_t r, unsigned voff, int x, bool ok) {  // column x of a field
        return bldo<T>(r, ok ? voff : kOOB, ok ? (unsigned)x * oc : 0u);
    };
    auto lap = [&](const T* q, const T* R) {
        return (C::K0 * q[4] + C::K1 * (q[5] + q[3]) + C::K2 * (q[6] + q[2]) + C::K3 * (q[7] + q[1]) +
                C::K4 * (q[8] + q[0])) * rdx +
               (C::K0 * R[0] + C::K1 * (R[1] + R[-1]) + C::K2 * (R[2] + R[-2]) + C::K3 * (R[3] + R[-3]) +
                C::K4 * (R[4] + R[-4])) * rdz;
    };
    // receiver injection: lane z == rz of column x takes trace sample k(x) (regular spread)
    const T* inj = a.inj + slot * a.trs + (long long)a.it * g.nr;
    const auto rI = win(inj, (unsigned)(g.nr * E));
    const bool rlane = z == g.rz;
    auto rk = [&](int x) -> int {
        if (!rlane || g.dr_v <= 0) return -1;
        const int d = x - sh.r0_v;
        if (d < 0 || d % g.dr_v != 0 || d / g.dr_v >= sh.nr2) return -1;
        const int k = sh.r02 + d / g.dr_v;
        return (k >= 0 && k < g.nr) ? k : -1;
    };
    const bool rec_rows = g.rz >= z0 && g.rz < z0 + 64;
    // Lean tiles: every column of the stencil reach inside [4, pnx-4) (x0 >= 8, x0 + XC + 8 <=
    // pnx) and receivers (if the tile holds their row) at every grid point: in-range fetches
    // without column masks, LDS sponge pairs and an unrolled loop without scalar branches.
    if ((!rec_rows || g.dr_v == 1) && x0 >= 8 && x0 + XC + 8 <= g.pnx) {
        __shared__ T sxp[2 * XC];
        if (lane < 2 * XC) sxp[lane] = a.sfx[2 * x0 + lane];
        const unsigned orz = rlane ? 0u : kOOB;
        struct XCol {
            T a5, m5, a, b, g, t, ha, xl, xh, s;
        };
        // REC: the tile holds the receiver row; R^T g_t, the receiver sample on lane z == rz (0
        // elsewhere and off the receivers), is fetched with the column's other operands, one
        // column ahead: a load issued after them would make every column wait for all of them
        auto lean = [&](auto rec) {
            constexpr bool REC = decltype(rec)::value;
            auto get = [&](int xx) -> XCol {
                XCol c;
                c.a5 = bldo<T>(rA, ozi, (unsigned)(xx + 4) * oc);
                c.m5 = bldo<T>(rM, ozi, (unsigned)(xx + 4) * oc);
                c.a = bldo<T>(rA, oz, (unsigned)xx * oc);
                c.b = bldo<T>(rB, oz, (unsigned)xx * oc);
                c.g = bldo<T>(rG, oz, (unsigned)xx * oc);
                c.t = bldo_nt<T>(rF, oz, (unsigned)xx * oc);  // the stored term, read once
                c.ha = bldo<T>(rM, ozhi, (unsigned)xx * oc) * bldo<T>(rA, ozhi, (unsigned)xx * oc);
                c.xl = sxp[2 * (xx - x0)];
                c.xh = sxp[2 * (xx - x0) + 1];
                if constexpr (REC) {
                    const int d = xx - sh.r0_v, kk = sh.r02 + d;
                    const bool rx = d >= 0 && d < sh.nr2 && kk >= 0 && kk < g.nr;
                    c.s = bldo<T>(rI, rx ? orz : kOOB, rx ? (unsigned)kk * E : 0u);
                } else {
                    c.s = T(0);
                }
                return c;
            };
            T qa[9];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const unsigned so = (unsigned)(x0 - 4 + k) * oc;
                qa[k] = bldo<T>(rM, ozi, so) * bldo<T>(rA, ozi, so);
            }
            __syncthreads();
            XCol pend = get(x0);
#pragma unroll
            for (int i = 0; i < XC; ++i) {
                const int x = x0 + i;
                const XCol cur = pend;
                if (i + 1 < XC) pend = get(x + 1);
                qa[8] = cur.m5 * cur.a5;
                rowa[i & 1][lane + 4] = qa[4];
                rowa[i & 1][lzh] = cur.ha;
                __syncthreads();
                const T lt = lap(qa, &rowa[i & 1][lane + 4]);
                const T qn = 2 * cur.a - sponge(cur.b, zl, zh, cur.xl, cur.xh) + lt + cur.s;
                bsto(rB, oz, (unsigned)x * oc, sponge(qn, zl, zh, cur.xl, cur.xh));
                bsto(rG, oz, (unsigned)x * oc, cur.g + cur.t * cur.a);
#pragma unroll
                for (int k = 0; k < 8; ++k) qa[k] = qa[k + 1];
            }
        };
        if (rec_rows) lean(std::true_type{});
        else lean(std::false_type{});
        return;
    }
    T qa[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int xx = x0 - 4 + k;
        const bool okp = xv(xx) && xin(xx);
        qa[k] = ldc(rM, ozi, xx, okp) * ldc(rA, ozi, xx, okp);
    }
    T av = ldc(rA, oz, x0, true), bv = ldc(rB, oz, x0, true);
    T gv = ldc(rG, oz, x0, true), tv = ldc(rF, oz, x0, true);
    T ha;
    {
        const bool okp = xin(x0);
        ha = ldc(rM, ozhi, x0, okp) * ldc(rA, ozhi, x0, okp);
    }
    T xl = a.sfx[2 * x0], xh = a.sfx[2 * x0 + 1];
    int k0 = rk(x0);
    T iv = bld<T>(rI, (unsigned)(k0 < 0 ? 0 : k0) * E, k0 >= 0);
    int buf = 0;
    for (int x = x0; x < x1; ++x) {
        const int xn = x + 1, x5 = x + 5;
        const bool nv = xn < x1;
        const bool ok5 = nv && x5 < g.pnx;
        const bool okp5 = ok5 && xin(x5);
        const T na5 = ldc(rA, ozi, x5, okp5), nm5 = ldc(rM, ozi, x5, okp5);
        const T nav = ldc(rA, oz, xn, nv);
        const T nbv = ldc(rB, oz, xn, nv);
        const T ngv = ldc(rG, oz, xn, nv);
        const T ntv = ldc(rF, oz, xn, nv);
        const bool okh = nv && xin(xn);
        const T nha = ldc(rM, ozhi, xn, okh) * ldc(rA, ozhi, xn, okh);
        const int xs = xn < g.pnx ? xn : g.pnx - 1;
        const T nxl = a.sfx[2 * xs], nxh = a.sfx[2 * xs + 1];
        const int kn = nv ? rk(xn) : -1;
        const T niv = bld<T>(rI, (unsigned)(kn < 0 ? 0 : kn) * E, kn >= 0);
        rowa[buf][lane + 4] = qa[4];
        if (lane < 8) rowa[buf][lzh] = ha;
        __syncthreads();
        if (zok) {
            // L^T (m q): the symmetric stencil over the masked product, at every cell
            const T lt = lap(qa, &rowa[buf][lane + 4]);
            T src = T(0);
            if (rlane) {
                if (g.dr_v > 0) {
                    src = src + iv;  // iv is 0 off the receivers
                } else if (x == sh.r0_v) {
                    for (int ir = 0; ir < sh.nr2; ++ir)
                        if (sh.r02 + ir >= 0 && sh.r02 + ir < g.nr) src = src + inj[sh.r02 + ir];
                }
            }
            T qn = 2 * av - sponge(bv, zl, zh, xl, xh) + lt + src;
            bsto(rB, oz, (unsigned)x * oc, sponge(qn, zl, zh, xl, xh));
            bsto(rG, oz, (unsigned)x * oc, gv + tv * av);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) qa[k] = qa[k + 1];
        qa[8] = nm5 * na5;
        av = nav; bv = nbv; gv = ngv; tv = ntv; ha = nha; xl = nxl; xh = nxh; iv = niv;
        buf ^= 1;
    }
}

// ---------------------------------------------------------------- exact adjoint, two steps per launch
// Temporal blocking of k_adj_exact for batched sweeps (f32): one launch takes (q_{t+1}, q_{t+2}) to
// (q_{t-1}, q_t), written to two other buffers, and adds both steps' dJ/dm terms in one
// read-modify-write of the accumulator.  The tiles are k_fwd2's: 56 owned z rows x W columns;
// step 1 (q_t) runs on all 64 lanes four columns ahead of step 2 (q_{t-1}) on the owned rows, and
// covers the tile grown by 4 cells on every side (receiver injection included).  Every cell of
// both steps uses k_adj_exact's operations in the same order.  HBM traffic per two steps: q_{t+1},
// q_{t+2}, m, the two forward terms and dJ/dm read, q_t, q_{t-1} and dJ/dm written: 36 B per
// cell against 56 B for two single steps.
template <typename T>
struct Adj2Args {
    const T* A;        // q_{t+1}
    const T* B;        // q_{t+2}
    T* C;              // out: q_t
    T* D;              // out: q_{t-1}
    const T* m;
    const T* sfz;
    const T* sfx;
    const ShotInfo* shots;
    const T* inj;      // residual traces [slot][nt][nr]
    long long trs;
    const T* F1;       // the forward terms of steps t and t-1 on the full field [slot][fs]
    const T* F2;
    long long fsl;
    T* gm;             // dJ/dm accumulator [slot][fs]
    int it;            // step t; the launch makes steps t and t-1
};

template <typename T, int W>
__global__ __launch_bounds__(64) void k_adj2(FGeo<T> g, Adj2Args<T> a, TileMap tm)
{
#pragma clang fp contract(fast)
    using C = Coef<T>;
    int slot, tz, xc;
    if (!tile_of(tm, slot, tz, xc)) return;
    const int lane = threadIdx.x;
    const int zo = tz * kOwnZ;
    const int z0 = zo - 4;
    const int z = z0 + lane;
    const int x0 = xc * W;
    __shared__ T rowP[2][128];  // m q_{t+1} column with its z halo (step 1)
    __shared__ T rowQ[2][72];   // m q_t column, lane l at l + 4 (step 2)
    const ShotInfo sh = a.shots[slot];
    const int L = g.ldz;
    const int bias = 4 * L + 4;
    const unsigned wb = (unsigned)((g.fs + 2 * bias) * (long long)sizeof(T));
    constexpr unsigned E = sizeof(T);
    const auto rA = win(a.A + slot * g.fs - bias, wb), rB = win(a.B + slot * g.fs - bias, wb),
               rM = win(a.m - bias, wb), rC = win(a.C + slot * g.fs - bias, wb),
               rD = win(a.D + slot * g.fs - bias, wb), rF1 = win(a.F1 + slot * a.fsl - bias, wb),
               rF2 = win(a.F2 + slot * a.fsl - bias, wb), rG = win(a.gm + slot * g.fs - bias, wb);
    const T rdx = T(1) / g.dx2, rdz = T(1) / g.dz2;
    const bool zok = z >= 0 && z < g.pnz;
    const bool zin = z >= 4 && z < g.pnz - 4;
    const bool own = zok && lane >= 4 && lane < 4 + kOwnZ;
    const T zl = zok ? a.sfz[2 * z] : T(1), zh = zok ? a.sfz[2 * z + 1] : T(1);
    const int zhz = lane < 4 ? z0 - 4 + lane : z0 + 60 + lane;  // z halo of lanes 0..7
    const int lzh = lane < 4 ? lane : 64 + lane;
    const bool zhv = lane < 8 && zhz >= 0 && zhz < g.pnz;
    const bool zhin = zhv && zhz >= 4 && zhz < g.pnz - 4;
    const unsigned oz = zok ? (unsigned)(bias + z) * E : kOOB;
    const unsigned ozhi = zhin ? (unsigned)(bias + zhz) * E : kOOB;
    const unsigned oo = own ? oz : kOOB;
    const unsigned oc = (unsigned)L * E;
    auto xv = [&](int x) { return x >= 0 && x < g.pnx; };
    auto xin = [&](int x) { return x >= 4 && x < g.pnx - 4; };
    auto ld = [&](__amdgpu_buffer_rsrc_t r, unsigned voff, int x, bool ok) {
        return bldo<T>(r, ok ? voff : kOOB, ok ? (unsigned)x * oc : 0u);
    };
    auto lap = [&](const T* q, const T* R) {
        return (C::K0 * q[4] + C::K1 * (q[5] + q[3]) + C::K2 * (q[6] + q[2]) + C::K3 * (q[7] + q[1]) +
                C::K4 * (q[8] + q[0])) * rdx +
               (C::K0 * R[0] + C::K1 * (R[1] + R[-1]) + C::K2 * (R[2] + R[-2]) + C::K3 * (R[3] + R[-3]) +
                C::K4 * (R[4] + R[-4])) * rdz;
    };
    // R^T g at lane z == rz of column x for step it (0 elsewhere)
    const bool rlane = z == g.rz;
    auto inj = [&](int x, int it) -> T {
        if (!rlane || !xv(x)) return T(0);
        const T* r = a.inj + slot * a.trs + (long long)it * g.nr;
        T v = T(0);
        if (g.dr_v > 0) {
            const int d = x - sh.r0_v;
            if (d >= 0 && d % g.dr_v == 0 && d / g.dr_v < sh.nr2) {
                const int k = sh.r02 + d / g.dr_v;
                if (k >= 0 && k < g.nr) v = v + r[k];
            }
        } else if (x == sh.r0_v) {
            for (int ir = 0; ir < sh.nr2; ++ir)
                if (sh.r02 + ir >= 0 && sh.r02 + ir < g.nr) v = v + r[sh.r02 + ir];
        }
        return v;
    };
    // Lean tiles: every column the step-1 stencil reaches inside [4, pnx-4) and receivers (if the
    // window holds their row) at every grid point: unmasked fetches, LDS sponge pairs, the receiver
    // samples through masked buffer loads, no scalar branches in the unrolled loop
    const bool rwin = g.rz >= z0 && g.rz < z0 + 64;
    if (x0 >= 12 && x0 + W + 12 <= g.pnx && (!rwin || g.dr_v == 1)) {
        __shared__ T sxp[2 * (W + 8)];  // sponge pairs of columns x0-4 .. x0+W+3
        if (lane < 2 * (W + 8) / 2) {
            sxp[2 * lane] = a.sfx[2 * (x0 - 4 + lane)];
            sxp[2 * lane + 1] = a.sfx[2 * (x0 - 4 + lane) + 1];
        }
        const unsigned orz = rlane ? 0u : kOOB;
        const auto rI1 = win(a.inj + slot * a.trs + (long long)a.it * g.nr, (unsigned)(g.nr * E));
        const auto rI2 = win(a.inj + slot * a.trs + (long long)(a.it - 1) * g.nr, (unsigned)(g.nr * E));
        auto src = [&](__amdgpu_buffer_rsrc_t r, int x) -> T {  // receiver k = r02 + (x - r0_v)
            const int d = x - sh.r0_v, k = sh.r02 + d;
            const bool ok = rwin && d >= 0 && d < sh.nr2 && k >= 0 && k < g.nr;
            return T(0) + bldo<T>(r, ok ? orz : kOOB, ok ? (unsigned)k * E : 0u);
        };
        T ra[9], mq[9], P[9], Q[9], qt[5];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const unsigned so = (unsigned)(x0 - 8 + k) * oc;
            ra[k] = bldo<T>(rA, oz, so);
            mq[k] = bldo<T>(rM, oz, so);
            P[k] = keep_if(mq[k] * ra[k], zin);
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) Q[k] = T(0);
#pragma unroll
        for (int k = 0; k < 5; ++k) qt[k] = T(0);
        struct LCol {
            T a4, m4, b, ha, i1;
        };
        auto get = [&](int c) -> LCol {
            LCol o;
            const unsigned so = (unsigned)c * oc, s4 = (unsigned)(c + 4) * oc;
            o.a4 = bldo<T>(rA, oz, s4);
            o.m4 = bldo<T>(rM, oz, s4);
            o.b = bldo<T>(rB, oz, so);
            o.ha = bldo<T>(rM, ozhi, so) * bldo<T>(rA, ozhi, so);
            o.i1 = src(rI1, c);
            return o;
        };
        __syncthreads();
        LCol pend = get(x0 - 4);
#pragma unroll 1
        for (int i = 0; i < W + 8; ++i) {
            const int c = x0 - 4 + i;
            const LCol cur = pend;
            if (i + 1 < W + 8) pend = get(c + 1);
            ra[8] = cur.a4;
            mq[8] = cur.m4;
            P[8] = keep_if(cur.m4 * cur.a4, zin);
            rowP[i & 1][lane + 4] = P[4];
            rowP[i & 1][lzh] = cur.ha;
            __syncthreads();
            const T cx = sxp[2 * i], cxh = sxp[2 * i + 1];
            const T lt = lap(P, &rowP[i & 1][lane + 4]);
            const T q1 = 2 * ra[4] - sponge(cur.b, zl, zh, cx, cxh) + lt + cur.i1;
            const T x1 = keep_if(sponge(q1, zl, zh, cx, cxh), zok);
            if (i >= 4 && i < W + 4) bsto(rC, oo, (unsigned)c * oc, x1);
            qt[4] = x1;
            Q[8] = keep_if(mq[4] * x1, zin);
            if (i >= 8) {
                const int s = c - 4;
                rowQ[i & 1][lane + 4] = Q[4];
                __syncthreads();
                const T sx = sxp[2 * (i - 4)], sxh = sxp[2 * (i - 4) + 1];
                const T lt2 = lap(Q, &rowQ[i & 1][lane + 4]);
                const T q2 = 2 * qt[0] - sponge(ra[0], zl, zh, sx, sxh) + lt2 + src(rI2, s);
                const unsigned so = (unsigned)s * oc;
                bsto(rD, oo, so, sponge(q2, zl, zh, sx, sxh));
                const T gv = bldo<T>(rG, oo, so), f1 = bldo<T>(rF1, oo, so), f2 = bldo<T>(rF2, oo, so);
                const T g1 = gv + f1 * ra[0];
                bsto(rG, oo, so, g1 + f2 * qt[0]);
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) { ra[k] = ra[k + 1]; mq[k] = mq[k + 1]; P[k] = P[k + 1]; Q[k] = Q[k + 1]; }
#pragma unroll
            for (int k = 0; k < 4; ++k) qt[k] = qt[k + 1];
            __builtin_amdgcn_sched_barrier(0);  // keep each column's loads one column ahead, no further
        }
        return;
    }
    // queues while step 1 runs at column c: raw q_{t+1} (ra), m (mq) and the masked product m q_{t+1}
    // (P) at c-4 .. c+4; the masked product m q_t (Q) at c-8 .. c and raw q_t (qt) at c-4 .. c
    T ra[9], mq[9], P[9], Q[9], qt[5];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int x = x0 - 8 + k;
        ra[k] = ld(rA, oz, x, xv(x));
        mq[k] = ld(rM, oz, x, xv(x));
        P[k] = (zin && xin(x)) ? mq[k] * ra[k] : T(0);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) Q[k] = T(0);
#pragma unroll
    for (int k = 0; k < 5; ++k) qt[k] = T(0);
    struct Col {
        T a4, m4, b, ha;
    };
    auto fetch = [&](int c) -> Col {
        Col o;
        o.a4 = ld(rA, oz, c + 4, xv(c + 4));
        o.m4 = ld(rM, oz, c + 4, xv(c + 4));
        o.b = ld(rB, oz, c, xv(c));
        o.ha = ld(rM, ozhi, c, xin(c)) * ld(rA, ozhi, c, xin(c));
        return o;
    };
    Col pend = fetch(x0 - 4);
#pragma unroll 1
    for (int i = 0; i < W + 8; ++i) {
        const int c = x0 - 4 + i;
        const Col cur = pend;
        if (i + 1 < W + 8) pend = fetch(c + 1);
        ra[8] = cur.a4;
        mq[8] = cur.m4;
        P[8] = (zin && xin(c + 4)) ? cur.m4 * cur.a4 : T(0);
        // ---- step 1 (q_t) at column c, every lane
        rowP[i & 1][lane + 4] = P[4];
        rowP[i & 1][lzh] = cur.ha;
        __syncthreads();
        const bool cv = xv(c);
        const T cx = cv ? a.sfx[2 * c] : T(1), cxh = cv ? a.sfx[2 * c + 1] : T(1);
        const T lt = lap(P, &rowP[i & 1][lane + 4]);
        const T q1 = 2 * ra[4] - sponge(cur.b, zl, zh, cx, cxh) + lt + inj(c, a.it);
        const T x1 = (zok && cv) ? sponge(q1, zl, zh, cx, cxh) : T(0);
        const bool cown = c >= x0 && c < x0 + W && cv;
        if (cown) bsto(rC, oo, (unsigned)c * oc, x1);
        qt[4] = x1;
        Q[8] = (zin && xin(c)) ? mq[4] * x1 : T(0);
        // ---- step 2 (q_{t-1}) at column s = c - 4 on the owned rows, with both dJ/dm terms
        if (i >= 8) {
            const int s = c - 4;
            rowQ[i & 1][lane + 4] = Q[4];
            __syncthreads();
            const bool sv = xv(s);
            const T sx = sv ? a.sfx[2 * s] : T(1), sxh = sv ? a.sfx[2 * s + 1] : T(1);
            const T lt2 = lap(Q, &rowQ[i & 1][lane + 4]);
            const T q2 = 2 * qt[0] - sponge(ra[0], zl, zh, sx, sxh) + lt2 + inj(s, a.it - 1);
            if (sv) {
                const unsigned so = (unsigned)s * oc;
                bsto(rD, oo, so, sponge(q2, zl, zh, sx, sxh));
                const T gv = bldo<T>(rG, oo, so), f1 = bldo<T>(rF1, oo, so), f2 = bldo<T>(rF2, oo, so);
                const T g1 = gv + f1 * ra[0];
                bsto(rG, oo, so, g1 + f2 * qt[0]);
            }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) { ra[k] = ra[k + 1]; mq[k] = mq[k + 1]; P[k] = P[k + 1]; Q[k] = Q[k + 1]; }
#pragma unroll
        for (int k = 0; k < 4; ++k) qt[k] = qt[k + 1];
        __builtin_amdgcn_sched_barrier(0);
    }
}

// ---------------------------------------------------------------- small kernels
template <typename T>
__global__ void k_fill(T* p, long long n, T v)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}

// pad2d (wave_fwiutil.c:123-147): clamp-to-edge replication of the model into the padded grid;
// writes the padded velocity and m = (v*v)*dt^2
template <typename T>
__global__ void k_pad_model(FGeo<T> g, const T* __restrict__ vel, T* __restrict__ vpad, T* __restrict__ m, T dt2)
{
    const int z = blockIdx.x * blockDim.x + threadIdx.x, x = blockIdx.y;
    if (z >= g.pnz) return;
    const int sz = min(max(z - g.nb, 0), g.nz - 1), sx = min(max(x - g.nb, 0), g.nx - 1);
    const T v = vel[(long long)sx * g.nz + sz];
    const long long j = (long long)x * g.ldz + z;
    vpad[j] = v;
    m[j] = v * v * dt2;
}

// temp = v; temp = temp*temp*temp; temp = -2./temp  (wave_fwigradient.c:225-227), f32
__global__ void k_tmpv(int n, const float* __restrict__ vel, float* __restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float t = vel[i];
    t = t * t * t;
    out[i] = -2. / t;
}
__global__ void k_tmpv_d(int n, const double* __restrict__ vel, float* __restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float t = (float)vel[i];
    t = t * t * t;
    out[i] = -2. / t;
}

// residual r = d - p (misfit terms 0.5 r^2 into per-block partials), then r *= W
// layouts: obs [shot][nt][nr], pred/res [slot][nt][nr], W [nt][nr]
template <typename T>
__global__ void k_residual(int nslot, long long per, const T* __restrict__ obs, const T* __restrict__ pred,
                           const T* __restrict__ W, T* __restrict__ res, const int* __restrict__ shot_of_slot,
                           int weighted, int sign, double* __restrict__ partial)
{
    __shared__ double red[256];
    double acc = 0.0;
    const long long n = per * nslot;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int s = (int)(i / per);
        const long long k = i - (long long)s * per;
        T r = obs[(long long)shot_of_slot[s] * per + k] - pred[i];
        acc += 0.5 * (double)r * (double)r;
        if (weighted) r *= W[k];
        res[i] = sign > 0 ? r : -r;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// deterministic f64 dot / max-abs reductions (two-stage, fixed order)
template <typename T>
__global__ void k_dot_partial(long long n, const T* __restrict__ a, const T* __restrict__ b, double* __restrict__ partial)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        acc += (double)a[i] * (double)b[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

template <typename T>
__global__ void k_absmax_partial(long long n, const T* __restrict__ a, double* __restrict__ partial)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        acc = fmax(acc, fabs((double)a[i]));
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_sum_partials(int n, const double* __restrict__ partial, double* __restrict__ out, int is_max)
{
    if (threadIdx.x || blockIdx.x) return;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s = is_max ? fmax(s, partial[i]) : s + partial[i];
    *out = s;
}

// An order-independent 64-bit fingerprint of n values: the sum (mod 2^64) over i of the splitmix64
// finaliser of (bits_i xor i * golden-ratio constant), added up by the blocks in any order into
// *out, which the caller clears first.  Any change to a value changes it except with probability
// about 2^-64.
__global__ void k_zero_u64(unsigned long long* p) { *p = 0; }

template <typename T>
__global__ __launch_bounds__(256) void k_fingerprint(long long n, const T* __restrict__ v, unsigned long long* out)
{
    using u64 = unsigned long long;
    __shared__ u64 red[256];
    u64 h = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        u64 b;
        if constexpr (sizeof(T) == 8) b = __builtin_bit_cast(u64, v[i]);
        else b = (u64)__builtin_bit_cast(unsigned, v[i]);
        u64 x = b ^ ((u64)i * 0x9E3779B97F4A7C15ull);
        x ^= x >> 30; x *= 0xBF58476D1CE4E5B9ull;
        x ^= x >> 27; x *= 0x94D049BB133111EBull;
        x ^= x >> 31;
        h += x;
    }
    red[threadIdx.x] = h;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

int fingerprint(pw_fwi* p, const void* d_vel, unsigned long long* out);

// sum per-slot interior gradients into the total, in slot order
template <typename T>
__global__ void k_add_slots(long long n, int nslot, const T* __restrict__ g, long long gs, T* __restrict__ tot)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T v = tot[i];
    for (int s = 0; s < nslot; ++s) v = v + g[s * gs + i];
    tot[i] = v;
}

// exact mode: dJ/dv on the model grid = sum over the cells pad2d replicates from each model cell
// of dJ/dm * 2 v dt^2, over all slots
template <typename T>
__global__ void k_gm_fold_z(FGeo<T> g, int nslot, const T* __restrict__ gm, const T* __restrict__ vpad, T dt2,
                            double* __restrict__ t)
{
    // t[x][iz] = sum over the padded rows that replicate model row iz (pad2d: the first and last
    // rows absorb the sponge) of sum_slots dJ/dm * 2 v dt^2, for every padded column x
    const int iz = blockIdx.x * blockDim.x + threadIdx.x, x = blockIdx.y;
    if (iz >= g.nz) return;
    const int za = iz == 0 ? 0 : iz + g.nb, zb = iz == g.nz - 1 ? g.pnz : iz + g.nb + 1;
    double acc = 0.0;
    for (int z = za; z < zb; ++z) {
        const long long j = (long long)x * g.ldz + z;
        double s = 0.0;
        for (int k = 0; k < nslot; ++k) s += (double)gm[k * g.fs + j];
        acc += s * 2.0 * (double)vpad[j] * (double)dt2;
    }
    t[(long long)x * g.nz + iz] = acc;
}

template <typename T>
__global__ void k_gm_fold_x(FGeo<T> g, const double* __restrict__ t, T* __restrict__ gv)
{
    const int iz = blockIdx.x * blockDim.x + threadIdx.x, ix = blockIdx.y;
    if (iz >= g.nz) return;
    const int xa = ix == 0 ? 0 : ix + g.nb, xb = ix == g.nx - 1 ? g.pnx : ix + g.nb + 1;
    double acc = 0.0;
    for (int x = xa; x < xb; ++x) acc += t[(long long)x * g.nz + iz];
    gv[(long long)ix * g.nz + iz] = (T)((double)gv[(long long)ix * g.nz + iz] + acc);
}

// adjoint triangle smoothing of one line (np_smooth2, wave_triangle.c:177-238) in the reference's
// f32 order; one thread per line
__global__ void k_tri_smooth2(float* __restrict__ f, int nlines, int line_stride, int d, int n, int nb,
                              float* __restrict__ work)
{
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= nlines) return;
    const int o = li * line_stride;
    const int np = n + 2 * nb;
    float* tmp = work + (long long)li * np;
    float wt = 1.0 / (nb * nb);
    for (int i = 0; i < np; ++i) tmp[i] = 0;
    float a1 = -wt, a2 = 2. * wt, a3 = -wt;
    for (int i = 0; i < n; ++i) tmp[i] += a1 * f[o + i * d];
    for (int i = 0; i < n; ++i) tmp[nb + i] += a2 * f[o + i * d];
    for (int i = 0; i < n; ++i) tmp[2 * nb + i] += a3 * f[o + i * d];
    float t = 0.;
    for (int i = 0; i < np; ++i) { t += tmp[i]; tmp[i] = t; }
    t = 0.;
    for (int i = np - 1; i >= 0; --i) { t += tmp[i]; tmp[i] = t; }
    for (int i = 0; i < n; ++i) f[o + i * d] = tmp[i + nb];
    for (int j = nb + n; j < np; j += n) {
        for (int i = 0; i < n && i < np - j; ++i) f[o + (n - 1 - i) * d] += tmp[j + i];
        j += n;
        for (int i = 0; i < n && i < np - j; ++i) f[o + i * d] += tmp[j + i];
    }
    for (int j = nb; j >= 0; j -= n) {
        for (int i = 0; i < n && i < j; ++i) f[o + i * d] += tmp[j - 1 - i];
        j -= n;
        for (int i = 0; i < n && i < j; ++i) f[o + (n - 1 - i) * d] += tmp[j - 1 - i];
    }
}

__global__ void k_scale_f(long long n, float* __restrict__ f, float s)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] *= s;
}

// L-BFGS vector kernels (x, y in T; scalars are f64 from the host)
template <typename T>
__global__ void k_axpby(long long n, double a, const T* __restrict__ x, double b, const T* __restrict__ y, T* __restrict__ out)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (T)(a * (double)x[i] + b * (double)y[i]);
}

template <typename T>
__global__ void k_clip(long long n, T* __restrict__ x, T lo, T hi)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        T v = x[i];
        if (v < lo) v = lo;
        if (v > hi) v = hi;
        x[i] = v;
    }
}

// out[c][r] = in[r][c]
template <typename T>
__global__ void k_transpose_t(const T* __restrict__ in, T* __restrict__ out, int R, int C)
{
    __shared__ T t[32][33];
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

template <typename S, typename D>
__global__ void k_convert(long long n, const S* __restrict__ s, D* __restrict__ d)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (D)s[i];
}

// ---------------------------------------------------------------- host geometry (preparation)
struct Acq {
    int nz, nx, nt, ns, nb, acqui_type, interval, niter, nr;
    float dz, dx, z0, x0, dt, t0, ds, s0, coef, f0, dr, r0;
    int padnz, padnx, sz, rz, ds_v, s0_v, dr_v;
    std::vector<int> r0_v, r02, nr2;
    std::vector<float> bc;
};

int acq_from_pararray(const float* p, Acq& a, std::string& err)
{
    a.nz = p[0]; a.nx = p[1]; a.dz = p[2]; a.dx = p[3]; a.z0 = p[4]; a.x0 = p[5];
    a.nt = p[6]; a.dt = p[7]; a.t0 = p[8]; a.ns = p[10]; a.ds = p[11]; a.s0 = p[12];
    a.nb = p[14]; a.coef = p[15]; a.f0 = p[16]; a.acqui_type = p[17]; a.interval = p[18]; a.niter = p[19];
    // hard-coded by pfwic (pywave/src/pfwi.c:796-806): sz = 5, nr = nx, dr = dx, r0 = x0, rz = 1
    a.nr = a.nx; a.dr = a.dx; a.r0 = a.x0;
    if (a.nz < 1 || a.nx < 1 || a.nt < 2 || a.ns < 1 || a.interval < 1 || !(a.dx > 0) || !(a.dz > 0) || !(a.dt > 0)) {
        err = "pararray: invalid dimensions (nz, nx, nt, ns, interval must be positive)";
        return PW_ERR_ARG;
    }
    if (a.nb < 4) {
        err = "pararray: nb must be >= 4 (receivers and sources must stay off the Laplacian's 4-cell rim)";
        return PW_ERR_ARG;
    }
    const int nb = a.nb;
    a.bc.resize(nb);
    for (int i = 0; i < nb; ++i) {
        float tmp = a.coef * (nb - i);
        a.bc[i] = expf(-tmp * tmp);
    }
    a.padnx = a.nx + 2 * nb;
    a.padnz = a.nz + 2 * nb;
    a.ds_v = a.ds / a.dx + 0.5;
    a.s0_v = a.s0 / a.dx + 0.5 + nb;
    a.sz = 5 + nb;
    a.dr_v = a.dr / a.dx + 0.5;
    a.rz = 1 + nb;
    a.r0_v.resize(a.ns); a.r02.resize(a.ns); a.nr2.resize(a.ns);
    float xend = a.x0 + (a.nx - 1) * a.dx;
    for (int i = 0; i < a.ns; ++i) {
        if (a.acqui_type == 1) {
            a.r0_v[i] = (a.r0 - a.x0) / a.dx + 0.5 + nb;
            a.r02[i] = 0;
            a.nr2[i] = a.nr;
        } else {
            float sx = a.s0 + a.ds * i;
            float rbegin = (sx + a.r0 < a.x0) ? a.x0 : sx + a.r0;
            float rend = sx + a.r0 + (a.nr - 1) * a.dr;
            rend = (rend < xend) ? rend : xend;
            a.r0_v[i] = rbegin / a.dx + 0.5 + nb;
            a.r02[i] = (rbegin - sx - a.r0) / a.dx + 0.5;
            a.nr2[i] = (rend - rbegin) / a.dx + 1.5;
        }
    }
    return PW_OK;
}

// one line of the adjoint triangle smoothing (host copy for the source map)
void tri_line_host(int o, int d, int n, int nb, float* x)
{
    const int np = n + 2 * nb;
    std::vector<float> tmp(np, 0.f);
    float wt = 1.0 / (nb * nb);
    float a1 = -wt, a2 = 2. * wt, a3 = -wt;
    for (int i = 0; i < n; ++i) tmp[i] += a1 * x[o + i * d];
    for (int i = 0; i < n; ++i) tmp[nb + i] += a2 * x[o + i * d];
    for (int i = 0; i < n; ++i) tmp[2 * nb + i] += a3 * x[o + i * d];
    float t = 0.;
    for (int i = 0; i < np; ++i) { t += tmp[i]; tmp[i] = t; }
    t = 0.;
    for (int i = np - 1; i >= 0; --i) { t += tmp[i]; tmp[i] = t; }
    for (int i = 0; i < n; ++i) x[o + i * d] = tmp[i + nb];
    for (int j = nb + n; j < np; j += n) {
        for (int i = 0; i < n && i < np - j; ++i) x[o + (n - 1 - i) * d] += tmp[j + i];
        j += n;
        for (int i = 0; i < n && i < np - j; ++i) x[o + i * d] += tmp[j + i];
    }
    for (int j = nb; j >= 0; j -= n) {
        for (int i = 0; i < n && i < j; ++i) x[o + i * d] += tmp[j - 1 - i];
        j -= n;
        for (int i = 0; i < n && i < j; ++i) x[o + (n - 1 - i) * d] += tmp[j - 1 - i];
    }
}

// source_map with rect 2 x 2 (pfwi.c:805-806) -> its non-zero patch
int source_patch(const Acq& a, int sx, int rect, ShotInfo& sh, float* patch, std::string& err)
{
    const int pnz = a.padnz, pnx = a.padnx;
    if (sx - rect + 1 < 4 || sx + rect - 1 >= pnx - 4 || a.sz - rect + 1 < 4 || a.sz + rect - 1 >= pnz - 4) {
        err = "source map support reaches the padded grid's 4-cell rim (shot at x index " + std::to_string(sx) + ")";
        return PW_ERR_ARG;
    }
    std::vector<float> rr((size_t)pnx * pnz, 0.f);
    rr[(size_t)sx * pnz + a.sz] = 1.;
    if (rect > 1) {
        for (int j = 0; j < pnx; ++j) tri_line_host(j * pnz, 1, pnz, rect, rr.data());
        for (int j = 0; j < pnz; ++j) tri_line_host(j, pnz, pnx, rect, rr.data());
    }
    const int r = rect - 1;
    sh.sx0 = sx - r; sh.sz0 = a.sz - r; sh.pw = 2 * r + 1; sh.ph = 2 * r + 1;
    for (int i = 0; i < PMAX * PMAX; ++i) patch[i] = 0.f;
    for (int px = 0; px < sh.pw; ++px)
        for (int pz = 0; pz < sh.ph; ++pz)
            patch[px * PMAX + pz] = rr[(size_t)(sh.sx0 + px) * pnz + sh.sz0 + pz];
    return PW_OK;
}

void residual_weights(const Acq& a, std::vector<float>& w /* [ir][it] */)
{
    const int nt = a.nt, nr = a.nr;
    // fwic's full windows (pfwi.c:847-860) through gradient_init (wave_fwigradient.c:75-80)
    float wt1 = a.t0, wt2 = a.t0 + (nt - 1) * a.dt;
    float woff1 = a.r0, woff2 = a.r0 + (nr - 1) * a.dr;
    int wtn1 = (wt1 - a.t0) / a.dt + 0.5, wtn2 = (wt2 - a.t0) / a.dt + 0.5;
    int woffn1 = (woff1 - a.r0) / a.dr + 0.5, woffn2 = (woff2 - a.r0) / a.dr + 0.5;
    float sw[10];
    for (int it = 0; it < 10; ++it) sw[it] = sinf(0.5 * PI_D * (it + 1) / 11.);
    w.assign((size_t)nt * nr, 0.f);
    auto W = [&](int ir, int it) -> float& { return w[(size_t)ir * nt + it]; };
    for (int ir = woffn1; ir <= woffn2; ++ir)
        for (int it = wtn1; it <= wtn2; ++it) W(ir, it) = 1.;
    for (int ir = 0; ir < 10; ++ir)
        for (int it = wtn1; it <= wtn2; ++it) {
            W(woffn1 + ir, it) *= sw[ir];
            W(woffn2 - ir, it) *= sw[ir];
        }
    for (int it = 0; it < 10; ++it)
        for (int ir = woffn1; ir <= woffn2; ++ir) {
            W(ir, wtn1 + it) *= sw[it];
            W(ir, wtn2 - it) *= sw[it];
        }
}

inline int nblk(long long n, int bs = 256) { return (int)std::min<long long>((n + bs - 1) / bs, 1 << 20); }

}  // namespace

// ==================================================================================== plan
struct pw_fwi {
    pw_ctx* ctx = nullptr;
    int prec = 4;  // 4 = f32, 8 = f64
    Acq acq;
    int nslot = 1;        // shots per launch
    int K = 0;            // checkpoint segment length (steps)
    int ldz = 0;
    long long fs = 0;     // field elements
    long long nzx = 0;
    // buffers (void*; typed by prec)
    void *vel = nullptr, *vpad = nullptr, *m = nullptr, *sfz = nullptr, *sfx = nullptr, *wav = nullptr;
    void *X = nullptr, *Y = nullptr, *A = nullptr, *B = nullptr;
    void *rec = nullptr, *res = nullptr, *obs = nullptr, *W = nullptr;
    void *ckpt = nullptr, *ckpt_snap = nullptr, *seg = nullptr, *gm = nullptr, *gslot = nullptr, *gtot = nullptr;
    void *patch = nullptr;
    void* fsrc = nullptr;   // passive source field [ns][nt][nx][nz] (plan precision) or null
    bool passive = false;   // shots are the source field (gradient_pas_av), not point sources
    int cur_b0 = 0;         // first shot of the batch set by set_batch
    std::string active_err; // why the point-source geometry is unusable (empty: usable)
    bool visco = false;     // visco-acoustic modelling and the gradient_v imaging gradient
    // tau, taus on the padded grid; forward memory variable R0/R1 and backward RB0/RB1 per slot;
    // the forward one at every checkpoint [segment][slot][fs]
    float *tau_pad = nullptr, *taus_pad = nullptr, *R0 = nullptr, *R1 = nullptr;
    float *RB0 = nullptr, *RB1 = nullptr, *ckpt_r = nullptr;
    std::vector<ShotInfo> shot_active;
    float* tmpv = nullptr;
    float* work = nullptr;  // smoothing workspace
    ShotInfo* shots = nullptr;
    int* slot_shot = nullptr;
    double* partial = nullptr;
    double* dscalar = nullptr;
    double* gfold = nullptr;  // dJ/dv folded over padded rows: [padnx][nz]
    int nseg = 0;
    long long seg_cap = 0;  // elements of the segment buffer
    bool have_wav = false, have_obs = false;
    float scaling = 0.f;    // first-call gradient scaling (0 = not yet set)
    std::vector<float> patch_host;  // per shot PMAX*PMAX
    std::vector<ShotInfo> shot_host;
    double last_fcost = 0.0;
    // checkpoints left by pw_fwi_forward_dev for the model at dev_ckpt_vel (all shots in one
    // batch), reused by the next pw_fwi_adjoint_dev on the same model; any other call that
    // uploads a model or changes the set-up clears the flag
    bool dev_ckpt_valid = false;
    const void* dev_ckpt_vel = nullptr;
    unsigned long long* dhash = nullptr;  // fingerprints of the forward_dev and adjoint_dev models
    size_t budget = 0;  // bytes allowed for checkpoints + segments
    // tile widths of the step kernels (forward, visco forward, exact adjoint, reference adjoint,
    // visco reference adjoint), picked by tune_widths for the plan's grid and batch
    int w_fwd = kDefaultWidth, w_fwdv = kDefaultWidth, w_exact = kDefaultWidth, w_ref = kDefaultWidth,
        w_refv = kDefaultWidth;
    int w_fwdt = kDefaultWidth;  // forward steps that store their term for the exact adjoint
    int w_fwd2 = 0;       // two-step forward tiles' width (0: single steps are faster or not allowed)
    int w_adj2 = 0;       // two-step exact adjoint's width (0: single steps)
    int max_fused = 2;    // forward steps one launch may advance (1: one kernel per time step)
    bool force_fused = false;  // two steps per launch whether or not they time faster
    std::vector<float> tune_log;  // (kernel, width, ms per launch) of the last timing sweep
    bool widths_tuned = false;
    bool widths_forced = false;  // pw_fwi_set_widths: no timing sweep
    void *X2 = nullptr, *Y2 = nullptr;  // the two-step forward's output fields (allocated on first use)
    void *A2 = nullptr, *B2 = nullptr;  // the two-step adjoint's output fields (allocated on first use)
    void *adj_sa = nullptr, *adj_sb = nullptr;  // the adjoint fields free during a backward sweep
    double t_forward_ms = 0, t_backward_ms = 0;
    // device-entry-point timing: events around the last forward_dev (0, 1) and adjoint_dev (2, 3),
    // read by pw_fwi_timing
    hipEvent_t ev_dev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool dev_timed_fwd = false, dev_timed_adj = false;
};

namespace {

size_t esz(const pw_fwi* p) { return p->prec == 8 ? 8 : 4; }

template <typename T>
FGeo<T> geo(const pw_fwi* p)
{
    FGeo<T> g{};
    const Acq& a = p->acq;
    g.pnz = a.padnz; g.pnx = a.padnx; g.ldz = p->ldz;
    g.nz = a.nz; g.nx = a.nx; g.nb = a.nb;
    g.nt = a.nt; g.nr = a.nr; g.rz = a.rz; g.dr_v = a.dr_v;
    if (sizeof(T) == 4) {
        float dx2 = a.dx * a.dx, dz2 = a.dz * a.dz;  // wave_fwigradient.c:60-61
        g.dx2 = dx2; g.dz2 = dz2;
        float wdt = a.dt * a.interval;
        float wdt2 = wdt * wdt;
        g.wdt2 = wdt2;
    } else {
        g.dx2 = (T)a.dx * (T)a.dx; g.dz2 = (T)a.dz * (T)a.dz;
        double wdt = (double)a.dt * a.interval;
        g.wdt2 = wdt * wdt;
    }
    g.fs = p->fs;
    return g;
}

template <typename T>
T dt2_of(const pw_fwi* p)
{
    if (sizeof(T) == 4) {
        float d = p->acq.dt * p->acq.dt;
        return (T)d;
    }
    return (T)p->acq.dt * (T)p->acq.dt;
}

void fwi_free(pw_fwi* p)
{
    void* bufs[] = {p->vel, p->vpad, p->m, p->sfz, p->sfx, p->wav, p->X, p->Y, p->A, p->B, p->rec, p->res,
                    p->obs, p->W, p->ckpt, p->ckpt_snap, p->seg, p->gm, p->gslot, p->gtot, p->patch,
                    p->tmpv, p->work, p->shots, p->slot_shot, p->partial, p->dscalar, p->gfold, p->fsrc,
                    p->tau_pad, p->taus_pad, p->R0, p->R1, p->RB0, p->RB1, p->ckpt_r, p->dhash, p->X2, p->Y2,
                    p->A2, p->B2};
    for (void* b : bufs)
        if (b) hipFree(b);
    for (hipEvent_t ev : p->ev_dev)
        if (ev) hipEventDestroy(ev);
}

#define PW_HIPS(call)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess) {                                                                         \
            p->ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);                            \
            return PW_ERR_HIP;                                                                          \
        }                                                                                               \
    } while (0)

int dmalloc(pw_fwi* p, void** ptr, size_t bytes)
{
    if (hipMalloc(ptr, bytes ? bytes : 8) != hipSuccess) {
        (void)hipGetLastError();  // an out-of-memory result must not surface as a later launch error
        *ptr = nullptr;
        p->ctx->err = "pw_fwi: device allocation of " + std::to_string(bytes) + " bytes failed";
        return PW_ERR_NOMEM;
    }
    return PW_OK;
}

// (re)size checkpoint and segment storage for K-step segments
int plan_storage(pw_fwi* p, int K)
{
    const Acq& a = p->acq;
    const size_t e = esz(p);
    if (K <= 0) {
        // one segment when everything fits the budget (no recomputation), else 50 steps; if the
        // one-segment storage cannot be allocated after all, 50-step segments
        size_t seg_all = (size_t)a.nt * p->nslot * p->fs * e;
        const int k_all = seg_all + 2 * p->nslot * p->fs * e <= p->budget ? a.nt : 50;
        const std::string err = p->ctx->err;
        if (k_all != 50 && plan_storage(p, k_all) == PW_OK) return PW_OK;
        p->ctx->err = err;
        K = 50;
    }
    K = std::max(1, std::min(K, a.nt));
    if (K % a.interval) K += a.interval - K % a.interval;  // segments start at imaging steps
    int nseg = (a.nt + K - 1) / K;
    if (p->ckpt) { hipFree(p->ckpt); p->ckpt = nullptr; }
    if (p->ckpt_snap) { hipFree(p->ckpt_snap); p->ckpt_snap = nullptr; }
    if (p->seg) { hipFree(p->seg); p->seg = nullptr; }
    if (p->ckpt_r) { hipFree(p->ckpt_r); p->ckpt_r = nullptr; }
    int rc = dmalloc(p, &p->ckpt, (size_t)nseg * 2 * p->nslot * p->fs * e);
    if (!rc && p->visco) rc = dmalloc(p, (void**)&p->ckpt_r, (size_t)nseg * p->nslot * p->fs * 4);
    if (!rc) rc = dmalloc(p, &p->ckpt_snap, (size_t)nseg * p->nslot * p->nzx * e);
    // segment buffer: full fields for the exact adjoint (K of them) or interior snapshots for
    // the imaging condition (K/I + 2); the larger of the two
    long long full_elems = (long long)K * p->nslot * p->fs;
    long long snap_elems = (long long)(K / a.interval + 2) * p->nslot * p->nzx;
    p->seg_cap = std::max(full_elems, snap_elems);
    if (!rc) rc = dmalloc(p, &p->seg, (size_t)p->seg_cap * e);
    if (rc) return rc;
    p->K = K;
    p->nseg = nseg;
    return PW_OK;
}

template <typename T>
int upload_model(pw_fwi* p, const T* d_vel_src, bool from_device)
{
    pw_ctx* ctx = p->ctx;
    p->dev_ckpt_valid = false;
    hipStream_t s = ctx->s();
    const Acq& a = p->acq;
    if (from_device)
        PW_HIPS(hipMemcpyAsync(p->vel, d_vel_src, p->nzx * sizeof(T), hipMemcpyDeviceToDevice, s));
    FGeo<T> g = geo<T>(p);
    hipLaunchKernelGGL(k_pad_model<T>, dim3((a.padnz + 255) / 256, a.padnx), dim3(256), 0, s, g, (const T*)p->vel,
                       (T*)p->vpad, (T*)p->m, dt2_of<T>(p));
    if (sizeof(T) == 4)
        hipLaunchKernelGGL(k_tmpv, dim3(nblk(p->nzx)), dim3(256), 0, s, (int)p->nzx, (const float*)p->vel, p->tmpv);
    else
        hipLaunchKernelGGL(k_tmpv_d, dim3(nblk(p->nzx)), dim3(256), 0, s, (int)p->nzx, (const double*)p->vel, p->tmpv);
    PW_HIPS(hipGetLastError());
    return PW_OK;
}

int upload_vel_host(pw_fwi* p, const float* vel)
{
    hipStream_t s = p->ctx->s();
    if (p->prec == 4) {
        PW_HIPS(hipMemcpyAsync(p->vel, vel, p->nzx * 4, hipMemcpyHostToDevice, s));
        return upload_model<float>(p, nullptr, false);
    }
    std::vector<double> v(vel, vel + p->nzx);
    PW_HIPS(hipMemcpyAsync(p->vel, v.data(), p->nzx * 8, hipMemcpyHostToDevice, s));
    PW_HIPS(hipStreamSynchronize(s));
    return upload_model<double>(p, nullptr, false);
}

// per-batch shot tables
int set_batch(pw_fwi* p, int b0, int nb_)
{
    if (!p->passive && !p->active_err.empty()) PW_FAIL(p->ctx, PW_ERR_ARG, p->active_err);
    p->cur_b0 = b0;
    hipStream_t s = p->ctx->s();
    std::vector<ShotInfo> sh(p->nslot);
    std::vector<int> sos(p->nslot, 0);
    for (int k = 0; k < nb_; ++k) {
        sh[k] = p->shot_host[b0 + k];
        sos[k] = b0 + k;
    }
    PW_HIPS(hipMemcpyAsync(p->shots, sh.data(), sizeof(ShotInfo) * p->nslot, hipMemcpyHostToDevice, s));
    PW_HIPS(hipMemcpyAsync(p->slot_shot, sos.data(), sizeof(int) * p->nslot, hipMemcpyHostToDevice, s));
    const size_t e = esz(p);
    if (p->prec == 4) {
        PW_HIPS(hipMemcpyAsync(p->patch, p->patch_host.data() + (size_t)b0 * PMAX * PMAX,
                               sizeof(float) * nb_ * PMAX * PMAX, hipMemcpyHostToDevice, s));
    } else {
        std::vector<double> pd(p->patch_host.begin() + (size_t)b0 * PMAX * PMAX,
                               p->patch_host.begin() + (size_t)(b0 + nb_) * PMAX * PMAX);
        PW_HIPS(hipMemcpyAsync(p->patch, pd.data(), e * pd.size(), hipMemcpyHostToDevice, s));
        PW_HIPS(hipStreamSynchronize(s));
    }
    PW_HIPS(hipStreamSynchronize(s));
    return PW_OK;
}

template <typename T>
StepArgs<T> base_args(pw_fwi* p)
{
    StepArgs<T> a{};
    a.m = (const T*)p->m;
    a.sfz = (const T*)p->sfz;
    a.sfx = (const T*)p->sfx;
    a.shots = p->shots;
    a.patch = (const T*)p->patch;
    a.wav = (const T*)p->wav;
    a.trs = (long long)p->acq.nt * p->acq.nr;
    a.snaps = p->nzx;
    a.fulls = p->fs;
    a.fss = (long long)p->acq.nt * p->nzx;
    return a;
}

// the batch's passive source field at step it (slot 0), or null for point sources
template <typename T>
const T* step_src(const pw_fwi* p, int it)
{
    if (!p->passive) return nullptr;
    return (const T*)p->fsrc + ((size_t)p->cur_b0 * p->acq.nt + it) * (size_t)p->nzx;
}

template <typename T>
int fill(pw_fwi* p, void* ptr, long long n, T v)
{
    hipLaunchKernelGGL(k_fill<T>, dim3(nblk(n)), dim3(256), 0, p->ctx->s(), (T*)ptr, n, v);
    PW_HIPS(hipGetLastError());
    return PW_OK;
}

// launches of the step kernels at a tile width (one of kWidths)
template <typename T, bool VISCO>
void launch_fwd(int W, const FGeo<T>& g, const StepArgs<T>& a, int pnz, int pnx, int nb_, hipStream_t s)
{
    const TileMap tm = tile_map(pnz, pnx, nb_, W);
    if (W == 4) hipLaunchKernelGGL((k_fwd<T, VISCO, 4>), tile_grid(tm), dim3(64), 0, s, g, a, tm);
    else if (W == 16) hipLaunchKernelGGL((k_fwd<T, VISCO, 16>), tile_grid(tm), dim3(64), 0, s, g, a, tm);
    else hipLaunchKernelGGL((k_fwd<T, VISCO, 8>), tile_grid(tm), dim3(64), 0, s, g, a, tm);
}

// two-step forward tiles: 56 owned rows x W columns
template <typename T>
void launch_fwd2(int W, const FGeo<T>& g, const Fwd2Args<T>& a, int pnz, int pnx, int nb_, hipStream_t s)
{
    TileMap tm;
    tm.ntz = (pnz + kOwnZ - 1) / kOwnZ;
    tm.nxc = (pnx + W - 1) / W;
    tm.ns = nb_;
    tm.nblk = tm.ntz * tm.nxc * nb_;
    if (W == 8) hipLaunchKernelGGL((k_fwd2<T, 8>), tile_grid(tm), dim3(64), 0, s, g, a, tm);
    else if (W == 32) hipLaunchKernelGGL((k_fwd2<T, 32>), tile_grid(tm), dim3(64), 0, s, g, a, tm);
    else hipLaunchKernelGGL((k_fwd2<T, 16>), tile_grid(tm), dim3(64), 0, s, g, a, tm);
}

void launch_adj2(int W, const FGeo<float>& g, const Adj2Args<float>& a, int pnz, int pnx, int nb_, hipStream_t s)
{
    TileMap tm;
    tm.ntz = (pnz + kOwnZ - 1) / kOwnZ;
    tm.nxc = (pnx + W - 1) / W;
    tm.ns = nb_;
    tm.nblk = tm.ntz * tm.nxc * nb_;
    if (W == 8) hipLaunchKernelGGL((k_adj2<float, 8>), tile_grid(tm), dim3(64), 0, s, g, a, tm);
    else if (W == 32) hipLaunchKernelGGL((k_adj2<float, 32>), tile_grid(tm), dim3(64), 0, s, g, a, tm);
    else hipLaunchKernelGGL((k_adj2<float, 16>), tile_grid(tm), dim3(64), 0, s, g, a, tm);
}

template <typename T>
void launch_exact(int W, const FGeo<T>& g, const StepArgs<T>& a, const ExArgs<T>& ex, int pnz, int pnx, int nb_,
                  hipStream_t s)
{
    const TileMap tm = tile_map(pnz, pnx, nb_, W);
    if (W == 4) hipLaunchKernelGGL((k_adj_exact<T, 4>), tile_grid(tm), dim3(64), 0, s, g, a, ex, tm);
    else if (W == 16) hipLaunchKernelGGL((k_adj_exact<T, 16>), tile_grid(tm), dim3(64), 0, s, g, a, ex, tm);
    else hipLaunchKernelGGL((k_adj_exact<T, 8>), tile_grid(tm), dim3(64), 0, s, g, a, ex, tm);
}

template <bool VISCO>
void launch_ref(int W, const FGeo<float>& g, const StepArgs<float>& a, const ImgArgs<float>& im, const ViscoBack& vb,
                int pnz, int pnx, int nb_, hipStream_t s)
{
    const TileMap tm = tile_map(pnz, pnx, nb_, W);
    if (W == 4) hipLaunchKernelGGL((k_adj_ref<VISCO, 4>), tile_grid(tm), dim3(64), 0, s, g, a, im, vb, tm);
    else if (W == 16) hipLaunchKernelGGL((k_adj_ref<VISCO, 16>), tile_grid(tm), dim3(64), 0, s, g, a, im, vb, tm);
    else hipLaunchKernelGGL((k_adj_ref<VISCO, 8>), tile_grid(tm), dim3(64), 0, s, g, a, im, vb, tm);
}

// three timing events, released on every return path
struct EventSet {
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    hipError_t create()
    {
        for (auto& x : e)
            if (hipError_t r = hipEventCreate(&x)) return r;
        return hipSuccess;
    }
    ~EventSet()
    {
        for (auto x : e)
            if (x) hipEventDestroy(x);
    }
};

// whether forward sweeps advance two steps per launch (acoustic point-source plans whose tuner
// found the two-step kernel faster, unless limited to one kernel per time step)
bool fused2(const pw_fwi* p) { return p->max_fused >= 2 && p->w_fwd2 > 0 && !p->visco && !p->passive && p->X2; }

// whether exact-adjoint sweeps advance two steps per launch (f32 plans)
bool fused2_adj(const pw_fwi* p) { return p->max_fused >= 2 && p->w_adj2 > 0 && p->prec == 4 && p->A2; }

Adj2Args<float> adj2_args(pw_fwi* p, const float* A, const float* B, float* C, float* D, const float* F1,
                          const float* F2, int it)
{
    Adj2Args<float> r{};
    r.A = A; r.B = B; r.C = C; r.D = D;
    r.m = (const float*)p->m;
    r.sfz = (const float*)p->sfz;
    r.sfx = (const float*)p->sfx;
    r.shots = p->shots;
    r.inj = (const float*)p->res;
    r.trs = (long long)p->acq.nt * p->acq.nr;
    r.F1 = F1; r.F2 = F2;
    r.fsl = p->fs;
    r.gm = (float*)p->gm;
    r.it = it;
    return r;
}

template <typename T>
Fwd2Args<T> fwd2_args(pw_fwi* p, const T* X, const T* Y, T* C, T* D, int it)
{
    Fwd2Args<T> f{};
    f.X = X; f.Y = Y; f.C = C; f.D = D;
    f.m = (const T*)p->m;
    f.sfz = (const T*)p->sfz;
    f.sfx = (const T*)p->sfx;
    f.shots = p->shots;
    f.patch = (const T*)p->patch;
    f.wav = (const T*)p->wav;
    f.trs = (long long)p->acq.nt * p->acq.nr;
    f.snaps = p->nzx;
    f.terms = p->fs;
    f.it = it;
    return f;
}

// Milliseconds per launch of fn (two warm-up launches, then `reps` timed ones).
template <typename F>
int time_launches(pw_fwi* p, F fn, int reps, float* ms)
{
    hipStream_t s = p->ctx->s();
    EventSet ev;
    PW_HIPS(ev.create());
    fn();
    fn();
    PW_HIPS(hipEventRecord(ev.e[0], s));
    for (int k = 0; k < reps; ++k) fn();
    PW_HIPS(hipEventRecord(ev.e[1], s));
    PW_HIPS(hipGetLastError());
    PW_HIPS(hipEventSynchronize(ev.e[1]));
    PW_HIPS(hipEventElapsedTime(ms, ev.e[0], ev.e[1]));
    *ms /= reps;
    return PW_OK;
}

// Tile-width autotuner of the step kernels: times each width in kWidths for the kernels this plan
// runs, on its own grid, batch and buffers, and keeps the fastest per kernel.  It runs before a
// sweep sets its fields up (every sweep clears what the timing launches write: the fields, traces,
// adjoint and gradient accumulators), on the first batch's shots.
template <typename T>
int tune_widths(pw_fwi* p)
{
    const Acq& a = p->acq;
    hipStream_t s = p->ctx->s();
    const int nb_ = std::min(p->nslot, a.ns);
    int rc = set_batch(p, 0, nb_);
    if (rc) return rc;
    const FGeo<T> g = geo<T>(p);
    const int reps = 8;
    p->tune_log.clear();
    auto note = [&](int kind, int W, float ms) {
        p->tune_log.push_back((float)kind);
        p->tune_log.push_back((float)W);
        p->tune_log.push_back(ms);
    };
    int kind = 0;  // 0 forward, 1 visco forward, 2 exact adjoint, 3 reference adjoint, 4 visco reference adjoint,
                   // 5 two-step forward, 6 two-step exact adjoint, 7 forward storing its term
    auto pick = [&](auto launch, int* best) -> int {
        float tbest = 0.f;
        for (int W : kWidths) {
            float ms = 0.f;
            if (int r = time_launches(p, [&] { launch(W); }, reps, &ms)) return r;
            note(kind, W, ms);
            if (W == kWidths[0] || ms < tbest) { tbest = ms; *best = W; }
        }
        return PW_OK;
    };
    const bool visco = p->visco && sizeof(T) == 4;
    StepArgs<T> f = base_args<T>(p);
    f.X = (const T*)p->X; f.Y = (T*)p->Y; f.it = 0;
    f.fsrc = step_src<T>(p, 0);
    f.rec = (T*)p->rec;
    if (visco) {
        f.tau = p->tau_pad; f.taus = p->taus_pad; f.r1 = p->R0; f.r2 = p->R1;
        f.idt = (float)(1. / (double)a.dt);
        kind = 1;
        rc = pick([&](int W) { launch_fwd<T, true>(W, g, f, a.padnz, a.padnx, nb_, s); }, &p->w_fwdv);
    } else {
        kind = 0;
        rc = pick([&](int W) { launch_fwd<T, false>(W, g, f, a.padnz, a.padnx, nb_, s); }, &p->w_fwd);
        // the exact gradient's forward steps also store their term (4 B more per cell, written
        // once): timed on their own, as the extra stream moves the best width
        if (!rc) {
            StepArgs<T> ft = f;
            ft.full = (T*)p->seg;
            ft.fulls = p->fs;
            kind = 7;
            rc = pick([&](int W) { launch_fwd<T, false>(W, g, ft, a.padnz, a.padnx, nb_, s); }, &p->w_fwdt);
        }
    }
    if (rc) return rc;
    // two steps per launch: kept when a launch takes less than two single steps
    p->w_fwd2 = 0;
    if (p->max_fused >= 2 && !visco && !p->passive && a.nt >= 2) {
        const long long fld = (long long)p->nslot * p->fs;
        if (!p->X2 && (rc = dmalloc(p, &p->X2, fld * sizeof(T)))) return rc;
        if (!p->Y2 && (rc = dmalloc(p, &p->Y2, fld * sizeof(T)))) return rc;
        float t1 = 0.f;
        const int w1 = p->w_fwd;
        if ((rc = time_launches(p, [&] { launch_fwd<T, false>(w1, g, f, a.padnz, a.padnx, nb_, s); }, reps, &t1)))
            return rc;
        Fwd2Args<T> f2 = fwd2_args<T>(p, (const T*)p->X, (const T*)p->Y, (T*)p->X2, (T*)p->Y2, 0);
        f2.rec = (T*)p->rec;
        float tbest = 0.f;
        int wbest = 0;
        for (int W : {8, 16, 32}) {
            float ms = 0.f;
            if ((rc = time_launches(p, [&] { launch_fwd2<T>(W, g, f2, a.padnz, a.padnx, nb_, s); }, reps, &ms)))
                return rc;
            note(5, W, ms);
            if (!wbest || ms < tbest) { tbest = ms; wbest = W; }
        }
        if (tbest < 2.f * t1 || p->force_fused) p->w_fwd2 = wbest;
    }
    StepArgs<T> b = base_args<T>(p);
    b.X = (const T*)p->A; b.Y = (T*)p->B; b.it = a.nt - 1;
    b.fsrc = step_src<T>(p, a.nt - 1);
    b.inj = (const T*)p->res;
    if (!visco) {
        ExArgs<T> ex{(const T*)p->seg, p->fs, (T*)p->gm};
        kind = 2;
        rc = pick([&](int W) { launch_exact<T>(W, g, b, ex, a.padnz, a.padnx, nb_, s); }, &p->w_exact);
        if (rc) return rc;
        // two adjoint steps per launch (f32): kept when a launch takes less than two single steps
        p->w_adj2 = 0;
        if constexpr (sizeof(T) == 4) {
            if (p->max_fused >= 2 && a.nt >= 2) {
                const long long fld = (long long)p->nslot * p->fs;
                if (!p->A2 && (rc = dmalloc(p, &p->A2, fld * 4))) return rc;
                if (!p->B2 && (rc = dmalloc(p, &p->B2, fld * 4))) return rc;
                float t1 = 0.f;
                const int w1 = p->w_exact;
                if ((rc = time_launches(p, [&] { launch_exact<T>(w1, g, b, ex, a.padnz, a.padnx, nb_, s); }, reps,
                                        &t1)))
                    return rc;
                // the two terms: the first two fields of the segment buffer (it holds K >= 2 of them)
                const float* F = (const float*)p->seg;
                const bool two_terms = p->seg_cap >= 2 * fld;
                Adj2Args<float> r = adj2_args(p, (const float*)p->A, (const float*)p->B, (float*)p->A2,
                                              (float*)p->B2, F, two_terms ? F + fld : F, a.nt - 1);
                float tbest = 0.f;
                int wbest = 0;
                for (int W : {8, 16, 32}) {
                    float ms = 0.f;
                    if ((rc = time_launches(p, [&] { launch_adj2(W, g, r, a.padnz, a.padnx, nb_, s); }, reps, &ms)))
                        return rc;
                    note(6, W, ms);
                    if (!wbest || ms < tbest) { tbest = ms; wbest = W; }
                }
                if (tbest < 2.f * t1 || p->force_fused) p->w_adj2 = wbest;
            }
        }
    }
    if constexpr (sizeof(T) == 4) {
        const long long ss = (long long)p->nslot * p->nzx;
        ImgArgs<T> im{};  // an imaging step: three snapshots, the -2/v^3 field and a gradient
        im.w_prev = (const T*)p->seg; im.w_cur = (const T*)p->seg + ss; im.w_next = (const T*)p->seg + 2 * ss;
        im.snaps = p->nzx;
        im.tmpv = p->tmpv;
        im.grad = (T*)p->gslot;
        im.gstride = p->nzx;
        if (visco) {
            ViscoBack vb{p->tau_pad, p->taus_pad, p->RB0, p->RB1, (float)(1. / (double)a.dt)};
            kind = 4;
            rc = pick([&](int W) { launch_ref<true>(W, g, b, im, vb, a.padnz, a.padnx, nb_, s); }, &p->w_refv);
        } else {
            kind = 3;
            rc = pick([&](int W) { launch_ref<false>(W, g, b, im, ViscoBack{}, a.padnz, a.padnx, nb_, s); }, &p->w_ref);
        }
        if (rc) return rc;
    }
    p->widths_tuned = true;
    return PW_OK;
}

// What decides the timing sweep's outcome: precision, grid, shots per launch, the kernels the
// plan runs and the acquisition (which tiles hold sources and receivers and so run the general
// loop).  Plans with equal keys on one context share one sweep's widths.
std::vector<long long> tune_key(const pw_fwi* p)
{
    const Acq& a = p->acq;
    std::vector<long long> k = {2, (long long)p->prec, p->nslot, a.padnz, a.padnx, a.nt, a.nb, a.sz, a.rz,
                                a.ns, a.acqui_type, a.ds_v, a.s0_v, a.dr_v, p->visco, p->passive,
                                p->max_fused, p->force_fused, p->seg_cap >= 2LL * p->nslot * p->fs};
    k.insert(k.end(), a.r0_v.begin(), a.r0_v.end());
    k.insert(k.end(), a.nr2.begin(), a.nr2.end());
    return k;
}

int ensure_widths(pw_fwi* p)
{
    if (p->widths_tuned) return PW_OK;
    const std::vector<long long> key = tune_key(p);
    auto hit = p->ctx->tuned2d.find(key);
    if (hit != p->ctx->tuned2d.end()) {
        const std::vector<int>& w = hit->second.first;
        p->w_fwd = w[0]; p->w_fwdv = w[1]; p->w_exact = w[2]; p->w_ref = w[3]; p->w_refv = w[4];
        p->w_fwd2 = w[5]; p->w_adj2 = w[6]; p->w_fwdt = w[7];
        const long long fld = p->nslot * p->fs;
        int rc = PW_OK;
        if (p->w_fwd2 && !p->X2) rc = dmalloc(p, &p->X2, fld * esz(p));
        if (!rc && p->w_fwd2 && !p->Y2) rc = dmalloc(p, &p->Y2, fld * esz(p));
        if (!rc && p->w_adj2 && !p->A2) rc = dmalloc(p, &p->A2, fld * 4);
        if (!rc && p->w_adj2 && !p->B2) rc = dmalloc(p, &p->B2, fld * 4);
        if (rc) return rc;
        p->tune_log = hit->second.second;  // the sweep of the earlier plan that chose them
        p->widths_tuned = true;
        return PW_OK;
    }
    int rc = p->prec == 4 ? tune_widths<float>(p) : tune_widths<double>(p);
    if (!rc)
        p->ctx->tuned2d[key] = {{p->w_fwd, p->w_fwdv, p->w_exact, p->w_ref, p->w_refv, p->w_fwd2, p->w_adj2,
                                 p->w_fwdt},
                                p->tune_log};
    return rc;
}

// forward sweep over one batch: traces into rec; checkpoints (and the last segment's
// snapshots when 'store_last') when 'ckpt'
template <typename T>
int forward_batch(pw_fwi* p, int nb_, bool ckpt, int mode, bool store_last)
{
    const Acq& a = p->acq;
    hipStream_t s = p->ctx->s();
    FGeo<T> g = geo<T>(p);
    const long long fld = (long long)p->nslot * p->fs;
    int rc = fill<T>(p, p->X, fld, T(0));
    if (!rc) rc = fill<T>(p, p->Y, fld, T(0));
    if (!rc) rc = fill<T>(p, p->rec, (long long)p->nslot * a.nt * a.nr, T(0));
    if (rc) return rc;
    T* X = (T*)p->X;
    T* Y = (T*)p->Y;
    const bool visco = p->visco && sizeof(T) == 4;
    float* Ra = p->R0;
    float* Rb = p->R1;
    if (visco) {
        if ((rc = fill<float>(p, Ra, fld, 0.f)) || (rc = fill<float>(p, Rb, fld, 0.f))) return rc;
    }
    const int I = a.interval, K = p->K;
    const int last = p->nseg - 1, a_last = last * K;
    T* Sa = (T*)p->X2;  // the two-step launches' output fields
    T* Sb = (T*)p->Y2;
    // what step it stores besides its level: an interior snapshot, or its term (exact adjoint)
    auto io = [&](int it, T** snap, T** full) {
        *snap = nullptr;
        *full = nullptr;
        if (!ckpt) return;
        if (mode == 0 && (it + I) % K == 0 && it + I < a.nt)  // snapshot just before the next segment
            *snap = (T*)p->ckpt_snap + (size_t)((it + I) / K) * p->nslot * p->nzx;
        if (store_last && it >= a_last) {
            if (mode == 1) *full = (T*)p->seg + (size_t)(it - a_last) * fld;
            else if (it % I == 0) *snap = (T*)p->seg + (size_t)((it - a_last) / I + 1) * p->nslot * p->nzx;
        }
    };
    for (int it = 0; it < a.nt;) {
        if (ckpt && it % K == 0) {
            const int sgi = it / K;
            T* c = (T*)p->ckpt + (size_t)sgi * 2 * fld;
            PW_HIPS(hipMemcpyAsync(c, X, fld * sizeof(T), hipMemcpyDeviceToDevice, s));
            PW_HIPS(hipMemcpyAsync(c + fld, Y, fld * sizeof(T), hipMemcpyDeviceToDevice, s));
            if (visco)  // the memory variable of the same level as X
                PW_HIPS(hipMemcpyAsync(p->ckpt_r + (size_t)sgi * fld, Ra, fld * 4, hipMemcpyDeviceToDevice, s));
        }
        // two steps in one launch unless a checkpoint falls between them
        if (fused2(p) && it + 1 < a.nt && !(ckpt && (it + 1) % K == 0)) {
            Fwd2Args<T> f = fwd2_args<T>(p, X, Y, Sa, Sb, it);
            f.rec = (T*)p->rec;
            io(it, &f.snap1, &f.term1);
            io(it + 1, &f.snap2, &f.term2);
            launch_fwd2<T>(p->w_fwd2, g, f, a.padnz, a.padnx, nb_, s);
            PW_HIPS(hipGetLastError());
            T* oX = X;
            T* oY = Y;
            X = Sb; Y = Sa; Sa = oX; Sb = oY;
            it += 2;
            continue;
        }
        StepArgs<T> s_ = base_args<T>(p);
        s_.X = X; s_.Y = Y; s_.it = it;
        s_.fsrc = step_src<T>(p, it);
        s_.rec = (T*)p->rec;
        io(it, &s_.snap, &s_.full);
        if (visco) {
            s_.tau = p->tau_pad; s_.taus = p->taus_pad; s_.r1 = Ra; s_.r2 = Rb;
            s_.idt = (float)(1. / (double)p->acq.dt);
        }
        if (visco) launch_fwd<T, true>(p->w_fwdv, g, s_, a.padnz, a.padnx, nb_, s);
        else launch_fwd<T, false>(s_.full ? p->w_fwdt : p->w_fwd, g, s_, a.padnz, a.padnx, nb_, s);
        PW_HIPS(hipGetLastError());
        std::swap(X, Y);
        if (visco) std::swap(Ra, Rb);
        ++it;
    }
    return PW_OK;
}

// recompute segment sg from its checkpoint into the segment buffer
template <typename T>
int recompute_segment(pw_fwi* p, int nb_, int sg, int mode)
{
    const Acq& a = p->acq;
    hipStream_t s = p->ctx->s();
    FGeo<T> g = geo<T>(p);
    const long long fld = (long long)p->nslot * p->fs;
    const int I = a.interval, K = p->K;
    const int a0 = sg * K, b0 = std::min(a0 + K, a.nt);
    T* c = (T*)p->ckpt + (size_t)sg * 2 * fld;
    PW_HIPS(hipMemcpyAsync(p->X, c, fld * sizeof(T), hipMemcpyDeviceToDevice, s));
    PW_HIPS(hipMemcpyAsync(p->Y, c + fld, fld * sizeof(T), hipMemcpyDeviceToDevice, s));
    T* X = (T*)p->X;
    T* Y = (T*)p->Y;
    const bool visco = p->visco && sizeof(T) == 4;
    float* Ra = p->R0;
    float* Rb = p->R1;
    if (visco) PW_HIPS(hipMemcpyAsync(Ra, p->ckpt_r + (size_t)sg * fld, fld * 4, hipMemcpyDeviceToDevice, s));
    // imaging needs snapshots up to b0 (inclusive) for segments that are not the last
    const int t_end = (mode == 0 && b0 < a.nt) ? b0 + 1 : b0;
    T* Sa = (T*)p->X2;
    T* Sb = (T*)p->Y2;
    auto io = [&](int it, T** snap, T** full) {
        *snap = nullptr;
        *full = nullptr;
        if (mode == 1) *full = (T*)p->seg + (size_t)(it - a0) * fld;
        else if (it % I == 0) *snap = (T*)p->seg + (size_t)((it - a0) / I + 1) * p->nslot * p->nzx;
    };
    for (int it = a0; it < t_end;) {
        if (fused2(p) && it + 1 < t_end) {
            Fwd2Args<T> f = fwd2_args<T>(p, X, Y, Sa, Sb, it);
            io(it, &f.snap1, &f.term1);
            io(it + 1, &f.snap2, &f.term2);
            launch_fwd2<T>(p->w_fwd2, g, f, a.padnz, a.padnx, nb_, s);
            PW_HIPS(hipGetLastError());
            T* oX = X;
            T* oY = Y;
            X = Sb; Y = Sa; Sa = oX; Sb = oY;
            it += 2;
            continue;
        }
        StepArgs<T> s_ = base_args<T>(p);
        s_.X = X; s_.Y = Y; s_.it = it;
        s_.fsrc = step_src<T>(p, it);
        io(it, &s_.snap, &s_.full);
        if (visco) {
            s_.tau = p->tau_pad; s_.taus = p->taus_pad; s_.r1 = Ra; s_.r2 = Rb;
            s_.idt = (float)(1. / (double)a.dt);
            launch_fwd<T, true>(p->w_fwdv, g, s_, a.padnz, a.padnx, nb_, s);
            std::swap(Ra, Rb);
        } else {
            launch_fwd<T, false>(s_.full ? p->w_fwdt : p->w_fwd, g, s_, a.padnz, a.padnx, nb_, s);
        }
        PW_HIPS(hipGetLastError());
        std::swap(X, Y);
        ++it;
    }
    return PW_OK;
}

// exact-adjoint sweep of segment sg with two steps per launch (an odd step runs alone); the two
// free adjoint fields p->adj_sa, p->adj_sb rotate with A and B
int adjoint_segment2(pw_fwi* p, int nb_, int sg, float** Ap, float** Bp)
{
    const Acq& a = p->acq;
    hipStream_t s = p->ctx->s();
    FGeo<float> g = geo<float>(p);
    const long long fld = (long long)p->nslot * p->fs;
    const int K = p->K;
    const int a0 = sg * K, b0 = std::min(a0 + K, a.nt);
    float* A = *Ap;
    float* B = *Bp;
    float* Sa = (float*)p->adj_sa;
    float* Sb = (float*)p->adj_sb;
    const float* seg = (const float*)p->seg;
    for (int it = b0 - 1; it >= a0;) {
        if (it - 1 >= a0) {
            Adj2Args<float> r = adj2_args(p, A, B, Sa, Sb, seg + (size_t)(it - a0) * fld,
                                          seg + (size_t)(it - 1 - a0) * fld, it);
            launch_adj2(p->w_adj2, g, r, a.padnz, a.padnx, nb_, s);
            PW_HIPS(hipGetLastError());
            float* oA = A;
            float* oB = B;
            A = Sb; B = Sa; Sa = oA; Sb = oB;  // (q_{t-1}, q_t) lead the next launch
            it -= 2;
            continue;
        }
        StepArgs<float> s_ = base_args<float>(p);
        s_.X = A; s_.Y = B; s_.it = it;
        s_.fsrc = step_src<float>(p, it);
        s_.inj = (const float*)p->res;
        ExArgs<float> ex{seg + (size_t)(it - a0) * fld, p->fs, (float*)p->gm};
        launch_exact<float>(p->w_exact, g, s_, ex, a.padnz, a.padnx, nb_, s);
        PW_HIPS(hipGetLastError());
        std::swap(A, B);
        --it;
    }
    *Ap = A;
    *Bp = B;
    p->adj_sa = Sa;
    p->adj_sb = Sb;
    return PW_OK;
}

// backward sweep for segment sg: steps b0-1 .. a0
template <typename T>
int adjoint_segment(pw_fwi* p, int nb_, int sg, int mode, T** Ap, T** Bp, T* grad_ref, long long gstride)
{
    if constexpr (sizeof(T) == 4) {
        if (mode == 1 && fused2_adj(p)) return adjoint_segment2(p, nb_, sg, Ap, Bp);
    }
    const Acq& a = p->acq;
    hipStream_t s = p->ctx->s();
    FGeo<T> g = geo<T>(p);
    const long long fld = (long long)p->nslot * p->fs;
    const int I = a.interval, K = p->K;
    const int a0 = sg * K, b0 = std::min(a0 + K, a.nt);
    const int wnt = (a.nt - 1) / I + 1;
    T* A = *Ap;
    T* B = *Bp;
    if (mode == 0 && sg > 0) {  // snapshot at a0 - I into slot 0
        PW_HIPS(hipMemcpyAsync(p->seg, (T*)p->ckpt_snap + (size_t)sg * p->nslot * p->nzx,
                               (size_t)p->nslot * p->nzx * sizeof(T), hipMemcpyDeviceToDevice, s));
    }
    for (int it = b0 - 1; it >= a0; --it) {
        StepArgs<T> s_ = base_args<T>(p);
        s_.X = A; s_.Y = B; s_.it = it;
        s_.fsrc = step_src<T>(p, it);
        s_.inj = (const T*)p->res;
        if (mode == 0) {
            ImgArgs<T> im{};
            const int wit = it / I;
            if (it % I == 0 && wit != wnt - 1 && wit != 0) {
                const long long ss = (long long)p->nslot * p->nzx;
                const int k = (it - a0) / I + 1;  // slot index of snapshot 'it'
                im.w_prev = (const T*)p->seg + (size_t)(k - 1) * ss;
                im.w_cur = (const T*)p->seg + (size_t)k * ss;
                im.w_next = (const T*)p->seg + (size_t)(k + 1) * ss;
                im.snaps = p->nzx;
                im.tmpv = p->tmpv;
                im.grad = grad_ref;
                im.gstride = gstride;
            }
            if constexpr (sizeof(T) == 4) {  // the reference mode runs in f32 (checked by the caller)
                if (p->visco) {
                    ViscoBack vb{p->tau_pad, p->taus_pad, p->RB0, p->RB1, (float)(1. / (double)a.dt)};
                    launch_ref<true>(p->w_refv, g, s_, im, vb, a.padnz, a.padnx, nb_, s);
                    std::swap(p->RB0, p->RB1);
                } else {
                    launch_ref<false>(p->w_ref, g, s_, im, ViscoBack{}, a.padnz, a.padnx, nb_, s);
                }
            } else {
                return PW_ERR_ARG;
            }
        } else {
            ExArgs<T> ex{};
            ex.F = (const T*)p->seg + (size_t)(it - a0) * fld;
            ex.fsl = p->fs;
            ex.gm = (T*)p->gm;
            launch_exact<T>(p->w_exact, g, s_, ex, a.padnz, a.padnx, nb_, s);
        }
        PW_HIPS(hipGetLastError());
        std::swap(A, B);
    }
    *Ap = A;
    *Bp = B;
    return PW_OK;
}

template <typename T>
int reduce_sum(pw_fwi* p, double* out_host)
{
    (void)sizeof(T);
    hipStream_t s = p->ctx->s();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1), 0, s, 1024, p->partial, p->dscalar, 0);
    PW_HIPS(hipGetLastError());
    PW_HIPS(hipMemcpyAsync(out_host, p->dscalar, 8, hipMemcpyDeviceToHost, s));
    PW_HIPS(hipStreamSynchronize(s));
    return PW_OK;
}

// residual for the batch: mode 0: windowed d - p (reference); mode 1: p - d (gradient of 0.5||p-d||^2)
template <typename T>
int residual_batch(pw_fwi* p, int nb_, int mode, double* fc)
{
    const Acq& a = p->acq;
    hipStream_t s = p->ctx->s();
    const long long per = (long long)a.nt * a.nr;
    PW_HIPS(hipMemsetAsync(p->partial, 0, 1024 * 8, s));
    hipLaunchKernelGGL(k_residual<T>, dim3(1024), dim3(256), 0, s, nb_, per, (const T*)p->obs, (const T*)p->rec,
                       (const T*)p->W, (T*)p->res, (const int*)p->slot_shot, mode == 0 ? 1 : 0, mode == 0 ? 1 : -1,
                       p->partial);
    PW_HIPS(hipGetLastError());
    double v = 0;
    int rc = reduce_sum<T>(p, &v);
    *fc += v;
    return rc;
}

// full gradient over all shots.  mode 0 -> float grad in p->gtot (interior, reference order);
// mode 1 -> T grad in p->gtot.  fcost: FP64 misfit (mode 0: 0.5||d-p||^2 unweighted as the
// reference reports; mode 1 the same).
template <typename T>
int gradient_all(pw_fwi* p, int mode, double* fcost)
{
    const Acq& a = p->acq;
    hipStream_t s = p->ctx->s();
    if (!p->have_wav || !p->have_obs) PW_FAIL(p->ctx, PW_ERR_STATE, "gradient: wavelet or observed data not set");
    if (mode == 0 && p->prec != 4) PW_FAIL(p->ctx, PW_ERR_ARG, "gradient: the reference mode runs in f32");
    if (p->visco && mode != 0)
        PW_FAIL(p->ctx, PW_ERR_ARG, "gradient: visco-acoustic plans take the reference gradient (gradient_v) only");
    const long long fld = (long long)p->nslot * p->fs;
    int rc = 0;
    double fc = 0.0;
    if (mode == 0) rc = fill<float>(p, p->gtot, p->nzx, 0.f);
    else rc = fill<T>(p, p->gtot, p->nzx, T(0));
    if (rc) return rc;
    EventSet ev;
    PW_HIPS(ev.create());
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1], e2 = ev.e[2];
    float fwd_ms = 0, bwd_ms = 0;
    const bool store_last = true;
    for (int b0 = 0; b0 < a.ns; b0 += p->nslot) {
        const int nb_ = std::min(p->nslot, a.ns - b0);
        if ((rc = set_batch(p, b0, nb_))) return rc;
        PW_HIPS(hipEventRecord(e0, s));
        if ((rc = forward_batch<T>(p, nb_, true, mode, store_last))) return rc;
        if ((rc = residual_batch<T>(p, nb_, mode, &fc))) return rc;
        PW_HIPS(hipEventRecord(e1, s));
        if ((rc = fill<T>(p, p->A, fld, T(0))) || (rc = fill<T>(p, p->B, fld, T(0)))) return rc;
        if (p->visco && sizeof(T) == 4 &&
            ((rc = fill<float>(p, p->RB0, fld, 0.f)) || (rc = fill<float>(p, p->RB1, fld, 0.f))))
            return rc;
        T* grad_ref = nullptr;
        long long gstride = 0;
        if (mode == 0) {
            if (p->nslot == 1) {
                grad_ref = (T*)p->gtot;  // one shot per launch: accumulate in the reference's order
            } else {
                if ((rc = fill<T>(p, p->gslot, (long long)p->nslot * p->nzx, T(0)))) return rc;
                grad_ref = (T*)p->gslot;
                gstride = p->nzx;
            }
        } else {
            if ((rc = fill<T>(p, p->gm, fld, T(0)))) return rc;
        }
        T* A = (T*)p->A;
        T* B = (T*)p->B;
        p->adj_sa = p->A2;
        p->adj_sb = p->B2;
        for (int sg = p->nseg - 1; sg >= 0; --sg) {
            if (!(store_last && sg == p->nseg - 1))
                if ((rc = recompute_segment<T>(p, nb_, sg, mode))) return rc;
            if ((rc = adjoint_segment<T>(p, nb_, sg, mode, &A, &B, grad_ref, gstride))) return rc;
        }
        if (mode == 0 && p->nslot > 1) {
            hipLaunchKernelGGL(k_add_slots<T>, dim3(nblk(p->nzx)), dim3(256), 0, s, p->nzx, nb_, (const T*)p->gslot,
                               p->nzx, (T*)p->gtot);
        } else if (mode == 1) {
            FGeo<T> g = geo<T>(p);
            hipLaunchKernelGGL(k_gm_fold_z<T>, dim3((a.nz + 255) / 256, a.padnx), dim3(256), 0, s, g, nb_,
                               (const T*)p->gm, (const T*)p->vpad, dt2_of<T>(p), p->gfold);
            hipLaunchKernelGGL(k_gm_fold_x<T>, dim3((a.nz + 255) / 256, a.nx), dim3(256), 0, s, g,
                               (const double*)p->gfold, (T*)p->gtot);
        }
        PW_HIPS(hipGetLastError());
        PW_HIPS(hipEventRecord(e2, s));
        PW_HIPS(hipEventSynchronize(e2));
        float m1 = 0, m2 = 0;
        hipEventElapsedTime(&m1, e0, e1);
        hipEventElapsedTime(&m2, e1, e2);
        fwd_ms += m1;
        bwd_ms += m2;
    }
    p->t_forward_ms = fwd_ms;
    p->dev_timed_fwd = p->dev_timed_adj = false;
    p->t_backward_ms = bwd_ms;
    *fcost = fc;
    p->last_fcost = fc;
    return PW_OK;
}

// reference post-processing of the mode-0 gradient: first-call scaling, smoothing, *scaling
int finish_reference_gradient(pw_fwi* p, float* scaling_io)
{
    const Acq& a = p->acq;
    hipStream_t s = p->ctx->s();
    float* g = (float*)p->gtot;
    float sc = scaling_io ? *scaling_io : p->scaling;
    if (!(sc > 0.f)) {
        hipLaunchKernelGGL(k_absmax_partial<float>, dim3(1024), dim3(256), 0, s, p->nzx, g, p->partial);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1), 0, s, 1024, p->partial, p->dscalar, 1);
        PW_HIPS(hipGetLastError());
        double dmax = 0;
        PW_HIPS(hipMemcpyAsync(&dmax, p->dscalar, 8, hipMemcpyDeviceToHost, s));
        PW_HIPS(hipStreamSynchronize(s));
        float fd = (float)dmax;  // max |g| is exact in f32
        sc = 0.1 / fd;
        p->scaling = sc;
    }
    if (scaling_io) *scaling_io = sc;
    // gradient_smooth2 with rect 3 x 3 (pfwi.c:858-859): along z for every column, then along x
    const int r = 3;
    hipLaunchKernelGGL(k_tri_smooth2, dim3((a.nx + 63) / 64), dim3(64), 0, s, g, a.nx, a.nz, 1, a.nz, r, p->work);
    hipLaunchKernelGGL(k_tri_smooth2, dim3((a.nz + 63) / 64), dim3(64), 0, s, g, a.nz, 1, a.nz, a.nx, r, p->work);
    hipLaunchKernelGGL(k_scale_f, dim3(nblk(p->nzx)), dim3(256), 0, s, p->nzx, g, sc);
    PW_HIPS(hipGetLastError());
    return PW_OK;
}

// Result watchdog of the gradient entry points: a blown-up run (time step beyond the CFL limit)
// leaves an Inf/NaN misfit or gradient.  The L-BFGS line search calls gradient_all directly and
// handles such trial models as the reference does (its Wolfe tests fail on them).
int finite_result(pw_fwi* p, double fc, const char* what)
{
    if (!std::isfinite(fc))
        PW_FAIL(p->ctx, PW_ERR_NONFINITE,
                std::string(what) + ": the misfit is Inf or NaN (the run blew up; is the time step within the CFL limit?)");
    return pw::check_finite(p->ctx, p->gtot, (size_t)p->nzx, (int)esz(p), what);
}

int fingerprint(pw_fwi* p, const void* d_vel, unsigned long long* out)
{
    hipStream_t s = p->ctx->s();
    const unsigned blocks = (unsigned)std::min<long long>(256, (p->nzx + 255) / 256);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, out);
    if (p->prec == 4)
        hipLaunchKernelGGL(k_fingerprint<float>, dim3(blocks), dim3(256), 0, s, p->nzx, (const float*)d_vel, out);
    else
        hipLaunchKernelGGL(k_fingerprint<double>, dim3(blocks), dim3(256), 0, s, p->nzx, (const double*)d_vel, out);
    PW_HIPS(hipGetLastError());
    return PW_OK;
}

template <typename T>
int dot_dev(pw_fwi* p, const void* x, const void* y, double* out)
{
    hipStream_t s = p->ctx->s();
    hipLaunchKernelGGL(k_dot_partial<T>, dim3(1024), dim3(256), 0, s, p->nzx, (const T*)x, (const T*)y, p->partial);
    PW_HIPS(hipGetLastError());
    return reduce_sum<T>(p, out);
}

}  // namespace

// ==================================================================================== C ABI
extern "C" {

pw_fwi* pw_fwi_create(pw_ctx* ctx, const float* pararray, int batch, int precision)
{
    if (!ctx || !pararray) return nullptr;
    if (precision != 4 && precision != 8) {
        ctx->err = "pw_fwi_create: precision must be 4 (f32) or 8 (f64)";
        return nullptr;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) {
        ctx->err = "hipSetDevice failed";
        return nullptr;
    }
    pw_fwi* p = new pw_fwi;
    p->ctx = ctx;
    p->prec = precision;
    std::string err;
    if (acq_from_pararray(pararray, p->acq, err)) {
        ctx->err = err;
        delete p;
        return nullptr;
    }
    Acq& a = p->acq;
    p->nslot = std::max(1, std::min(batch <= 0 ? a.ns : batch, a.ns));
    p->ldz = (a.padnz + 31) / 32 * 32;
    p->fs = (long long)p->ldz * a.padnx;
    p->nzx = (long long)a.nz * a.nx;
    if ((double)(p->fs + 8 * p->ldz + 8) * precision >= 2147483648.0) {  // 32-bit buffer windows per field
        ctx->err = "pw_fwi_create: padded model too large for one field window (2 GiB)";
        delete p;
        return nullptr;
    }
    // geometry per shot
    p->shot_host.resize(a.ns);
    p->patch_host.assign((size_t)a.ns * PMAX * PMAX, 0.f);
    // point-source geometry; an unusable one is reported by the first call that would use it
    // (a passive source field set with pw_fwi_set_source_field replaces it)
    for (int is = 0; is < a.ns; ++is) {
        ShotInfo& sh = p->shot_host[is];
        const int sx = a.s0_v + is * a.ds_v;
        if (source_patch(a, sx, 2, sh, &p->patch_host[(size_t)is * PMAX * PMAX], err)) {
            if (p->active_err.empty()) p->active_err = "pw_fwi: " + err;
            sh.pw = sh.ph = 0;
        }
        sh.r0_v = a.r0_v[is]; sh.r02 = a.r02[is]; sh.nr2 = a.nr2[is]; sh.shot = is;
        int rlast = sh.r0_v + (sh.nr2 - 1) * a.dr_v;
        if (sh.nr2 > 0 && (sh.r0_v < 4 || rlast >= a.padnx - 4)) {
            if (p->active_err.empty()) p->active_err = "pw_fwi: receivers reach the padded grid's 4-cell rim";
            sh.nr2 = 0;
        }
    }
    p->shot_active = p->shot_host;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)8 << 30;
    // checkpoint and segment storage may take most of the device memory free at creation (the
    // forward fields of every step stay in HBM when they fit: no recomputation)
    p->budget = free_b / 10 * 8;
    const size_t e = precision;
    const long long fld = (long long)p->nslot * p->fs;
    const long long tr = (long long)a.nt * a.nr;
    int rc = 0;
    rc = rc ? rc : dmalloc(p, &p->vel, p->nzx * e);
    rc = rc ? rc : dmalloc(p, &p->vpad, p->fs * e);
    rc = rc ? rc : dmalloc(p, &p->m, p->fs * e);
    rc = rc ? rc : dmalloc(p, &p->sfz, 2 * (size_t)a.padnz * e);
    rc = rc ? rc : dmalloc(p, &p->sfx, 2 * (size_t)a.padnx * e);
    rc = rc ? rc : dmalloc(p, &p->wav, a.nt * e);
    rc = rc ? rc : dmalloc(p, &p->X, fld * e);
    rc = rc ? rc : dmalloc(p, &p->Y, fld * e);
    rc = rc ? rc : dmalloc(p, &p->A, fld * e);
    rc = rc ? rc : dmalloc(p, &p->B, fld * e);
    rc = rc ? rc : dmalloc(p, &p->rec, p->nslot * tr * e);
    rc = rc ? rc : dmalloc(p, &p->res, p->nslot * tr * e);
    rc = rc ? rc : dmalloc(p, &p->obs, (size_t)a.ns * tr * e);
    rc = rc ? rc : dmalloc(p, &p->W, tr * e);
    rc = rc ? rc : dmalloc(p, &p->gm, fld * e);
    rc = rc ? rc : dmalloc(p, &p->gslot, (size_t)p->nslot * p->nzx * e);
    rc = rc ? rc : dmalloc(p, &p->gtot, p->nzx * 8);
    rc = rc ? rc : dmalloc(p, (void**)&p->gfold, (size_t)a.nz * a.padnx * 8);
    rc = rc ? rc : dmalloc(p, &p->patch, (size_t)p->nslot * PMAX * PMAX * e);
    rc = rc ? rc : dmalloc(p, (void**)&p->tmpv, p->nzx * 4);
    rc = rc ? rc : dmalloc(p, (void**)&p->work, (size_t)std::max(a.nz, a.nx) * (std::max(a.nz, a.nx) + 8) * 4);
    rc = rc ? rc : dmalloc(p, (void**)&p->shots, sizeof(ShotInfo) * p->nslot);
    rc = rc ? rc : dmalloc(p, (void**)&p->slot_shot, sizeof(int) * p->nslot);
    rc = rc ? rc : dmalloc(p, (void**)&p->partial, 1024 * 8);
    rc = rc ? rc : dmalloc(p, (void**)&p->dscalar, 8);
    rc = rc ? rc : dmalloc(p, (void**)&p->dhash, 16);
    if (!rc) rc = plan_storage(p, 0);
    if (rc) {
        fwi_free(p);
        delete p;
        return nullptr;
    }
    // sponge pairs, residual weights
    std::vector<double> sz(2 * (size_t)a.padnz, 1.0), sx(2 * (size_t)a.padnx, 1.0);
    for (int i = 0; i < a.padnz; ++i) {
        if (i < a.nb) sz[2 * i] = a.bc[i];
        if (i >= a.padnz - a.nb) sz[2 * i + 1] = a.bc[a.padnz - i - 1];
    }
    for (int i = 0; i < a.padnx; ++i) {
        if (i < a.nb) sx[2 * i] = a.bc[i];
        if (i >= a.padnx - a.nb) sx[2 * i + 1] = a.bc[a.padnx - i - 1];
    }
    std::vector<float> w;
    residual_weights(a, w);
    std::vector<float> wt((size_t)a.nt * a.nr);  // [it][ir]
    for (int ir = 0; ir < a.nr; ++ir)
        for (int it = 0; it < a.nt; ++it) wt[(size_t)it * a.nr + ir] = w[(size_t)ir * a.nt + it];
    hipStream_t s = ctx->s();
    bool ok;
    if (precision == 4) {
        std::vector<float> fz(sz.begin(), sz.end()), fx(sx.begin(), sx.end());
        ok = hipMemcpyAsync(p->sfz, fz.data(), fz.size() * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
             hipMemcpyAsync(p->sfx, fx.data(), fx.size() * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
             hipMemcpyAsync(p->W, wt.data(), wt.size() * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
    } else {
        std::vector<double> wd(wt.begin(), wt.end());
        ok = hipMemcpyAsync(p->sfz, sz.data(), sz.size() * 8, hipMemcpyHostToDevice, s) == hipSuccess &&
             hipMemcpyAsync(p->sfx, sx.data(), sx.size() * 8, hipMemcpyHostToDevice, s) == hipSuccess &&
             hipMemcpyAsync(p->W, wd.data(), wd.size() * 8, hipMemcpyHostToDevice, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
    }
    if (!ok) {
        ctx->err = "pw_fwi_create: upload failed";
        fwi_free(p);
        delete p;
        return nullptr;
    }
    return p;
}

void pw_fwi_destroy(pw_fwi* p)
{
    if (!p) return;
    hipSetDevice(p->ctx->device);
    hipStreamSynchronize(p->ctx->s());
    fwi_free(p);
    delete p;
}

int pw_fwi_set_checkpointing(pw_fwi* p, int steps)
{
    if (p) p->dev_ckpt_valid = false;
    if (!p || steps < 0) return PW_ERR_ARG;
    return plan_storage(p, steps);
}

int pw_fwi_info(pw_fwi* p, int* ints /* 8 */)
{
    if (!p || !ints) return PW_ERR_ARG;
    const Acq& a = p->acq;
    ints[0] = a.padnz; ints[1] = a.padnx; ints[2] = p->nslot; ints[3] = p->K;
    ints[4] = p->nseg; ints[5] = a.ns; ints[6] = a.nt; ints[7] = a.nr;
    return PW_OK;
}

int pw_fwi_set_wavelet(pw_fwi* p, const float* wav)
{
    if (p) p->dev_ckpt_valid = false;
    if (!p || !wav) return PW_ERR_ARG;
    hipStream_t s = p->ctx->s();
    if (p->prec == 4) {
        PW_HIPS(hipMemcpyAsync(p->wav, wav, p->acq.nt * 4, hipMemcpyHostToDevice, s));
    } else {
        std::vector<double> w(wav, wav + p->acq.nt);
        PW_HIPS(hipMemcpyAsync(p->wav, w.data(), w.size() * 8, hipMemcpyHostToDevice, s));
    }
    PW_HIPS(hipStreamSynchronize(s));
    p->have_wav = true;
    return PW_OK;
}

// data: (nt, nr, ns) Fortran order = [is][ir][it]; stored time-major [is][it][ir]
int pw_fwi_set_observed(pw_fwi* p, const float* data)
{
    if (!p || !data) return PW_ERR_ARG;
    const Acq& a = p->acq;
    hipStream_t s = p->ctx->s();
    const long long per = (long long)a.nt * a.nr;
    std::vector<float> t((size_t)per);
    std::vector<double> td;
    for (int is = 0; is < a.ns; ++is) {
        const float* d = data + (size_t)is * per;
        for (int ir = 0; ir < a.nr; ++ir)
            for (int it = 0; it < a.nt; ++it) t[(size_t)it * a.nr + ir] = d[(size_t)ir * a.nt + it];
        if (p->prec == 4) {
            PW_HIPS(hipMemcpyAsync((float*)p->obs + (size_t)is * per, t.data(), per * 4, hipMemcpyHostToDevice, s));
            PW_HIPS(hipStreamSynchronize(s));
        } else {
            td.assign(t.begin(), t.end());
            PW_HIPS(hipMemcpyAsync((double*)p->obs + (size_t)is * per, td.data(), per * 8, hipMemcpyHostToDevice, s));
            PW_HIPS(hipStreamSynchronize(s));
        }
    }
    p->have_obs = true;
    return PW_OK;
}

}  // extern "C"

namespace {

// switch the plan to the passive source field already in p->fsrc (or back to point sources)
int use_source_field(pw_fwi* p, bool on)
{
    p->dev_ckpt_valid = false;
    if (!p->widths_forced) p->widths_tuned = false;  // the tiles that run the general loop change
    const Acq& a = p->acq;
    if (!on) {
        p->passive = false;
        p->shot_host = p->shot_active;
        return PW_OK;
    }
    if (p->visco) PW_FAIL(p->ctx, PW_ERR_ARG, "source field: the passive path is acoustic (pw_fwi_set_visco(p, NULL))");
    // gradient_pas_av (pywave/src/wave_fwigradient.c:710-713, :807-810): every shot records and
    // injects at r0_v[0] + ir*dr_v, ir < nr; no point source
    const int rlast = a.r0_v[0] + (a.nr - 1) * a.dr_v;
    if (a.r0_v[0] < 4 || rlast >= a.padnx - 4)
        PW_FAIL(p->ctx, PW_ERR_ARG, "pw_fwi_set_source_field: receivers reach the padded grid's 4-cell rim");
    for (int is = 0; is < a.ns; ++is) {
        ShotInfo& sh = p->shot_host[is];
        sh.sx0 = sh.sz0 = 0; sh.pw = sh.ph = 0;
        sh.r0_v = a.r0_v[0]; sh.r02 = 0; sh.nr2 = a.nr; sh.shot = is;
    }
    p->passive = true;
    return PW_OK;
}

int alloc_source_field(pw_fwi* p)
{
    if ((double)p->nzx * esz(p) >= 2147483648.0)
        PW_FAIL(p->ctx, PW_ERR_ARG, "pw_fwi_set_source_field: one source-field plane exceeds 2 GiB");
    if (!p->fsrc) return dmalloc(p, &p->fsrc, (size_t)p->acq.ns * p->acq.nt * p->nzx * esz(p));
    return PW_OK;
}

}  // namespace

extern "C" {

// passive source field (pfwi mode 4, gradient_pas_init's ww3): src (nz, nx, nt, ns) Fortran order,
// i.e. [is][it][ix][iz]; NULL returns the plan to its point sources
int pw_fwi_set_source_field(pw_fwi* p, const float* src)
{
    if (!p) return PW_ERR_ARG;
    if (!src) return use_source_field(p, false);
    int rc = alloc_source_field(p);
    if (rc) return rc;
    const Acq& a = p->acq;
    hipStream_t s = p->ctx->s();
    const size_t per = (size_t)a.nt * p->nzx;
    if (p->prec == 4) {
        PW_HIPS(hipMemcpyAsync(p->fsrc, src, per * a.ns * 4, hipMemcpyHostToDevice, s));
    } else {
        std::vector<double> d(per);
        for (int is = 0; is < a.ns; ++is) {
            for (size_t i = 0; i < per; ++i) d[i] = src[is * per + i];
            PW_HIPS(hipMemcpyAsync((double*)p->fsrc + is * per, d.data(), per * 8, hipMemcpyHostToDevice, s));
            PW_HIPS(hipStreamSynchronize(s));
        }
    }
    PW_HIPS(hipStreamSynchronize(s));
    return use_source_field(p, true);
}

// the same from device memory of the plan's precision (stream-ordered, no host synchronisation)
int pw_fwi_set_source_field_dev(pw_fwi* p, const void* d_src)
{
    if (!p) return PW_ERR_ARG;
    if (!d_src) return use_source_field(p, false);
    int rc = alloc_source_field(p);
    if (rc) return rc;
    PW_HIPS(hipMemcpyAsync(p->fsrc, d_src, (size_t)p->acq.ns * p->acq.nt * p->nzx * esz(p),
                           hipMemcpyDeviceToDevice, p->ctx->s()));
    return use_source_field(p, true);
}

// visco-acoustic plan with the quality factor q (nz, nx) Fortran order and the pararray's f0: the
// forward follows forward_modeling (pywave/src/wave_fwimodeling.c:116-220), the reference gradient
// and L-BFGS follow gradient_v (pywave/src/wave_fwigradient.c:264-452, fwi() with media != 1);
// f32 plans; the exact gradient and the device entry points refuse it.  NULL returns to acoustic.
int pw_fwi_set_visco(pw_fwi* p, const float* q)
{
    if (!p) return PW_ERR_ARG;
    p->dev_ckpt_valid = false;
    if (!p->widths_forced) p->widths_tuned = false;  // other kernels run
    if (!q) {
        p->visco = false;
        return PW_OK;
    }
    if (p->prec != 4) PW_FAIL(p->ctx, PW_ERR_ARG, "pw_fwi_set_visco: visco-acoustic modelling runs in f32");
    if (p->passive) PW_FAIL(p->ctx, PW_ERR_ARG, "pw_fwi_set_visco: the passive path is acoustic (drop the source field)");
    const Acq& a = p->acq;
    const long long nzx = p->nzx;
    // tau, taus (preparation, pywave/src/wave_fwiutil.c:91-97): f32 with the 2 pi f0 q product in double
    const double PI = 3.14159265358979323846264338328;
    std::vector<float> ta(nzx), ts(nzx);
    for (long long i = 0; i < nzx; ++i) {
        const float tmp = sqrtf(q[i] * q[i] + 1);
        const float taue = (tmp + 1) / (2. * PI * a.f0 * q[i]);
        ts[i] = (tmp - 1) / (2. * PI * a.f0 * q[i]);
        ta[i] = taue / ts[i] - 1;
    }
    // padded with the edge values (pad2d, wave_fwiutil.c:123-147), z stride ldz
    std::vector<float> tap((size_t)p->fs, 0.f), tsp((size_t)p->fs, 1.f);
    for (int ix = 0; ix < a.padnx; ++ix) {
        const int sx = std::min(std::max(ix - a.nb, 0), a.nx - 1);
        for (int iz = 0; iz < a.padnz; ++iz) {
            const int sz = std::min(std::max(iz - a.nb, 0), a.nz - 1);
            tap[(size_t)ix * p->ldz + iz] = ta[(size_t)sx * a.nz + sz];
            tsp[(size_t)ix * p->ldz + iz] = ts[(size_t)sx * a.nz + sz];
        }
    }
    int rc = PW_OK;
    const long long fld = (long long)p->nslot * p->fs;
    if (!p->tau_pad) rc = dmalloc(p, (void**)&p->tau_pad, p->fs * 4);
    if (!rc && !p->taus_pad) rc = dmalloc(p, (void**)&p->taus_pad, p->fs * 4);
    if (!rc && !p->R0) rc = dmalloc(p, (void**)&p->R0, fld * 4);
    if (!rc && !p->R1) rc = dmalloc(p, (void**)&p->R1, fld * 4);
    if (!rc && !p->RB0) rc = dmalloc(p, (void**)&p->RB0, fld * 4);
    if (!rc && !p->RB1) rc = dmalloc(p, (void**)&p->RB1, fld * 4);
    if (!rc && !p->ckpt_r) rc = dmalloc(p, (void**)&p->ckpt_r, (size_t)p->nseg * fld * 4);
    if (rc) return rc;
    hipStream_t s = p->ctx->s();
    PW_HIPS(hipMemcpyAsync(p->tau_pad, tap.data(), tap.size() * 4, hipMemcpyHostToDevice, s));
    PW_HIPS(hipMemcpyAsync(p->taus_pad, tsp.data(), tsp.size() * 4, hipMemcpyHostToDevice, s));
    PW_HIPS(hipStreamSynchronize(s));
    p->visco = true;
    return PW_OK;
}

// forward_modeling_a: data_out (nt, nr, ns) Fortran order
int pw_fwi_forward(pw_fwi* p, const float* vel, float* data_out)
{
    if (!p || !vel || !data_out) return PW_ERR_ARG;
    if (!p->have_wav) PW_FAIL(p->ctx, PW_ERR_STATE, "forward: wavelet not set");
    const Acq& a = p->acq;
    hipStream_t s = p->ctx->s();
    int rc = upload_vel_host(p, vel);
    if (!rc) rc = ensure_widths(p);
    if (rc) return rc;
    const long long per = (long long)a.nt * a.nr;
    std::vector<float> t((size_t)per * p->nslot);
    std::vector<double> td;
    EventSet ev;
    PW_HIPS(ev.create());
    float fwd_ms = 0;
    for (int b0 = 0; b0 < a.ns; b0 += p->nslot) {
        const int nb_ = std::min(p->nslot, a.ns - b0);
        if ((rc = set_batch(p, b0, nb_))) return rc;
        PW_HIPS(hipEventRecord(ev.e[0], s));
        rc = p->prec == 4 ? forward_batch<float>(p, nb_, false, 0, false) : forward_batch<double>(p, nb_, false, 0, false);
        if (rc) return rc;
        PW_HIPS(hipEventRecord(ev.e[1], s));
        PW_HIPS(hipEventSynchronize(ev.e[1]));
        float m = 0;
        PW_HIPS(hipEventElapsedTime(&m, ev.e[0], ev.e[1]));
        fwd_ms += m;
        if ((rc = pw::check_finite(p->ctx, p->rec, (size_t)nb_ * per, (int)esz(p), "fwi forward: modelled data")))
            return rc;
        if (p->prec == 4) {
            PW_HIPS(hipMemcpyAsync(t.data(), p->rec, (size_t)nb_ * per * 4, hipMemcpyDeviceToHost, s));
            PW_HIPS(hipStreamSynchronize(s));
        } else {
            td.resize((size_t)nb_ * per);
            PW_HIPS(hipMemcpyAsync(td.data(), p->rec, (size_t)nb_ * per * 8, hipMemcpyDeviceToHost, s));
            PW_HIPS(hipStreamSynchronize(s));
            for (size_t i = 0; i < td.size(); ++i) t[i] = (float)td[i];
        }
        for (int k = 0; k < nb_; ++k) {
            float* d = data_out + (size_t)(b0 + k) * per;
            const float* src = t.data() + (size_t)k * per;
            for (int it = 0; it < a.nt; ++it)
                for (int ir = 0; ir < a.nr; ++ir) d[(size_t)ir * a.nt + it] = src[(size_t)it * a.nr + ir];
        }
    }
    p->t_forward_ms = fwd_ms;
    p->dev_timed_fwd = p->dev_timed_adj = false;
    p->t_backward_ms = 0;
    return PW_OK;
}

// one gradient over all shots.  mode 0: gradient_av (scaling_io: in <= 0 -> first call, out the
// scaling used; NULL -> the plan remembers its first-call scaling).  mode 1: exact adjoint of
// 0.5||F(v)-d||^2.  fcost: FP64 misfit 0.5 sum (d-p)^2.
int pw_fwi_gradient(pw_fwi* p, const float* vel, int mode, float* grad, double* fcost, float* scaling_io)
{
    if (!p || !vel || !grad || (mode != 0 && mode != 1)) return PW_ERR_ARG;
    hipStream_t s = p->ctx->s();
    int rc = upload_vel_host(p, vel);
    if (!rc) rc = ensure_widths(p);
    if (rc) return rc;
    double fc = 0;
    if (p->prec == 4) rc = gradient_all<float>(p, mode, &fc);
    else rc = gradient_all<double>(p, mode, &fc);
    if (rc) return rc;
    if ((rc = finite_result(p, fc, "fwi gradient"))) return rc;
    if (mode == 0) {
        if ((rc = finish_reference_gradient(p, scaling_io))) return rc;
        PW_HIPS(hipMemcpyAsync(grad, p->gtot, p->nzx * 4, hipMemcpyDeviceToHost, s));
    } else if (p->prec == 4) {
        PW_HIPS(hipMemcpyAsync(grad, p->gtot, p->nzx * 4, hipMemcpyDeviceToHost, s));
    } else {
        std::vector<double> gd(p->nzx);
        PW_HIPS(hipMemcpyAsync(gd.data(), p->gtot, p->nzx * 8, hipMemcpyDeviceToHost, s));
        PW_HIPS(hipStreamSynchronize(s));
        for (long long i = 0; i < p->nzx; ++i) grad[i] = (float)gd[i];
    }
    PW_HIPS(hipStreamSynchronize(s));
    if (fcost) *fcost = fc;
    return PW_OK;
}

// f64 gradient output (for the finite-difference check)
int pw_fwi_gradient_f64(pw_fwi* p, const double* vel, int mode, double* grad, double* fcost)
{
    if (!p || !vel || !grad || p->prec != 8 || mode != 1) return PW_ERR_ARG;
    hipStream_t s = p->ctx->s();
    PW_HIPS(hipMemcpyAsync(p->vel, vel, p->nzx * 8, hipMemcpyHostToDevice, s));
    int rc = upload_model<double>(p, nullptr, false);
    if (!rc) rc = ensure_widths(p);
    double fc = 0;
    if (!rc) rc = gradient_all<double>(p, mode, &fc);
    if (!rc) rc = finite_result(p, fc, "fwi gradient");
    if (rc) return rc;
    PW_HIPS(hipMemcpyAsync(grad, p->gtot, p->nzx * 8, hipMemcpyDeviceToHost, s));
    PW_HIPS(hipStreamSynchronize(s));
    if (fcost) *fcost = fc;
    return PW_OK;
}

// misfit only (forward sweep + FP64 reduction), f64 model input
int pw_fwi_misfit_f64(pw_fwi* p, const double* vel, double* fcost)
{
    if (!p || !vel || !fcost) return PW_ERR_ARG;
    if (!p->have_wav || !p->have_obs) PW_FAIL(p->ctx, PW_ERR_STATE, "misfit: wavelet or observed data not set");
    hipStream_t s = p->ctx->s();
    int rc;
    if (p->prec == 8) {
        PW_HIPS(hipMemcpyAsync(p->vel, vel, p->nzx * 8, hipMemcpyHostToDevice, s));
        rc = upload_model<double>(p, nullptr, false);
    } else {
        std::vector<float> v(vel, vel + p->nzx);
        rc = upload_vel_host(p, v.data());
    }
    if (!rc) rc = ensure_widths(p);
    if (rc) return rc;
    double fc = 0;
    const Acq& a = p->acq;
    for (int b0 = 0; b0 < a.ns; b0 += p->nslot) {
        const int nb_ = std::min(p->nslot, a.ns - b0);
        if ((rc = set_batch(p, b0, nb_))) return rc;
        rc = p->prec == 4 ? forward_batch<float>(p, nb_, false, 1, false) : forward_batch<double>(p, nb_, false, 1, false);
        if (!rc) rc = p->prec == 4 ? residual_batch<float>(p, nb_, 1, &fc) : residual_batch<double>(p, nb_, 1, &fc);
        if (rc) return rc;
    }
    if (!std::isfinite(fc))
        PW_FAIL(p->ctx, PW_ERR_NONFINITE, "fwi misfit: Inf or NaN (the run blew up; is the time step within the CFL limit?)");
    *fcost = fc;
    return PW_OK;
}

// device-pointer variants for torch (T = the plan precision); traces [ns][nt][nr] time-major
int pw_fwi_forward_dev(pw_fwi* p, const void* d_vel, void* d_traces)
{
    if (!p || !d_vel || !d_traces) return PW_ERR_ARG;
    if (p->visco) PW_FAIL(p->ctx, PW_ERR_ARG, "forward_dev: visco-acoustic modelling is host-only (pw_fwi_forward)");
    if (!p->have_wav) PW_FAIL(p->ctx, PW_ERR_STATE, "forward: wavelet not set");
    const Acq& a = p->acq;
    hipStream_t s = p->ctx->s();
    int rc = p->prec == 4 ? upload_model<float>(p, (const float*)d_vel, true) : upload_model<double>(p, (const double*)d_vel, true);
    if (!rc) rc = ensure_widths(p);
    if (rc) return rc;
    if (!p->ev_dev[0])
        for (hipEvent_t& ev : p->ev_dev) PW_HIPS(hipEventCreate(&ev));
    PW_HIPS(hipEventRecord(p->ev_dev[0], s));
    const long long per = (long long)a.nt * a.nr;
    const size_t e = esz(p);
    // one batch holds every shot: keep its checkpoints for the adjoint of the same model
    const bool keep = p->nslot >= a.ns;
    for (int b0 = 0; b0 < a.ns; b0 += p->nslot) {
        const int nb_ = std::min(p->nslot, a.ns - b0);
        if ((rc = set_batch(p, b0, nb_))) return rc;
        rc = p->prec == 4 ? forward_batch<float>(p, nb_, keep, 1, keep) : forward_batch<double>(p, nb_, keep, 1, keep);
        if (rc) return rc;
        PW_HIPS(hipMemcpyAsync((char*)d_traces + (size_t)b0 * per * e, p->rec, (size_t)nb_ * per * e,
                               hipMemcpyDeviceToDevice, s));
    }
    if (keep && (rc = fingerprint(p, d_vel, p->dhash))) return rc;  // of the model the checkpoints belong to
    PW_HIPS(hipEventRecord(p->ev_dev[1], s));
    p->dev_timed_fwd = true;
    p->dev_ckpt_valid = keep;
    p->dev_ckpt_vel = d_vel;
    return PW_OK;
}

// gradient of <g, F(v)> w.r.t. v for traces cotangent g [ns][nt][nr] (exact adjoint)
int pw_fwi_adjoint_dev(pw_fwi* p, const void* d_vel, const void* d_gtraces, void* d_gvel)
{
    if (!p || !d_vel || !d_gtraces || !d_gvel) return PW_ERR_ARG;
    if (!p->have_wav) PW_FAIL(p->ctx, PW_ERR_STATE, "adjoint: wavelet not set");
    if (p->visco) PW_FAIL(p->ctx, PW_ERR_ARG, "adjoint_dev: the exact adjoint is acoustic (visco plans: reference gradient)");
    const Acq& a = p->acq;
    hipStream_t s = p->ctx->s();
    if (!p->ev_dev[0])
        for (hipEvent_t& ev : p->ev_dev) PW_HIPS(hipEventCreate(&ev));
    PW_HIPS(hipEventRecord(p->ev_dev[2], s));
    // checkpoints of this very model from the preceding forward_dev (one batch): no forward sweep.
    // The same pointer is not enough (a caller may have changed the model in place): the model's
    // fingerprint must equal the one taken by forward_dev.
    bool reuse = p->dev_ckpt_valid && p->dev_ckpt_vel == d_vel && p->nslot >= a.ns;
    if (reuse) {
        if (int r = fingerprint(p, d_vel, p->dhash + 1)) return r;
        unsigned long long h[2] = {0, 1};
        PW_HIPS(hipMemcpyAsync(h, p->dhash, sizeof(h), hipMemcpyDeviceToHost, s));
        PW_HIPS(hipStreamSynchronize(s));
        reuse = h[0] == h[1];
    }
    int rc = 0;
    if (!reuse)
        rc = p->prec == 4 ? upload_model<float>(p, (const float*)d_vel, true)
                          : upload_model<double>(p, (const double*)d_vel, true);
    if (!rc && !reuse) rc = ensure_widths(p);
    if (rc) return rc;
    const size_t e = esz(p);
    const long long per = (long long)a.nt * a.nr;
    const long long fld = (long long)p->nslot * p->fs;
    if (p->prec == 4) rc = fill<float>(p, p->gtot, p->nzx, 0.f);
    else rc = fill<double>(p, p->gtot, p->nzx, 0.0);
    if (rc) return rc;
    for (int b0 = 0; b0 < a.ns; b0 += p->nslot) {
        const int nb_ = std::min(p->nslot, a.ns - b0);
        if ((rc = set_batch(p, b0, nb_))) return rc;
        // the cotangent is the injected signal of the adjoint sweep
        PW_HIPS(hipMemcpyAsync(p->res, (const char*)d_gtraces + (size_t)b0 * per * e, (size_t)nb_ * per * e,
                               hipMemcpyDeviceToDevice, s));
        if (p->prec == 4) {
            if (!reuse) rc = forward_batch<float>(p, nb_, true, 1, true);
            if (!rc) rc = fill<float>(p, p->A, fld, 0.f);
            if (!rc) rc = fill<float>(p, p->B, fld, 0.f);
            if (!rc) rc = fill<float>(p, p->gm, fld, 0.f);
            float *A = (float*)p->A, *B = (float*)p->B;
            p->adj_sa = p->A2;
            p->adj_sb = p->B2;
            for (int sg = p->nseg - 1; sg >= 0 && !rc; --sg) {
                if (sg != p->nseg - 1) rc = recompute_segment<float>(p, nb_, sg, 1);
                if (!rc) rc = adjoint_segment<float>(p, nb_, sg, 1, &A, &B, nullptr, 0);
            }
            if (!rc) {
                FGeo<float> g = geo<float>(p);
                hipLaunchKernelGGL(k_gm_fold_z<float>, dim3((a.nz + 255) / 256, a.padnx), dim3(256), 0, s, g, nb_,
                               (const float*)p->gm, (const float*)p->vpad, dt2_of<float>(p), p->gfold);
            hipLaunchKernelGGL(k_gm_fold_x<float>, dim3((a.nz + 255) / 256, a.nx), dim3(256), 0, s, g,
                               (const double*)p->gfold, (float*)p->gtot);
            }
        } else {
            if (!reuse) rc = forward_batch<double>(p, nb_, true, 1, true);
            if (!rc) rc = fill<double>(p, p->A, fld, 0.0);
            if (!rc) rc = fill<double>(p, p->B, fld, 0.0);
            if (!rc) rc = fill<double>(p, p->gm, fld, 0.0);
            double *A = (double*)p->A, *B = (double*)p->B;
            p->adj_sa = p->A2;
            p->adj_sb = p->B2;
            for (int sg = p->nseg - 1; sg >= 0 && !rc; --sg) {
                if (sg != p->nseg - 1) rc = recompute_segment<double>(p, nb_, sg, 1);
                if (!rc) rc = adjoint_segment<double>(p, nb_, sg, 1, &A, &B, nullptr, 0);
            }
            if (!rc) {
                FGeo<double> g = geo<double>(p);
                hipLaunchKernelGGL(k_gm_fold_z<double>, dim3((a.nz + 255) / 256, a.padnx), dim3(256), 0, s, g, nb_,
                               (const double*)p->gm, (const double*)p->vpad, dt2_of<double>(p), p->gfold);
            hipLaunchKernelGGL(k_gm_fold_x<double>, dim3((a.nz + 255) / 256, a.nx), dim3(256), 0, s, g,
                               (const double*)p->gfold, (double*)p->gtot);
            }
        }
        if (rc) return rc;
        PW_HIPS(hipGetLastError());
    }
    PW_HIPS(hipMemcpyAsync(d_gvel, p->gtot, p->nzx * e, hipMemcpyDeviceToDevice, s));
    PW_HIPS(hipEventRecord(p->ev_dev[3], s));
    p->dev_timed_adj = true;
    p->dev_ckpt_valid = false;  // the sweep reused the segment buffer: a second adjoint recomputes
    return PW_OK;
}

int pw_fwi_widths(pw_fwi* p, int* w)
{
    if (!p || !w) return PW_ERR_ARG;
    w[0] = p->w_fwd; w[1] = p->w_fwdv; w[2] = p->w_exact; w[3] = p->w_ref; w[4] = p->w_refv;
    w[5] = p->widths_tuned ? 1 : 0;
    w[6] = p->w_fwd2;
    w[7] = p->w_adj2;
    w[8] = p->w_fwdt;
    return PW_OK;
}

int pw_fwi_tuning_log(pw_fwi* p, float* out, int cap)
{
    if (!p) return PW_ERR_ARG;
    const int n = (int)p->tune_log.size() / 3;
    for (int i = 0; i < std::min(n, cap) * 3 && out; ++i) out[i] = p->tune_log[i];
    return n;
}

// force the step kernels' tile widths: w[0] forward, w[1] visco forward, w[2] exact adjoint, w[3]
// reference adjoint, w[4] visco reference adjoint (4, 8 or 16 each), one step per launch; NULL
// returns to the plan's timing sweep
int pw_fwi_set_widths(pw_fwi* p, const int* w)
{
    if (!p) return PW_ERR_ARG;
    if (!w) {
        p->widths_tuned = p->widths_forced = false;
        return PW_OK;
    }
    for (int i = 0; i < 5; ++i) {
        bool ok = false;
        for (int W : kWidths) ok |= w[i] == W;
        if (!ok) PW_FAIL(p->ctx, PW_ERR_ARG, "pw_fwi_set_widths: widths must be 4, 8 or 16");
    }
    p->w_fwd = w[0]; p->w_fwdv = w[1]; p->w_exact = w[2]; p->w_ref = w[3]; p->w_refv = w[4];
    p->w_fwdt = w[0];
    p->w_fwd2 = p->w_adj2 = 0;
    p->tune_log.clear();
    p->widths_tuned = p->widths_forced = true;
    return PW_OK;
}

int pw_fwi_set_steps_per_launch(pw_fwi* p, int n)
{
    if (!p || !(n == 1 || n == 2 || n == -2)) return PW_ERR_ARG;
    p->max_fused = n < 0 ? -n : n;
    p->force_fused = n < 0;
    if (!p->widths_forced) p->widths_tuned = false;
    return PW_OK;
}

int pw_fwi_timing(pw_fwi* p, double* forward_ms, double* backward_ms)
{
    if (!p) return PW_ERR_ARG;
    // the device entry points' spans (waits for them), when they ran after the host-array calls
    float ms = 0.f;
    if (p->dev_timed_fwd) {
        PW_HIPS(hipEventSynchronize(p->ev_dev[1]));
        PW_HIPS(hipEventElapsedTime(&ms, p->ev_dev[0], p->ev_dev[1]));
        p->t_forward_ms = ms;
        p->dev_timed_fwd = false;
    }
    if (p->dev_timed_adj) {
        PW_HIPS(hipEventSynchronize(p->ev_dev[3]));
        PW_HIPS(hipEventElapsedTime(&ms, p->ev_dev[2], p->ev_dev[3]));
        p->t_backward_ms = ms;
        p->dev_timed_adj = false;
    }
    if (forward_ms) *forward_ms = p->t_forward_ms;
    if (backward_ms) *backward_ms = p->t_backward_ms;
    return PW_OK;
}

// fwi() (pywave/src/wave_fwi.c:24-119) with the Wolfe line search of wave_fwilbfgs.c:108-165 and
// the two-loop recursion of :38-95; vectors stay on the device and every dot product is an f64
// reduction.  vinv: niter * nz*nx; returns the number of stored iterates in *nstored (slots after
// an early stop are left untouched).
int pw_fwi_lbfgs(pw_fwi* p, const float* vel0, int niter, float* vinv, int* nstored)
{
    if (!p || !vel0 || !vinv || niter < 0) return PW_ERR_ARG;
    if (p->prec != 4) PW_FAIL(p->ctx, PW_ERR_ARG, "lbfgs: the reference optimiser runs in f32");
    hipStream_t s = p->ctx->s();
    const long long n = p->nzx;
    const int npair = 20, nls = 20;
    const double c1 = (float)1e-4, c2 = (float)0.9, factor = 10, conv = (float)0.01;
    const float v1 = 0.f, v2 = 10.f;
    float *x = nullptr, *g = nullptr, *dir = nullptr, *xk = nullptr, *q = nullptr, *sk = nullptr, *yk = nullptr;
    size_t bytes = n * 4;
    int rc = 0;
    rc = rc ? rc : dmalloc(p, (void**)&x, bytes);
    rc = rc ? rc : dmalloc(p, (void**)&g, bytes);
    rc = rc ? rc : dmalloc(p, (void**)&dir, bytes);
    rc = rc ? rc : dmalloc(p, (void**)&xk, bytes);
    rc = rc ? rc : dmalloc(p, (void**)&q, bytes);
    rc = rc ? rc : dmalloc(p, (void**)&sk, bytes * npair);
    rc = rc ? rc : dmalloc(p, (void**)&yk, bytes * npair);
    rc = rc ? rc : ensure_widths(p);
    auto cleanup = [&]() {
        float* bs[] = {x, g, dir, xk, q, sk, yk};
        for (float* b : bs)
            if (b) hipFree(b);
    };
    if (rc) { cleanup(); return rc; }
    auto grad_at = [&](const float* dx, double* f) -> int {
        // gradient of the current device model (copied to the plan's model buffer)
        PW_HIPS(hipMemcpyAsync(p->vel, dx, bytes, hipMemcpyDeviceToDevice, s));
        int r = upload_model<float>(p, nullptr, false);
        if (!r) r = gradient_all<float>(p, 0, f);
        if (!r) r = finish_reference_gradient(p, nullptr);
        if (!r) PW_HIPS(hipMemcpyAsync(g, p->gtot, bytes, hipMemcpyDeviceToDevice, s));
        return r;
    };
    auto axpby = [&](double al, const float* xa, double be, const float* yb, float* out) {
        hipLaunchKernelGGL(k_axpby<float>, dim3(nblk(n)), dim3(256), 0, s, n, al, xa, be, yb, out);
    };
    auto dot = [&](const float* a_, const float* b_, double* o) { return dot_dev<float>(p, a_, b_, o); };
    int stored = 0;
    do {
        if ((rc = hipMemcpyAsync(x, vel0, bytes, hipMemcpyHostToDevice, s) != hipSuccess ? PW_ERR_HIP : 0)) break;
        double fk = 0;
        p->scaling = 0.f;  // gradient_init: first call sets the scaling
        if ((rc = grad_at(x, &fk))) break;
        const double f0 = fk;
        int ipair = 0;
        double alpha = 1.0;
        for (int iter = 0; iter < niter; ++iter) {
            if (p->passive && iter % 5 == 0) alpha = 1.0;  // pfwi: optpar->repeat = 5 (wave_fwi.c:282)
            if (iter == 0) {
                axpby(-1.0, g, 0.0, g, dir);
            } else {
                // lbfgs_update: s = x - s_last, y = g - y_last
                float* sl = sk + (size_t)(ipair - 1) * n;
                float* yl = yk + (size_t)(ipair - 1) * n;
                axpby(1.0, x, -1.0, sl, sl);
                axpby(1.0, g, -1.0, yl, yl);
                // lbfgs_direction
                double ns_ = 0, ny_ = 0;
                if ((rc = dot(sl, sl, &ns_)) || (rc = dot(yl, yl, &ny_))) break;
                if (ns_ == 0.0 || ny_ == 0.0) {
                    axpby(-1.0, g, 0.0, g, dir);
                } else {
                    std::vector<double> rho(ipair), al(ipair);
                    PW_HIPS(hipMemcpyAsync(q, g, bytes, hipMemcpyDeviceToDevice, s));
                    for (int i = ipair - 1; i >= 0 && !rc; --i) {
                        float* si = sk + (size_t)i * n;
                        float* yi = yk + (size_t)i * n;
                        double t = 0;
                        if ((rc = dot(yi, si, &t))) break;
                        rho[i] = 1.0 / t;
                        if ((rc = dot(si, q, &t))) break;
                        al[i] = rho[i] * t;
                        axpby(1.0, q, -al[i], yi, q);
                    }
                    if (rc) break;
                    double yy = 0;
                    if ((rc = dot(yl, yl, &yy))) break;
                    const double gamma = 1.0 / yy / rho[ipair - 1];
                    axpby(gamma, q, 0.0, q, dir);
                    for (int i = 0; i < ipair && !rc; ++i) {
                        float* si = sk + (size_t)i * n;
                        float* yi = yk + (size_t)i * n;
                        double t = 0;
                        if ((rc = dot(yi, dir, &t))) break;
                        const double beta = t * rho[i];
                        axpby(1.0, dir, al[i] - beta, si, dir);
                    }
                    if (rc) break;
                    axpby(-1.0, dir, 0.0, dir, dir);
                }
            }
            // lbfgs_save
            if (ipair < npair) {
                PW_HIPS(hipMemcpyAsync(sk + (size_t)ipair * n, x, bytes, hipMemcpyDeviceToDevice, s));
                PW_HIPS(hipMemcpyAsync(yk + (size_t)ipair * n, g, bytes, hipMemcpyDeviceToDevice, s));
                ipair++;
            } else {
                for (int i = 0; i < npair - 1; ++i) {
                    PW_HIPS(hipMemcpyAsync(sk + (size_t)i * n, sk + (size_t)(i + 1) * n, bytes, hipMemcpyDeviceToDevice, s));
                    PW_HIPS(hipMemcpyAsync(yk + (size_t)i * n, yk + (size_t)(i + 1) * n, bytes, hipMemcpyDeviceToDevice, s));
                }
                PW_HIPS(hipMemcpyAsync(sk + (size_t)(npair - 1) * n, x, bytes, hipMemcpyDeviceToDevice, s));
                PW_HIPS(hipMemcpyAsync(yk + (size_t)(npair - 1) * n, g, bytes, hipMemcpyDeviceToDevice, s));
            }
            // Wolfe line search
            PW_HIPS(hipMemcpyAsync(xk, x, bytes, hipMemcpyDeviceToDevice, s));
            double m1 = 0;
            if ((rc = dot(g, dir, &m1))) break;
            const double m2 = m1 * c2;
            m1 = m1 * c1;
            double a1 = 0., a2 = 0., fcost = 0.;
            int flag = 0, i;
            for (i = 0; i < nls; ++i) {
                axpby(1.0, xk, alpha, dir, x);
                hipLaunchKernelGGL(k_clip<float>, dim3(nblk(n)), dim3(256), 0, s, n, x, v1, v2);
                if ((rc = grad_at(x, &fcost))) break;
                double m3 = 0;
                if ((rc = dot(g, dir, &m3))) break;
                if (fcost <= fk + alpha * m1 && m3 >= m2) {
                    fk = fcost;
                    flag = 0;
                    break;
                } else if (fcost > fk + alpha * m1) {
                    a2 = alpha;
                    alpha = 0.5 * (a1 + a2);
                } else {
                    a1 = alpha;
                    if (a2 == 0.) alpha *= factor;
                    else alpha = 0.5 * (a1 + a2);
                }
            }
            if (rc) break;
            if (i == nls) flag = fcost <= fk ? 1 : 2;
            if (flag == 2) break;
            if (fk / f0 < conv) break;
            PW_HIPS(hipMemcpyAsync(vinv + (size_t)iter * n, x, bytes, hipMemcpyDeviceToHost, s));
            PW_HIPS(hipStreamSynchronize(s));
            stored++;
        }
    } while (0);
    hipStreamSynchronize(s);
    cleanup();
    if (nstored) *nstored = stored;
    return rc;
}

// pfwicfun.forward_modeling_ac (pywave/src/pfwi.c:27-495); q is accepted and unused (acoustic)
int pw_forward_modeling_ac(pw_ctx* ctx, const float* vel, const float* q, const float* wav,
                           const float* pararray, float* out, long long out_len)
{
    (void)q;
    if (!ctx || !vel || !wav || !pararray || !out) return PW_ERR_ARG;
    Acq a;
    std::string err;
    if (acq_from_pararray(pararray, a, err)) PW_FAIL(ctx, PW_ERR_ARG, err);
    if (out_len != (long long)a.nt * a.nx * a.ns) PW_FAIL(ctx, PW_ERR_ARG, "forward_modeling_ac: output length");
    pw_fwi* p = pw_fwi_create(ctx, pararray, 0, 4);
    if (!p) return PW_ERR_ARG;
    int rc = pw_fwi_set_wavelet(p, wav);
    if (!rc) rc = pw_fwi_forward(p, vel, out);
    pw_fwi_destroy(p);
    return rc;
}

// visco-acoustic counterpart of forward_modeling_ac: the forward_modelingc that pywave/pfwi.py:72
// calls for media=2 (the reference's extension does not define it) on forward_modeling
// (pywave/src/wave_fwimodeling.c:116-220); q is the quality factor
int pw_forward_modeling_c(pw_ctx* ctx, const float* vel, const float* q, const float* wav,
                          const float* pararray, float* out, long long out_len)
{
    if (!ctx || !vel || !q || !wav || !pararray || !out) return PW_ERR_ARG;
    Acq a;
    std::string err;
    if (acq_from_pararray(pararray, a, err)) PW_FAIL(ctx, PW_ERR_ARG, err);
    if (out_len != (long long)a.nt * a.nx * a.ns) PW_FAIL(ctx, PW_ERR_ARG, "forward_modelingc: output length");
    pw_fwi* p = pw_fwi_create(ctx, pararray, 0, 4);
    if (!p) return PW_ERR_ARG;
    int rc = pw_fwi_set_wavelet(p, wav);
    if (!rc) rc = pw_fwi_set_visco(p, q);
    if (!rc) rc = pw_fwi_forward(p, vel, out);
    pw_fwi_destroy(p);
    return rc;
}

// pfwicfun.fwic (pywave/src/pfwi.c:497-1028): returns vinv (niter * nz*nx); slots after an early
// stop are zero (the reference leaves them uninitialised)
int pw_fwic(pw_ctx* ctx, const float* vel, const float* q, const float* wav, const float* data,
            const float* pararray, float* out, long long out_len)
{
    (void)q;
    if (!ctx || !vel || !wav || !data || !pararray || !out) return PW_ERR_ARG;
    Acq a;
    std::string err;
    if (acq_from_pararray(pararray, a, err)) PW_FAIL(ctx, PW_ERR_ARG, err);
    if (out_len != (long long)a.nz * a.nx * a.niter) PW_FAIL(ctx, PW_ERR_ARG, "fwic: output length");
    std::memset(out, 0, sizeof(float) * out_len);
    pw_fwi* p = pw_fwi_create(ctx, pararray, 0, 4);
    if (!p) return PW_ERR_ARG;
    int rc = pw_fwi_set_wavelet(p, wav);
    if (!rc) rc = pw_fwi_set_observed(p, data);
    int k = 0;
    if (!rc) rc = pw_fwi_lbfgs(p, vel, a.niter, out, &k);
    pw_fwi_destroy(p);
    return rc;
}

// pfwicfun.pfwic (pywave/src/pfwi.c:1507-2034): passive FWI for velocity with a known source
// field (paspar onlyvel).  datasrc = [data (nt*nx*ns) | src (nz*nx*nt*ns)] when pararray's inv
// flag is set, else src alone with zero data (pfwi.c:1753-1766).  out: vinv (niter * nz*nx); slots
// after an early stop are zero.
int pw_pfwic(pw_ctx* ctx, const float* vel, const float* q, const float* wav, const float* datasrc,
             const float* pararray, float* out, long long out_len, long long datasrc_len)
{
    (void)q;
    if (!ctx || !vel || !wav || !datasrc || !pararray || !out) return PW_ERR_ARG;
    Acq a;
    std::string err;
    if (acq_from_pararray(pararray, a, err)) PW_FAIL(ctx, PW_ERR_ARG, err);
    if (out_len != (long long)a.nz * a.nx * a.niter) PW_FAIL(ctx, PW_ERR_ARG, "pfwic: output length");
    const bool inv = pararray[9] != 0.f;
    const long long nd = (long long)a.nt * a.nx * a.ns, nsrc = (long long)a.nz * a.nx * a.nt * a.ns;
    if (datasrc_len != (inv ? nd : 0) + nsrc) PW_FAIL(ctx, PW_ERR_ARG, "pfwic: datasrc length");
    std::memset(out, 0, sizeof(float) * out_len);
    pw_fwi* p = pw_fwi_create(ctx, pararray, 0, 4);
    if (!p) return PW_ERR_ARG;
    int rc = pw_fwi_set_wavelet(p, wav);
    std::vector<float> zero;
    if (!rc) {
        if (inv) {
            rc = pw_fwi_set_observed(p, datasrc);
        } else {
            zero.assign((size_t)nd, 0.f);
            rc = pw_fwi_set_observed(p, zero.data());
        }
    }
    if (!rc) rc = pw_fwi_set_source_field(p, datasrc + (inv ? nd : 0));
    int k = 0;
    if (!rc) rc = pw_fwi_lbfgs(p, vel, a.niter, out, &k);
    pw_fwi_destroy(p);
    return rc;
}

}  // extern "C"
