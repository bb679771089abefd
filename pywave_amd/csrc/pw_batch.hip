// pw_batch.hip -- many afd2d shots of uneven grid size and length advanced together.
//
// The reference models one shot per afd2d call (pywave/afd2d.py:4-90 -> pywave/src/wave_fdm.c:
// 397-686).  Here every shot keeps its own padded fields in one pooled allocation; shots are
// ordered by decreasing nt so the shots still running at step it are a prefix, and one launch per
// step covers the flat list of 64 x XC tiles of all running shots (a tile looks up its shot in a
// table; the tile width XC is timed on the batch itself before its first run).  The step loop is
// captured once into a hipGraph and replayed.  Per cell the arithmetic
// is that of the single-shot kernel (pw_afd.hip), so every shot is bit-identical to afd2d.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "pw_internal.h"

namespace {

constexpr float C0 = -30. / 12., C1 = +16. / 12., C2 = -1. / 12.;  // pywave/src/wave_fdm.c:15
constexpr int kBatchWidths[] = {2, 4, 8, 16};  // x columns streamed per tile, as timed
constexpr int kBatchDefaultWidth = 4;

struct Shot {
    int nz, nx, nz2, nx2, ldz, nb, nz1, nx1;
    float idz, idx, cc;
    int nt;
    long long off;      // element offset of this shot's fields in the pools
    long long fz_off;   // offset into the damping pools (z pairs), x pairs follow
    int ns, src_off;    // sources: centres at src_pool[3*src_off ...], wavelets at wav_pool[src_off*.. ]
    long long wav_off;
    long long rec_off;  // traces [nt][nx1] at rec_pool + rec_off
    int ntz, nxc, tile0;
};

struct BatchArgs {
    const Shot* __restrict__ shots;
    const int* __restrict__ tile_shot;
    float* pool_a;
    float* pool_b;
    const float* __restrict__ vv;
    const float2* __restrict__ damp;  // per shot: nz2 z pairs then nx2 x pairs
    const int* __restrict__ src;
    const float* __restrict__ wav;
    float* __restrict__ rec;
    int it;
};

constexpr unsigned kOOB = 0x80000000u;  // out-of-range lane offset: load returns 0, store dropped

__device__ __forceinline__ __amdgpu_buffer_rsrc_t field_rsrc(const float* p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)bytes, 0x00020000);
}

// One 64 x XC tile of one shot per block (one wave), streamed along x with a 5-column register
// queue and an LDS z row.  Loads for column x+1 (prev, vv, z halo, x damping pair) and the centre
// of column x+3 are issued before column x is computed.  Field accesses go through one buffer
// resource per shot field: the lane's cell in the column is a fixed vector offset, the column a
// scalar offset; lanes outside the shot's updated domain (zero in every field), and every lane
// of a column outside it, get an out-of-range offset, so such cells are neither read nor
// written and need no branches.
template <int XC>
__global__ __launch_bounds__(64) void k_afd2d_batch(BatchArgs a)
{
    const int tile = blockIdx.x;
    const Shot sh = a.shots[a.tile_shot[tile]];
    const int lt = tile - sh.tile0;
    const int tz = lt % sh.ntz, xc = lt / sh.ntz;
    const int lane = threadIdx.x, z0 = tz * 64, z = z0 + lane;
    const int x0 = xc * XC, x1 = min(x0 + XC, sh.nx);
    // ping-pong by step parity: the field holding C_n
    const bool even = (a.it & 1) == 0;
    const float* cur = (even ? a.pool_a : a.pool_b) + sh.off;
    float* pv = (even ? a.pool_b : a.pool_a) + sh.off;
    const float* vv = a.vv + sh.off;
    const float2* __restrict__ dz_ = a.damp + sh.fz_off;
    const float2* __restrict__ dx_ = dz_ + sh.nz2;
    __shared__ float row[2][68];
    const int L = sh.ldz;
    const int bias = 2 * L + 2;  // resources start 2 columns + 2 cells before the field
    const unsigned wbytes = (unsigned)(((long long)L * sh.nx2 + 2 * bias) * 4);
    const auto rc = field_rsrc(cur - bias, wbytes), rp = field_rsrc(pv - bias, wbytes),
               rv = field_rsrc(vv - bias, wbytes);
    const bool zact = z < sh.nz;
    const bool zin = z >= 2 && z < sh.nz2 - 2;
    const float2 fzz = zact ? dz_[z] : make_float2(1.f, 1.f);
    const int zh = lane < 2 ? z0 - 2 + lane : z0 + 62 + lane;  // z halo of lanes 0..3
    // lanes 0..3 stage the z halo, the others rewrite their own centre (no branch)
    const int lzh = lane < 4 ? (lane < 2 ? lane : 64 + lane) : lane + 2;
    // lane offsets are fixed; the column goes in the scalar offset; a column outside the shot
    // gives every lane the out-of-range offset
    const unsigned oz = zact ? (unsigned)(bias + z) * 4u : kOOB;
    const unsigned ozh = (lane < 4 && zh >= 0 && zh < sh.nz) ? (unsigned)(bias + zh) * 4u : kOOB;
    const unsigned ocol = (unsigned)L * 4u;
    auto xv = [&](int x) { return x >= 0 && x < sh.nx; };
    auto ld = [](__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    };
    bool blk_src = false;
    for (int s = 0; s < sh.ns; ++s) {
        const int* cz = a.src + 3 * (sh.src_off + s);
        blk_src |= cz[0] >= z0 - 1 && cz[0] <= z0 + 64 && cz[1] >= x0 - 1 && cz[1] <= x1;
    }
    const bool rtile = sh.nb >= z0 && sh.nb < z0 + 64;  // the receiver row is in this tile
    // Lean tiles: no source or receiver rows and every column of the stencil reach inside the
    // updated domain (x0 >= 2, x0 + XC + 2 <= nx): in-range fetches, the tile's damping pairs in
    // LDS and an unrolled loop without scalar branches.
    if (!blk_src && !rtile && x0 >= 2 && x0 + XC + 2 <= sh.nx) {
        __shared__ float2 sxp[XC];
        if (lane < XC) sxp[lane] = dx_[x0 + lane];
        float q[5];
#pragma unroll
        for (int k = 0; k < 4; ++k) q[k] = ld(rc, oz, (unsigned)(x0 - 2 + k) * ocol);
        __syncthreads();
        float nq = ld(rc, oz, (unsigned)(x0 + 2) * ocol);
        float pr = ld(rp, oz, (unsigned)x0 * ocol), vc = ld(rv, oz, (unsigned)x0 * ocol);
        float hz = ld(rc, ozh, (unsigned)x0 * ocol);
#pragma unroll
        for (int i = 0; i < XC; ++i) {
            const int x = x0 + i;
            q[4] = nq;
            const float cpr = pr, cvc = vc, chz = hz;
            if (i + 1 < XC) {
                const unsigned sn = (unsigned)(x + 1) * ocol;
                nq = ld(rc, oz, (unsigned)(x + 3) * ocol);
                pr = ld(rp, oz, sn);
                vc = ld(rv, oz, sn);
                hz = ld(rc, ozh, sn);
            }
            row[i & 1][lane + 2] = q[2];
            row[i & 1][lzh] = lane < 4 ? chz : q[2];
            __syncthreads();
            const float u = q[2];
            const float* R = &row[i & 1][lane + 2];
            float lap = C0 * u * sh.cc
                      + C1 * (q[1] + q[3]) * sh.idx
                      + C2 * (q[0] + q[4]) * sh.idx
                      + C1 * (R[-1] + R[1]) * sh.idz
                      + C2 * (R[-2] + R[2]) * sh.idz;
            lap = __builtin_bit_cast(float, __builtin_bit_cast(unsigned, lap) & (zin ? ~0u : 0u));
            const float2 fx = sxp[i];
            float p = cpr;
            p = p * fzz.x; p = p * fzz.y; p = p * fx.x; p = p * fx.y;
            float c = u + (u - p);
            c = c + lap * cvc;
            c = c * fzz.x; c = c * fzz.y; c = c * fx.x; c = c * fx.y;
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, c), rp, oz, (unsigned)x * ocol, 0);
            q[0] = q[1]; q[1] = q[2]; q[2] = q[3]; q[3] = q[4];
        }
        return;
    }
    float q[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const int xx = x0 - 2 + k;
        q[k] = ld(rc, xv(xx) ? oz : kOOB, xv(xx) ? (unsigned)xx * ocol : 0u);
    }
    float pr = ld(rp, oz, (unsigned)x0 * ocol);
    float vc = ld(rv, oz, (unsigned)x0 * ocol);
    float hz = ld(rc, ozh, (unsigned)x0 * ocol);
    float2 fxx = dx_[x0];
    int buf = 0;
    for (int x = x0; x < x1; ++x) {
        const int xn = x + 1;
        const bool nv = xn < x1;
        const bool n3 = nv && x + 3 < sh.nx;
        // column validity is uniform: it selects the lane offset (one VALU op per load) rather
        // than the resource (several scalar ops), the scalar side being the busier one here
        const unsigned sn = nv ? (unsigned)xn * ocol : 0u, s3 = n3 ? (unsigned)(x + 3) * ocol : 0u;
        const float nq = ld(rc, n3 ? oz : kOOB, s3);
        const float npr = ld(rp, nv ? oz : kOOB, sn);
        const float nvc = ld(rv, nv ? oz : kOOB, sn);
        const float nhz = ld(rc, nv ? ozh : kOOB, sn);
        const float2 nfx = dx_[min(xn, sh.nx2 - 1)];  // xn may equal nx == nx2 on the last column
        row[buf][lane + 2] = q[2];
        row[buf][lzh] = lane < 4 ? hz : q[2];
        __syncthreads();
        const float u = q[2];
        const float* R = &row[buf][lane + 2];
        float lap = C0 * u * sh.cc
                  + C1 * (q[1] + q[3]) * sh.idx
                  + C2 * (q[0] + q[4]) * sh.idx
                  + C1 * (R[-1] + R[1]) * sh.idz
                  + C2 * (R[-2] + R[2]) * sh.idz;
        lap = (zin && x >= 2 && x < sh.nx2 - 2) ? lap : 0.f;
        float p = pr;
        p = p * fzz.x; p = p * fzz.y; p = p * fxx.x; p = p * fxx.y;
        float c = u + (u - p);
        c = c + lap * vc;
        if (blk_src) {
            for (int s = 0; s < sh.ns; ++s) {
                const int* cz = a.src + 3 * (sh.src_off + s);
                int ddz = z - cz[0], ddx = x - cz[1];
                int az_ = ddz < 0 ? -ddz : ddz, ax_ = ddx < 0 ? -ddx : ddx;
                if (az_ <= 1 && ax_ <= 1)
                    c = c + (vc * a.wav[sh.wav_off + (long long)s * sh.nt + a.it]) / (float)(az_ + ax_ + 1);
            }
        }
        c = c * fzz.x; c = c * fzz.y; c = c * fxx.x; c = c * fxx.y;
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, c), rp, oz, (unsigned)x * ocol, 0);
        if (rtile && x >= sh.nb && x < sh.nb + sh.nx1) {
            if (z == sh.nb) a.rec[sh.rec_off + (long long)a.it * sh.nx1 + (x - sh.nb)] = c;
        }
        q[0] = q[1]; q[1] = q[2]; q[2] = q[3]; q[3] = q[4]; q[4] = nq;
        pr = npr; vc = nvc; hz = nhz; fxx = nfx;
        buf ^= 1;
    }
}

// clamp-expanded vv = ((v*v)*dt)*dt for one shot
__global__ void k_batch_vv(Shot sh, const float* __restrict__ vel, float* __restrict__ vv, float dt)
{
    const int z = blockIdx.x * blockDim.x + threadIdx.x, x = blockIdx.y;
    if (z >= sh.nz) return;
    const int sz = min(max(z - sh.nb, 0), sh.nz1 - 1), sx = min(max(x - sh.nb, 0), sh.nx1 - 1);
    const float v = vel[(size_t)sx * sh.nz1 + sz];
    vv[sh.off + (long long)x * sh.ldz + z] = v * v * dt * dt;
}

__global__ void k_zero(float* p, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = 0.f;
}

}  // namespace

struct pw_batch {
    pw_ctx* ctx = nullptr;
    int n = 0;
    std::vector<Shot> shots;    // sorted by decreasing nt
    std::vector<int> order;     // sorted position -> caller's shot index
    std::vector<int> where;     // caller's shot index -> sorted position
    std::vector<float> dts;
    std::vector<int> tiles_active;  // number of active tiles at step it (piecewise by shot)
    long long pool = 0, rec_pool = 0, damp_pool = 0;
    int max_nt = 0, total_tiles = 0;
    float *d_a = nullptr, *d_b = nullptr, *d_vv = nullptr, *d_rec = nullptr, *d_wav = nullptr;
    float2* d_damp = nullptr;
    int *d_src = nullptr, *d_tile_shot = nullptr;
    Shot* d_shots = nullptr;
    std::vector<int> src_host;    // 3 per source
    std::vector<float> wav_host;  // per source nt
    std::vector<int> ns_of;
    std::vector<char> have;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool dirty = true;
    int xc = kBatchDefaultWidth;  // tile width of the current layout
    int xc_forced = 0;            // 0: timed before the first run; else this width
    bool tuned = false;
    int max_tiles = 0;            // tiles at the narrowest width (size of the tile -> shot table)
    std::vector<float> tune_log;  // (width, ms per launch) of the timing sweep
};

namespace {

void drop_graph(pw_batch* p)
{
    if (p->exec) hipGraphExecDestroy(p->exec);
    if (p->graph) hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
}

void batch_free(pw_batch* p)
{
    drop_graph(p);
    void* b[] = {p->d_a, p->d_b, p->d_vv, p->d_rec, p->d_wav, p->d_damp, p->d_src, p->d_tile_shot, p->d_shots};
    for (void* x : b)
        if (x) hipFree(x);
}

// Tile layout for width xc: per-shot tile counts and first tiles, the tile -> shot table and the
// active tile count per step (shots sorted by decreasing nt, so the running shots are a prefix).
void layout(pw_batch* p, int xc, std::vector<int>& tile_shot)
{
    int tile0 = 0;
    for (Shot& s : p->shots) {
        s.nxc = pw::ceil_div(s.nx, xc);
        s.tile0 = tile0;
        tile0 += s.ntz * s.nxc;
    }
    p->total_tiles = tile0;
    tile_shot.assign(tile0, 0);
    for (int k = 0; k < p->n; ++k)
        for (int t = 0; t < p->shots[k].ntz * p->shots[k].nxc; ++t) tile_shot[p->shots[k].tile0 + t] = k;
    // shots [0, k) run at step it; k shrinks as it grows (shot 0 runs to the last step)
    p->tiles_active.assign(p->max_nt, 0);
    for (int it = 0, k = p->n; it < p->max_nt; ++it) {
        while (k > 1 && p->shots[k - 1].nt <= it) --k;
        p->tiles_active[it] = k == p->n ? tile0 : p->shots[k].tile0;
    }
    p->xc = xc;
}

// every tile launched at step it belongs to a shot that has step it (its traces and wavelet reach it)
bool layout_consistent(const pw_batch* p, const std::vector<int>& tile_shot)
{
    for (int it = 0; it < p->max_nt; ++it) {
        const int n = p->tiles_active[it];
        if (n < 1 || n > p->total_tiles || p->shots[tile_shot[n - 1]].nt <= it) return false;
        if (n < p->total_tiles && p->shots[tile_shot[n]].nt > it) return false;
    }
    return true;
}

// switch the device tables to width xc (the captured step loop holds the old grid sizes)
int apply_layout(pw_batch* p, int xc)
{
    std::vector<int> tile_shot;
    layout(p, xc, tile_shot);
    if (!layout_consistent(p, tile_shot)) PW_FAIL(p->ctx, PW_ERR_STATE, "batch2d: inconsistent tile layout");
    drop_graph(p);
    hipStream_t st = p->ctx->s();
    PW_HIP(p->ctx, hipMemcpyAsync(p->d_tile_shot, tile_shot.data(), sizeof(int) * tile_shot.size(),
                                  hipMemcpyHostToDevice, st));
    PW_HIP(p->ctx, hipMemcpyAsync(p->d_shots, p->shots.data(), sizeof(Shot) * p->n, hipMemcpyHostToDevice, st));
    PW_HIP(p->ctx, hipStreamSynchronize(st));
    return PW_OK;
}

BatchArgs step_args(const pw_batch* p, int it)
{
    BatchArgs a{};
    a.shots = p->d_shots; a.tile_shot = p->d_tile_shot;
    a.pool_a = p->d_a; a.pool_b = p->d_b; a.vv = p->d_vv; a.damp = p->d_damp;
    a.src = p->d_src; a.wav = p->d_wav; a.rec = p->d_rec; a.it = it;
    return a;
}

void launch_step(int xc, int tiles, const BatchArgs& a, hipStream_t st)
{
    switch (xc) {
        case 2: hipLaunchKernelGGL(k_afd2d_batch<2>, dim3(tiles), dim3(64), 0, st, a); break;
        case 8: hipLaunchKernelGGL(k_afd2d_batch<8>, dim3(tiles), dim3(64), 0, st, a); break;
        case 16: hipLaunchKernelGGL(k_afd2d_batch<16>, dim3(tiles), dim3(64), 0, st, a); break;
        default: hipLaunchKernelGGL(k_afd2d_batch<4>, dim3(tiles), dim3(64), 0, st, a); break;
    }
}

// What decides the timing sweep's outcome: the shots' grids and lengths, the absorbing width and
// the source centres (tiles holding a source run the general loop).
std::vector<long long> tune_key(const pw_batch* p)
{
    std::vector<long long> k = {1, p->n, p->shots.empty() ? 0 : p->shots[0].nb, p->xc_forced};
    for (const Shot& s : p->shots) {
        k.push_back(s.nz);
        k.push_back(s.nx);
        k.push_back(s.nt);
        k.push_back(s.ns);
    }
    k.insert(k.end(), p->src_host.begin(), p->src_host.end());
    return k;
}

// Tile-width sweep: each width in kBatchWidths laid out and timed on the first time step (every
// shot running) after clearing the fields; the fastest is kept.  The run that follows clears
// the fields again and rewrites every trace sample the timing launches wrote.
int tune_width(pw_batch* p)
{
    const std::vector<long long> key = tune_key(p);
    auto hit = p->ctx->tuned2d.find(key);
    if (hit != p->ctx->tuned2d.end()) {
        p->tune_log = hit->second.second;
        p->tuned = true;
        return apply_layout(p, hit->second.first[0]);
    }
    hipStream_t st = p->ctx->s();
    p->tune_log.clear();
    int best = p->xc_forced ? p->xc_forced : kBatchDefaultWidth;
    if (!p->xc_forced) {
        hipEvent_t e0, e1;
        PW_HIP(p->ctx, hipEventCreate(&e0));
        PW_HIP(p->ctx, hipEventCreate(&e1));
        float tbest = 0.f;
        int rc = PW_OK;
        for (int W : kBatchWidths) {
            if ((rc = apply_layout(p, W))) break;
            hipLaunchKernelGGL(k_zero, dim3(2048), dim3(256), 0, st, p->d_a, p->pool);
            hipLaunchKernelGGL(k_zero, dim3(2048), dim3(256), 0, st, p->d_b, p->pool);
            const int reps = 8;
            // every launch is time step 0, which every shot has (traces row 0, wavelet sample 0)
            const BatchArgs a0 = step_args(p, 0);
            for (int r = 0; r < 2; ++r) launch_step(W, p->tiles_active[0], a0, st);
            hipEventRecord(e0, st);
            for (int r = 0; r < reps; ++r) launch_step(W, p->tiles_active[0], a0, st);
            hipEventRecord(e1, st);
            float ms = 0.f;
            if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                p->ctx->err = "batch2d: width timing launch failed";
                rc = PW_ERR_HIP;
                break;
            }
            ms /= reps;
            p->tune_log.push_back((float)W);
            p->tune_log.push_back(ms);
            if (W == kBatchWidths[0] || ms < tbest) { tbest = ms; best = W; }
        }
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        if (rc) return rc;
    }
    int rc = apply_layout(p, best);
    if (rc) return rc;
    p->tuned = true;
    p->ctx->tuned2d[key] = {{best}, p->tune_log};
    return PW_OK;
}

}  // namespace

extern "C" {

// dims: n*3 ints (nz, nx, nt) per shot; d: n*3 floats (dz, dx, dt); nbt, ct shared
pw_batch* pw_batch2d_create(pw_ctx* ctx, int n, const int* dims, const float* d, int nbt, float ct)
{
    if (!ctx || n < 1 || !dims || !d || nbt < 0) return nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess) { ctx->err = "hipSetDevice failed"; return nullptr; }
    pw_batch* p = new pw_batch;
    p->ctx = ctx;
    p->n = n;
    p->order.resize(n);
    std::iota(p->order.begin(), p->order.end(), 0);
    std::stable_sort(p->order.begin(), p->order.end(), [&](int i, int j) { return dims[3 * i + 2] > dims[3 * j + 2]; });
    p->where.resize(n);
    for (int k = 0; k < n; ++k) p->where[p->order[k]] = k;
    std::vector<float> w(std::max(nbt, 1));
    pw::abc_weights(nbt, ct, w.data());
    std::vector<float2> damp;
    long long off = 0, roff = 0;
    for (int k = 0; k < n; ++k) {
        const int i = p->order[k];
        const int nz1 = dims[3 * i], nx1 = dims[3 * i + 1], nt = dims[3 * i + 2];
        const float dz = d[3 * i], dx = d[3 * i + 1], dt = d[3 * i + 2];
        if (nz1 < 1 || nx1 < 1 || nt < 1 || !(dz > 0) || !(dx > 0) || !(dt > 0)) {
            ctx->err = "pw_batch2d_create: invalid dimensions for shot " + std::to_string(i);
            delete p;
            return nullptr;
        }
        Shot s{};
        s.nz1 = nz1; s.nx1 = nx1; s.nb = nbt; s.nt = nt;
        s.nz = nz1 + 2 * nbt; s.nx = nx1 + 2 * nbt;
        pw::padded_sizes(s.nz, s.nx, 1, &s.nz2, &s.nx2, nullptr);
        s.ldz = pw::ceil_div(s.nz2, 32) * 32;
        s.idz = 1 / (dz * dz);
        s.idx = 1 / (dx * dx);
        s.cc = s.idx + s.idz;
        s.off = off;
        off += (long long)s.ldz * s.nx2;
        s.fz_off = (long long)damp.size();
        std::vector<float> fz, fx;
        pw::damping_pairs(s.nz2, nbt, nbt, w.data(), w.data(), fz);
        pw::damping_pairs(s.nx2, nbt, nbt, w.data(), w.data(), fx);
        for (int t = 0; t < s.nz2; ++t) damp.push_back(make_float2(fz[2 * t], fz[2 * t + 1]));
        for (int t = 0; t < s.nx2; ++t) damp.push_back(make_float2(fx[2 * t], fx[2 * t + 1]));
        s.rec_off = roff;
        roff += (long long)nt * nx1;
        s.ntz = pw::ceil_div(s.nz, 64);
        p->shots.push_back(s);
        p->dts.push_back(dt);
        p->max_nt = std::max(p->max_nt, nt);
    }
    p->pool = off;
    p->rec_pool = roff;
    p->damp_pool = (long long)damp.size();
    p->ns_of.assign(n, 0);
    p->have.assign(n, 0);
    std::vector<int> tile_shot;
    layout(p, kBatchWidths[0], tile_shot);  // the narrowest width has the most tiles
    p->max_tiles = p->total_tiles;
    layout(p, kBatchDefaultWidth, tile_shot);
    if (!layout_consistent(p, tile_shot)) {
        ctx->err = "pw_batch2d_create: inconsistent tile layout";
        delete p;
        return nullptr;
    }
    bool ok = hipMalloc(&p->d_a, p->pool * 4) == hipSuccess && hipMalloc(&p->d_b, p->pool * 4) == hipSuccess &&
              hipMalloc(&p->d_vv, p->pool * 4) == hipSuccess && hipMalloc(&p->d_rec, p->rec_pool * 4) == hipSuccess &&
              hipMalloc(&p->d_damp, p->damp_pool * sizeof(float2)) == hipSuccess &&
              hipMalloc(&p->d_tile_shot, sizeof(int) * p->max_tiles) == hipSuccess &&
              hipMalloc(&p->d_shots, sizeof(Shot) * n) == hipSuccess;
    hipStream_t st = ctx->s();
    ok = ok && hipMemcpyAsync(p->d_damp, damp.data(), damp.size() * sizeof(float2), hipMemcpyHostToDevice, st) == hipSuccess &&
         hipMemcpyAsync(p->d_tile_shot, tile_shot.data(), sizeof(int) * tile_shot.size(), hipMemcpyHostToDevice, st) == hipSuccess &&
         hipMemsetAsync(p->d_vv, 0, p->pool * 4, st) == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
    if (!ok) {
        ctx->err = "pw_batch2d_create: device allocation failed";
        batch_free(p);
        delete p;
        return nullptr;
    }
    return p;
}

void pw_batch2d_destroy(pw_batch* p)
{
    if (!p) return;
    hipSetDevice(p->ctx->device);
    hipStreamSynchronize(p->ctx->s());
    batch_free(p);
    delete p;
}

// shot i: vel (nz*nx, z fastest), src_pack [sx.. sz.. f.. t.. A..] (5*ns), wavelet (ns*nt) or NULL
int pw_batch2d_set_shot(pw_batch* p, int i, const float* vel, int ns, const float* src_pack, const float* wavelet)
{
    if (!p || i < 0 || i >= p->n || !vel || ns < 0 || (ns > 0 && !src_pack)) return PW_ERR_ARG;
    pw_ctx* ctx = p->ctx;
    const int k = p->where[i];
    Shot& s = p->shots[k];
    const float dt = p->dts[k];
    float* d_vel = nullptr;
    PW_HIP(ctx, hipMalloc(&d_vel, (size_t)s.nz1 * s.nx1 * 4));
    hipStream_t st = ctx->s();
    hipError_t e = hipMemcpyAsync(d_vel, vel, (size_t)s.nz1 * s.nx1 * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_batch_vv, dim3(pw::ceil_div(s.nz, 256), s.nx), dim3(256), 0, st, s, d_vel, p->d_vv, dt);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    hipFree(d_vel);
    if (e != hipSuccess) PW_FAIL(ctx, PW_ERR_HIP, std::string("batch set_shot: ") + hipGetErrorString(e));
    // sources of this shot (validated as in pw_afd_set_sources)
    std::vector<int> cen(3 * ns);
    std::vector<float> wav((size_t)ns * s.nt);
    for (int j = 0; j < ns; ++j) {
        int sx = (int)src_pack[j], sz = (int)src_pack[ns + j];
        float f0 = src_pack[2 * ns + j], t0 = src_pack[3 * ns + j], A = src_pack[4 * ns + j];
        int gz = sz + s.nb, gx = sx + s.nb;
        if (gz - 1 < 0 || gz + 1 >= s.nz || gx - 1 < 0 || gx + 1 >= s.nx)
            PW_FAIL(ctx, PW_ERR_ARG, "batch set_shot: source stencil leaves the padded domain");
        cen[3 * j] = gz; cen[3 * j + 1] = gx; cen[3 * j + 2] = 0;
        if (wavelet) {
            std::memcpy(&wav[(size_t)j * s.nt], wavelet + (size_t)j * s.nt, 4 * s.nt);
        } else {
            std::string err;
            int rc = pw::ricker(s.nt, dt, f0, t0, A, &wav[(size_t)j * s.nt], &err);
            if (rc) PW_FAIL(ctx, rc, "batch set_shot: " + err);
        }
    }
    // per-shot source blocks are stored in sorted order; rebuild the pools
    p->ns_of[k] = ns;
    std::vector<std::vector<int>> cs(p->n);
    std::vector<std::vector<float>> ws(p->n);
    for (int kk = 0; kk < p->n; ++kk) {  // keep the other shots' entries
        const Shot& t = p->shots[kk];
        if (kk != k && t.ns > 0) {
            cs[kk].assign(p->src_host.begin() + 3 * t.src_off, p->src_host.begin() + 3 * (t.src_off + t.ns));
            ws[kk].assign(p->wav_host.begin() + t.wav_off, p->wav_host.begin() + t.wav_off + (long long)t.ns * t.nt);
        }
    }
    cs[k] = cen;
    ws[k] = wav;
    p->src_host.clear();
    p->wav_host.clear();
    for (int kk = 0; kk < p->n; ++kk) {
        Shot& t = p->shots[kk];
        t.ns = kk == k ? ns : t.ns;
        t.src_off = (int)(p->src_host.size() / 3);
        t.wav_off = (long long)p->wav_host.size();
        p->src_host.insert(p->src_host.end(), cs[kk].begin(), cs[kk].end());
        p->wav_host.insert(p->wav_host.end(), ws[kk].begin(), ws[kk].end());
    }
    p->have[k] = 1;
    p->dirty = true;
    return PW_OK;
}

static int batch_upload(pw_batch* p)
{
    pw_ctx* ctx = p->ctx;
    hipStream_t st = ctx->s();
    if (p->d_src) { hipFree(p->d_src); p->d_src = nullptr; }
    if (p->d_wav) { hipFree(p->d_wav); p->d_wav = nullptr; }
    PW_HIP(ctx, hipMalloc(&p->d_src, std::max<size_t>(4, p->src_host.size() * 4)));
    PW_HIP(ctx, hipMalloc(&p->d_wav, std::max<size_t>(4, p->wav_host.size() * 4)));
    if (!p->src_host.empty())
        PW_HIP(ctx, hipMemcpyAsync(p->d_src, p->src_host.data(), p->src_host.size() * 4, hipMemcpyHostToDevice, st));
    if (!p->wav_host.empty())
        PW_HIP(ctx, hipMemcpyAsync(p->d_wav, p->wav_host.data(), p->wav_host.size() * 4, hipMemcpyHostToDevice, st));
    PW_HIP(ctx, hipMemcpyAsync(p->d_shots, p->shots.data(), sizeof(Shot) * p->n, hipMemcpyHostToDevice, st));
    PW_HIP(ctx, hipStreamSynchronize(st));
    drop_graph(p);
    p->dirty = false;
    return PW_OK;
}

static void batch_enqueue_steps(pw_batch* p, hipStream_t st)
{
    // zero the fields with a fill kernel (graph-safe), then one launch per step
    hipLaunchKernelGGL(k_zero, dim3(2048), dim3(256), 0, st, p->d_a, p->pool);
    hipLaunchKernelGGL(k_zero, dim3(2048), dim3(256), 0, st, p->d_b, p->pool);
    for (int it = 0; it < p->max_nt; ++it) launch_step(p->xc, p->tiles_active[it], step_args(p, it), st);
}

// use_graph != 0: capture the step loop once and replay it
int pw_batch2d_run(pw_batch* p, int use_graph, double* ms)
{
    if (!p) return PW_ERR_ARG;
    pw_ctx* ctx = p->ctx;
    for (int k = 0; k < p->n; ++k)
        if (!p->have[k]) PW_FAIL(ctx, PW_ERR_STATE, "batch run: a shot has no model/sources");
    if (p->dirty) {
        int rc = batch_upload(p);
        if (rc) return rc;
        p->tuned = false;  // sources moved: the tiles that run the general loop may differ
    }
    if (!p->tuned) {
        int rc = tune_width(p);
        if (rc) return rc;
    }
    hipStream_t st = ctx->s();
    hipEvent_t e0, e1;
    PW_HIP(ctx, hipEventCreate(&e0));
    PW_HIP(ctx, hipEventCreate(&e1));
    if (use_graph && !p->exec) {
        PW_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        batch_enqueue_steps(p, st);
        PW_HIP(ctx, hipStreamEndCapture(st, &p->graph));
        PW_HIP(ctx, hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    }
    PW_HIP(ctx, hipEventRecord(e0, st));
    if (use_graph) PW_HIP(ctx, hipGraphLaunch(p->exec, st));
    else batch_enqueue_steps(p, st);
    PW_HIP(ctx, hipGetLastError());
    PW_HIP(ctx, hipEventRecord(e1, st));
    PW_HIP(ctx, hipEventSynchronize(e1));
    float t = 0.f;
    PW_HIP(ctx, hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return PW_OK;
}

// traces of shot i: (nt, nx) Fortran order (time fastest), as afd2d's dout
int pw_batch2d_traces(pw_batch* p, int i, float* out, long long len)
{
    if (!p || i < 0 || i >= p->n || !out) return PW_ERR_ARG;
    const Shot& s = p->shots[p->where[i]];
    if (len != (long long)s.nt * s.nx1) PW_FAIL(p->ctx, PW_ERR_ARG, "batch traces: length mismatch");
    std::vector<float> t((size_t)len);
    hipStream_t st = p->ctx->s();
    if (int rc = pw::check_finite(p->ctx, p->d_rec + s.rec_off, (size_t)len, 4, "batch2d: receiver traces")) return rc;
    PW_HIP(p->ctx, hipMemcpyAsync(t.data(), p->d_rec + s.rec_off, len * 4, hipMemcpyDeviceToHost, st));
    PW_HIP(p->ctx, hipStreamSynchronize(st));
    for (int it = 0; it < s.nt; ++it)
        for (int ix = 0; ix < s.nx1; ++ix) out[(size_t)ix * s.nt + it] = t[(size_t)it * s.nx1 + ix];
    return PW_OK;
}

// tile width: 0 (default) times each width on the batch before its next run and keeps the
// fastest; 2, 4, 8 or 16 uses that width
int pw_batch2d_set_width(pw_batch* p, int xc)
{
    if (!p) return PW_ERR_ARG;
    bool ok = xc == 0;
    for (int W : kBatchWidths) ok |= xc == W;
    if (!ok) PW_FAIL(p->ctx, PW_ERR_ARG, "batch set_width: width must be 0, 2, 4, 8 or 16");
    p->xc_forced = xc;
    p->tuned = false;
    return PW_OK;
}

// tile width of the last run (or of the next, before the first); log (width, ms per launch) pairs of
// the timing sweep that chose it into out[2*cap]; returns the width
int pw_batch2d_width(pw_batch* p, float* log, int cap, int* nlog)
{
    if (!p) return PW_ERR_ARG;
    const int n = (int)p->tune_log.size() / 2;
    for (int i = 0; i < std::min(n, cap) * 2 && log; ++i) log[i] = p->tune_log[i];
    if (nlog) *nlog = n;
    return p->xc;
}

// total cell-updates of one run (updated domains x nt, summed over shots)
long long pw_batch2d_cell_updates(pw_batch* p)
{
    if (!p) return 0;
    long long c = 0;
    for (const Shot& s : p->shots) c += (long long)s.nz * s.nx * s.nt;
    return c;
}

}  // extern "C"
