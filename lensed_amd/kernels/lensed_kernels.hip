// Device kernels of the likelihood path, compiled at run time with hipRTC
// after the generated model code (namespace lcl: compute_<k>, set_params_<k>)
// Quadrature nodes come in a buffer of float4 (dx, dy, w_value, w_error),
// already scaled by the pixel scale, as the node buffers of src/lensed.c:813-822.
//
// Launch structure per evaluation (one launch each, any number of systems
// and candidates per launch):
//   lensed_set_params  parameters -> object data blocks      (src/nested.c:77)
//   lensed_render      quadrature ray-shooting -> model image (kernel/lensed.cl:9-38)
//   lensed_conv_chi2   PSF convolution + chi^2 partial sums   (kernel/lensed.cl:41-103)
//   lensed_finalize    FP64 combine -> logL = -chi^2/2        (src/nested.c:101-115)
// Work is a flat list of items strided over a persistent grid; each item
// belongs to one (system, candidate, tile), found from the descriptors.

#ifndef LENSED_DWORDS
#error "generated model code missing"
#endif
#ifndef LENSED_FOR_PSFS
#define LENSED_FOR_PSFS(X)
#endif
#include "epl_series.h"

typedef unsigned int u32;

// system owning a work item: last descriptor whose first item <= item
template <int WHICH>
__device__ inline int lensed_find_sys(const lensed_desc* __restrict__ sys, int nsys, int item) {
    int lo = 0, hi = nsys - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        int first = WHICH == 0 ? sys[mid].item_r : WHICH == 1 ? sys[mid].item_c : WHICH == 2 ? sys[mid].item_s
                  : WHICH == 3 ? sys[mid].item_f : WHICH == 4 ? sys[mid].item_t : sys[mid].item_k;
        if (first <= item) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Work of a persistent grid split with the XCDs in mind: workgroups go to
// the eight XCDs in turn (block b on XCD b % 8) and each XCD has its own L2,
// so each XCD takes a contiguous eighth of the work, cut into one
// contiguous run per block [first, end): neighbouring items (one
// candidate's, sharing its EPL table) then meet in one L2, and consecutive
// items go to one block, which keeps what they share (the staged table)
// instead of each block fetching it again.
__device__ inline void lensed_xcd_run(int n, int& first, int& end) {
    const int g = (int)gridDim.x, b = (int)blockIdx.x;
    int lo = 0, hi = n, nb = g, k = b;
    if (g >= 8) {
        const int x = b & 7;
        lo = (int)((long long)n * x / 8);
        hi = (int)((long long)n * (x + 1) / 8);
        nb = (g - x + 7) >> 3;  // blocks on XCD x
        k = b >> 3;
    }
    first = lo + (int)((long long)(hi - lo) * k / nb);
    end = lo + (int)((long long)(hi - lo) * (k + 1) / nb);
}

// --------------------------------------------------------------- set_params
// One thread per candidate: derived object constants from the physical
// parameters, written as a data block of LENSED_DWORDS words; with an EPL
// table, also the lens-frame pair (F q_n, S q_n) of every quadrature node
// offset, which the render kernel adds to the pixel's own part.
template <int K>
__device__ inline void lensed_node_pairs(const lensed_desc& d, u32* blk, const float4* __restrict__ quad,
                                         float4* __restrict__ npre, int cand)
{
    const float4* q = quad + d.quad_off;
    float4* out = npre + d.npre_off + (long long)cand * LENSED_NPRE_STRIDE(d.nq);
    for (int n = 0; n < d.nq; ++n) {
        lcl::float4 v = lcl::epl_pre_model<K>(blk, lcl::float2(q[n].x, q[n].y), false);
        out[n] = make_float4(v.x, v.y, v.z, v.w);
    }
    if (d.cull) {  // the candidate's constants of the pixel cull (lensed_frame_cull)
        lcl::float4 cc[LENSED_NPRE_CULL];
        lcl::cull_consts_model<K>(blk, cc);
        for (int i = 0; i < LENSED_NPRE_CULL; ++i) out[d.nq + i] = make_float4(cc[i].x, cc[i].y, cc[i].z, cc[i].w);
    }
}

__device__ __attribute__((always_inline)) inline void lensed_set_params_item(
    const lensed_desc* __restrict__ sys, int nsys, int nitems, int item, const float* __restrict__ params,
    u32* __restrict__ blocks, const float4* __restrict__ quad, float4* __restrict__ npre, int* __restrict__ live,
    int* __restrict__ owner, int layout)
{
    if (item >= nitems) return;
    if (item == 0) {
        if (live) live[0] = 0;  // the cull pass's list of lit squares starts empty
        // owner[0]: the launch layout that last wrote the model images;
        // owner[1]: whether that is this one, so that the cull pass may trust
        // its record of which squares hold nonzero values
        owner[1] = owner[0] == layout ? 1 : 0;
        owner[0] = layout;
    }
    int s = lensed_find_sys<2>(sys, nsys, item);
    const lensed_desc& d = sys[s];
    int cand = item - d.item_s;
    u32 blk[LENSED_DWORDS];
#pragma unroll
    for (int i = 0; i < LENSED_DWORDS; ++i) blk[i] = 0u;
    const float* p = params + d.param_off + (long long)cand * d.npar;
    switch (d.model) {
#define LENSED_SET_CASE(K)                                                \
    case K:                                                               \
        lcl::set_params_##K(blk, p);                                      \
        if (d.table_mode) lensed_node_pairs<K>(d, blk, quad, npre, cand); \
        break;
        LENSED_FOR_MODELS(LENSED_SET_CASE)
#undef LENSED_SET_CASE
        default: break;
    }
    u32* out = blocks + d.block_off + (long long)cand * d.dwords;
#pragma unroll
    for (int i = 0; i < LENSED_DWORDS; ++i)
        if (i < d.dwords) out[i] = blk[i];
}

extern "C" __global__ __launch_bounds__(256) void lensed_set_params(
    const lensed_desc* __restrict__ sys, int nsys, int nitems,
    const float* __restrict__ params, u32* __restrict__ blocks, const float4* __restrict__ quad,
    float4* __restrict__ npre, int* __restrict__ live, int* __restrict__ owner, int layout)
{
    lensed_set_params_item(sys, nsys, nitems, (int)(blockIdx.x * blockDim.x + threadIdx.x), params, blocks, quad, npre,
                           live, owner, layout);
}

// ------------------------------------------------------------------ render
// Pixel of this thread in 8 x 8 square w (0..3) of render item `tile`:
// items are 16 x 16 blocks (compact squares, so that a wave's pixels lie
// wholly inside or outside the region a source lights, for the pixel cull);
// false outside the image.
__device__ inline bool lensed_render_pixel(const lensed_desc& d, int tile, int w, int& px, int& py) {
    const int ntx = (d.width + LENSED_RENDER_SIDE - 1) / LENSED_RENDER_SIDE;
    const int by = tile / ntx, bx = tile - by * ntx;
    const int l = (int)threadIdx.x & 63;
    px = bx * LENSED_RENDER_SIDE + (w & 1) * 8 + (l & 7);
    py = by * LENSED_RENDER_SIDE + (w >> 1) * 8 + (l >> 3);
    return px < d.width && py < d.height;
}

// One thread per pixel (px, py; on: the thread has one); the 2-D quadrature
// loop runs in the thread with the node table in scalar registers and the
// candidate's object block held in SGPRs (wave-uniform), so the per-ray work
// is pure VALU.
template <int K, bool ERR>
__device__ inline void lensed_render_item(const lensed_desc& d, int cand, int px, int py, bool on, bool restage,
                                          const u32* __restrict__ blocks, const float4* __restrict__ quad,
                                          const float4* __restrict__ tables, const float4* __restrict__ npre,
                                          float4* tab_lds, float* __restrict__ model, float* __restrict__ errp)
{
    u32 blk[LENSED_DWORDS];
    const u32* gb = blocks + d.block_off + (long long)cand * d.dwords;
#pragma unroll
    for (int i = 0; i < LENSED_DWORDS; ++i) blk[i] = i < d.dwords ? __builtin_amdgcn_readfirstlane(gb[i]) : 0u;
    const float4* q = quad + d.quad_off;
    if (d.table_mode && restage) {  // the candidate's EPL angular table, staged in LDS
        __syncthreads();  // the previous candidate's table is no longer read
        const float4* src = tables + d.table_off + (long long)cand * LENSED_EPL_TABLE_F4;
        for (int i = threadIdx.x; i < LENSED_EPL_TABLE_F4; i += LENSED_RENDER_BS) tab_lds[i] = src[i];
        __syncthreads();
    }
    if (on) {
        const int k = py * d.width + px;
        // pixel centre as kernel/lensed.cl:25, then one ray per node
        const float cx = d.pcs[0] + d.pcs[2] * (float)px;
        const float cy = d.pcs[1] + d.pcs[3] * (float)py;
        float f = 0.0f, e = 0.0f;
        if (d.table_mode) {
            // F (p - c) + F q_n: the lens-frame pair of each ray from the
            // pixel's part and the node's (from set_params, scalar loads)
            const lcl::float4 base = lcl::epl_pre_model<K>(blk, lcl::float2(cx, cy), true);
            const float4* np_c = npre + d.npre_off + (long long)cand * LENSED_NPRE_STRIDE(d.nq);
            for (int n = 0; n < d.nq; ++n) {
                const float4 node = q[n];
                const float4 np = np_c[n];
                const lcl::float4 pre(base.x + np.x, base.y + np.y, base.z + np.z, base.w + np.w);
                float v = lcl::compute_model<K>(blk, lcl::float2(cx + node.x, cy + node.y),
                                                (const lcl::float4*)tab_lds, pre);
                f += node.z * v;
                if (ERR) e += node.w * v;
            }
        } else {
            for (int n = 0; n < d.nq; ++n) {
                const float4 node = q[n];
                float v = lcl::compute_model<K>(blk, lcl::float2(cx + node.x, cy + node.y), nullptr,
                                                lcl::float4(0.0f, 0.0f, 0.0f, 0.0f));
                f += node.z * v;
                if (ERR) e += node.w * v;
            }
        }
        long long o = d.model_off + (long long)cand * d.npix + k;
        model[o] = f;
        if (ERR) errp[o] = e;
    }
}

// eight waves per SIMD (at most 64 VGPRs): the ray loop's latencies
// (transcendentals, LDS lookups) are hidden by the other waves
extern "C" __global__ __launch_bounds__(LENSED_RENDER_BS) __attribute__((amdgpu_waves_per_eu(8, 8))) void lensed_render(
    const lensed_desc* __restrict__ sys, int nsys, int nitems,
    const u32* __restrict__ blocks, const float4* __restrict__ quad, const float4* __restrict__ tables,
    const float4* __restrict__ npre, const int* __restrict__ live, const int* __restrict__ trace,
    float* __restrict__ model, float* __restrict__ errp, int list_max)
{
    extern __shared__ float4 tab_lds[];
    // after a cull pass that left fewer than list_max quads lit, only the
    // lit 4 x 4 quads it listed, sixteen to an entry, all of one candidate
    // (live[1..live[0]]: quad = (item * 4 + w) * 4 + q, -1 for none), four to
    // each wave, a quad to each 16 lanes; otherwise every item, square w to
    // wave w, a culled system's dark squares (zeroed by the cull) skipped by
    // their flag.  (Walking a list that holds nearly every square measured
    // much slower than walking the items: 304 against 221 ms per launch of 64
    // systems whose sources never reach exact zero; list_max is three quads in
    // four by default.)
    const int wave = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    const bool listed = live && live[0] < list_max;
    int first, end;
    lensed_xcd_run(listed ? live[0] / LENSED_LIST_ENTRY : nitems, first, end);
    int staged = -1;  // first item of the candidate whose table is in LDS
    for (int j = first; j < end; ++j) {
        int item = j, q = -1;
        if (listed) {
            const int* e = live + 1 + (long long)LENSED_LIST_ENTRY * j;
            item = e[0] >> 4;                    // the entry's candidate is that of its first quad
            q = e[4 * wave + (lane >> 4)];       // this lane's quad
        }
        int s = lensed_find_sys<0>(sys, nsys, item);
        const lensed_desc& d = sys[s];
        int local = item - d.item_r;
        int cand = local / d.ntile_r;
        int px = 0, py = 0;
        bool on;
        if (listed) {
            // pixel (lane % 16) of quad q % 4 of square (q / 4) % 4 of render item q / 16
            const int ntx = (d.width + LENSED_RENDER_SIDE - 1) / LENSED_RENDER_SIDE;
            const int tile = (q >> 4) - d.item_r - cand * d.ntile_r, by = tile / ntx, bx = tile - by * ntx;
            const int w = (q >> 2) & 3, p = lane & 15;
            px = bx * LENSED_RENDER_SIDE + (w & 1) * 8 + (q & 1) * 4 + (p & 3);
            py = by * LENSED_RENDER_SIDE + (w >> 1) * 8 + ((q >> 1) & 1) * 4 + (p >> 2);
            on = q >= 0 && px < d.width && py < d.height;
        } else {
            on = true;
            if (live && d.cull)  // flags exist for the culled systems' squares only
                on = __builtin_amdgcn_readfirstlane(trace[(long long)j * 4 + wave]) != 0;
            on = on && lensed_render_pixel(d, j - d.item_r - cand * d.ntile_r, wave, px, py);
        }
        const int key = d.item_r + cand * d.ntile_r;  // block-uniform
        const bool restage = key != staged;
        staged = key;
        switch (d.model) {
#define LENSED_RENDER_CASE(K)                                                                                \
    case K:                                                                                                  \
        if (errp)                                                                                            \
            lensed_render_item<K, true>(d, cand, px, py, on, restage, blocks, quad, tables, npre, tab_lds,   \
                                        model, errp);                                                        \
        else                                                                                                 \
            lensed_render_item<K, false>(d, cand, px, py, on, restage, blocks, quad, tables, npre, tab_lds,  \
                                         model, errp);                                                       \
        break;
            LENSED_FOR_MODELS(LENSED_RENDER_CASE)
#undef LENSED_RENDER_CASE
            default: break;
        }
    }
}

// -------------------------------------------------------------------- cull
// Before the render: for each 8 x 8 square of each render item (a wave's
// pixels in the render's walk), and for each 4 x 4 quad of a square, whether
// any of its pixels can receive light (lensed_frame_cull in epl_table.h
// bounds where a pixel's rays land).  Dark quads get their zero model values
// here (where they may be missing) and are left out of the render; the lit
// ones go to a list that the render's blocks share out evenly, four quads to
// a wave (lit quads bunch in the middle of the image, so a fixed
// item-to-block mapping would leave the blocks of the edge tiles idle, and
// quads rather than squares trace about 8 % fewer rays at C3: the edge of
// the lit region cuts through squares).  One flag per square (`trace`, its
// mask of lit quads) tells the convolution which tiles are all dark and the
// render's walk which squares to trace.
// A cull work item is LENSED_CULL_GROUP render items of one candidate: 256
// squares, 64 to each wave, which settles its own squares without waiting
// for the others; a block takes consecutive items and stages a candidate's
// table once for all.  Only systems with d.cull (a tabulated EPL lens and one
// framed source with a cutoff) are tested; the quads of the others are all
// listed.
static_assert(LENSED_RENDER_BS == 256 && LENSED_CULL_GROUP == 64, "a cull item gives each wave 64 squares");

// lanes (pixels y * 8 + x of a square) of its quad q = 2 * (y / 4) + x / 4
#define LENSED_QUAD_LANES(q) (0x0F0F0F0Full << (((q) & 1) * 4 + ((q) >> 1) * 32))

// The squares of one wave of a cull item: lane i stands for square
// 64 * wave + i of the item (render item g0 + square / 4, its square
// square % 4) and returns that square's mask of lit quads.  x0, y0: the
// square's first pixel (lane-wise, for the caller's zeroing).
template <int K>
__device__ inline int lensed_cull_wave(const lensed_desc& d, int cand, int g0, int ng, int wave,
                                       const u32* __restrict__ blocks, const float4* __restrict__ npre,
                                       const float4* tab_lds, int* slots, int& x0, int& y0)
{
    u32 blk[LENSED_DWORDS];
    const u32* gb = blocks + d.block_off + (long long)cand * d.dwords;
#pragma unroll
    for (int i = 0; i < LENSED_DWORDS; ++i) blk[i] = i < d.dwords ? __builtin_amdgcn_readfirstlane(gb[i]) : 0u;
    const float4* ccg = npre + d.npre_off + (long long)cand * LENSED_NPRE_STRIDE(d.nq) + d.nq;
    lcl::float4 cc[LENSED_NPRE_CULL];
#pragma unroll
    for (int i = 0; i < LENSED_NPRE_CULL; ++i) cc[i] = lcl::float4(ccg[i].x, ccg[i].y, ccg[i].z, ccg[i].w);
    const lcl::float4* tab = (const lcl::float4*)tab_lds;
    const float ox = d.pcs[0], oy = d.pcs[1], sx = d.pcs[2], sy = d.pcs[3], qrad = d.qrad;
    const int W = d.width, H = d.height;
    const int ntx = (W + LENSED_RENDER_SIDE - 1) / LENSED_RENDER_SIDE;
    const int lane = (int)threadIdx.x & 63, sq = wave * 64 + lane, t = sq >> 2, w = sq & 3;
    // coarse pass, lane = square: the bound at the square's centre with rho
    // grown by the square's half-diagonal holds for every ray of the square,
    // so a square found dark here needs no per-pixel test.  A square whose
    // own centre ray lands within the cutoff is kept whole without the
    // per-pixel test (most such squares are lit throughout, and keeping a
    // quad is always exact: the render then computes its values, 0 where
    // they are 0)
    bool dark_sq = true, lit_sq = false;
    x0 = y0 = 0;
    if (t < ng) {
        const int tile = g0 + t, by = tile / ntx, bx = tile - by * ntx;
        x0 = bx * LENSED_RENDER_SIDE + (w & 1) * 8;
        y0 = by * LENSED_RENDER_SIDE + (w >> 1) * 8;
        const float px = (float)x0 + 3.5f, py = (float)y0 + 3.5f;
        const float rho = 3.5f * sqrtf(sx * sx + sy * sy) + qrad;
        const lcl::float4 base = lcl::epl_pre_model<K>(blk, lcl::float2(ox + sx * px, oy + sy * py), true);
        // squares past the image edge hold no pixel: dark
        dark_sq = x0 >= W || y0 >= H || lcl::cull_model<K>(blk, base, tab, cc, rho);
        lit_sq = !dark_sq && !lcl::cull_model<K>(blk, base, tab, cc, 0.0f);
    }
    int quads = lit_sq ? 15 : 0;
    const unsigned long long todo = __ballot(!dark_sq && !lit_sq);
    const int nu = __popcll(todo);
    if (nu == 0) return quads;
    // the unsettled squares (their lanes, in order) through the wave's LDS
    // slots: one wave writes and reads them, and its LDS accesses complete in
    // order; the compiler must keep that order
    asm volatile("" ::: "memory");
    if ((todo >> lane) & 1ull) slots[__popcll(todo & ((1ull << lane) - 1ull))] = lane;
    asm volatile("" ::: "memory");
    // quad pass, lane = (unsettled square, quad), sixteen squares at a time:
    // the bound at a quad's centre with rho grown by the quad's half-diagonal
    // settles its dark quads (about half of them at C3); the per-pixel pass,
    // lane = (open quad, pixel), then takes the open quads four at a time.  A
    // quad is lit where one of its pixels is, as from the square's pixels.
    const float rq = 1.5f * sqrtf(sx * sx + sy * sy) + qrad;
    for (int first = 0; first < nu; first += 16) {
        const int k = first + (lane >> 2), q = lane & 3;
        const bool valid = k < nu;
        const int sl = valid ? slots[k] : 0;  // the square's lane
        const int qx0 = __builtin_amdgcn_ds_bpermute(sl << 2, x0) + (q & 1) * 4;
        const int qy0 = __builtin_amdgcn_ds_bpermute(sl << 2, y0) + (q >> 1) * 4;
        bool open = false;
        if (valid && qx0 < W && qy0 < H) {
            const float px = (float)qx0 + 1.5f, py = (float)qy0 + 1.5f;
            const lcl::float4 base = lcl::epl_pre_model<K>(blk, lcl::float2(ox + sx * px, oy + sy * py), true);
            open = !lcl::cull_model<K>(blk, base, tab, cc, rq);
        }
        unsigned long long mq = __ballot(open);
        while (mq) {
            int b[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                b[j] = mq ? __builtin_ctzll(mq) : -1;
                mq &= mq - 1ull;  // (0 stays 0)
            }
            const int j = lane >> 4, pp = lane & 15;
            int bx[4], by[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                bx[i] = b[i] < 0 ? 0x3fffffff : __builtin_amdgcn_readlane(qx0, b[i]);
                by[i] = b[i] < 0 ? 0x3fffffff : __builtin_amdgcn_readlane(qy0, b[i]);
            }
            const int px = (j == 0 ? bx[0] : j == 1 ? bx[1] : j == 2 ? bx[2] : bx[3]) + (pp & 3);
            const int py = (j == 0 ? by[0] : j == 1 ? by[1] : j == 2 ? by[2] : by[3]) + (pp >> 2);
            bool l = false;
            if (px < W && py < H) {
                const float cx = ox + sx * (float)px, cy = oy + sy * (float)py;
                const lcl::float4 base = lcl::epl_pre_model<K>(blk, lcl::float2(cx, cy), true);
                l = !lcl::cull_model<K>(blk, base, tab, cc, qrad);
            }
            const unsigned long long m = __ballot(l);
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (b[i] >= 0 && ((m >> (16 * i)) & 0xFFFFull)) {
                    const int owner = __builtin_amdgcn_readlane(sl, b[i]);
                    if (lane == owner) quads |= 1 << (b[i] & 3);
                }
        }
    }
    return quads;
}

// live[0] = number of listed quads (zeroed by lensed_set_params), then the
// quads (square * 4 + q, square = item * 4 + w; -1 pads an entry), sixteen to
// an entry, all sixteen of one candidate: the render gives a block one entry
// and each wave four quads, so no wave of the render idles on a dark quad
// next to a lit one.  A block gathers its entries in LDS and reserves room
// for its whole entries with one atomic per chunk, so the counter sees a few
// thousand atomics per launch rather than one per entry.  An entry is padded
// only where the candidate changes and at the block's end (a chunk's partial
// last entry waits for the next chunk), so the list holds at most the quads
// plus fifteen pads per block: the host sizes it so.
#define LENSED_CULL_CHUNK 4096
extern "C" __global__ __launch_bounds__(LENSED_RENDER_BS) void lensed_cull(
    const lensed_desc* __restrict__ sys, int nsys, int nitems,
    const u32* __restrict__ blocks, const float4* __restrict__ tables, const float4* __restrict__ npre,
    int* __restrict__ trace, int* __restrict__ live, float* __restrict__ model, float* __restrict__ errp,
    const int* __restrict__ owner, int* __restrict__ written)
{
    extern __shared__ float4 tab_lds[];
    __shared__ int chunk[LENSED_CULL_CHUNK];
    __shared__ int wave_n[2][LENSED_RENDER_BS / 64];
    __shared__ int cull_slots[LENSED_RENDER_BS];  // each wave's 64: its unsettled squares
    __shared__ int at;
    // written[square]: the square's quads whose model values may be nonzero
    // (a render wrote them).  Trusted when this layout wrote the model images
    // last and no error plane is asked for (that plane is written only by
    // such launches); otherwise every dark quad is zeroed.
    const bool trust = owner[1] != 0 && !errp;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int lane = (int)threadIdx.x & 63;
    int* slots = cull_slots + wave * 64;
    // the chunk's fill, the same in every thread (what is listed is
    // block-uniform)
    int n = 0, parity = 0;
    auto flush = [&]() {
        const int whole = n & ~(LENSED_LIST_ENTRY - 1);
        __syncthreads();  // the chunk is complete
        if (threadIdx.x == 0) at = whole ? atomicAdd(live, whole) : 0;
        __syncthreads();
        for (int i = threadIdx.x; i < whole; i += LENSED_RENDER_BS) live[1 + at + i] = chunk[i];
        const int rest = n - whole;  // a partial entry moves to the front
        const int keep = (int)threadIdx.x < rest ? chunk[whole + threadIdx.x] : 0;
        __syncthreads();  // every read of the chunk and of `at` is done
        if ((int)threadIdx.x < rest) chunk[threadIdx.x] = keep;
        n = rest;
    };
    auto pad = [&]() {
        const int p = (LENSED_LIST_ENTRY - (n & (LENSED_LIST_ENTRY - 1))) & (LENSED_LIST_ENTRY - 1);
        if ((int)threadIdx.x < p) chunk[n + threadIdx.x] = -1;
        n += p;
    };
    // each wave's lit quads (lane i: the quad mask of square 64 * wave + i
    // of the item starting at square0), square by square, after those of the
    // waves before it; the waves' counts meet in LDS (two sets, so one
    // barrier per item suffices: a set is written again only after every
    // wave has passed the next item's barrier)
    auto list_item = [&](int square0, int quads) {
        if (n + LENSED_CULL_GROUP * 16 + LENSED_LIST_ENTRY - 1 > LENSED_CULL_CHUNK) flush();
        const unsigned long long m0 = __ballot(quads & 1), m1 = __ballot(quads & 2);
        const unsigned long long m2 = __ballot(quads & 4), m3 = __ballot(quads & 8);
        const unsigned long long below = (1ull << lane) - 1ull;
        const int mine = __popcll(m0) + __popcll(m1) + __popcll(m2) + __popcll(m3);
        if (lane == 0) wave_n[parity][wave] = mine;
        __syncthreads();
        int before = 0, total = 0;
#pragma unroll
        for (int v = 0; v < LENSED_RENDER_BS / 64; ++v) {
            const int c = wave_n[parity][v];
            before += v < wave ? c : 0;
            total += c;
        }
        int at_q = n + before + __popcll(m0 & below) + __popcll(m1 & below) + __popcll(m2 & below) +
                   __popcll(m3 & below);
        const int qid = (square0 + wave * 64 + lane) * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if ((quads >> q) & 1) chunk[at_q++] = qid + q;
        n += total;
        parity ^= 1;
    };
    int staged = -1;  // first item of the candidate whose table is in LDS and whose quads the chunk ends with
    int first, end;
    lensed_xcd_run(nitems, first, end);
    for (int k = first; k < end; ++k) {
        int s = lensed_find_sys<5>(sys, nsys, k);
        const lensed_desc& d = sys[s];
        const int groups = (d.ntile_r + LENSED_CULL_GROUP - 1) / LENSED_CULL_GROUP;
        const int local = k - d.item_k;
        const int cand = local / groups;
        const int g0 = (local - cand * groups) * LENSED_CULL_GROUP;
        const int ng = min(LENSED_CULL_GROUP, d.ntile_r - g0);
        const int item0 = d.item_r + cand * d.ntile_r + g0;
        const int key = d.item_r + cand * d.ntile_r;  // block-uniform
        const int sq = wave * 64 + lane;              // this lane's square of the item
        if (key != staged) {
            pad();  // entries never mix candidates
            if (d.cull) {
                __syncthreads();  // no wave reads the previous candidate's table any more
                const float4* src = tables + d.table_off + (long long)cand * LENSED_EPL_TABLE_F4;
                for (int i = threadIdx.x; i < LENSED_EPL_TABLE_F4; i += LENSED_RENDER_BS) tab_lds[i] = src[i];
                __syncthreads();
            }
        }
        staged = key;
        if (!d.cull) {
            list_item(4 * item0, sq < 4 * ng ? 15 : 0);
            continue;
        }
        int quads = 0, x0 = 0, y0 = 0;
        switch (d.model) {
#define LENSED_CULL_CASE(K)                                                                         \
    case K:                                                                                         \
        quads = lensed_cull_wave<K>(d, cand, g0, ng, wave, blocks, npre, tab_lds, slots, x0, y0);    \
        break;
            LENSED_FOR_MODELS(LENSED_CULL_CASE)
#undef LENSED_CULL_CASE
            default: break;
        }
        // the squares' flags (for the convolution, the render's walk and the
        // next launch), one coalesced store per wave; the dark quads that may
        // hold nonzero values get their zeros, a square at a time, lane =
        // pixel
        int zero = 0;
        if (sq < 4 * ng) {
            const long long sqi = (long long)item0 * (LENSED_RENDER_BS / 64) + sq;
            trace[sqi] = quads;
            zero = ~quads & (trust ? written[sqi] : 15) & 15;
            written[sqi] = quads;
        }
        unsigned long long todo = __ballot(zero != 0);
        const int my_quad = ((lane >> 5) << 1) | ((lane >> 2) & 1);  // this lane's quad as a pixel of a square
        while (todo) {
            const int b = __builtin_ctzll(todo);
            todo &= todo - 1ull;
            const int px = __builtin_amdgcn_readlane(x0, b) + (lane & 7);
            const int py = __builtin_amdgcn_readlane(y0, b) + (lane >> 3);
            const int zq = __builtin_amdgcn_readlane(zero, b);
            if (((zq >> my_quad) & 1) && px < d.width && py < d.height) {
                const long long o = d.model_off + (long long)cand * d.npix + (long long)py * d.width + px;
                model[o] = 0.0f;
                if (errp) errp[o] = 0.0f;
            }
        }
        list_item(4 * item0, quads);
    }
    pad();
    if (n > 0) flush();
}

// -------------------------------------------------------------- EPL tables
// One block per candidate: thread i evaluates the tabulated factor W and
// dW/dp (the tenth-order series times the radial factor, epl_series.h) at
// node p = i / NT in FP64 from the candidate's (q, t), thread 0 also at
// p = 1; the nodes meet in LDS, and thread i turns nodes i and i + 1 into
// the cubic Hermite coefficients of row i (each node evaluated once, not
// once per interval end); the last row is the constant value at p = 1.
// Block b is the candidate whose rows start at row item b * (NT + 1) (the
// systems' rows follow each other).  With `only` set, just the flagged
// candidates (the emulator's fallback).
static_assert(LENSED_EPL_NT == 256, "a table block has a thread per interval");
__device__ __attribute__((always_inline)) inline void lensed_epl_tables_block(
    const lensed_desc* __restrict__ sys, int nsys, int nitems, int b, const float* __restrict__ params,
    float4* __restrict__ tables, const int* __restrict__ only)
{
    __shared__ double nx[LENSED_EPL_ROWS], ny[LENSED_EPL_ROWS], ndx[LENSED_EPL_ROWS], ndy[LENSED_EPL_ROWS];
    const int row0 = b * LENSED_EPL_ROWS;
    if (row0 >= nitems) return;
    int s = lensed_find_sys<4>(sys, nsys, row0);
    const lensed_desc& d = sys[s];
    const int cand = (row0 - d.item_t) / LENSED_EPL_ROWS;
    if (only && !only[d.table_off / LENSED_EPL_TABLE_F4 + cand]) return;  // the whole block
    const float* p = params + d.param_off + (long long)cand * d.npar;
    const double q = p[d.epl_q], t = p[d.epl_t];
    const int i = (int)threadIdx.x;
    lensed_epl_node(q, t, (double)i / LENSED_EPL_NT, nx[i], ny[i], ndx[i], ndy[i]);
    if (i == 0)
        lensed_epl_node(q, t, (double)LENSED_EPL_NT / LENSED_EPL_NT, nx[LENSED_EPL_NT], ny[LENSED_EPL_NT],
                        ndx[LENSED_EPL_NT], ndy[LENSED_EPL_NT]);
    __syncthreads();
    float4* out = tables + d.table_off + (long long)cand * LENSED_EPL_TABLE_F4;
    out[LENSED_EPL_AT(0, i)] = lensed_hermite(nx[i], ndx[i], nx[i + 1], ndx[i + 1]);
    out[LENSED_EPL_AT(1, i)] = lensed_hermite(ny[i], ndy[i], ny[i + 1], ndy[i + 1]);
    if (i == 0) {  // the end point p = 1 as a constant row
        out[LENSED_EPL_AT(0, LENSED_EPL_NT)] = float4((float)nx[LENSED_EPL_NT], 0.0f, 0.0f, 0.0f);
        out[LENSED_EPL_AT(1, LENSED_EPL_NT)] = float4((float)ny[LENSED_EPL_NT], 0.0f, 0.0f, 0.0f);
    }
}

extern "C" __global__ __launch_bounds__(LENSED_EPL_NT) void lensed_epl_tables(
    const lensed_desc* __restrict__ sys, int nsys, int nitems,
    const float* __restrict__ params, float4* __restrict__ tables, const int* __restrict__ only)
{
    lensed_epl_tables_block(sys, nsys, nitems, (int)blockIdx.x, params, tables, only);
}

// lensed_set_params and lensed_epl_tables (every candidate) in one launch:
// the two read the parameters and write different buffers, so the first
// set_blocks blocks are set_params' blocks and the rest the tables' blocks,
// run side by side (set_params' one thread per candidate is latency-bound,
// the tables' block per candidate fills the device meanwhile).  The same
// code per thread as the two launches, so the same bits.
extern "C" __global__ __launch_bounds__(LENSED_EPL_NT) void lensed_set_params_tables(
    const lensed_desc* __restrict__ sys, int nsys, int nitems,
    const float* __restrict__ params, u32* __restrict__ blocks, const float4* __restrict__ quad,
    float4* __restrict__ npre, int* __restrict__ live, int* __restrict__ owner, int layout,
    int ntable, float4* __restrict__ tables, int set_blocks)
{
    const int b = (int)blockIdx.x;
    if (b < set_blocks)
        lensed_set_params_item(sys, nsys, nitems, b * LENSED_EPL_NT + (int)threadIdx.x, params, blocks, quad, npre,
                               live, owner, layout);
    else
        lensed_epl_tables_block(sys, nsys, ntable, b - set_blocks, params, tables, nullptr);
}

// --------------------------------------------------------------- conv+chi2
// Convolution with the flipped PSF and clamp-to-edge borders, as
// kernel/lensed.cl:56-103: out(y, x) = sum_{j,i} psf[j][i] *
// in(clamp(y + PH-1-PH/2 - j), clamp(x + PW-1-PW/2 - i)).  An output tile of
// 64 x (4*RY) pixels plus halo is staged in LDS; each thread produces RY rows
// of one column.  PSFs that are numerically a sum of few outer products
// (sep_rank > 0, factors from the host SVD) go through row and column passes;
// the others are summed directly.  The chi^2 term w (m - d)^2 is formed per
// pixel in FP32 (kernel/lensed.cl:50-51) and summed with compensation over
// fixed blocks of 8 rows x 64 columns: TwoSum down each column, then a TwoSum
// tree across the 64 lanes, so each block yields the FP32 pair (sum, error)
// of its terms' total; blocks are combined in FP64 in a fixed order.  The
// block is the same whatever the tile height, so every launch shape gives
// the same bits for the same candidate.
__device__ inline double lensed_wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline void lensed_two_sum_add(float& s, float& c, float x) {
#pragma clang fp reassociate(off) contract(off)
    float t = s + x;  // TwoSum(s, x): t + e == s + x exactly
    float bb = t - s;
    float e = (s - (t - bb)) + (x - bb);
    s = t;
    c += e;
}

// The chi^2 term of one pixel, w (m - d)^2 in FP32 (kernel/lensed.cl:50-51),
// 0 for a lane past the image's right edge; the sweep and the dark partials
// share it, so that a block with a model of 0 gives the same bits in both.
__device__ inline float lensed_chi_term(float m, float dat, float wt, bool in)
{
#pragma clang fp reassociate(off)
    const float dm = m - dat;
    return in ? wt * dm * dm : 0.0f;
}

// the value of lane ^ o: a swizzle within each 32-lane half for o < 32
// (bitmask mode: and 0x1f, or 0, xor o), a permute across the halves for 32;
// no lane arithmetic, unlike a general shuffle
template <int O>
__device__ inline float lensed_lane_xor(float v) {
    if (O == 32) {
        const int lane = (int)__lane_id();
        return __int_as_float(__builtin_amdgcn_ds_bpermute((lane ^ 32) << 2, __float_as_int(v)));
    }
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), (O << 10) | 0x1f));
}

__device__ inline double lensed_wave_csum(float s, float c) {
#pragma clang fp reassociate(off) contract(off)
#pragma unroll
    for (int lv = 0; lv < 6; ++lv) {
        const int o = 32 >> lv;
        float s2, c2;
        switch (o) {
            case 32: s2 = lensed_lane_xor<32>(s), c2 = lensed_lane_xor<32>(c); break;
            case 16: s2 = lensed_lane_xor<16>(s), c2 = lensed_lane_xor<16>(c); break;
            case 8: s2 = lensed_lane_xor<8>(s), c2 = lensed_lane_xor<8>(c); break;
            case 4: s2 = lensed_lane_xor<4>(s), c2 = lensed_lane_xor<4>(c); break;
            case 2: s2 = lensed_lane_xor<2>(s), c2 = lensed_lane_xor<2>(c); break;
            default: s2 = lensed_lane_xor<1>(s), c2 = lensed_lane_xor<1>(c); break;
        }
        float t = s + s2;  // TwoSum(s, s2): t + e == s + s2 exactly
        float bb = t - s;
        float e = (s - (t - bb)) + (s2 - bb);
        s = t;
        c = (c + c2) + e;
    }
    return (double)s + (double)c;
}

// Stage the model tile plus halo (clamped to the image) in LDS; row stride
// LENSED_TILE_STRIDE(PW) keeps rows 16-byte aligned for vector reads.  A
// thread owns columns lx and lx + 64 (the halo is narrower than 64) of
// every fourth row; the loads of eight rows are issued together before any
// is stored, so their latencies overlap.
template <int RY>
__device__ inline void lensed_conv_load_tile(const lensed_desc& d, const float* __restrict__ m, float* tile,
                                             int tx0, int ty0, int lx, int ly, int PW, int PH)
{
    constexpr int TX = LENSED_CONV_TX, TY = LENSED_CONV_TY, NB = 8;
    const int W = d.width, H = d.height;
    const int cw = TX + PW - 1, ch = TY * RY + PH - 1;
    const int stride = LENSED_TILE_STRIDE(PW);
    const int ox = tx0 - PW / 2, oy = ty0 - PH / 2;
    const int c0 = min(max(ox + lx, 0), W - 1), c1 = min(max(ox + lx + TX, 0), W - 1);
    const bool second = lx + TX < cw;
    for (int r0 = ly; r0 < ch; r0 += NB * TY) {
        float v0[NB], v1[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const int r = r0 + k * TY;
            v0[k] = v1[k] = 0.0f;
            if (r < ch) {
                const float* src = m + min(max(oy + r, 0), H - 1) * W;
                v0[k] = src[c0];
                if (second) v1[k] = src[c1];
            }
        }
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const int r = r0 + k * TY;
            if (r < ch) {
                tile[r * stride + lx] = v0[k];
                if (second) tile[r * stride + lx + TX] = v1[k];
            }
        }
    }
}

// PSF of any size, sizes from the descriptor
template <int RY>
__device__ inline void lensed_conv_generic(const lensed_desc& d, const float* __restrict__ psf, float* tile,
                                           int lx, int ly, float (&acc)[RY])
{
    constexpr int TX = LENSED_CONV_TX, TH = LENSED_CONV_TY * RY, BS = LENSED_CONV_BS;
    const int PW = d.psfw, PH = d.psfh;
    const int stride = LENSED_TILE_STRIDE(PW), ch = TH + PH - 1;
    if (d.sep_rank > 0) {
        // row-pass results, ch rows: over the tile itself for rank 1 (a row
        // is read and then written by one wave, whose LDS accesses complete
        // in order), after it for higher ranks, which read the tile again
        const bool inplace = d.sep_rank == 1;
        float* tmp = inplace ? tile : tile + stride * ch;
        const int ts = inplace ? stride : TX;
        for (int k = 0; k < d.sep_rank; ++k) {
            const float* cf = psf + d.sep_off + (long long)k * (PH + PW);
            const float* rf = cf + PH;
            for (int idx = threadIdx.x; idx < ch * TX; idx += BS) {
                int r = idx / TX, x = idx - r * TX;
                const float* src = tile + r * stride + x + PW - 1;
                float sum = 0.0f;
                for (int i = 0; i < PW; ++i) sum += rf[i] * src[-i];
                tmp[r * ts + x] = sum;
            }
            __syncthreads();
            for (int j = 0; j < PH; ++j) {
                const float w = cf[j];
                const float* col = tmp + (ly * RY + PH - 1 - j) * ts + lx;
#pragma unroll
                for (int r = 0; r < RY; ++r) acc[r] += w * col[r * ts];
            }
            __syncthreads();
        }
    } else {
        const float* p = psf + d.psf_off;
        for (int j = 0; j < PH; ++j) {
            const float* row = tile + (ly * RY + PH - 1 - j) * stride + lx + PW - 1;
            for (int i = 0; i < PW; ++i) {
                const float w = p[j * PW + i];
#pragma unroll
                for (int r = 0; r < RY; ++r) acc[r] += w * row[r * stride - i];
            }
        }
        __syncthreads();
    }
}

// PSF size fixed at compile time: taps unrolled into scalar registers and
// every LDS value reused from VGPRs.  Separable: the row pass computes four
// adjacent outputs per thread from one vector read of 4 + PW - 1 values;
// the column pass RY outputs from RY + PH - 1 values.  Direct: per PSF
// column, RY + PH - 1 values feed RY x PH multiply-adds (taps stored
// transposed so a column's taps are contiguous).
// One rank of the separable convolution: row pass into `tmp` (row stride
// TS), then the column pass into acc.  With TMP_IN_TILE (rank 1) the row
// pass writes over the tile itself: a row is read and then written by one
// wave, whose LDS accesses complete in order.
template <int RY, int PW, int PH, bool TMP_IN_TILE>
__device__ inline void lensed_conv_sep_rank(const float* __restrict__ cf, float* tile, int lx, int ly,
                                            float (&acc)[RY])
{
    constexpr int TX = LENSED_CONV_TX, TH = LENSED_CONV_TY * RY, BS = LENSED_CONV_BS;
    constexpr int CH = TH + PH - 1, STRIDE = LENSED_TILE_STRIDE(PW), NC = RY + PH - 1;
    constexpr int NX = 4, NV = (NX + PW - 1 + 3) / 4;
    constexpr int TS = TMP_IN_TILE ? STRIDE : TX;
    float* tmp = TMP_IN_TILE ? tile : tile + STRIDE * CH;
    const float* rf = cf + PH;
    for (int r = threadIdx.x / (TX / NX); r < CH; r += BS / (TX / NX)) {
        const int x0 = (threadIdx.x % (TX / NX)) * NX;
        const float4* src = (const float4*)(tile + r * STRIDE + x0);
        float v[4 * NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            float4 f = src[q];
            v[4 * q] = f.x; v[4 * q + 1] = f.y; v[4 * q + 2] = f.z; v[4 * q + 3] = f.w;
        }
        // the chain's order written out (first tap a product, then fused
        // multiply-adds in tap order): contraction may not pick another
        // pairing, which the column sweep repeats bit for bit
        float o[NX];
#pragma unroll
        for (int x = 0; x < NX; ++x) o[x] = rf[0] * v[x + PW - 1];
#pragma unroll
        for (int i = 1; i < PW; ++i)
#pragma unroll
            for (int x = 0; x < NX; ++x) o[x] = __builtin_fmaf(rf[i], v[x + PW - 1 - i], o[x]);
        *(float4*)(tmp + r * TS + x0) = make_float4(o[0], o[1], o[2], o[3]);
    }
    __syncthreads();
    float col[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) col[q] = tmp[(ly * RY + q) * TS + lx];
#pragma unroll
    for (int j = 0; j < PH; ++j)
#pragma unroll
        for (int r = 0; r < RY; ++r) acc[r] = __builtin_fmaf(cf[j], col[r + PH - 1 - j], acc[r]);
    __syncthreads();
}

template <int RY, int PW, int PH, bool SEP>
__device__ inline void lensed_conv_fixed(const lensed_desc& d, const float* __restrict__ psf, float* tile,
                                         int lx, int ly, float (&acc)[RY])
{
    constexpr int STRIDE = LENSED_TILE_STRIDE(PW), NC = RY + PH - 1;
    if (SEP) {
        if (d.sep_rank == 1) {
            lensed_conv_sep_rank<RY, PW, PH, true>(psf + d.sep_off, tile, lx, ly, acc);
        } else {
            for (int k = 0; k < d.sep_rank; ++k)
                lensed_conv_sep_rank<RY, PW, PH, false>(psf + d.sep_off + (long long)k * (PH + PW), tile, lx, ly,
                                                        acc);
        }
    } else {
        const float* pt = psf + d.psft_off;  // [PW][PH]
#pragma unroll 1
        for (int i = 0; i < PW; ++i) {
            float col[NC];
#pragma unroll
            for (int q = 0; q < NC; ++q) col[q] = tile[(ly * RY + q) * STRIDE + lx + PW - 1 - i];
#pragma unroll
            for (int j = 0; j < PH; ++j)
#pragma unroll
                for (int r = 0; r < RY; ++r) acc[r] += pt[i * PH + j] * col[r + PH - 1 - j];
        }
        __syncthreads();
    }
}

template <int RY>
__device__ inline void lensed_conv_item(const lensed_desc& d, int cand, int t,
                                        const float* __restrict__ model, const float* __restrict__ image,
                                        const float* __restrict__ weight, const float* __restrict__ psf,
                                        double* __restrict__ partials, float* __restrict__ conv_out,
                                        float* __restrict__ chi2_out, const int* __restrict__ trace, float* tile,
                                        double* red, float* rows8)
{
#pragma clang fp reassociate(off)
    constexpr int TX = LENSED_CONV_TX, TH = LENSED_CONV_TY * RY;
    const int lx = threadIdx.x % TX;
    const int ly = threadIdx.x / TX;
    const int W = d.width, H = d.height;
    const int tx0 = (t % d.tiles_x) * TX;
    const int ty0 = (t / d.tiles_x) * TH;
    const float* m = model + d.model_off + (long long)cand * d.npix;
    const float* img = image + d.image_off;
    const float* wgt = weight + d.image_off;
    const int gx = tx0 + lx;
    float acc[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) acc[r] = 0.0f;
    // data and weights of this thread's pixels, requested once the tile is
    // staged, so that their loads overlap the convolution
    float dat[RY], wt[RY];
    auto fetch = [&]() {
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int gy = ty0 + ly * RY + r;
            dat[r] = wt[r] = 0.0f;
            if (gx < W && gy < H) {
                dat[r] = img[gy * W + gx];
                wt[r] = wgt[gy * W + gx];
            }
        }
    };
    // after a pixel cull: if every render item under the tile's input
    // region (the tile and its PSF halo, clamped to the image as the tile
    // load clamps) was found dark, the model there is exactly 0 and so is
    // the convolution: chi^2 needs only data and weights
    bool dark = false;
    if (trace && d.cull && d.psfw > 0) {
        const int ox = tx0 - d.psfw / 2, oy = ty0 - d.psfh / 2;
        const int c0 = min(max(ox, 0), W - 1), c1 = min(max(ox + TX + d.psfw - 2, 0), W - 1);
        const int r0 = min(max(oy, 0), H - 1), r1 = min(max(oy + TH + d.psfh - 2, 0), H - 1);
        const int ntx = (W + LENSED_RENDER_SIDE - 1) / LENSED_RENDER_SIDE;
        const int bx0 = c0 / LENSED_RENDER_SIDE, nbx = c1 / LENSED_RENDER_SIDE - bx0 + 1;
        const int by0 = r0 / LENSED_RENDER_SIDE, nby = r1 / LENSED_RENDER_SIDE - by0 + 1;
        const int4* flags = (const int4*)trace + d.item_r + (long long)cand * d.ntile_r;
        int lit = 0;
        for (int i = threadIdx.x; i < nbx * nby; i += LENSED_CONV_BS) {
            const int by = by0 + i / nbx, bx = bx0 + i % nbx;
            const int4 f = flags[by * ntx + bx];
            lit |= f.x | f.y | f.z | f.w;
        }
        dark = __syncthreads_or(lit) == 0;
    }
    if (dark) {
        fetch();
    } else if (d.psfw > 0) {
        lensed_conv_load_tile<RY>(d, m, tile, tx0, ty0, lx, ly, d.psfw, d.psfh);
        __syncthreads();
        fetch();
        bool done = false;
#define LENSED_PSF_CASE(PW_, PH_, SEP_)                                                  \
        if (!done && d.psfw == PW_ && d.psfh == PH_ && (d.sep_rank > 0) == SEP_) {       \
            lensed_conv_fixed<RY, PW_, PH_, SEP_>(d, psf, tile, lx, ly, acc);            \
            done = true;                                                                 \
        }
        LENSED_FOR_PSFS(LENSED_PSF_CASE)
#undef LENSED_PSF_CASE
        if (!done) lensed_conv_generic<RY>(d, psf, tile, lx, ly, acc);
    } else {
        fetch();
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            int gy = ty0 + ly * RY + r;
            acc[r] = (gx < W && gy < H) ? m[gy * W + gx] : 0.0f;
        }
    }
    float chi[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        const int gy = ty0 + ly * RY + r;  // wave-uniform
        chi[r] = 0.0f;
        if (gx < W && gy < H) {
            int kk = gy * W + gx;
            float dm = acc[r] - dat[r];
            chi[r] = wt[r] * dm * dm;
            if (conv_out) conv_out[d.model_off + (long long)cand * d.npix + kk] = acc[r];
            if (chi2_out) chi2_out[d.model_off + (long long)cand * d.npix + kk] = chi[r];
        }
    }
    // canonical 8-row blocks: a thread of the tall tile owns a column of
    // RY / 8 blocks; the short tile's 4 threads pass their rows through LDS
    double* part = partials + d.partial_off + (long long)cand * d.npart + t % d.tiles_x;
    if (RY % 8 == 0) {
#pragma unroll
        for (int g = 0; g < RY / 8; ++g) {
            float s = 0.0f, c = 0.0f;
#pragma unroll
            for (int r = 0; r < 8; ++r) lensed_two_sum_add(s, c, chi[8 * g + r]);
            const double blk = lensed_wave_csum(s, c);
            const int block_row = (ty0 + ly * RY) / 8 + g;
            if (lx == 0 && block_row * 8 < H) part[(long long)block_row * d.tiles_x] = blk;
        }
    } else {
        float s = 0.0f, c = 0.0f;
#pragma unroll
        for (int r = 0; r < RY; ++r) rows8[(ly * RY + r) * TX + lx] = chi[r];
        __syncthreads();
        if (ly == 0) {
            for (int r = 0; r < 8; ++r) lensed_two_sum_add(s, c, rows8[r * TX + lx]);
            const double blk = lensed_wave_csum(s, c);
            const int block_row = ty0 / 8;
            if (lx == 0 && block_row * 8 < H) part[(long long)block_row * d.tiles_x] = blk;
        }
        __syncthreads();  // rows8 is reused by the next item
    }
    (void)red;
}

// Two item orders.  order 0: candidate-fastest, items strided over the
// persistent grid, so every block sees a mix of tiles (dark tiles after a
// pixel cull cost less).  order 1: tile-fastest within a candidate, down
// each 64-column strip of the image, and each block a contiguous run of its
// XCD's eighth of the items (lensed_xcd_run): consecutive tiles of a block
// share their PSF halo rows, which the previous tile just brought into the
// XCD's L2.
template <int RY>
__device__ inline void lensed_conv_chi2_body(const lensed_desc* __restrict__ sys, int nsys, int nitems,
                                             const float* __restrict__ model, const float* __restrict__ image,
                                             const float* __restrict__ weight, const float* __restrict__ psf,
                                             double* __restrict__ partials, float* __restrict__ conv_out,
                                             float* __restrict__ chi2_out, const int* __restrict__ trace, int order)
{
    extern __shared__ float4 tile4[];
    float* tile = (float*)tile4;
    __shared__ double red[LENSED_CONV_BS / 64];
    __shared__ float rows8[8 * LENSED_CONV_TX];
    auto run = [&](int item) {
        int s = lensed_find_sys<1>(sys, nsys, item);
        const lensed_desc& d = sys[s];
        int local = item - d.item_c;
        int t, cand;
        if (order == 0) {
            t = local / d.ncand;
            cand = local - t * d.ncand;
        } else {
            cand = local / d.ntile_c;
            const int k = local - cand * d.ntile_c, tiles_y = d.ntile_c / d.tiles_x;
            const int tx = k / tiles_y, ty = k - tx * tiles_y;
            t = ty * d.tiles_x + tx;
        }
        lensed_conv_item<RY>(d, cand, t, model, image, weight, psf, partials, conv_out, chi2_out, trace, tile, red,
                             rows8);
    };
    if (order == 0) {
        for (int item = blockIdx.x; item < nitems; item += gridDim.x) run(item);
    } else {
        int first, end;
        lensed_xcd_run(nitems, first, end);
        for (int item = first; item < end; ++item) run(item);
    }
}

// six waves per SIMD (at most 80 VGPRs): with the rank-1 row pass over the
// tile (18 KiB of LDS for a 21 x 21 PSF) six blocks fit a CU
extern "C" __global__ __launch_bounds__(LENSED_CONV_BS) __attribute__((amdgpu_waves_per_eu(6))) void lensed_conv_chi2(
    const lensed_desc* __restrict__ sys, int nsys, int nitems,
    const float* __restrict__ model, const float* __restrict__ image,
    const float* __restrict__ weight, const float* __restrict__ psf,
    double* __restrict__ partials, float* __restrict__ conv_out, float* __restrict__ chi2_out,
    const int* __restrict__ trace, int order)
{
    lensed_conv_chi2_body<LENSED_CONV_RY>(sys, nsys, nitems, model, image, weight, psf, partials, conv_out,
                                          chi2_out, trace, order);
}

extern "C" __global__ __launch_bounds__(LENSED_CONV_BS) void lensed_conv_chi2_small(
    const lensed_desc* __restrict__ sys, int nsys, int nitems,
    const float* __restrict__ model, const float* __restrict__ image,
    const float* __restrict__ weight, const float* __restrict__ psf,
    double* __restrict__ partials, float* __restrict__ conv_out, float* __restrict__ chi2_out,
    const int* __restrict__ trace, int order)
{
    lensed_conv_chi2_body<LENSED_CONV_RY_SMALL>(sys, nsys, nitems, model, image, weight, psf, partials,
                                                conv_out, chi2_out, trace, order);
}

// 64-row tiles (16 rows per thread): a third less halo per output row than
// the 32-row tile, at four waves per SIMD for the larger register tile
extern "C" __global__ __launch_bounds__(LENSED_CONV_BS) __attribute__((amdgpu_waves_per_eu(4))) void lensed_conv_chi2_tall(
    const lensed_desc* __restrict__ sys, int nsys, int nitems,
    const float* __restrict__ model, const float* __restrict__ image,
    const float* __restrict__ weight, const float* __restrict__ psf,
    double* __restrict__ partials, float* __restrict__ conv_out, float* __restrict__ chi2_out,
    const int* __restrict__ trace, int order)
{
    lensed_conv_chi2_body<LENSED_CONV_RY_TALL>(sys, nsys, nitems, model, image, weight, psf, partials,
                                               conv_out, chi2_out, trace, order);
}

// ------------------------------------------------------ conv+chi2, sweep
// Rank-1 separable PSFs of at most 21 x 21, images up to 512 rows (the
// host checks both): one wave per 64-column strip
// of one candidate's image sweeps the strip from top to bottom.  Per input
// row, the row pass of the strip's 64 columns, from the row and its PSF halo
// staged in the wave's LDS row, goes into a window of the last PH row-pass
// values per lane (registers; the sweep is unrolled PH rows at a time, so
// every window index is static).  Per output row, the column pass over the
// window, the chi^2 term, and every 8 rows the compensated block sum.  Each
// output is the tile kernel's FMA chain (row taps in order, then column taps
// in order, clamp-to-edge rows and columns), and the chi^2 blocks are the
// same 8 x 64 blocks with the same lane order, so both kernels give the same
// bits.  Rows above and below the image are its first and last rows, whose
// row pass is computed once.  After a pixel cull, an input row whose 8-row
// band is dark over the strip and its halo is exactly 0: it is neither
// loaded nor row-passed; an output row whose whole window is dark is exactly
// 0 without the column pass.  Nothing is reread: each row of the model is
// loaded once per strip (its halo columns also by the neighbouring strips,
// which run beside it in the same XCD), against the tile kernel's halo rows.
#define LENSED_SWEEP_ROW 96  // floats per wave's staged row: 64 + PW - 1 for PW <= LENSED_SWEEP_MAX_PSF

template <int PW, int PH, bool OUT>
__device__ inline void lensed_conv_sweep_item(const lensed_desc& d, int cand, int strip,
                                              const float* __restrict__ model, const float* __restrict__ image,
                                              const float* __restrict__ weight, const float* __restrict__ psf,
                                              double* __restrict__ partials, float* __restrict__ conv_out,
                                              float* __restrict__ chi2_out, const int* __restrict__ trace,
                                              const double* __restrict__ dark, float* lrow)
{
#pragma clang fp reassociate(off)
    constexpr int A = PH - 1 - PH / 2;  // output row y takes row-pass rows y + A down to y + A - (PH - 1)
    constexpr int LW = 64 + PW - 1;     // the staged row: the strip and its PSF halo
    static_assert(LW <= LENSED_SWEEP_ROW, "PSF wider than the staged row");
    const int lane = (int)threadIdx.x & 63;
    const int W = d.width, H = d.height;
    const int x0 = strip * 64, gx = x0 + lane;
    const float* m = model + d.model_off + (long long)cand * d.npix;
    const float* img = image + d.image_off;
    const float* wgt = weight + d.image_off;
    const float* cf = psf + d.sep_off;  // column factor [PH], then row factor [PW]
    const float* rf = cf + PH;
    // the two staged columns of this lane (lane and lane + 64 of the row),
    // clamped to the image as the tile load clamps; the second is loaded by
    // every lane (a clamped column, cached) and stored by those it belongs to
    const unsigned ca = 4u * (unsigned)min(max(x0 - PW / 2 + lane, 0), W - 1);  // byte offsets in a row
    const bool has_b = lane + 64 < LW;
    const unsigned cb = 4u * (unsigned)min(max(x0 - PW / 2 + lane + 64, 0), W - 1);
    const unsigned xd = 4u * (unsigned)min(gx, W - 1);  // data column (lanes past the image edge keep nothing)
    // lit 8-row bands (bit b: rows 8b .. 8b + 7) over the columns the strip's
    // row pass reads; lane b ORs the flags of band b's squares
    unsigned long long lit = ~0ull;
    int band = 1;  // lane b: band b's flag (read per row with one readlane)
    if (trace && d.cull) {
        const int c0 = max(x0 - PW / 2, 0), c1 = min(x0 + 63 + (PW - 1 - PW / 2), W - 1);
        int any = 0;
        if (lane * 8 < H) {
            const int ntx = (W + LENSED_RENDER_SIDE - 1) / LENSED_RENDER_SIDE;
            // squares of a render item: x offset (w & 1) * 8, y offset (w >> 1) * 8
            const int* f = trace + 4 * ((long long)d.item_r + (long long)cand * d.ntile_r + (lane >> 1) * ntx);
            const int sub = (lane & 1) * 2;
            for (int sx = c0 >> 3; sx <= (c1 >> 3); ++sx) any |= f[4 * (sx >> 1) + (sx & 1) + sub];
        }
        lit = __ballot(any != 0);
        band = any != 0;
    }
    double* part = partials + d.partial_off + (long long)cand * d.npart + strip;
    float rfs[PW], cfs[PH];
#pragma unroll
    for (int i = 0; i < PW; ++i) rfs[i] = rf[i];
#pragma unroll
    for (int j = 0; j < PH; ++j) cfs[j] = cf[j];
    const bool last_lit = (lit >> ((H - 1) >> 3)) & 1ull;  // the rows below the image repeat the last one
    // An output block (8 rows) whose window rows are all dark is the data's
    // own chi^2 partial, the same for every candidate: dark[] holds it,
    // computed once from the data (lensed_dark_partials, the same operations
    // as a dark block here), and the sweep runs only from the first block
    // with a lit window row to the last.  Block b's windows read input rows
    // 8b + A - (PH - 1) .. 8b + 7 + A (clamped), in bands b + DLO .. b + DHI.
    const int nblk = (H + 7) >> 3;
    const bool skip = !OUT && dark && trace && d.cull;
    int first = 0, stop = nblk;  // output blocks [first, stop) are swept
    if (skip) {
        constexpr int DLO = -((PH / 2 + 7) / 8), DHI = (7 + A) / 8;
        unsigned long long need = 0;
#pragma unroll
        for (int o = DLO; o <= DHI; ++o) need |= o < 0 ? lit << -o : lit >> o;
        need &= nblk == 64 ? ~0ull : (1ull << nblk) - 1ull;
        first = need ? __builtin_ctzll(need) : 0;
        stop = need ? 64 - __builtin_clzll(need) : 0;
    }
    // (wave-uniform, and known so: the sweep's row counters stay scalar)
    first = __builtin_amdgcn_readfirstlane(first);
    stop = __builtin_amdgcn_readfirstlane(stop);
    if (skip && lane < nblk && (lane < first || lane >= stop))
        part[(long long)lane * d.tiles_x] = dark[d.dark_off + (long long)lane * d.tiles_x + strip];
    const int y_first = 8 * first;                          // first output row swept
    const int yi0 = max(y_first + A - (PH - 1), 0);         // first input row its window reads
    const int yi_end = stop > first ? min(8 * stop, H) + A : yi0;  // input rows [yi0, yi_end)
    // Loads, issued four iterations ahead in iteration order so that the
    // counter waits keep four iterations in flight; every address is a fixed
    // base plus a wave-uniform row offset plus the lane's byte offset.  Model
    // row yf of the staged columns, or row 0 where yf is dark or past the
    // image (cached; the values are not used), so every iteration issues the
    // same loads; data and weight of output row yd, clamped to the image
    // (rows outside it are not used).  Row yf's lit flag (the last row's for
    // rows below the image) goes into a queue of the four rows in flight, so
    // that the row pass reads it there instead of testing the band mask again.
    const char* mb = (const char*)m;
    const char* db = (const char*)img;
    const char* wb = (const char*)wgt;
    const unsigned rowb = 4u * (unsigned)W;
    unsigned moff = (unsigned)yi0 * rowb;  // byte offset of row yf
    int yf = yi0, yd = yi0 - A;            // the next fetches' rows
    unsigned litq = 0;    // bit i: lit flag of the i-th oldest row in flight
    auto at = [](const char* base, unsigned off) { return *(const float*)(base + off); };
    auto fetch = [&](float& va, float& vb, float& dv, float& wv) {
        const bool inside = yf < H, lf = inside && __builtin_amdgcn_readlane(band, min(yf >> 3, 63)) != 0;
        const unsigned mo = lf ? moff : 0u;
        const unsigned doff = (unsigned)min(max(yd, 0), H - 1) * rowb;
        va = at(mb, mo + ca);
        vb = at(mb, mo + cb);
        dv = at(db, doff + xd);
        wv = at(wb, doff + xd);
        litq = (litq >> 1) | ((inside ? lf : last_lit) ? 8u : 0u);
        moff += rowb;
        ++yf;
        ++yd;
    };
    float win[PH];
    float last = 0.0f;  // the row pass of the last image row (the rows below it)
    float s = 0.0f, c = 0.0f;
    int left = 8;  // output rows still to come in the current 8-row block
    int age = PH;  // input rows since the last lit one (a window is lit while under PH)
    float a0, b0, d0, w0, a1, b1, d1, w1, a2, b2, d2, w2, a3, b3, d3, w3;
    fetch(a0, b0, d0, w0);
    fetch(a1, b1, d1, w1);
    fetch(a2, b2, d2, w2);
    fetch(a3, b3, d3, w3);
    for (int y0 = yi0; y0 < yi_end; y0 += PH) {
#pragma unroll
        for (int k = 0; k < PH; ++k) {
            const int yi = y0 + k;  // input row (wave-uniform)
            if (yi < yi_end) {
                const float va = a0, vb = b0, dat = d0, wt = w0;
                const bool lit_in = litq & 1u;  // row yi's flag, the oldest in flight
                a0 = a1, b0 = b1, d0 = d1, w0 = w1;
                a1 = a2, b1 = b2, d1 = d2, w1 = w2;
                a2 = a3, b2 = b3, d2 = d3, w2 = w3;
                fetch(a3, b3, d3, w3);
                const int y = yi - A;  // output row
                age = lit_in ? 0 : age + 1;
                float r = last;
                if (yi < H) {
                    r = 0.0f;
                    if (lit_in) {
                        // the tile kernel's chain: the first tap a product, then
                        // fused multiply-adds in tap order (written out, so that
                        // contraction cannot pair the terms differently)
                        // one wave writes and reads the row, and its LDS accesses
                        // complete in order; the compiler must keep that order
                        // (it sees one lane's distinct addresses, not the others')
                        asm volatile("" ::: "memory");
                        lrow[lane] = va;
                        if (has_b) lrow[lane + 64] = vb;
                        asm volatile("" ::: "memory");
                        r = rfs[0] * lrow[lane + PW - 1];
#pragma unroll
                        for (int i = 1; i < PW; ++i) r = __builtin_fmaf(rfs[i], lrow[lane + PW - 1 - i], r);
                    }
                    last = r;
                    if (k == 0 && yi == 0) {  // (only the first unrolled row can be row 0)
#pragma unroll
                        for (int q = 0; q < PH; ++q) win[q] = r;  // rows above the image are row 0
                    }
                }
                win[k] = r;
                if (y >= y_first) {
                    // the window (input rows yi - PH + 1 .. yi, clamped) holds a
                    // lit row unless the last lit one is PH or more rows back
                    float acc = 0.0f;
                    if (age < PH) {
#pragma unroll
                        for (int j = 0; j < PH; ++j) acc = __builtin_fmaf(cfs[j], win[(k - j + PH) % PH], acc);
                    }
                    const float chi = lensed_chi_term(acc, dat, wt, gx < W);
                    if (OUT && gx < W) {
                        const long long kk = d.model_off + (long long)cand * d.npix + (long long)y * W + gx;
                        if (conv_out) conv_out[kk] = acc;
                        if (chi2_out) chi2_out[kk] = chi;
                    }
                    lensed_two_sum_add(s, c, chi);
                    if (--left == 0) {
                        const double blk = lensed_wave_csum(s, c);
                        if (lane == 0) part[(long long)(y >> 3) * d.tiles_x] = blk;
                        s = c = 0.0f;
                        left = 8;
                    }
                }
            }
        }
    }
    if (left < 8) {
        // a last block short of 8 rows: the rows past the last one count as
        // 0, as in the tile kernel's last block
        for (int z = 0; z < left; ++z) lensed_two_sum_add(s, c, 0.0f);
        const double blk = lensed_wave_csum(s, c);
        if (lane == 0) part[(long long)((H - 1) >> 3) * d.tiles_x] = blk;
    }
}

// One wave per item (a 64-column strip of one candidate), four per block.
// Blocks go to the eight XCDs in turn; XCD x takes the x-th contiguous share
// of the blocks, so the strips of a candidate run side by side in one L2,
// which holds the halo columns that neighbouring strips both read.  The
// dumper's planes have a kernel of their own (the stores would otherwise
// share the loglike kernel's registers and make it spill).
template <bool OUT>
__device__ inline void lensed_conv_sweep_body(const lensed_desc* __restrict__ sys, int nsys, int nitems,
                                              const float* __restrict__ model, const float* __restrict__ image,
                                              const float* __restrict__ weight, const float* __restrict__ psf,
                                              double* __restrict__ partials, float* __restrict__ conv_out,
                                              float* __restrict__ chi2_out, const int* __restrict__ trace,
                                              const double* __restrict__ dark)
{
    __shared__ float rows[4 * LENSED_SWEEP_ROW];
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    int blk = b;
    if (nb >= 8) {
        const int x = b & 7, q = nb >> 3, r = nb & 7;
        blk = x * q + min(x, r) + (b >> 3);
    }
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);  // wave-uniform, and known so
    const int item = blk * 4 + wave;
    if (item >= nitems) return;
    const int s = lensed_find_sys<1>(sys, nsys, item);
    const lensed_desc& d = sys[s];
    const int local = item - d.item_c;
    const int cand = local / d.tiles_x, strip = local - cand * d.tiles_x;
    float* lrow = rows + wave * LENSED_SWEEP_ROW;
    // (the host launches the sweep only for layouts whose PSFs are all
    // separable and at most LENSED_SWEEP_MAX_PSF on a side; the others are
    // not compiled into it)
#define LENSED_SWEEP_CASE(PW_, PH_, SEP_)                                                                 \
    if constexpr (SEP_ && PW_ <= LENSED_SWEEP_MAX_PSF && PH_ <= LENSED_SWEEP_MAX_PSF) {                 \
        if (d.psfw == PW_ && d.psfh == PH_) {                                                           \
            lensed_conv_sweep_item<PW_, PH_, OUT>(d, cand, strip, model, image, weight, psf, partials,  \
                                                  conv_out, chi2_out, trace, dark, lrow);               \
            return;                                                                                     \
        }                                                                                               \
    }
    LENSED_FOR_PSFS(LENSED_SWEEP_CASE)
#undef LENSED_SWEEP_CASE
}

extern "C" __global__ __launch_bounds__(256) void lensed_conv_chi2_sweep(
    const lensed_desc* __restrict__ sys, int nsys, int nitems,
    const float* __restrict__ model, const float* __restrict__ image,
    const float* __restrict__ weight, const float* __restrict__ psf,
    double* __restrict__ partials, float* __restrict__ conv_out, float* __restrict__ chi2_out,
    const int* __restrict__ trace, int order, const double* __restrict__ dark)
{
    lensed_conv_sweep_body<false>(sys, nsys, nitems, model, image, weight, psf, partials, nullptr, nullptr, trace,
                                  dark);
    (void)conv_out, (void)chi2_out, (void)order;
}

extern "C" __global__ __launch_bounds__(256) void lensed_conv_chi2_sweep_out(
    const lensed_desc* __restrict__ sys, int nsys, int nitems,
    const float* __restrict__ model, const float* __restrict__ image,
    const float* __restrict__ weight, const float* __restrict__ psf,
    double* __restrict__ partials, float* __restrict__ conv_out, float* __restrict__ chi2_out,
    const int* __restrict__ trace, int order, const double* __restrict__ dark)
{
    // (every output is stored: no block is taken from the dark partials)
    lensed_conv_sweep_body<true>(sys, nsys, nitems, model, image, weight, psf, partials, conv_out, chi2_out, trace,
                                 nullptr);
    (void)order, (void)dark;
}

// The chi^2 partial of every 8-row x 64-column block of one system's data
// for a model of 0, once per data set: the sweep takes it for the blocks a
// candidate leaves dark.  The operations of a dark block in the sweep: the
// term w (0 - d)^2 per pixel (lensed_chi_term), TwoSum down each column in
// row order, rows past the image's last as 0, the compensated wave sum.
// One wave per block; block b = row block * tiles_x + strip, as in a
// candidate's partials.
extern "C" __global__ __launch_bounds__(64) void lensed_dark_partials(
    const float* __restrict__ image, const float* __restrict__ weight, int W, int H, int tiles_x,
    double* __restrict__ dark)
{
    const int b = (int)blockIdx.x, rb = b / tiles_x, strip = b - rb * tiles_x;
    const int lane = (int)threadIdx.x & 63, gx = strip * 64 + lane;
    const int xd = min(gx, W - 1);
    const int y0 = rb * 8, y1 = min(y0 + 8, H);
    float s = 0.0f, c = 0.0f;
    for (int y = y0; y < y1; ++y) {
        const long long o = (long long)y * W + xd;
        lensed_two_sum_add(s, c, lensed_chi_term(0.0f, image[o], weight[o], gx < W));
    }
    for (int y = y1; y < y0 + 8; ++y) lensed_two_sum_add(s, c, 0.0f);
    const double v = lensed_wave_csum(s, c);
    if (lane == 0) dark[b] = v;
}

// ---------------------------------------------------------------- finalize
// One wave per candidate: FP64 sum of the row-segment partials in a fixed order,
// logL = -chi^2/2 (src/nested.c:115).
extern "C" __global__ __launch_bounds__(256) void lensed_finalize(
    const lensed_desc* __restrict__ sys, int nsys, int nitems,
    const double* __restrict__ partials, double* __restrict__ logl)
{
    int item = blockIdx.x * 4 + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (item >= nitems) return;
    int s = lensed_find_sys<3>(sys, nsys, item);
    const lensed_desc& d = sys[s];
    int cand = item - d.item_f;
    const double* p = partials + d.partial_off + (long long)cand * d.npart;
    double acc = 0.0;
    for (int t = lane; t < d.npart; t += 64) acc += p[t];
    acc = lensed_wave_sum(acc);
    if (lane == 0) logl[d.out_off + cand] = -0.5 * acc;
}
