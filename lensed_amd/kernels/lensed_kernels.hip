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

typedef unsigned int u32;

// system owning a work item: last descriptor whose first item <= item
template <int WHICH>
__device__ inline int lensed_find_sys(const lensed_desc* __restrict__ sys, int nsys, int item) {
    int lo = 0, hi = nsys - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        int first = WHICH == 0 ? sys[mid].item_r : WHICH == 1 ? sys[mid].item_c : WHICH == 2 ? sys[mid].item_s : sys[mid].item_f;
        if (first <= item) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// --------------------------------------------------------------- set_params
// One thread per candidate: derived object constants from the physical
// parameters, written as a data block of LENSED_DWORDS words.
extern "C" __global__ __launch_bounds__(256) void lensed_set_params(
    const lensed_desc* __restrict__ sys, int nsys, int nitems,
    const float* __restrict__ params, u32* __restrict__ blocks)
{
    int item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= nitems) return;
    int s = lensed_find_sys<2>(sys, nsys, item);
    const lensed_desc& d = sys[s];
    int cand = item - d.item_s;
    u32 blk[LENSED_DWORDS];
#pragma unroll
    for (int i = 0; i < LENSED_DWORDS; ++i) blk[i] = 0u;
    const float* p = params + d.param_off + (long long)cand * d.npar;
    switch (d.model) {
#define LENSED_SET_CASE(K) case K: lcl::set_params_##K(blk, p); break;
        LENSED_FOR_MODELS(LENSED_SET_CASE)
#undef LENSED_SET_CASE
        default: break;
    }
    u32* out = blocks + d.block_off + (long long)cand * d.dwords;
#pragma unroll
    for (int i = 0; i < LENSED_DWORDS; ++i)
        if (i < d.dwords) out[i] = blk[i];
}

// ------------------------------------------------------------------ render
// One thread per pixel; the 2-D quadrature loop runs in the thread with the
// node table in scalar registers and the candidate's object block held in
// SGPRs (wave-uniform), so the per-ray work is pure VALU.
template <int K>
__device__ inline void lensed_render_item(const lensed_desc& d, int cand, int tile,
                                          const u32* __restrict__ blocks, const float4* __restrict__ quad,
                                          float* __restrict__ model, float* __restrict__ errp)
{
    u32 blk[LENSED_DWORDS];
    const u32* gb = blocks + d.block_off + (long long)cand * d.dwords;
#pragma unroll
    for (int i = 0; i < LENSED_DWORDS; ++i) blk[i] = i < d.dwords ? __builtin_amdgcn_readfirstlane(gb[i]) : 0u;
    int k = tile * LENSED_RENDER_BS + (int)threadIdx.x;
    if (k >= d.npix) return;
    int py = k / d.width;
    int px = k - py * d.width;
    // pixel centre as kernel/lensed.cl:25, then one ray per node
    const float cx = d.pcs[0] + d.pcs[2] * (float)px;
    const float cy = d.pcs[1] + d.pcs[3] * (float)py;
    const float4* q = quad + d.quad_off;
    float f = 0.0f, e = 0.0f;
    for (int n = 0; n < d.nq; ++n) {
        const float4 node = q[n];
        float v = lcl::compute_model<K>(blk, lcl::float2(cx + node.x, cy + node.y));
        f += node.z * v;
        e += node.w * v;
    }
    long long o = d.model_off + (long long)cand * d.npix + k;
    model[o] = f;
    if (errp) errp[o] = e;
}

extern "C" __global__ __launch_bounds__(LENSED_RENDER_BS) void lensed_render(
    const lensed_desc* __restrict__ sys, int nsys, int nitems,
    const u32* __restrict__ blocks, const float4* __restrict__ quad,
    float* __restrict__ model, float* __restrict__ errp)
{
    for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
        int s = lensed_find_sys<0>(sys, nsys, item);
        const lensed_desc& d = sys[s];
        int local = item - d.item_r;
        int cand = local / d.ntile_r;
        int tile = local - cand * d.ntile_r;
        switch (d.model) {
#define LENSED_RENDER_CASE(K) case K: lensed_render_item<K>(d, cand, tile, blocks, quad, model, errp); break;
            LENSED_FOR_MODELS(LENSED_RENDER_CASE)
#undef LENSED_RENDER_CASE
            default: break;
        }
    }
}

// --------------------------------------------------------------- conv+chi2
// Direct convolution with the flipped PSF and clamp-to-edge borders, as
// kernel/lensed.cl:56-103: out(y, x) = sum_{j,i} psf[j][i] *
// in(clamp(y + PH-1-PH/2 - j), clamp(x + PW-1-PW/2 - i)).  A 64x32 output
// tile plus halo is staged in LDS; each thread produces 8 rows of one column.
// The chi^2 term w (m - d)^2 is formed per pixel in FP32 (kernel/lensed.cl:50-51),
// accumulated per thread with Kahan compensation and combined in FP64.
__device__ inline double lensed_wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

extern "C" __global__ __launch_bounds__(LENSED_CONV_BS) void lensed_conv_chi2(
    const lensed_desc* __restrict__ sys, int nsys, int nitems,
    const float* __restrict__ model, const float* __restrict__ image,
    const float* __restrict__ weight, const float* __restrict__ psf,
    double* __restrict__ partials, float* __restrict__ conv_out, float* __restrict__ chi2_out)
{
#pragma clang fp reassociate(off)
    extern __shared__ float tile[];
    __shared__ double red[LENSED_CONV_BS / 64];
    const int lx = threadIdx.x % LENSED_CONV_TX;
    const int ly = threadIdx.x / LENSED_CONV_TX;
    for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
        int s = lensed_find_sys<1>(sys, nsys, item);
        const lensed_desc& d = sys[s];
        int local = item - d.item_c;
        int cand = local / d.ntile_c;
        int t = local - cand * d.ntile_c;
        const int W = d.width, H = d.height;
        const int tx0 = (t % d.tiles_x) * LENSED_CONV_TX;
        const int ty0 = (t / d.tiles_x) * LENSED_CONV_TH;
        const float* m = model + d.model_off + (long long)cand * d.npix;
        const float* img = image + d.image_off;
        const float* wgt = weight + d.image_off;
        const int gx = tx0 + lx;
        float acc[LENSED_CONV_RY];
        if (d.psfw > 0) {
            const int PW = d.psfw, PH = d.psfh;
            const int cw = LENSED_CONV_TX + PW - 1, ch = LENSED_CONV_TH + PH - 1;
            const int ox = tx0 - PW / 2, oy = ty0 - PH / 2;
            for (int idx = threadIdx.x; idx < cw * ch; idx += LENSED_CONV_BS) {
                int r = idx / cw, c = idx - r * cw;
                int yy = min(max(oy + r, 0), H - 1);
                int xx = min(max(ox + c, 0), W - 1);
                tile[idx] = m[yy * W + xx];
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < LENSED_CONV_RY; ++r) acc[r] = 0.0f;
            const float* p = psf + d.psf_off;
            for (int j = 0; j < PH; ++j) {
                const float* row = tile + (ly * LENSED_CONV_RY + PH - 1 - j) * cw + lx + PW - 1;
                for (int i = 0; i < PW; ++i) {
                    const float w = p[j * PW + i];
#pragma unroll
                    for (int r = 0; r < LENSED_CONV_RY; ++r) acc[r] += w * row[r * cw - i];
                }
            }
            __syncthreads();
        } else {
#pragma unroll
            for (int r = 0; r < LENSED_CONV_RY; ++r) {
                int gy = ty0 + ly * LENSED_CONV_RY + r;
                acc[r] = (gx < W && gy < H) ? m[gy * W + gx] : 0.0f;
            }
        }
        float sum = 0.0f, comp = 0.0f;
#pragma unroll
        for (int r = 0; r < LENSED_CONV_RY; ++r) {
            int gy = ty0 + ly * LENSED_CONV_RY + r;
            if (gx < W && gy < H) {
                int kk = gy * W + gx;
                float dm = acc[r] - img[kk];
                float chi = wgt[kk] * dm * dm;
                float yk = chi - comp;
                float tk = sum + yk;
                comp = (tk - sum) - yk;
                sum = tk;
                if (conv_out) conv_out[d.model_off + (long long)cand * d.npix + kk] = acc[r];
                if (chi2_out) chi2_out[d.model_off + (long long)cand * d.npix + kk] = chi;
            }
        }
        double v = lensed_wave_sum((double)sum - (double)comp);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            double tot = 0.0;
#pragma unroll
            for (int w = 0; w < LENSED_CONV_BS / 64; ++w) tot += red[w];
            partials[d.partial_off + (long long)cand * d.ntile_c + t] = tot;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- finalize
// One wave per candidate: FP64 sum of the tile partials in a fixed order,
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
    const double* p = partials + d.partial_off + (long long)cand * d.ntile_c;
    double acc = 0.0;
    for (int t = lane; t < d.ntile_c; t += 64) acc += p[t];
    acc = lensed_wave_sum(acc);
    if (lane == 0) logl[d.out_off + cand] = -0.5 * acc;
}
