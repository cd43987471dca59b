// EPL emulator kernels: training data from the exact series, FP32 training
// (forward, backward, Adam), BF16-MFMA inference of the angular-factor table
// nodes with FP32 accumulation, the held-out check that certifies the
// network per (q, t) cell once per weights, and the per-candidate choice
// between the network and the exact series.  Network and weight layout:
// emulator.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/lensed.h"
#include "../kernels/desc.h"
#include "../kernels/epl_series.h"
#include "emulator.h"

namespace lsd {
void set_error(const std::string& msg);
}

namespace lsd {
namespace emu {

#define EMU_HIP(x)                                                             \
    do {                                                                       \
        hipError_t e_ = (x);                                                   \
        if (e_ != hipSuccess) {                                                \
            set_error(std::string(#x) + ": " + hipGetErrorString(e_));         \
            return -1;                                                         \
        }                                                                      \
    } while (0)

// ------------------------------------------------------------------ data

__device__ inline unsigned long long mix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

static inline unsigned long long mix64_host(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ inline double uniform01(unsigned long long seed, long long i, int dim) {
    return (double)(mix64(mix64(seed) ^ (unsigned long long)(i * 4 + dim)) >> 11) * (1.0 / 9007199254740992.0);
}

// seeded samples: q ~ U[0.3, 1], t ~ U[0.5, 1.5], psi ~ U[0, pi/2]; target
// Omega - (cos psi, sin psi) from the FP64 series
__global__ void k_make_data(long long n, unsigned long long seed, float* __restrict__ X, float* __restrict__ Y) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double q = 0.3 + 0.7 * uniform01(seed, i, 0);
    double t = 0.5 + uniform01(seed, i, 1);
    double psi = 1.57079632679489661923 * uniform01(seed, i, 2);
    double ox, oy, dx, dy;
    lensed_epl_omega(q, t, psi, ox, oy, dx, dy);
    double c = cos(psi), s = sin(psi);
    X[4 * i] = (float)((q - kQ0) / kQS);
    X[4 * i + 1] = (float)((t - kT0) / kTS);
    X[4 * i + 2] = (float)c;
    X[4 * i + 3] = (float)s;
    Y[2 * i] = (float)(ox - c);
    Y[2 * i + 1] = (float)(oy - s);
}

// minibatch b of an epoch: sample (start + i * stride) mod n, stride odd so
// every epoch visits each sample once (n a power of two)
__global__ void k_gather(int rows, long long n, long long start, long long stride, const float* __restrict__ X,
                         const float* __restrict__ Y, float* __restrict__ xb, float* __restrict__ yb) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    long long j = (start + (long long)i * stride) & (n - 1);
    for (int k = 0; k < kIn; ++k) xb[i * kIn + k] = X[j * kIn + k];
    for (int k = 0; k < kOut; ++k) yb[i * kOut + k] = Y[j * kOut + k];
}

// ------------------------------------------------------------- FP32 GEMM
// C[M][N] (+)= sum_k A(m, k) B(k, n), A and B addressed by strides, 64 x 64
// tiles, 4 x 4 outputs per thread, K staged in LDS in steps of 16.
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const float* __restrict__ A, long long sAm,
                                              long long sAk, const float* __restrict__ B, long long sBk,
                                              long long sBn, float* __restrict__ C, long long ldc, int accumulate) {
    __shared__ float As[16][64 + 1];
    __shared__ float Bs[16][64 + 1];
    const int tm = threadIdx.x / 16, tn = threadIdx.x % 16;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += 16) {
        for (int i = threadIdx.x; i < 64 * 16; i += 256) {
            int mm = i / 16, kk = i % 16;
            int gm = m0 + mm, gk = k0 + kk;
            As[kk][mm] = (gm < M && gk < K) ? A[gm * sAm + gk * sAk] : 0.0f;
        }
        for (int i = threadIdx.x; i < 16 * 64; i += 256) {
            int kk = i / 64, nn = i % 64;
            int gk = k0 + kk, gn = n0 + nn;
            Bs[kk][nn] = (gk < K && gn < N) ? B[gk * sBk + gn * sBn] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][tm * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tn * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int gm = m0 + tm * 4 + i, gn = n0 + tn * 4 + j;
            if (gm < M && gn < N) {
                float* c = C + gm * ldc + gn;
                *c = accumulate ? *c + acc[i][j] : acc[i][j];
            }
        }
}

__device__ inline float sigmoid(float z) { return 1.0f / (1.0f + expf(-z)); }

// Z += bias; H = silu(Z)
__global__ void k_bias_silu(int rows, int cols, float* __restrict__ Z, const float* __restrict__ b,
                            float* __restrict__ H) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * cols) return;
    float z = Z[i] + b[i % cols];
    Z[i] = z;
    H[i] = z * sigmoid(z);
}

// output layer: pred = Z + b; dZ = d(mean squared error)/dZ; loss partials
__global__ void k_out_grad(int rows, float* __restrict__ Z, const float* __restrict__ b, const float* __restrict__ Y,
                           float* __restrict__ dZ, double* __restrict__ loss) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (i < rows * kOut) {
        float p = Z[i] + b[i % kOut];
        Z[i] = p;
        float d = p - Y[i];
        dZ[i] = 2.0f * d / (float)(rows * kOut);
        l = (double)d * d;
    }
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
    if ((threadIdx.x & 63) == 0 && loss) atomicAdd(loss, l);
}

// dZ = dH * silu'(Z), in place on dH
__global__ void k_silu_back(int n, float* __restrict__ dH, const float* __restrict__ Z) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float z = Z[i], s = sigmoid(z);
    dH[i] *= s * (1.0f + z * (1.0f - s));
}

// db[c] = sum over rows of dZ[r][c]
__global__ void k_colsum(int rows, int cols, const float* __restrict__ dZ, float* __restrict__ db) {
    int c = blockIdx.x;
    float s = 0.0f;
    for (int r = threadIdx.x; r < rows; r += blockDim.x) s += dZ[r * cols + c];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    __shared__ float part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) db[c] = part[0] + part[1] + part[2] + part[3];
}

__global__ void k_adam(int n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1 - b1) * gi;
    float vi = b2 * v[i] + (1 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Workspace of one forward/backward pass over `rows` samples.
struct Net {
    int rows = 0;
    float *w = nullptr, *g = nullptr;
    float *xb = nullptr, *yb = nullptr;
    float *z1 = nullptr, *h1 = nullptr, *z2 = nullptr, *h2 = nullptr, *z3 = nullptr, *h3 = nullptr, *z4 = nullptr;
    float *d4 = nullptr, *dh = nullptr, *dh2 = nullptr;
    double* loss = nullptr;
    int alloc(int r) {
        rows = r;
        size_t H = (size_t)r * kHidden;
        EMU_HIP(hipMalloc(&w, kParams * 4));
        EMU_HIP(hipMalloc(&g, kParams * 4));
        EMU_HIP(hipMalloc(&xb, (size_t)r * kIn * 4));
        EMU_HIP(hipMalloc(&yb, (size_t)r * kOut * 4));
        for (float** p : {&z1, &h1, &z2, &h2, &z3, &h3, &dh, &dh2}) EMU_HIP(hipMalloc(p, H * 4));
        EMU_HIP(hipMalloc(&z4, (size_t)r * kOut * 4));
        EMU_HIP(hipMalloc(&d4, (size_t)r * kOut * 4));
        EMU_HIP(hipMalloc(&loss, 8));
        return 0;
    }
    void release() {
        for (void* p : {(void*)w, (void*)g, (void*)xb, (void*)yb, (void*)z1, (void*)h1, (void*)z2, (void*)h2,
                        (void*)z3, (void*)h3, (void*)z4, (void*)d4, (void*)dh, (void*)dh2, (void*)loss})
            if (p) (void)hipFree(p);
    }
};

static void gemm(hipStream_t st, int M, int N, int K, const float* A, long long sAm, long long sAk, const float* B,
                 long long sBk, long long sBn, float* C, long long ldc, int acc) {
    dim3 grid((N + 63) / 64, (M + 63) / 64);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, st, M, N, K, A, sAm, sAk, B, sBk, sBn, C, ldc, acc);
}

static inline dim3 blocks(long long n) { return dim3((unsigned)((n + 255) / 256)); }

// forward pass over net.xb; z4 = prediction without bias until k_out_grad
static void forward(hipStream_t st, Net& net) {
    const int R = net.rows, Hd = kHidden;
    const float* w = net.w;
    // layer l: Z[R][out] = X[R][in] W^T  (W row-major [out][in])
    gemm(st, R, Hd, kIn, net.xb, kIn, 1, w + kW1, 1, kIn, net.z1, Hd, 0);
    hipLaunchKernelGGL(k_bias_silu, blocks((long long)R * Hd), dim3(256), 0, st, R, Hd, net.z1, w + kB1, net.h1);
    gemm(st, R, Hd, Hd, net.h1, Hd, 1, w + kW2, 1, Hd, net.z2, Hd, 0);
    hipLaunchKernelGGL(k_bias_silu, blocks((long long)R * Hd), dim3(256), 0, st, R, Hd, net.z2, w + kB2, net.h2);
    gemm(st, R, Hd, Hd, net.h2, Hd, 1, w + kW3, 1, Hd, net.z3, Hd, 0);
    hipLaunchKernelGGL(k_bias_silu, blocks((long long)R * Hd), dim3(256), 0, st, R, Hd, net.z3, w + kB3, net.h3);
    gemm(st, R, kOut, Hd, net.h3, Hd, 1, w + kW4, 1, Hd, net.z4, kOut, 0);
}

// backward pass after k_out_grad filled d4; gradients into net.g
static void backward(hipStream_t st, Net& net) {
    const int R = net.rows, Hd = kHidden;
    const float* w = net.w;
    float* g = net.g;
    // layer 4: dW4 = d4^T h3, db4, dh3 = d4 W4
    gemm(st, kOut, Hd, R, net.d4, 1, kOut, net.h3, Hd, 1, g + kW4, Hd, 0);
    hipLaunchKernelGGL(k_colsum, dim3(kOut), dim3(256), 0, st, R, kOut, net.d4, g + kB4);
    gemm(st, R, Hd, kOut, net.d4, kOut, 1, w + kW4, Hd, 1, net.dh, Hd, 0);
    hipLaunchKernelGGL(k_silu_back, blocks((long long)R * Hd), dim3(256), 0, st, R * Hd, net.dh, net.z3);
    // layer 3
    gemm(st, Hd, Hd, R, net.dh, 1, Hd, net.h2, Hd, 1, g + kW3, Hd, 0);
    hipLaunchKernelGGL(k_colsum, dim3(Hd), dim3(256), 0, st, R, Hd, net.dh, g + kB3);
    gemm(st, R, Hd, Hd, net.dh, Hd, 1, w + kW3, Hd, 1, net.dh2, Hd, 0);
    hipLaunchKernelGGL(k_silu_back, blocks((long long)R * Hd), dim3(256), 0, st, R * Hd, net.dh2, net.z2);
    // layer 2
    gemm(st, Hd, Hd, R, net.dh2, 1, Hd, net.h1, Hd, 1, g + kW2, Hd, 0);
    hipLaunchKernelGGL(k_colsum, dim3(Hd), dim3(256), 0, st, R, Hd, net.dh2, g + kB2);
    gemm(st, R, Hd, Hd, net.dh2, Hd, 1, w + kW2, Hd, 1, net.dh, Hd, 0);
    hipLaunchKernelGGL(k_silu_back, blocks((long long)R * Hd), dim3(256), 0, st, R * Hd, net.dh, net.z1);
    // layer 1
    gemm(st, Hd, kIn, R, net.dh, 1, Hd, net.xb, kIn, 1, g + kW1, kIn, 0);
    hipLaunchKernelGGL(k_colsum, dim3(Hd), dim3(256), 0, st, R, Hd, net.dh, g + kB1);
}

// ------------------------------------------------------ BF16 MFMA inference
// gfx950's v_mfma_f32_32x32x16_bf16, wave64: A (32 x 16) lane l holds
// A[l % 32][8 (l / 32) + 0..7], B (16 x 32) lane l holds B[8 (l / 32) + 0..7][l % 32],
// D (32 x 32) lane l register v holds D[8 (v / 4) + 4 (l / 32) + v % 4][l % 32].
// A layer computes Z^T = W H^T for 32 samples per wave: M = 128 outputs in
// 4 tiles, K = 128 inputs in 8 steps of 16.  The K order within a step is
// permuted so that a layer's accumulators are the next layer's B operands
// as they stand: element i of K step S held by lane half h is input unit
// 16 S + 8 (i / 4) + 4 h + i % 4, which is where output tile S / 2 left
// that unit (register 8 (S % 2) + i of the lane); the weights are packed in
// the same order, so the products are those of the plain order.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// bf16x8 operands per packed layer: 4 M-tiles x 8 K steps x 64 lanes
constexpr int kLayerOperands = 4 * 8 * 64;

__device__ inline float silu_d(float z, float& dsilu) {
    // the hardware reciprocal (about 1 ulp) rather than an IEEE division,
    // a ten-instruction sequence in this unit's precise-math build
    float s = __builtin_amdgcn_rcpf(1.0f + __expf(-z));
    dsilu = s * (1.0f + z * (1.0f - s));
    return z * s;
}

// Output rows of an MFMA 32x32 tile held by a lane: element v of M-tile mt
// is unit 32 mt + 8 (v / 4) + 4 (lane / 32) + v % 4 of sample lane % 32.  Its
// elements 0..7 and 8..15 are the lane's B operands of K steps 2 mt and
// 2 mt + 1 of the next layer, so a layer's output is packed into the next
// layer's operands one M-tile at a time and only the operands (never a
// whole layer of FP32 activations) stay live: the kernel then fits several
// waves per SIMD to hide the weight loads.
__device__ inline int mlp_unit(int mt, int v, int lane) { return 32 * mt + 8 * (v / 4) + 4 * (lane / 32) + (v % 4); }

// activations z (primal) and dz (tangent) of M-tile mt -> silu, packed as
// the bf16 B operands hb / tb of K steps 2 mt, 2 mt + 1 (round to nearest
// even, gfx950's two-value conversion); h and dh keep the FP32 values for a
// following FP32 layer
__device__ inline void pack_tile(const float (&z)[16], const float (&dz)[16], int mt, bf16x8 (&hb)[8],
                                 bf16x8 (&tb)[8], float (&h)[16], float (&dh)[16]) {
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        float ds;
        h[v] = silu_d(z[v], ds);
        dh[v] = ds * dz[v];
    }
    typedef float fl8 __attribute__((ext_vector_type(8)));
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        const fl8 a = {h[8 * half], h[8 * half + 1], h[8 * half + 2], h[8 * half + 3],
                       h[8 * half + 4], h[8 * half + 5], h[8 * half + 6], h[8 * half + 7]};
        const fl8 b = {dh[8 * half], dh[8 * half + 1], dh[8 * half + 2], dh[8 * half + 3],
                       dh[8 * half + 4], dh[8 * half + 5], dh[8 * half + 6], dh[8 * half + 7]};
        hb[2 * mt + half] = __builtin_convertvector(a, bf16x8);
        tb[2 * mt + half] = __builtin_convertvector(b, bf16x8);
    }
}

// M-tile mt of a hidden 128 x 128 layer in MFMA, primal and tangent (no bias
// on the tangent)
__device__ inline void mfma_tile(const bf16x8* __restrict__ wl, const float* __restrict__ bias, const bf16x8 (&hb)[8],
                                 const bf16x8 (&tb)[8], int mt, float (&z)[16], float (&dz)[16], int lane) {
    f32x16 acc, tac;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = tac[v] = 0.0f;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        bf16x8 a = wl[(mt * 8 + s) * 64 + lane];
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, hb[s], acc, 0, 0, 0);
        tac = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, tb[s], tac, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        z[v] = acc[v] + bias[mlp_unit(mt, v, lane)];
        dz[v] = tac[v];
    }
}

// the network on 32 samples per wave (lanes l and l + 32 share sample l % 32):
// output y and its derivative dy along the input direction dx
__device__ inline void mlp_wave(const float* __restrict__ ws, const bf16x8* __restrict__ w2, const bf16x8* __restrict__ w3,
                                const float (&x)[4], const float (&dx)[4], float (&y)[2], float (&dy)[2], int lane) {
    bf16x8 hb[8], tb[8], nb[8], nt[8];
    float z[16], dz[16], h[16], dh[16];
    // layer 1 in FP32 (K = 4), straight into layer 2's operands
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int o = mlp_unit(mt, v, lane);
            const float* wr = ws + kW1 + o * kIn;
            z[v] = ws[kB1 + o] + wr[0] * x[0] + wr[1] * x[1] + wr[2] * x[2] + wr[3] * x[3];
            dz[v] = wr[0] * dx[0] + wr[1] * dx[1] + wr[2] * dx[2] + wr[3] * dx[3];
        }
        pack_tile(z, dz, mt, hb, tb, h, dh);
    }
    // layer 2 in MFMA, into layer 3's operands
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        mfma_tile(w2, ws + kB2, hb, tb, mt, z, dz, lane);
        pack_tile(z, dz, mt, nb, nt, h, dh);
    }
    // layer 3 in MFMA; layer 4 in FP32 on its tiles as they come (each lane
    // holds 64 of the 128 inputs of its sample; summed over M-tiles, then
    // units, as one pass over the layer would)
    float acc[kOut], tac[kOut];
#pragma unroll
    for (int c = 0; c < kOut; ++c) acc[c] = tac[c] = 0.0f;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        mfma_tile(w3, ws + kB3, nb, nt, mt, z, dz, lane);
        pack_tile(z, dz, mt, hb, tb, h, dh);
#pragma unroll
        for (int c = 0; c < kOut; ++c)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const float wv = ws[kW4 + c * kHidden + mlp_unit(mt, v, lane)];
                acc[c] += wv * h[v];
                tac[c] += wv * dh[v];
            }
    }
#pragma unroll
    for (int c = 0; c < kOut; ++c) {
        acc[c] += __shfl_xor(acc[c], 32, 64);
        tac[c] += __shfl_xor(tac[c], 32, 64);
        y[c] = acc[c] + ws[kB4 + c];
        dy[c] = tac[c];
    }
}

// raw inference for tests: X[n][4], dX[n][4] -> Y[n][2], dY[n][2]
__global__ __launch_bounds__(256) void k_mlp_raw(int n, const float* __restrict__ wf, const bf16x8* __restrict__ wp,
                                                 const float* __restrict__ X, const float* __restrict__ dX,
                                                 float* __restrict__ Y, float* __restrict__ dY) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int i = wave * 32 + (lane & 31);
    const int ic = min(i, n - 1);
    float x[4], dx[4], y[2], dy[2];
    for (int k = 0; k < 4; ++k) {
        x[k] = X[ic * 4 + k];
        dx[k] = dX[ic * 4 + k];
    }
    mlp_wave(wf, wp, wp + kLayerOperands, x, dx, y, dy, lane);
    if (lane < 32 && i < n) {
        Y[2 * i] = y[0];
        Y[2 * i + 1] = y[1];
        dY[2 * i] = dy[0];
        dY[2 * i + 1] = dy[1];
    }
}

// table nodes: one sample per (candidate, node j = 0..NT); the network gives
// Omega and its derivative at psi(p_j), turned into the tabulated factor
// W = g^h Omega and its f-derivative as lensed_epl_node (epl_series.h) does
template <int WHICH>
__device__ inline int find_sys_t(const lensed_desc* __restrict__ sys, int nsys, int item) {
    int lo = 0, hi = nsys - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        int first = WHICH == 0 ? sys[mid].item_t : sys[mid].item_e;
        if (first <= item) lo = mid; else hi = mid - 1;
    }
    return lo;
}

constexpr int kNodes = LENSED_EPL_ROWS;  // per candidate

// table slot (candidate among those with a table) of node sample g
__device__ inline long long node_slot(const lensed_desc* __restrict__ sys, int nsys, int g, int* rem) {
    const int s = find_sys_t<1>(sys, nsys, g);
    const lensed_desc& d = sys[s];
    const int local = g - d.item_e, cand = local / kNodes;
    if (rem) *rem = local - cand * kNodes;
    return d.table_off / LENSED_EPL_TABLE_F4 + cand;
}

// two waves per SIMD (at most 256 VGPRs): one wave's weight loads hide
// behind the other's MFMA.  With `skip`, a wave whose samples all belong to
// candidates flagged for the exact series (a wave spans at most two
// candidates) does nothing.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_emu_nodes(const lensed_desc* __restrict__ sys, int nsys, int nitems,
                                                   const float* __restrict__ params, const float* __restrict__ wf,
                                                   const bf16x8* __restrict__ wp, float4* __restrict__ nodes,
                                                   const int* __restrict__ skip) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (wave * 32 >= nitems) return;  // wave-uniform
    if (skip && skip[node_slot(sys, nsys, wave * 32, nullptr)] &&
        skip[node_slot(sys, nsys, min(wave * 32 + 31, nitems - 1), nullptr)])
        return;  // wave-uniform: both candidates take the exact series
    const int g = wave * 32 + (lane & 31);
    const int gc = min(g, nitems - 1);
    const int s = find_sys_t<1>(sys, nsys, gc);
    const lensed_desc& d = sys[s];
    const int local = gc - d.item_e;
    const int cand = local / kNodes;
    const int rem = local - cand * kNodes;
    const float* p = params + d.param_off + (long long)cand * d.npar;
    const float q = p[d.epl_q], t = p[d.epl_t];
    const float pj = (float)rem / LENSED_EPL_NT;
    const float gj = pj * pj + (1.0f - pj) * (1.0f - pj), h = 0.5f * (1.0f - t);
    const float psi = atan2f(pj, 1.0f - pj);
    const float dpsi = 1.0f / gj / LENSED_EPL_NT;  // d psi / d f
    float sn, cs;
    sincosf(psi, &sn, &cs);
    float x[4] = {(q - kQ0) / kQS, (t - kT0) / kTS, cs, sn};
    float dx[4] = {0.0f, 0.0f, -sn * dpsi, cs * dpsi};
    float y[2], dy[2];
    mlp_wave(wf, wp, wp + kLayerOperands, x, dx, y, dy, lane);
    if (lane < 32 && g < nitems) {
        long long cb = d.table_off / LENSED_EPL_TABLE_F4 + cand;
        const float gh = powf(gj, h), lg = h * (4.0f * pj - 2.0f) / gj / LENSED_EPL_NT;  // d log(g^h) / df
        const float ox = cs + y[0], oy = sn + y[1];
        nodes[cb * kNodes + rem] =
            make_float4(gh * ox, gh * oy, gh * (lg * ox + dx[2] + dy[0]), gh * (lg * oy + dx[3] + dy[1]));
    }
}

// the (q, t) cell of the certification grid holding a candidate, or -1
// outside the network's training domain
__device__ __host__ inline int cell_of(float q, float t) {
    if (!(q >= kQ0 - kQS && q <= kQ0 + kQS && t >= kT0 - kTS && t <= kT0 + kTS)) return -1;
    int iq = (int)((q - (kQ0 - kQS)) / (2.0f * kQS) * kCellsQ), it = (int)((t - (kT0 - kTS)) / (2.0f * kTS) * kCellsT);
    iq = iq < 0 ? 0 : iq >= kCellsQ ? kCellsQ - 1 : iq;
    it = it < 0 ? 0 : it >= kCellsT ? kCellsT - 1 : it;
    return iq * kCellsT + it;
}

// per candidate with a table: 0 if its (q, t) lies in a certified cell (the
// network builds its table), 1 if not (the exact series does)
__global__ __launch_bounds__(256) void k_emu_select(const lensed_desc* __restrict__ sys, int nsys, int ncands,
                                                    const float* __restrict__ params, const int* __restrict__ cells,
                                                    int* __restrict__ flags) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= ncands) return;
    int s = 0;
    for (; s < nsys - 1; ++s)
        if (w < sys[s + 1].item_e / kNodes) break;
    const lensed_desc& d = sys[s];
    const int cand = w - d.item_e / kNodes;
    const float* p = params + d.param_off + (long long)cand * d.npar;
    const int c = cell_of(p[d.epl_q], p[d.epl_t]);
    flags[d.table_off / LENSED_EPL_TABLE_F4 + cand] = c >= 0 && cells[c] ? 0 : 1;
}

// nodes -> cubic Hermite coefficients per interval (the table layout of
// epl_table.h)
__global__ __launch_bounds__(256) void k_emu_assemble(const lensed_desc* __restrict__ sys, int nsys, int nitems,
                                                      const float4* __restrict__ nodes, float4* __restrict__ tables,
                                                      const int* __restrict__ skip) {
    int item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= nitems) return;
    int s = find_sys_t<0>(sys, nsys, item);
    const lensed_desc& d = sys[s];
    int local = item - d.item_t;
    int cand = local / LENSED_EPL_ROWS;
    int i = local - cand * LENSED_EPL_ROWS;
    long long cb = d.table_off / LENSED_EPL_TABLE_F4 + cand;
    if (skip && skip[cb]) return;  // the exact series builds this table
    float4 a = nodes[cb * kNodes + i];
    float4* out = tables + d.table_off + (long long)cand * LENSED_EPL_TABLE_F4;
    float4& ox = out[LENSED_EPL_AT(0, i)];
    float4& oy = out[LENSED_EPL_AT(1, i)];
    if (i == LENSED_EPL_NT) {  // the end point p = 1 as a constant row
        ox = float4(a.x, 0.0f, 0.0f, 0.0f);
        oy = float4(a.y, 0.0f, 0.0f, 0.0f);
        return;
    }
    float4 b = nodes[cb * kNodes + i + 1];
    ox = lensed_hermite(a.x, a.z, b.x, b.z);
    oy = lensed_hermite(a.y, a.w, b.y, b.w);
}

// held-out check, one wave per candidate: the table against the FP64 series
// at every interval midpoint (never a training or node point); flag the
// candidate for the exact fallback if the largest error exceeds eps
__global__ __launch_bounds__(256) void k_emu_check(const lensed_desc* __restrict__ sys, int nsys, int ncands,
                                                   const float* __restrict__ params, const float4* __restrict__ tables,
                                                   float eps, int* __restrict__ flags, float* __restrict__ errs) {
    int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (w >= ncands) return;
    // candidate w of the launch: walk the systems (few) to find its own
    int s = 0;
    for (; s < nsys - 1; ++s)
        if (w < sys[s + 1].item_e / kNodes) break;
    const lensed_desc& d = sys[s];
    int cand = w - d.item_e / kNodes;
    const float* p = params + d.param_off + (long long)cand * d.npar;
    double q = p[d.epl_q], t = p[d.epl_t];
    const float4* tab = tables + d.table_off + (long long)cand * LENSED_EPL_TABLE_F4;
    float err = 0.0f;
    for (int i = lane; i < LENSED_EPL_NT; i += 64) {
        double ox, oy, dx, dy;
        lensed_epl_node(q, t, (i + 0.5) / LENSED_EPL_NT, ox, oy, dx, dy);
        float4 cx = tab[LENSED_EPL_AT(0, i)], cy = tab[LENSED_EPL_AT(1, i)];
        float f = 0.5f;
        float tx = cx.x + f * (cx.y + f * (cx.z + f * cx.w));
        float ty = cy.x + f * (cy.y + f * (cy.z + f * cy.w));
        err = fmaxf(err, fmaxf(fabsf(tx - (float)ox), fabsf(ty - (float)oy)));
    }
    for (int o = 32; o > 0; o >>= 1) err = fmaxf(err, __shfl_xor(err, o, 64));
    if (lane == 0) {
        long long cb = d.table_off / LENSED_EPL_TABLE_F4 + cand;
        flags[cb] = err > eps ? 1 : 0;
        if (errs) errs[cb] = err;
    }
}

// host: bf16 copies of W2, W3 in MFMA operand order (the permuted K order
// of the x16 operands: K step S, lane half h, element i is input unit
// 16 S + 8 (i / 4) + 4 h + i % 4)
static std::vector<unsigned short> pack_bf16(const float* w) {
    std::vector<unsigned short> out(2 * (size_t)kLayerOperands * 8);
    auto bf = [](float f) {
        unsigned u;
        memcpy(&u, &f, 4);
        u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
        return (unsigned short)(u >> 16);
    };
    for (int layer = 0; layer < 2; ++layer) {
        const float* W = w + (layer == 0 ? kW2 : kW3);
        for (int mt = 0; mt < 4; ++mt)
            for (int S = 0; S < 8; ++S)
                for (int lane = 0; lane < 64; ++lane)
                    for (int i = 0; i < 8; ++i) {
                        const int o = 32 * mt + lane % 32, k = 16 * S + 8 * (i / 4) + 4 * (lane / 32) + i % 4;
                        out[(((size_t)layer * kLayerOperands + (mt * 8 + S) * 64 + lane) * 8) + i] =
                            bf(W[o * kHidden + k]);
                    }
    }
    return out;
}

}  // namespace emu

// ----------------------------------------------------------- host entry
int emu_upload(const float* weights, void** wf, void** wp) {
    std::vector<unsigned short> packed = emu::pack_bf16(weights);
    EMU_HIP(hipMalloc(wf, emu::kParams * 4));
    EMU_HIP(hipMalloc(wp, packed.size() * 2));
    EMU_HIP(hipMemcpy(*wf, weights, emu::kParams * 4, hipMemcpyHostToDevice));
    EMU_HIP(hipMemcpy(*wp, packed.data(), packed.size() * 2, hipMemcpyHostToDevice));
    return 0;
}

// Certification of the network per (q, t) cell, once per weights: tables
// built by the network at every corner and the centre of every cell are
// checked against the FP64 series at every interval midpoint (points the
// table's nodes never are); a cell is certified when all five of its points
// are within eps.  cells (kCellsQ x kCellsT, host) receives 1 for certified
// cells; returns their number, or -1.  *max_err: the largest error seen.
int emu_certify(const void* wf, const void* wp, float eps, int* cells, double* max_err) {
    using namespace emu;
    const int nq = kCellsQ + 1, nt = kCellsT + 1;
    std::vector<float> qt;
    for (int i = 0; i < nq; ++i)
        for (int j = 0; j < nt; ++j) {
            qt.push_back(kQ0 - kQS + 2.0f * kQS * i / kCellsQ);
            qt.push_back(kT0 - kTS + 2.0f * kTS * j / kCellsT);
        }
    for (int i = 0; i < kCellsQ; ++i)
        for (int j = 0; j < kCellsT; ++j) {
            qt.push_back(kQ0 - kQS + 2.0f * kQS * (i + 0.5f) / kCellsQ);
            qt.push_back(kT0 - kTS + 2.0f * kTS * (j + 0.5f) / kCellsT);
        }
    const int np = (int)qt.size() / 2;
    lensed_desc d;
    memset(&d, 0, sizeof d);
    d.ncand = np;
    d.npar = 2;
    d.epl_q = 0;
    d.epl_t = 1;
    d.table_mode = 1;
    lensed_desc* dd = nullptr;
    float *par = nullptr, *errs = nullptr;
    float4 *nodes = nullptr, *tables = nullptr;
    int* flags = nullptr;
    hipStream_t st = nullptr;
    int rc = -1;
    std::vector<float> e(np);
    if (hipMalloc(&dd, sizeof d) == hipSuccess && hipMalloc(&par, qt.size() * 4) == hipSuccess &&
        hipMalloc(&nodes, (size_t)np * kNodes * 16) == hipSuccess &&
        hipMalloc(&tables, (size_t)np * LENSED_EPL_TABLE_F4 * 16) == hipSuccess &&
        hipMalloc(&flags, (size_t)np * 4) == hipSuccess && hipMalloc(&errs, (size_t)np * 4) == hipSuccess &&
        hipMemcpy(dd, &d, sizeof d, hipMemcpyHostToDevice) == hipSuccess &&
        hipMemcpy(par, qt.data(), qt.size() * 4, hipMemcpyHostToDevice) == hipSuccess) {
        const int n_nodes = np * kNodes, waves = (n_nodes + 31) / 32;
        hipLaunchKernelGGL(k_emu_nodes, dim3((waves + 3) / 4), dim3(256), 0, st, dd, 1, n_nodes, par,
                           (const float*)wf, (const bf16x8*)wp, nodes, (const int*)nullptr);
        hipLaunchKernelGGL(k_emu_assemble, dim3((n_nodes + 255) / 256), dim3(256), 0, st, dd, 1, n_nodes,
                           (const float4*)nodes, tables, (const int*)nullptr);
        hipLaunchKernelGGL(k_emu_check, dim3((np + 3) / 4), dim3(256), 0, st, dd, 1, np, (const float*)par,
                           (const float4*)tables, eps, flags, errs);
        if (hipGetLastError() == hipSuccess &&
            hipMemcpy(e.data(), errs, (size_t)np * 4, hipMemcpyDeviceToHost) == hipSuccess)
            rc = 0;
    }
    for (void* p : {(void*)dd, (void*)par, (void*)nodes, (void*)tables, (void*)flags, (void*)errs})
        if (p) (void)hipFree(p);
    if (rc) {
        set_error("emulator certification failed");
        return -1;
    }
    int n = 0;
    double mx = 0;
    for (float x : e) mx = std::max(mx, (double)x);
    for (int i = 0; i < kCellsQ; ++i)
        for (int j = 0; j < kCellsT; ++j) {
            const float worst = std::max({e[i * nt + j], e[(i + 1) * nt + j], e[i * nt + j + 1],
                                          e[(i + 1) * nt + j + 1], e[nq * nt + i * kCellsT + j]});
            cells[i * kCellsT + j] = worst <= eps ? 1 : 0;
            n += cells[i * kCellsT + j];
        }
    if (max_err) *max_err = mx;
    return n;
}

// emulated tables for a launch: which candidates lie in certified cells,
// then network nodes -> Hermite coefficients for those; flags (1 = exact
// series) tell the exact-table kernel which tables it builds
int emu_launch_tables(hipStream_t st, const lensed_desc* d_desc, int nsys, int n_nodes, int n_intervals, int n_cands,
                      const float* params, const void* wf, const void* wp, void* nodes, void* tables, int* flags,
                      const int* cells) {
    using namespace emu;
    int waves = (n_nodes + 31) / 32;
    hipLaunchKernelGGL(k_emu_select, dim3((n_cands + 255) / 256), dim3(256), 0, st, d_desc, nsys, n_cands, params,
                       cells, flags);
    hipLaunchKernelGGL(k_emu_nodes, dim3((waves + 3) / 4), dim3(256), 0, st, d_desc, nsys, n_nodes, params,
                       (const float*)wf, (const bf16x8*)wp, (float4*)nodes, (const int*)flags);
    hipLaunchKernelGGL(k_emu_assemble, dim3((n_intervals + 255) / 256), dim3(256), 0, st, d_desc, nsys, n_intervals,
                       (const float4*)nodes, (float4*)tables, (const int*)flags);
    EMU_HIP(hipGetLastError());
    return 0;
}

}  // namespace lsd

using namespace lsd;
using namespace lsd::emu;

static int write_weights(const char* path, const std::vector<float>& w) {
    FILE* f = fopen(path, "wb");
    if (!f) {
        set_error(std::string("cannot write ") + path);
        return -1;
    }
    const char magic[8] = {'L', 'E', 'M', 'U', '0', '0', '0', '1'};
    unsigned n = (unsigned)w.size();
    fwrite(magic, 1, 8, f);
    fwrite(&n, 4, 1, f);
    fwrite(w.data(), 4, w.size(), f);
    fclose(f);
    return 0;
}

extern "C" {

int lensed_emulator_read(const char* path, float* weights) {
    FILE* f = fopen(path, "rb");
    if (!f) {
        set_error(std::string("cannot read emulator weights ") + (path ? path : "(null)"));
        return -1;
    }
    char magic[8];
    unsigned n = 0;
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "LEMU0001", 8) == 0 && fread(&n, 4, 1, f) == 1 &&
              n == (unsigned)kParams && fread(weights, 4, n, f) == n;
    fclose(f);
    if (!ok) set_error(std::string("not an emulator weights file: ") + path);
    return ok ? 0 : -1;
}

int lensed_emulator_train(const char* path, long long nsamples, int epochs, int batch, double lr,
                          unsigned long long seed, double* stats) {
    if (!path || nsamples <= 0 || (nsamples & (nsamples - 1)) || epochs <= 0 || batch <= 0 || batch > nsamples) {
        set_error("lensed_emulator_train: invalid arguments (nsamples must be a power of two)");
        return -1;
    }
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t st;
    EMU_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    float *X, *Y, *m, *v;
    EMU_HIP(hipMalloc(&X, (size_t)nsamples * kIn * 4));
    EMU_HIP(hipMalloc(&Y, (size_t)nsamples * kOut * 4));
    EMU_HIP(hipMalloc(&m, kParams * 4));
    EMU_HIP(hipMalloc(&v, kParams * 4));
    EMU_HIP(hipMemsetAsync(m, 0, kParams * 4, st));
    EMU_HIP(hipMemsetAsync(v, 0, kParams * 4, st));
    hipLaunchKernelGGL(k_make_data, blocks(nsamples), dim3(256), 0, st, nsamples, seed, X, Y);
    Net net;
    if (net.alloc(batch)) return -1;
    // He-style uniform initialisation, seeded
    std::vector<float> w(kParams, 0.0f);
    std::mt19937_64 rng(seed);
    auto init = [&](int off, int fan_in, int count) {
        std::uniform_real_distribution<float> u(-std::sqrt(3.0f / fan_in), std::sqrt(3.0f / fan_in));
        for (int i = 0; i < count; ++i) w[off + i] = u(rng);
    };
    init(kW1, kIn, kHidden * kIn);
    init(kW2, kHidden, kHidden * kHidden);
    init(kW3, kHidden, kHidden * kHidden);
    init(kW4, kHidden, kOut * kHidden);
    EMU_HIP(hipMemcpyAsync(net.w, w.data(), kParams * 4, hipMemcpyHostToDevice, st));
    const long long steps_per_epoch = nsamples / batch;
    const long long total = steps_per_epoch * epochs;
    long long step = 0;
    double last_loss = 0.0;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int e = 0; e < epochs; ++e) {
        long long stride = (long long)(mix64_host(seed + 1000 + e) | 1ull) & (nsamples - 1);
        stride |= 1;
        long long start = (long long)(mix64_host(seed + 2000 + e) & (unsigned long long)(nsamples - 1));
        EMU_HIP(hipMemsetAsync(net.loss, 0, 8, st));
        for (long long b = 0; b < steps_per_epoch; ++b, ++step) {
            hipLaunchKernelGGL(k_gather, blocks(batch), dim3(256), 0, st, batch, nsamples, start + b * batch * stride,
                               stride, X, Y, net.xb, net.yb);
            forward(st, net);
            hipLaunchKernelGGL(k_out_grad, blocks((long long)batch * kOut), dim3(256), 0, st, batch, net.z4,
                               net.w + kB4, net.yb, net.d4, net.loss);
            backward(st, net);
            double frac = (double)step / (double)total;
            float lrt = (float)(lr * (0.01 + 0.99 * 0.5 * (1 + std::cos(3.14159265358979 * frac))));
            float c1 = 1.0f - std::pow(b1, (float)(step + 1)), c2 = 1.0f - std::pow(b2, (float)(step + 1));
            hipLaunchKernelGGL(k_adam, blocks(kParams), dim3(256), 0, st, kParams, net.w, net.g, m, v, lrt, b1, b2,
                               eps, c1, c2);
        }
        double l = 0.0;
        EMU_HIP(hipMemcpyAsync(&l, net.loss, 8, hipMemcpyDeviceToHost, st));
        EMU_HIP(hipStreamSynchronize(st));
        last_loss = l / ((double)steps_per_epoch * batch * kOut);
    }
    EMU_HIP(hipMemcpyAsync(w.data(), net.w, kParams * 4, hipMemcpyDeviceToHost, st));
    EMU_HIP(hipStreamSynchronize(st));
    net.release();
    (void)hipFree(X);
    (void)hipFree(Y);
    (void)hipFree(m);
    (void)hipFree(v);
    (void)hipStreamDestroy(st);
    if (write_weights(path, w)) return -1;
    if (stats) {
        stats[0] = last_loss;
        stats[1] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return 0;
}

// FP32 forward + backward of the training kernels on n samples: mean
// squared error loss and its gradient (flat layout) for given weights
double lensed_emulator_loss_grad(const float* weights, const float* X, const float* Y, int n, float* grad) {
    hipStream_t st = nullptr;
    Net net;
    if (n <= 0 || net.alloc(n)) return NAN;
    double loss = 0.0;
    if (hipMemcpy(net.w, weights, kParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(net.xb, X, (size_t)n * kIn * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(net.yb, Y, (size_t)n * kOut * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(net.loss, 0, 8) != hipSuccess) {
        net.release();
        return NAN;
    }
    forward(st, net);
    hipLaunchKernelGGL(k_out_grad, blocks((long long)n * kOut), dim3(256), 0, st, n, net.z4, net.w + kB4, net.yb,
                       net.d4, net.loss);
    backward(st, net);
    if (hipMemcpy(grad, net.g, kParams * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&loss, net.loss, 8, hipMemcpyDeviceToHost) != hipSuccess)
        loss = NAN;
    net.release();
    return loss / ((double)n * kOut);
}

// the network at n inputs X[n][4] with tangent directions dX[n][4]:
// bf16 = 0 runs the FP32 training forward pass (dY not computed, left 0),
// bf16 = 1 the BF16-MFMA inference kernel (Y and dY)
int lensed_emulator_eval(const float* weights, const float* X, const float* dX, int n, int bf16, float* Y, float* dY) {
    if (n <= 0) return 0;
    if (!bf16) {
        Net net;
        if (net.alloc(n)) return -1;
        std::vector<float> z(n * kOut), b4(weights + kB4, weights + kB4 + kOut);
        EMU_HIP(hipMemcpy(net.w, weights, kParams * 4, hipMemcpyHostToDevice));
        EMU_HIP(hipMemcpy(net.xb, X, (size_t)n * kIn * 4, hipMemcpyHostToDevice));
        forward(nullptr, net);
        EMU_HIP(hipMemcpy(z.data(), net.z4, (size_t)n * kOut * 4, hipMemcpyDeviceToHost));
        net.release();
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < kOut; ++c) {
                Y[i * kOut + c] = z[i * kOut + c] + b4[c];
                if (dY) dY[i * kOut + c] = 0.0f;
            }
        return 0;
    }
    void *wf, *wp;
    float *dx, *ddx, *dy, *ddy;
    if (emu_upload(weights, &wf, &wp)) return -1;
    EMU_HIP(hipMalloc(&dx, (size_t)n * 16));
    EMU_HIP(hipMalloc(&ddx, (size_t)n * 16));
    EMU_HIP(hipMalloc(&dy, (size_t)n * 8));
    EMU_HIP(hipMalloc(&ddy, (size_t)n * 8));
    EMU_HIP(hipMemcpy(dx, X, (size_t)n * 16, hipMemcpyHostToDevice));
    EMU_HIP(hipMemcpy(ddx, dX, (size_t)n * 16, hipMemcpyHostToDevice));
    int waves = (n + 31) / 32;
    hipLaunchKernelGGL(k_mlp_raw, dim3((waves + 3) / 4), dim3(256), 0, nullptr, n, (const float*)wf,
                       (const bf16x8*)wp, dx, ddx, dy, ddy);
    EMU_HIP(hipMemcpy(Y, dy, (size_t)n * 8, hipMemcpyDeviceToHost));
    if (dY) EMU_HIP(hipMemcpy(dY, ddy, (size_t)n * 8, hipMemcpyDeviceToHost));
    for (void* p : {wf, wp, (void*)dx, (void*)ddx, (void*)dy, (void*)ddy}) (void)hipFree(p);
    return 0;
}

}  // extern "C"
