// Spike-proposal network kernels for gfx950: BF16 activations and weights, FP32
// accumulation on MFMA (v_mfma_f32_32x32x16_bf16), FP32 master weights.
//
// A layer is a GEMM: out[co][w] = sum_{k' < CI*9} Wt[co][k'] * X[k'][w] with the
// im2col operand X[ci*9 + k][w] = in[ci][w + k - 4] gathered on the fly from an
// LDS tile of the input (zero outside the batch).  One wave computes a 32 x 32
// (channel x column) tile; lane l holds A[l % 32][8 (l / 32) + e] and
// B[8 (l / 32) + e][l % 32], e = 0..7, and D element r of lane l is
// (row (r % 4) + 8 (r / 4) + 4 (l / 32), column l % 32).
// The data gradient of a layer is the same GEMM with the weights transposed and
// the taps reversed; the weight gradient is a GEMM over columns with per-slab
// partials summed in a fixed order (deterministic).
#include "lss_proposal.h"

#include <hip/hip_runtime.h>

namespace lss {

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int TW = PN_TILE + PN_KW - 1;      // input tile columns of a convolution workgroup

// ---------------------------------------------------------------------------
// input: group g, row ci = whitened channel 16 g - 4 + ci (zero off the probe)
// ---------------------------------------------------------------------------
struct PnInputArgs {
    const float* xw; int64_t xs; int xoff; int C; int W; int Wp; int G;
    __bf16* a0;
    const int16_t* const* active;
};
struct PnInputTable { int n; PnInputArgs a[PN_MAX_JOBS]; };

__global__ __launch_bounds__(256) void k_pn_input(PnInputTable tb) {
    const PnInputArgs& A = tb.a[blockIdx.z];
    if (A.active && !*A.active) return;
    const int w = blockIdx.x * 256 + threadIdx.x;
    const int row = blockIdx.y;                       // g * PN_CIN + ci
    if (w >= A.Wp || row >= A.G * PN_CIN) return;
    const float* xw = A.xw;
    const int64_t xs = A.xs;
    const int xoff = A.xoff, C = A.C, W = A.W, Wp = A.Wp;
    __bf16* a0 = A.a0;
    const int g = row / PN_CIN, ci = row - g * PN_CIN;
    const int ch = g * PN_GROUP - PN_HALO + ci;
    float v = 0.f;
    if (w < W && ch >= 0 && ch < C) v = xw[(int64_t)ch * xs + xoff + w];
    a0[(int64_t)row * Wp + w] = (__bf16)v;
}

// ---------------------------------------------------------------------------
// convolution layer, 32 output channels.  MODE 0: + bias, ReLU (forward).
// MODE 1: multiplied by (mask > 0) (data gradient through the previous ReLU).
// ---------------------------------------------------------------------------
struct PnConvArgs {
    const __bf16* in; const __bf16* wt; const float* bias; const __bf16* mask; __bf16* out;
    int W; int Wp; int G;
    const int16_t* const* active;
};
struct PnConvTable { int n; PnConvArgs a[PN_MAX_JOBS]; };

template <int CI, int MODE>
__global__ __launch_bounds__(256) void k_pn_conv(PnConvTable tb) {
    const PnConvArgs& A = tb.a[blockIdx.z];
    if (A.active && !*A.active) return;
    if ((int)blockIdx.y >= A.G || (int)blockIdx.x * PN_TILE >= A.W) return;
    const __bf16* __restrict__ in = A.in;
    const __bf16* __restrict__ wt = A.wt;
    const float* __restrict__ bias = A.bias;
    const __bf16* __restrict__ mask = A.mask;
    __bf16* __restrict__ out = A.out;
    const int W = A.W, Wp = A.Wp;
    constexpr int KR = CI * PN_KW;
    constexpr int KP = (KR + 15) / 16 * 16;
    constexpr int XR = TW + 8;
    __shared__ __attribute__((aligned(16))) __bf16 xs[CI * XR];
    __shared__ __attribute__((aligned(16))) __bf16 ws[PN_CH][KP + 8];
    const int g = blockIdx.y, w0 = blockIdx.x * PN_TILE;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    // the input rows from column w0 - 8 (a multiple of 8, as Wp is), 16 bytes per
    // load; columns outside the batch read as zero.  xs[ci][c] holds column w0 - 8 + c.
    const __bf16* src = in + (int64_t)g * CI * Wp;
    constexpr int N8 = XR / 8;
    static_assert(XR % 8 == 0 && XR >= TW + PN_PAD, "the staged window covers the taps");
    for (int i = tid; i < CI * N8; i += 256) {
        const int ci = i / N8, c8 = i - ci * N8;
        const int j = w0 - 8 + 8 * c8;
        uint4 u = make_uint4(0u, 0u, 0u, 0u);
        if (j >= 0 && j + 8 <= Wp) u = *reinterpret_cast<const uint4*>(src + (int64_t)ci * Wp + j);
        __bf16 v[8];
        __builtin_memcpy(v, &u, 16);
#pragma unroll
        for (int e = 0; e < 8; ++e)
            if (j + e >= W) v[e] = (__bf16)0.f;
        __builtin_memcpy(&xs[ci * XR + 8 * c8], v, 16);
    }
    for (int i = tid; i < PN_CH * (KP / 8); i += 256) {
        const int r = i / (KP / 8), c8 = i - r * (KP / 8);
        *reinterpret_cast<uint4*>(&ws[r][c8 * 8]) = *reinterpret_cast<const uint4*>(wt + r * KP + c8 * 8);
    }
    __syncthreads();
    const int h = lane >> 5, j = lane & 31;
    const int jc = wv * 32 + j;                  // this lane's output column within the tile
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
#pragma unroll
    for (int s = 0; s < KP / 16; ++s) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&ws[j][16 * s + 8 * h]);
        bf16x8 b;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            // reduction index of this element for either half of the wave (compile-time)
            const int k0 = 16 * s + e, k1 = 16 * s + 8 + e;
            const int o0 = (k0 / PN_KW) * XR + k0 % PN_KW, o1 = (k1 / PN_KW) * XR + k1 % PN_KW;
            const bool v0 = k0 < KR, v1 = k1 < KR;
            const bool valid = h ? v1 : v0;
            const int off = h ? (v1 ? o1 : 0) : (v0 ? o0 : 0);
            const __bf16 x = xs[off + jc + (8 - PN_PAD)];
            b[e] = valid ? x : (__bf16)0.f;
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    const int w = w0 + jc;
    if (w >= Wp) return;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int co = (r & 3) + 8 * (r >> 2) + 4 * h;
        const int64_t o = ((int64_t)g * PN_CH + co) * Wp + w;
        float v = acc[r];
        if (MODE == 0) {
            v = fmaxf(v + bias[co], 0.f);
        } else {
            v = (float)mask[o] > 0.f ? v : 0.f;
        }
        out[o] = w < W ? (__bf16)v : (__bf16)0.f;
    }
}

// ---------------------------------------------------------------------------
// head: logit z = b4 + sum_{ci,k} W4[ci][k] a3[ci][w + k - 4]; p = sigmoid(z)
// (HIP's device expf)
// ---------------------------------------------------------------------------
struct PnHeadArgs {
    const __bf16* a3; const float* params; float* z; float* p;
    int W; int Wp; int G;
    const int16_t* const* active;
};
struct PnHeadTable { int n; PnHeadArgs a[PN_MAX_JOBS]; };

// A workgroup takes 256 columns of one group: the 32 hidden rows of its window
// (columns w0 - 8 .. w0 + 263, zero outside 0..W-1) are staged in LDS as 16-byte
// loads, then each thread sums its column's
// taps in the same order as a direct loop (a tap outside the batch adds an exact
// zero, which changes no partial sum).
constexpr int PN_HEAD_COLS = 256;
constexpr int PN_HEAD_WIN = PN_HEAD_COLS + 16;      // from w0 - 8, a multiple of 8 columns

__global__ __launch_bounds__(256) void k_pn_head(PnHeadTable tb) {
    const PnHeadArgs& A = tb.a[blockIdx.z];
    if (A.active && !*A.active) return;
    if ((int)blockIdx.y >= A.G) return;
    const __bf16* __restrict__ a3 = A.a3;
    float* __restrict__ z = A.z;
    float* __restrict__ p = A.p;
    const int W = A.W, Wp = A.Wp;
    const int w0 = blockIdx.x * PN_HEAD_COLS, g = blockIdx.y;
    if (w0 >= Wp) return;                       // whole workgroup past this job's columns
    __shared__ float w4[PN_CH * PN_KW + 1];
    __shared__ __align__(16) float s_a[PN_CH][PN_HEAD_WIN];
    for (int i = threadIdx.x; i < PN_CH * PN_KW + 1; i += 256) w4[i] = A.params[PN_W4 + i];
    constexpr int N8 = PN_HEAD_WIN / 8;
    for (int i = threadIdx.x; i < PN_CH * N8; i += 256) {
        const int ci = i / N8, c8 = i - ci * N8;
        const int j = w0 - 8 + 8 * c8;          // a multiple of 8, as Wp is
        uint4 u = make_uint4(0u, 0u, 0u, 0u);
        if (j >= 0 && j + 8 <= Wp) u = *reinterpret_cast<const uint4*>(a3 + ((int64_t)g * PN_CH + ci) * Wp + j);
        __bf16 v[8];
        __builtin_memcpy(v, &u, 16);
#pragma unroll
        for (int e = 0; e < 8; ++e)
            if (j + e >= W) v[e] = (__bf16)0.f;     // columns past the batch read as zero
        float4* dst = reinterpret_cast<float4*>(&s_a[ci][8 * c8]);
        dst[0] = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
        dst[1] = make_float4((float)v[4], (float)v[5], (float)v[6], (float)v[7]);
    }
    __syncthreads();
    const int w = w0 + threadIdx.x;
    if (w >= Wp) return;
    float acc = w4[PN_CH * PN_KW];
    if (w < W) {
        for (int ci = 0; ci < PN_CH; ++ci) {
#pragma unroll
            for (int k = 0; k < PN_KW; ++k) acc += w4[ci * PN_KW + k] * s_a[ci][threadIdx.x + 4 + k];
        }
    }
    const int64_t o = (int64_t)g * Wp + w;
    z[o] = w < W ? acc : 0.f;
    p[o] = w < W ? 1.f / (1.f + expf(-acc)) : 0.f;
}

// ---------------------------------------------------------------------------
// weighted BCE with logits: dz = wgt (sigmoid(z) - y) / n, loss partial per block
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_pn_loss(const float* __restrict__ z, const uint8_t* __restrict__ y,
                                                 int G, int W, int Wp, float pos_weight, float* __restrict__ dz,
                                                 float* __restrict__ lpart) {
    __shared__ float red[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int n = G * W;
    float l = 0.f;
    if (i < G * Wp) {
        const int g = i / Wp, w = i - g * Wp;
        float d = 0.f;
        if (w < W) {
            const float zz = z[i];
            const float yy = y[(int64_t)g * Wp + w] ? 1.f : 0.f;
            const float sp_pos = fmaxf(zz, 0.f) + log1pf(expf(-fabsf(zz)));   // softplus(z)
            const float sp_neg = sp_pos - zz;                                  // softplus(-z)
            const float wgt = yy > 0.5f ? pos_weight : 1.f;
            l = pos_weight * yy * sp_neg + (1.f - yy) * sp_pos;
            d = wgt * (1.f / (1.f + expf(-zz)) - yy) / (float)n;
        }
        dz[i] = d;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = red[0] / (float)n;
}

// data gradient of the head through layer 3's ReLU
__global__ __launch_bounds__(256) void k_pn_head_back(const float* __restrict__ dz, const float* __restrict__ params,
                                                      const __bf16* __restrict__ a3, int W, int Wp,
                                                      __bf16* __restrict__ d3) {
    const int w = blockIdx.x * 256 + threadIdx.x;
    const int row = blockIdx.y;                       // g * 32 + ci
    if (w >= Wp) return;
    const int g = row / PN_CH, ci = row - g * PN_CH;
    float v = 0.f;
    const int64_t o = (int64_t)row * Wp + w;
    if (w < W && (float)a3[o] > 0.f) {
        for (int k = 0; k < PN_KW; ++k) {
            const int j = w - k + PN_PAD;
            if (j >= 0 && j < W) v += params[PN_W4 + ci * PN_KW + k] * dz[(int64_t)g * Wp + j];
        }
    }
    d3[o] = (__bf16)v;
}

// weight and bias gradient of the head: one workgroup per input channel (and one for the bias)
__global__ __launch_bounds__(256) void k_pn_head_wgrad(const float* __restrict__ dz, const __bf16* __restrict__ a3,
                                                       int G, int W, int Wp, float* __restrict__ grad) {
    __shared__ float red[PN_KW][256];
    const int ci = blockIdx.x;                        // PN_CH: the bias
    float acc[PN_KW];
#pragma unroll
    for (int k = 0; k < PN_KW; ++k) acc[k] = 0.f;
    for (int i = threadIdx.x; i < G * W; i += 256) {
        const int g = i / W, w = i - g * W;
        const float d = dz[(int64_t)g * Wp + w];
        if (ci == PN_CH) { acc[0] += d; continue; }
        const __bf16* row = a3 + ((int64_t)g * PN_CH + ci) * Wp;
#pragma unroll
        for (int k = 0; k < PN_KW; ++k) {
            const int j = w + k - PN_PAD;
            if (j >= 0 && j < W) acc[k] += d * (float)row[j];
        }
    }
#pragma unroll
    for (int k = 0; k < PN_KW; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int k = 0; k < PN_KW; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (ci == PN_CH) grad[PN_B4] = red[0][0];
        else for (int k = 0; k < PN_KW; ++k) grad[PN_W4 + ci * PN_KW + k] = red[k][0];
    }
}

// ---------------------------------------------------------------------------
// weight gradient of a convolution layer over one slab of PN_SLAB columns:
// part[co][ci*9 + k] = sum_w dy[co][w] x[ci][w + k - 4]; bpart[co] = sum_w dy[co][w]
// ---------------------------------------------------------------------------
template <int CI>
__global__ __launch_bounds__(256) void k_pn_wgrad(const __bf16* __restrict__ dy, const __bf16* __restrict__ x,
                                                  int W, int Wp, float* __restrict__ part,
                                                  float* __restrict__ bpart) {
    constexpr int KR = CI * PN_KW;
    constexpr int NCH = (KR + 31) / 32;
    constexpr int NP = NCH * 32;
    constexpr int YR = PN_SLAB + 8;
    constexpr int XR = PN_SLAB + PN_KW - 1 + 8;
    __shared__ __attribute__((aligned(16))) __bf16 ys[PN_CH * YR];
    __shared__ __attribute__((aligned(16))) __bf16 xs[CI * XR];
    __shared__ float bred[PN_CH][8];
    const int g = blockIdx.y, slab = blockIdx.x, w0 = slab * PN_SLAB;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int i = tid; i < PN_CH * PN_SLAB; i += 256) {
        const int co = i / PN_SLAB, c = i - co * PN_SLAB, w = w0 + c;
        ys[co * YR + c] = w < W ? dy[((int64_t)g * PN_CH + co) * Wp + w] : (__bf16)0.f;
    }
    for (int i = tid; i < CI * (PN_SLAB + PN_KW - 1); i += 256) {
        const int ci = i / (PN_SLAB + PN_KW - 1), c = i - ci * (PN_SLAB + PN_KW - 1), w = w0 - PN_PAD + c;
        xs[ci * XR + c] = (w >= 0 && w < W) ? x[((int64_t)g * CI + ci) * Wp + w] : (__bf16)0.f;
    }
    __syncthreads();
    const int pid = g * gridDim.x + slab;
    {   // bias partial: 8 threads per channel, fixed order
        const int co = tid >> 3, q = tid & 7;
        float s = 0.f;
        for (int c = q; c < PN_SLAB; c += 8) s += (float)ys[co * YR + c];
        bred[co][q] = s;
        __syncthreads();
        if (tid < PN_CH) {
            float t = 0.f;
            for (int k = 0; k < 8; ++k) t += bred[tid][k];
            bpart[(int64_t)pid * PN_CH + tid] = t;
        }
    }
    const int h = lane >> 5, j = lane & 31;
    for (int ch = wv; ch < NCH; ch += 4) {
        const int n = ch * 32 + j;                 // weight index ci*9 + k of this lane's column
        const bool nvalid = n < KR;
        const int ci = nvalid ? n / PN_KW : 0, tap = nvalid ? n % PN_KW : 0;
        const __bf16* xrow = xs + ci * XR + tap;
        f32x16 acc;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = 0.f;
#pragma unroll 4
        for (int s = 0; s < PN_SLAB / 16; ++s) {
            const int c0 = 16 * s + 8 * h;
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&ys[j * YR + c0]);   // A[co = j][w]
            bf16x8 b;
#pragma unroll
            for (int e = 0; e < 8; ++e) b[e] = nvalid ? xrow[c0 + e] : (__bf16)0.f;
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int co = (r & 3) + 8 * (r >> 2) + 4 * h;
            part[((int64_t)pid * PN_CH + co) * NP + n] = acc[r];
        }
    }
}

// sum of the weight-gradient partials in partial order (deterministic)
__global__ __launch_bounds__(256) void k_pn_reduce(const float* __restrict__ part, const float* __restrict__ bpart,
                                                   int nparts, int KR, int NP, float* __restrict__ gw,
                                                   float* __restrict__ gb) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < PN_CH * KR) {
        const int co = i / KR, n = i - co * KR;
        float s = 0.f;
        for (int q = 0; q < nparts; ++q) s += part[((int64_t)q * PN_CH + co) * NP + n];
        gw[i] = s;
    } else if (i < PN_CH * KR + PN_CH) {
        const int co = i - PN_CH * KR;
        float s = 0.f;
        for (int q = 0; q < nparts; ++q) s += bpart[(int64_t)q * PN_CH + co];
        gb[co] = s;
    }
}

// ---------------------------------------------------------------------------
// Adam on the FP32 master weights
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_pn_adam(float* __restrict__ w, const float* __restrict__ g,
                                                 float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                                 float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= PN_NPARAMS) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 operand copies of the weights (forward and data-gradient layouts)
__global__ __launch_bounds__(256) void k_pn_pack(PnWeights wt) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const float* P = wt.params;
    // forward: [co][KP] with k' = ci*9 + k; zero past CI*9
    if (i < PN_CH * PN_KP1) {
        const int co = i / PN_KP1, kk = i - co * PN_KP1;
        wt.wf1[i] = (__bf16)(kk < PN_CIN * PN_KW ? P[PN_W1 + co * PN_CIN * PN_KW + kk] : 0.f);
    }
    // inference, tap-major: column k * CI + ci holds W[co][ci][k]; zero past 9 * CI
    if (i < PN_CH * PN_KT1) {
        const int co = i / PN_KT1, kk = i - co * PN_KT1;
        const int k = kk / PN_CIN, ci = kk - k * PN_CIN;
        wt.wt1[i] = (__bf16)(k < PN_KW ? P[PN_W1 + (co * PN_CIN + ci) * PN_KW + k] : 0.f);
    }
    if (i < PN_CH * PN_KP) {
        const int co = i / PN_KP, kk = i - co * PN_KP;
        const int k = kk / PN_CH, ci = kk - k * PN_CH;
        wt.wt2[i] = (__bf16)P[PN_W2 + (co * PN_CH + ci) * PN_KW + k];
        wt.wt3[i] = (__bf16)P[PN_W3 + (co * PN_CH + ci) * PN_KW + k];
    }
    if (i < PN_CH * PN_KP) {
        const int r = i / PN_KP, kk = i - r * PN_KP;
        wt.wf2[i] = (__bf16)P[PN_W2 + r * PN_CH * PN_KW + kk];
        wt.wf3[i] = (__bf16)P[PN_W3 + r * PN_CH * PN_KW + kk];
        // data gradient: row r = ci, k' = co*9 + k holds W[co][ci][8 - k]
        const int co = kk / PN_KW, k = kk - co * PN_KW;
        const int src = (co * PN_CH + r) * PN_KW + (PN_KW - 1 - k);
        wt.wb2[i] = (__bf16)P[PN_W2 + src];
        wt.wb3[i] = (__bf16)P[PN_W3 + src];
    }
}

// ---------------------------------------------------------------------------
// Fused inference: one workgroup (8 waves) takes one group and PN_FT output
// columns and runs the input gather, the three convolutions and the head with
// every activation in LDS.  Each layer is computed over 256 columns (a 32-column
// MFMA tile per wave), 4 columns further left than the layer before and 8
// columns narrower in use, so layer l covers the taps layer l + 1 reads:
// layer l's column jc is batch column w0 - 16 + 4 l + jc, and its tap k of
// column jc reads column jc + k of the layer below.  Activations are stored
// [column][channel] with the reduction in tap-major order (k' = k * CI + ci), so
// the eight consecutive k' of an MFMA operand are eight channels of one column,
// one 16-byte LDS read.  Activations of columns outside the batch are zero (the
// 'same' zero padding of every layer), as in the per-layer kernels; the sums
// run in another order than there (tap-major), so results agree to FP32
// rounding and the rare BF16 storage flip it causes.  The next layer's weights
// are loaded into registers while the current layer runs.
// ---------------------------------------------------------------------------
constexpr int PN_FT = 224;                 // output columns per fused workgroup
constexpr int PN_FX = 8 * 32 + PN_KW - 1;  // staged columns of each layer's input (264)
constexpr int PN_FCP = PN_CH + 8;          // LDS pitch (channels) of a staged column
constexpr int PN_FWP = PN_KP + 8;          // LDS pitch of a weight row (the widest layer)

struct PnInferArgs {
    const float* xw; int64_t xs; int xoff; int C; int W; int Wp; int G;
    const __bf16* wt1; const __bf16* wt2; const __bf16* wt3; const float* params;
    float* z; float* p;
    const int16_t* const* active;
};
struct PnInferTable { int n; PnInferArgs a[PN_MAX_JOBS]; };

// one layer: out[jc][co] = relu(bias[co] + sum_{k, ci} w[co][k * CI + ci] in[jc + k][ci]),
// zero for batch columns base + jc outside [0, W)
template <int CI>
__device__ __forceinline__ void pn_fused_layer(const __bf16* in, __bf16* out, const __bf16* ws, const float* bias,
                                               int base, int W, int lane, int wv) {
    constexpr int KR = CI * PN_KW;
    constexpr int KP = (KR + 15) / 16 * 16;
    const int h = lane >> 5, j = lane & 31, jc = wv * 32 + j;
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
#pragma unroll
    for (int s = 0; s < KP / 16; ++s) {
        const int kq = 16 * s + 8 * h;              // first reduction index of this lane's eight
        const int k = kq / CI, ci0 = kq - k * CI;   // one tap, eight channels (CI is a multiple of 8)
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&ws[j * PN_FWP + kq]);
        bf16x8 b;
        if (kq < KR) {
            b = *reinterpret_cast<const bf16x8*>(&in[(jc + k) * PN_FCP + ci0]);
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) b[e] = (__bf16)0.f;
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    const int w = base + jc;
    const bool ok = w >= 0 && w < W;
#pragma unroll
    for (int g4 = 0; g4 < 4; ++g4) {
        __bf16 v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int co = 8 * g4 + 4 * h + e;           // D element 4 g4 + e of this lane
            v[e] = ok ? (__bf16)fmaxf(acc[4 * g4 + e] + bias[co], 0.f) : (__bf16)0.f;
        }
        uint2 u;
        __builtin_memcpy(&u, v, 8);
        *reinterpret_cast<uint2*>(&out[jc * PN_FCP + 8 * g4 + 4 * h]) = u;
    }
}

__global__ __launch_bounds__(512) void k_pn_fused(PnInferTable tb) {
    // grid (stream, column tile, group slot): consecutive workgroups are consecutive
    // streams, so with n | 8 streams stream j runs on the XCDs x = j mod n, where an
    // interleaved whitening launch left its whitened batch (in that XCD's L2).  A
    // workgroup runs the groups g = slot, slot + gridDim.z, ... in turn; the next
    // group's input rows and layer 1's weights are loaded into registers while the
    // current group's last layer runs.
    const PnInferArgs& A = tb.a[blockIdx.x];
    const int w0 = blockIdx.y * PN_FT, G = A.G, GZ = gridDim.z;
    int g = blockIdx.z;
    if (g >= G || w0 >= A.W) return;
    const bool idle = A.active && !*A.active;
    __shared__ __attribute__((aligned(16))) __bf16 bx[PN_FX * PN_FCP];
    __shared__ __attribute__((aligned(16))) __bf16 by[PN_FX * PN_FCP];
    __shared__ __attribute__((aligned(16))) __bf16 ws[PN_CH * PN_FWP];
    __shared__ float s_bias[3 * PN_CH];
    __shared__ __attribute__((aligned(16))) float s_w4[PN_KW * PN_CH + 1];   // [k][ci], then b4
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int W = A.W, C = A.C;
    static_assert(PN_CIN % 8 == 0 && PN_CH % 8 == 0, "channels come in 16-byte words of eight");
    constexpr int NIN = PN_FX * (PN_CIN / 8);            // items of eight input rows
    constexpr int NW1 = PN_CH * (PN_KT1 / 8);            // 16-byte words of layer 1's weights
    static_assert(NIN <= 1024 && NW1 <= 1024 && PN_KW * PN_CH < 512 && 3 * PN_CH <= 512,
                  "the prologue's items fit two per thread");
    // input rows ch = 16 g - 4 + ci of the whitened batch from column w0 - 16 (a thread
    // takes one column and eight consecutive rows; a wave's loads of one row are 64
    // consecutive columns), and layer 1's weights, into registers
    float xv[2][8];
    uint4 w1v[2];
    auto in_load = [&](int gg) __attribute__((always_inline)) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int i = tid + 512 * h;
            const int q = i / PN_FX, c = i - q * PN_FX;
            const int w = w0 - 16 + c;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int ch = gg * PN_GROUP - PN_HALO + 8 * q + e;
                xv[h][e] = (i < NIN && ch >= 0 && ch < C && w >= 0 && w < W) ? A.xw[(int64_t)ch * A.xs + A.xoff + w]
                                                                             : 0.f;
            }
            w1v[h] = i < NW1 ? reinterpret_cast<const uint4*>(A.wt1)[i] : make_uint4(0u, 0u, 0u, 0u);
        }
    };
    in_load(g);
    if (tid < 3 * PN_CH) {
        const int l = tid / PN_CH, co = tid - l * PN_CH;
        s_bias[tid] = A.params[(l == 0 ? PN_B1 : l == 1 ? PN_B2 : PN_B3) + co];
    }
    if (tid < PN_KW * PN_CH) {
        const int k = tid / PN_CH, ci = tid - k * PN_CH;
        s_w4[tid] = A.params[PN_W4 + ci * PN_KW + k];
    } else if (tid == PN_KW * PN_CH) {
        s_w4[tid] = A.params[PN_B4];
    }
    // the next layer's weights, in flight while the current layer runs
    constexpr int NW = PN_CH * PN_KP / 8;           // 16-byte words of a 32 x 288 layer
    constexpr int NWT = (NW + 511) / 512;
    uint4 wreg[NWT];
    auto wload = [&](const __bf16* src) {
#pragma unroll
        for (int q = 0; q < NWT; ++q) {
            const int i = tid + 512 * q;
            wreg[q] = i < NW ? reinterpret_cast<const uint4*>(src)[i] : make_uint4(0u, 0u, 0u, 0u);
        }
    };
    auto wstore = [&]() {
#pragma unroll
        for (int q = 0; q < NWT; ++q) {
            const int i = tid + 512 * q;
            if (i < NW) {
                const int r = i / (PN_KP / 8), c8 = i - r * (PN_KP / 8);
                *reinterpret_cast<uint4*>(&ws[r * PN_FWP + c8 * 8]) = wreg[q];
            }
        }
    };
    for (;;) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int i = tid + 512 * h;
            if (i < NIN) {
                const int q = i / PN_FX, c = i - q * PN_FX;
                __bf16 v[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = (__bf16)xv[h][e];
                uint4 u;
                __builtin_memcpy(&u, v, 16);
                *reinterpret_cast<uint4*>(&bx[c * PN_FCP + 8 * q]) = u;
            }
            if (i < NW1) {
                const int r = i / (PN_KT1 / 8), c8 = i - r * (PN_KT1 / 8);
                *reinterpret_cast<uint4*>(&ws[r * PN_FWP + c8 * 8]) = w1v[h];
            }
        }
        __syncthreads();
        if (idle) return;                             // uniform over the workgroup
        wload(A.wt2);
        pn_fused_layer<PN_CIN>(bx, by, ws, s_bias, w0 - 12, W, lane, wv);
        __syncthreads();
        wstore();
        __syncthreads();
        wload(A.wt3);
        pn_fused_layer<PN_CH>(by, bx, ws, s_bias + PN_CH, w0 - 8, W, lane, wv);
        __syncthreads();
        wstore();
        __syncthreads();
        const int gn = g + GZ;
        if (gn < G) in_load(gn);                      // the next group's input, during layer 3
        pn_fused_layer<PN_CH>(bx, by, ws, s_bias + 2 * PN_CH, w0 - 4, W, lane, wv);
        __syncthreads();
        // head: z = b4 + sum_k sum_ci W4[ci][k] a3[w + k - 4][ci] (layer 3's column jh + k)
        if (tid < PN_FT) {
            const int w = w0 + tid;
            if (w < A.Wp) {
                float acc = s_w4[PN_KW * PN_CH];
                if (w < W) {
#pragma unroll
                    for (int k = 0; k < PN_KW; ++k) {
                        const __bf16* col = &by[(tid + k) * PN_FCP];
#pragma unroll
                        for (int c8 = 0; c8 < PN_CH / 8; ++c8) {
                            const bf16x8 v = *reinterpret_cast<const bf16x8*>(col + 8 * c8);
#pragma unroll
                            for (int e = 0; e < 8; ++e) acc += s_w4[k * PN_CH + 8 * c8 + e] * (float)v[e];
                        }
                    }
                }
                const int64_t o = (int64_t)g * A.Wp + w;
                A.z[o] = w < W ? acc : 0.f;
                A.p[o] = w < W ? 1.f / (1.f + expf(-acc)) : 0.f;
            }
        }
        g = gn;
        if (g >= G) break;
        // the head reads by; the next group writes bx and ws, then the barrier above
        // its layer 1 (which writes by) orders it after every head read
    }
}

// ---------------------------------------------------------------------------
// gate: per 64-column block and group, the largest P within the block widened
// by `reach` columns on each side; a block with any group strictly between lo and hi falls back to full
// scoring, otherwise a template tile is scored when a group within one of its
// group span has P >= hi.
// ---------------------------------------------------------------------------
struct PnGateArgs {
    const float* p; int G; int W; int Wp;
    PnGate gt;
    const int16_t* const* active;
};
struct PnGateTable { int n; PnGateArgs a[PN_MAX_JOBS]; };

__global__ __launch_bounds__(256) void k_pn_gate(PnGateTable tb) {
    // grid (stream, column block): stream j on the XCDs x = j mod n, as its network's
    // workgroups and its projection tiles
    const PnGateArgs& A = tb.a[blockIdx.x];
    const bool idle = A.active && !*A.active;
    const float* __restrict__ p = A.p;
    const int G = A.G, W = A.W, Wp = A.Wp;
    const PnGate& gt = A.gt;
    if ((int)blockIdx.y * 64 >= W) return;
    __shared__ int s_state[64];                       // per group: 0 off, 1 uncertain, 2 on
    const int n = blockIdx.y;
    const int lo = max(0, n * 64 - gt.reach), hi = min(W, n * 64 + 64 + gt.reach);
    for (int g = threadIdx.x >> 6; g < G; g += 4) {  // one wave per group
        float mx = 0.f;
        for (int w = lo + (threadIdx.x & 63); w < hi; w += 64) mx = fmaxf(mx, p[(int64_t)g * Wp + w]);
        for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
        if ((threadIdx.x & 63) == 0) s_state[g] = mx >= gt.hi ? 2 : (mx > gt.lo ? 1 : 0);
    }
    __syncthreads();
    if (idle) return;                                 // stream idle this launch (uniform)
    bool unc = false;                                 // some group uncertain: the fallback
    for (int g = 0; g < G; ++g) unc = unc || s_state[g] == 1;
    int scored = 0;
    for (int m = threadIdx.x; m < gt.MT; m += 256) {
        bool on = unc;
        if (!on) {
            const int a = max(0, gt.mt_glo[m] - 1), b = min(G - 1, gt.mt_ghi[m] + 1);
            for (int g = a; g <= b && !on; ++g) on = s_state[g] == 2;
        }
        gt.mask[(int64_t)n * gt.MT + m] = on ? 1 : 0;
        scored += on;
    }
    for (int o = 32; o > 0; o >>= 1) scored += __shfl_xor(scored, o);
    if ((threadIdx.x & 63) == 0 && scored) atomicAdd(&gt.stats[0], (unsigned long long)scored);
    if (threadIdx.x == 0) {
        atomicAdd(&gt.stats[1], (unsigned long long)gt.MT);
        if (unc) atomicAdd(&gt.stats[2], 1ull);
        atomicAdd(&gt.stats[3], 1ull);
    }
}

}  // namespace

int pn_groups(int C) { return (C + PN_GROUP - 1) / PN_GROUP; }

size_t pn_part_floats(int G, int Wp) {
    const int slabs = (Wp + PN_SLAB - 1) / PN_SLAB;
    return (size_t)G * slabs * PN_CH * ((PN_CH * PN_KW + 31) / 32 * 32);
}

// All streams of a launch run each layer in one dispatch (grid z = stream).
void pn_forward_batch(const PnFwdJob* jobs, int n, hipStream_t s) {
    if (n <= 0) return;
    PnInputTable ti{}; PnConvTable t1{}, t2{}, t3{}; PnHeadTable th{};
    ti.n = t1.n = t2.n = t3.n = th.n = n;
    int Wpm = 0, Gm = 0, Wm = 0;
    for (int j = 0; j < n; ++j) {
        const PnFwdJob& J = jobs[j];
        const PnBuffers& b = J.b;
        ti.a[j] = PnInputArgs{J.xw, J.xs, J.xoff, J.C, J.W, b.Wp, b.G, b.a0, J.active};
        t1.a[j] = PnConvArgs{b.a0, J.wt.wf1, J.wt.params + PN_B1, nullptr, b.a1, J.W, b.Wp, b.G, J.active};
        t2.a[j] = PnConvArgs{b.a1, J.wt.wf2, J.wt.params + PN_B2, nullptr, b.a2, J.W, b.Wp, b.G, J.active};
        t3.a[j] = PnConvArgs{b.a2, J.wt.wf3, J.wt.params + PN_B3, nullptr, b.a3, J.W, b.Wp, b.G, J.active};
        th.a[j] = PnHeadArgs{b.a3, J.wt.params, b.z, b.p, J.W, b.Wp, b.G, J.active};
        Wpm = max(Wpm, b.Wp); Gm = max(Gm, b.G); Wm = max(Wm, J.W);
    }
    hipLaunchKernelGGL(k_pn_input, dim3((Wpm + 255) / 256, Gm * PN_CIN, n), dim3(256), 0, s, ti);
    const dim3 grid((Wm + PN_TILE - 1) / PN_TILE, Gm, n);
    hipLaunchKernelGGL((k_pn_conv<PN_CIN, 0>), grid, dim3(256), 0, s, t1);
    hipLaunchKernelGGL((k_pn_conv<PN_CH, 0>), grid, dim3(256), 0, s, t2);
    hipLaunchKernelGGL((k_pn_conv<PN_CH, 0>), grid, dim3(256), 0, s, t3);
    hipLaunchKernelGGL(k_pn_head, dim3((Wpm + PN_HEAD_COLS - 1) / PN_HEAD_COLS, Gm, n), dim3(256), 0, s, th);
}

void pn_infer_batch(const PnFwdJob* jobs, int n, hipStream_t s) {
    if (n <= 0) return;
    PnInferTable t{};
    t.n = n;
    int Gm = 0, Wm = 0;
    for (int j = 0; j < n; ++j) {
        const PnFwdJob& J = jobs[j];
        t.a[j] = PnInferArgs{J.xw, J.xs, J.xoff, J.C, J.W, J.b.Wp, J.b.G, J.wt.wt1, J.wt.wt2, J.wt.wt3,
                             J.wt.params, J.b.z, J.b.p, J.active};
        Gm = max(Gm, J.b.G);
        Wm = max(Wm, J.W);
    }
    // group slots: one round of two workgroups per CU for the 8-stream, 24-group shape
    const int gz = std::min(Gm, 8);
    hipLaunchKernelGGL(k_pn_fused, dim3(n, (Wm + PN_FT - 1) / PN_FT, gz), dim3(512), 0, s, t);
}

void pn_infer(const PnWeights& wt, const PnBuffers& b, const float* xw, int64_t xs, int xoff, int C, int W,
              hipStream_t s, const int16_t* const* active) {
    const PnFwdJob j{xw, xs, xoff, C, W, wt, b, active};
    pn_infer_batch(&j, 1, s);
}

void pn_gate_batch(const PnGateJob* jobs, int n, hipStream_t s) {
    if (n <= 0) return;
    PnGateTable tg{};
    tg.n = n;
    int Wm = 0;
    for (int j = 0; j < n; ++j) {
        tg.a[j] = PnGateArgs{jobs[j].b.p, jobs[j].b.G, jobs[j].W, jobs[j].b.Wp, jobs[j].gt, jobs[j].active};
        Wm = max(Wm, jobs[j].W);
    }
    hipLaunchKernelGGL(k_pn_gate, dim3(n, (Wm + 63) / 64), dim3(256), 0, s, tg);
}

void pn_forward(const PnWeights& wt, const PnBuffers& b, const float* xw, int64_t xs, int xoff, int C, int W,
                hipStream_t s, const int16_t* const* active) {
    const PnFwdJob j{xw, xs, xoff, C, W, wt, b, active};
    pn_forward_batch(&j, 1, s);
}

void pn_gate(const PnBuffers& b, const PnGate& gt, int W, hipStream_t s, const int16_t* const* active) {
    const PnGateJob j{b, gt, W, active};
    pn_gate_batch(&j, 1, s);
}

void pn_backward(const PnWeights& wt, const PnBuffers& b, const uint8_t* y, int W, float pos_weight, float* grad,
                 hipStream_t s) {
    const int G = b.G, Wp = b.Wp;
    const int nb = (G * Wp + 255) / 256;
    hipLaunchKernelGGL(k_pn_loss, dim3(nb), dim3(256), 0, s, b.z, y, G, W, Wp, pos_weight, b.dz, b.lpart);
    hipLaunchKernelGGL(k_pn_head_wgrad, dim3(PN_CH + 1), dim3(256), 0, s, b.dz, b.a3, G, W, Wp, grad);
    hipLaunchKernelGGL(k_pn_head_back, dim3((Wp + 255) / 256, G * PN_CH), dim3(256), 0, s, b.dz, wt.params, b.a3,
                       W, Wp, b.d3);
    const dim3 grid((W + PN_TILE - 1) / PN_TILE, G);
    // data gradients through layers 3 and 2 (masked by the ReLU of the layer below)
    PnConvTable c3{}, c2{};
    c3.n = c2.n = 1;
    c3.a[0] = PnConvArgs{b.d3, wt.wb3, nullptr, b.a2, b.d2, W, Wp, G, nullptr};
    c2.a[0] = PnConvArgs{b.d2, wt.wb2, nullptr, b.a1, b.d1, W, Wp, G, nullptr};
    hipLaunchKernelGGL((k_pn_conv<PN_CH, 1>), grid, dim3(256), 0, s, c3);
    hipLaunchKernelGGL((k_pn_conv<PN_CH, 1>), grid, dim3(256), 0, s, c2);
    // weight gradients of layers 3, 2, 1
    const int slabs = (W + PN_SLAB - 1) / PN_SLAB;
    const dim3 sg(slabs, G);
    const int nparts = slabs * G;
    const int NP32 = (PN_CH * PN_KW + 31) / 32 * 32, NP24 = (PN_CIN * PN_KW + 31) / 32 * 32;
    hipLaunchKernelGGL(k_pn_wgrad<PN_CH>, sg, dim3(256), 0, s, b.d3, b.a2, W, Wp, b.part, b.bpart);
    hipLaunchKernelGGL(k_pn_reduce, dim3((PN_CH * PN_CH * PN_KW + PN_CH + 255) / 256), dim3(256), 0, s, b.part,
                       b.bpart, nparts, PN_CH * PN_KW, NP32, grad + PN_W3, grad + PN_B3);
    hipLaunchKernelGGL(k_pn_wgrad<PN_CH>, sg, dim3(256), 0, s, b.d2, b.a1, W, Wp, b.part, b.bpart);
    hipLaunchKernelGGL(k_pn_reduce, dim3((PN_CH * PN_CH * PN_KW + PN_CH + 255) / 256), dim3(256), 0, s, b.part,
                       b.bpart, nparts, PN_CH * PN_KW, NP32, grad + PN_W2, grad + PN_B2);
    hipLaunchKernelGGL(k_pn_wgrad<PN_CIN>, sg, dim3(256), 0, s, b.d1, b.a0, W, Wp, b.part, b.bpart);
    hipLaunchKernelGGL(k_pn_reduce, dim3((PN_CH * PN_CIN * PN_KW + PN_CH + 255) / 256), dim3(256), 0, s, b.part,
                       b.bpart, nparts, PN_CIN * PN_KW, NP24, grad + PN_W1, grad + PN_B1);
}

void pn_adam(const PnWeights& wt, const float* grad, float* m, float* v, float lr, float beta1, float beta2,
             float eps, int step, hipStream_t s) {
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_pn_adam, dim3((PN_NPARAMS + 255) / 256), dim3(256), 0, s, wt.params, grad, m, v, lr, beta1,
                       beta2, eps, c1, c2);
}

void pn_pack(const PnWeights& wt, hipStream_t s) {
    hipLaunchKernelGGL(k_pn_pack, dim3((PN_CH * PN_KP + 255) / 256), dim3(256), 0, s, wt);
}

}  // namespace lss
