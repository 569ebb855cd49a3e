// HIP kernels of the online template matcher, written for gfx950 (CDNA4).
//
// Stage map (reference file:line -> kernel here):
//   whitening GEMM   OnlineSpikesV2.cu:797-801, myGPUhelpers.cu:722-738 -> k_whiten (FP32 MFMA,
//                    int16 -> f32 conversion and channel gather fused into the operand load)
//   6-PC conv        myGPUhelpers.cu:356-400, KS/template_matching.py:145 -> k_pcconv (LDS-staged)
//   projection       myGPUhelpers.cu:263-337, KS/template_matching.py:146 -> k_project (FP32 MFMA)
//   template norms   OnlineSpikesV2.cu:623-637 -> k_norms (FP64)
//   peel loop        OnlineSpikesV2.cu:903-1001, KS/template_matching.py:157-198 -> k_peel
//   PC features      OnlineSpikesV2.cu:113-188, KS/template_matching.py:50-51 -> k_peel epilogue
// (KS/ = src/Python/run_online_spikes/Lib/site-packages/kilosort/)

#include "lss_kernels.h"

#include <cmath>

namespace lss {

typedef float f32x16 __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------------------
// FP32 MFMA GEMM core.  A 256-thread workgroup owns a 64x64 output tile; its
// four waves each own a 32x32 quadrant computed with v_mfma_f32_32x32x2f32.
// Operand layouts of that instruction (lane l of 64):
//   A[i][k]: i = l % 32, k = l / 32      B[k][j]: j = l % 32, k = l / 32
//   D item r: row = (r % 4) + 8 * (r / 4) + 4 * (l / 32), column = l % 32
// Both operands sit in LDS as [k][m] / [k][n] rows, so the 32 lanes of one
// half-wave read 32 consecutive words.  The reduction runs only over the
// K-blocks listed for the tile (exact-zero blocks of A are skipped).
// ---------------------------------------------------------------------------
constexpr int LDS_PAD = 4;

struct GemmTile {
    float As[GEMM_BK][GEMM_BM + LDS_PAD];
    float Bs[GEMM_BK][GEMM_BN + LDS_PAD];
};

__device__ __forceinline__ f32x16 mfma_tile(const GemmTile& s, f32x16 acc, int lane, int wm, int wn) {
#pragma unroll
    for (int kk = 0; kk < GEMM_BK / 2; ++kk) {
        const float a = s.As[kk * 2 + (lane >> 5)][wm * 32 + (lane & 31)];
        const float b = s.Bs[kk * 2 + (lane >> 5)][wn * 32 + (lane & 31)];
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
    }
    return acc;
}

__device__ __forceinline__ void store_tile(float* __restrict__ C, int64_t ldc, int m0, int n0,
                                           const f32x16& acc, int lane, int wm, int wn) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = (r & 3) + 8 * (r >> 2) + 4 * (lane >> 5);
        const int j = lane & 31;
        C[(int64_t)(m0 + wm * 32 + i) * ldc + n0 + wn * 32 + j] = acc[r];
    }
}

// ---------------------------------------------------------------------------
// Batched launches: one launch serves up to MAX_JOBS independent jobs (probe
// streams).  Block ids are dealt to jobs by the start offsets in the batch
// descriptor (kernel arguments).
//
// XCD-aware tile order: the GPU deals consecutive workgroup ids round-robin to
// its 8 XCDs, each with its own L2.  Block k of a GEMM job goes to XCD k % 8;
// the job's tiles are ordered so that all M-tiles of one 64-column N-tile sit
// on the same XCD, so the shared B-operand column block is fetched from HBM
// once into that XCD's L2 instead of once per M-tile.  Jobs start at multiples
// of 8 and pad their N-tile count to a multiple of 8 (padding blocks exit).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int job_of(const int* start, int n, int b) {
    int j = 0;
    while (j + 1 < n && b >= start[j + 1]) ++j;
    return j;
}

__device__ __forceinline__ bool xcd_tile(int k, int mtiles, int ntiles, int& m, int& nt) {
    const int slot = k >> 3;
    nt = (k & 7) + 8 * (slot / mtiles);
    m = slot % mtiles;
    return nt < ntiles;
}

// The job and tile of this block.  Interleaved launches (bt.ilv): job j = block % n
// runs on the XCDs x = j (mod n) and deals its tiles over them as xcd_tile does over
// all eight (the M-tiles of one N-tile on one XCD).
template <class J>
__device__ __forceinline__ bool gemm_tile(const JobBatch<J>& bt, int& j, int& mt, int& nt) {
    if (bt.ilv) {
        const int n = bt.n, per = 8 / n;
        j = (int)blockIdx.x % n;
        const int k = (int)blockIdx.x / n;
        const int slot = k / per;
        nt = k % per + per * (slot / bt.job[j].mtiles);
        mt = slot % bt.job[j].mtiles;
        return nt < bt.job[j].ntiles;
    }
    j = job_of(bt.start, bt.n, blockIdx.x);
    return xcd_tile(blockIdx.x - bt.start[j], bt.job[j].mtiles, bt.job[j].ntiles, mt, nt);
}

// Xw[c][XOFF + w] = sum_c' Wr[c][c'] * x[w][cmap[c']]   (rows c < CP)
// Columns w >= W read zero input and so store zero: they are the right-hand
// zero padding the PC convolution reads.
//
// Each step stages four listed 16-deep K-blocks (64 of the reduction) in LDS,
// and the next step's operands are loaded into registers while the MFMAs run,
// so a tile waits for memory once per 64 of the reduction instead of once per
// 16.  Blocks are accumulated in list order, 8 MFMAs each, as a one-block step
// would.  The channel map is staged in LDS once per workgroup, so an int16
// sample load never waits for a map entry; with the identity map and aligned
// rows a thread reads its 16 channels as two 16-byte loads.
constexpr int WBK = 4 * GEMM_BK;     // reduction depth staged per step
// Channel maps up to this size are staged in LDS (larger ones are read from global
// memory).  Kept small: with it the whitening tiles take 37-39 KB of LDS, four
// workgroups per CU (a 2048-entry map left room for three).
constexpr int WCM = 512;

struct WhitenTile {
    float As[WBK][GEMM_BM + LDS_PAD];
    float Bs[WBK][GEMM_BN + LDS_PAD];
};

// One 16-deep block of the 64x64 tile: rows kb .. kb + 15 of the staged operands.
__device__ __forceinline__ f32x16 mfma_block(const WhitenTile& s, int kb, f32x16 acc, int lane, int wm, int wn) {
#pragma unroll
    for (int kk = 0; kk < GEMM_BK / 2; ++kk) {
        const float a = s.As[kb + kk * 2 + (lane >> 5)][wm * 32 + (lane & 31)];
        const float b = s.Bs[kb + kk * 2 + (lane >> 5)][wn * 32 + (lane & 31)];
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
    }
    return acc;
}

__device__ void reset_stream(const ResetBatch& b, int j);

__global__ __launch_bounds__(256) void k_whiten(WhitenLaunch wl) {
    __shared__ WhitenTile s;
    __shared__ int s_cm[WCM];
    __shared__ int s_kb[WCM / GEMM_BK];
    // the first workgroups also reset the batch's peel words, one stream each
    if ((int)blockIdx.x < wl.r.n) reset_stream(wl.r, blockIdx.x);
    const WhitenBatch& bt = wl.w;
    int j, mt, nt;
    if (!gemm_tile(bt, j, mt, nt)) return;
    const WhitenJob& jb = bt.job[j];
    const float* __restrict__ Wr = jb.Wr;
    const int16_t* __restrict__ x = jb.xptr ? *jb.xptr : jb.x;
    const float* __restrict__ xf = jb.xf;
    if (!x) return;                                             // stream idle this launch
    const int CP = jb.CP, W = jb.W, C = jb.C;
    const int64_t ld = jb.ld;
    const int m0 = mt * GEMM_BM, n0 = nt * GEMM_BN;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv >> 1, wn = wv & 1;
    const int q = tid & 3, ar = tid >> 2;        // listed block of the step, A row / B sample (int16 input)
    const bool cm_lds = !xf && CP <= WCM;
    if (cm_lds)
        for (int i = tid; i < CP; i += 256) s_cm[i] = jb.cmap[i];
    const int* __restrict__ cm = cm_lds ? s_cm : jb.cmap;
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    const int p0 = jb.kb_ptr[mt], p1 = jb.kb_ptr[mt + 1];
    // the tile's K-block list, staged too (indexed from p0)
    const bool kb_lds = p1 - p0 <= WCM / GEMM_BK;
    if (kb_lds)
        for (int i = tid; i < p1 - p0; i += 256) s_kb[i] = jb.kb_idx[p0 + i];
    auto kbi = [&](int p) { return kb_lds ? s_kb[p - p0] : jb.kb_idx[p]; };
    float4 ra[4];
    float rb[16];
    // registers of step pb (list entries pb .. pb + 3): this thread's A row segment
    // of block q, and 16 B-operand values
    auto gload = [&](int pb) {
        const bool have = pb + q < p1;
        const int k0 = have ? kbi(pb + q) * GEMM_BK : 0;
        const float4* ap = reinterpret_cast<const float4*>(Wr + (int64_t)(m0 + ar) * CP + k0);
#pragma unroll
        for (int i = 0; i < 4; ++i) ra[i] = have ? ap[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        if (xf) {
            // preprocessed input, channel-major: sample n0 + ar, channels k0 .. k0 + 15
            // of block q (a load instruction reads 16 consecutive samples of 4 rows)
            const int n = n0 + ar;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int c = k0 + i;
                rb[i] = (have && n < W && c < C) ? xf[(int64_t)c * jb.xfs + n] : 0.f;
            }
        } else {
            // int16 input, sample-major: sample n0 + ar, channels k0 .. k0 + 15 of block q
            const int n = n0 + ar;
            const int16_t* row = x + (int64_t)n * ld;
            const bool vec = have && n < W && jb.cmap_id && k0 + GEMM_BK <= C &&
                             ((reinterpret_cast<uintptr_t>(row + k0) & 15) == 0);
            if (vec) {
                const int4 v0 = reinterpret_cast<const int4*>(row + k0)[0];
                const int4 v1 = reinterpret_cast<const int4*>(row + k0)[1];
                const int w8[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    rb[2 * i] = (float)(int16_t)(w8[i] & 0xffff);
                    rb[2 * i + 1] = (float)(int16_t)((unsigned)w8[i] >> 16);
                }
            } else {
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int c = k0 + i;
                    rb[i] = (have && n < W && c < C) ? (float)row[cm[c]] : 0.f;
                }
            }
        }
    };
    __syncthreads();                             // the staged channel map and block list
    if (p0 < p1) gload(p0);
    for (int pb = p0; pb < p1; pb += 4) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            s.As[q * GEMM_BK + 4 * i + 0][ar] = ra[i].x;
            s.As[q * GEMM_BK + 4 * i + 1][ar] = ra[i].y;
            s.As[q * GEMM_BK + 4 * i + 2][ar] = ra[i].z;
            s.As[q * GEMM_BK + 4 * i + 3][ar] = ra[i].w;
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) s.Bs[q * GEMM_BK + i][ar] = rb[i];
        __syncthreads();
        if (pb + 4 < p1) gload(pb + 4);
        const int nq = min(4, p1 - pb);
        for (int qq = 0; qq < nq; ++qq) acc = mfma_block(s, qq * GEMM_BK, acc, lane, wm, wn);
    }
    store_tile(jb.Xw + XOFF, jb.XS, m0, n0, acc, lane, wm, wn);
}

// Split-BF16 whitening (the handle's precision is LSS_PREC_BF16X3): Wr carried as a
// BF16 pair (hi, lo) split at load, the input split as it is staged (an int16
// sample is exactly hi + lo), three BF16 MFMA products per FP32 product with
// FP32 accumulation.  Tile order, K-block lists, input forms and the output
// layout are k_whiten's; LDS holds both operands as [row][k].
__global__ __launch_bounds__(256) void k_whiten_x3(WhitenLaunch wl) {
    typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
    constexpr int KD = 4 * GEMM_BK, KPAD = KD + 8;
    __shared__ __attribute__((aligned(16))) __bf16 Ash[GEMM_BM][KPAD];
    __shared__ __attribute__((aligned(16))) __bf16 Asl[GEMM_BM][KPAD];
    __shared__ __attribute__((aligned(16))) __bf16 Bsh[GEMM_BN][KPAD];
    __shared__ __attribute__((aligned(16))) __bf16 Bsl[GEMM_BN][KPAD];
    __shared__ int s_cm[WCM];
    __shared__ int s_kb[WCM / GEMM_BK];
    if ((int)blockIdx.x < wl.r.n) reset_stream(wl.r, blockIdx.x);
    const WhitenBatch& bt = wl.w;
    int j, mt, nt;
    if (!gemm_tile(bt, j, mt, nt)) return;
    const WhitenJob& jb = bt.job[j];
    const int16_t* __restrict__ x = jb.xptr ? *jb.xptr : jb.x;
    const float* __restrict__ xf = jb.xf;
    if (!x) return;                                             // stream idle this launch
    const int CP = jb.CP, W = jb.W, C = jb.C;
    const int64_t ld = jb.ld;
    const int m0 = mt * GEMM_BM, n0 = nt * GEMM_BN;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv >> 1, wn = wv & 1;
    const int q = tid & 3, ar = tid >> 2;        // listed block of the step, A row / B sample
    const bool cm_lds = !xf && CP <= WCM;
    if (cm_lds)
        for (int i = tid; i < CP; i += 256) s_cm[i] = jb.cmap[i];
    const int* __restrict__ cm = cm_lds ? s_cm : jb.cmap;
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    const int p0 = jb.kb_ptr[mt], p1 = jb.kb_ptr[mt + 1];
    const bool kb_lds = p1 - p0 <= WCM / GEMM_BK;
    if (kb_lds)
        for (int i = tid; i < p1 - p0; i += 256) s_kb[i] = jb.kb_idx[p0 + i];
    auto kbi = [&](int p) { return kb_lds ? s_kb[p - p0] : jb.kb_idx[p]; };
    uint4 ah0, ah1, al0, al1;      // this thread's A row segment of block q: hi, lo
    float rb[16];                  // 16 input values: sample n0 + ar, channels of block q
    auto gload = [&](int pb) __attribute__((always_inline)) {
        const bool have = pb + q < p1;
        const int k0 = have ? kbi(pb + q) * GEMM_BK : 0;
        const uint4 z = make_uint4(0u, 0u, 0u, 0u);
        ah0 = ah1 = al0 = al1 = z;
        if (have) {
            const int64_t ao = (int64_t)(m0 + ar) * CP + k0;
            const uint4* sh = reinterpret_cast<const uint4*>(jb.Wrh + ao);
            const uint4* sl = reinterpret_cast<const uint4*>(jb.Wrl + ao);
            ah0 = sh[0]; ah1 = sh[1]; al0 = sl[0]; al1 = sl[1];
        }
        const int n = n0 + ar;
        if (xf) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int c = k0 + i;
                rb[i] = (have && n < W && c < C) ? xf[(int64_t)c * jb.xfs + n] : 0.f;
            }
        } else {
            const int16_t* row = x + (int64_t)n * ld;
            const bool vec = have && n < W && jb.cmap_id && k0 + GEMM_BK <= C &&
                             ((reinterpret_cast<uintptr_t>(row + k0) & 15) == 0);
            if (vec) {
                const int4 v0 = reinterpret_cast<const int4*>(row + k0)[0];
                const int4 v1 = reinterpret_cast<const int4*>(row + k0)[1];
                const int w8[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    rb[2 * i] = (float)(int16_t)(w8[i] & 0xffff);
                    rb[2 * i + 1] = (float)(int16_t)((unsigned)w8[i] >> 16);
                }
            } else {
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int c = k0 + i;
                    rb[i] = (have && n < W && c < C) ? (float)row[cm[c]] : 0.f;
                }
            }
        }
    };
    __syncthreads();                             // the staged channel map and block list
    if (p0 < p1) gload(p0);
    for (int pb = p0; pb < p1; pb += 4) {
        __syncthreads();
        *reinterpret_cast<uint4*>(&Ash[ar][q * 16]) = ah0;
        *reinterpret_cast<uint4*>(&Ash[ar][q * 16 + 8]) = ah1;
        *reinterpret_cast<uint4*>(&Asl[ar][q * 16]) = al0;
        *reinterpret_cast<uint4*>(&Asl[ar][q * 16 + 8]) = al1;
#pragma unroll
        for (int h8 = 0; h8 < 2; ++h8) {
            __bf16 hi[8], lo[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float v = rb[8 * h8 + e];
                hi[e] = (__bf16)v;
                lo[e] = (__bf16)(v - (float)hi[e]);
            }
            uint4 uh, ul;
            __builtin_memcpy(&uh, hi, 16);
            __builtin_memcpy(&ul, lo, 16);
            *reinterpret_cast<uint4*>(&Bsh[ar][q * 16 + 8 * h8]) = uh;
            *reinterpret_cast<uint4*>(&Bsl[ar][q * 16 + 8 * h8]) = ul;
        }
        __syncthreads();
        if (pb + 4 < p1) gload(pb + 4);
        const int nk = min(4, p1 - pb);
        for (int kk = 0; kk < nk; ++kk) {
            const int ko = kk * 16 + 8 * (lane >> 5);
            const bf16x8_t a_h = *reinterpret_cast<const bf16x8_t*>(&Ash[wm * 32 + (lane & 31)][ko]);
            const bf16x8_t a_l = *reinterpret_cast<const bf16x8_t*>(&Asl[wm * 32 + (lane & 31)][ko]);
            const bf16x8_t b_h = *reinterpret_cast<const bf16x8_t*>(&Bsh[wn * 32 + (lane & 31)][ko]);
            const bf16x8_t b_l = *reinterpret_cast<const bf16x8_t*>(&Bsl[wn * 32 + (lane & 31)][ko]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_l, b_h, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_h, b_l, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_h, b_h, acc, 0, 0, 0);
        }
    }
    store_tile(jb.Xw + XOFF, jb.XS, m0, n0, acc, lane, wm, wn);
}

// A projection tile's K-block list entries p0 .. p1 - 1, staged in LDS (indexed
// from p0) so that no operand load waits on a list entry; a longer list is read
// from global memory.  Ends with a barrier.
constexpr int PROJ_KB_LDS = 512;
__device__ __forceinline__ bool stage_kblocks(const int* __restrict__ kb_idx, int p0, int p1, int* s_kb, int tid) {
    const bool lds = p1 - p0 <= PROJ_KB_LDS;
    if (lds)
        for (int i = tid; i < p1 - p0; i += 256) s_kb[i] = kb_idx[p0 + i];
    __syncthreads();
    return lds;
}

__global__ __launch_bounds__(256) void k_project(ProjectBatch bt) {
    __shared__ GemmTile s;
    __shared__ int s_kb[PROJ_KB_LDS];
    int j, mt, nt;
    if (!gemm_tile(bt, j, mt, nt)) return;
    const ProjectJob& jb = bt.job[j];
    // the tile's flags and list bounds are loaded together (one memory latency)
    const bool gated = jb.gate && !jb.gate[nt * jb.gate_mt + mt];   // tile not proposed: not computed
    const bool idle = jb.active && !*jb.active;                     // stream idle this launch
    const bool kept = jb.act && !jb.act[nt * jb.mtiles + mt];       // BF16 mode: tile kept in BF16
    const int p0 = jb.kb_ptr[mt], p1 = jb.kb_ptr[mt + 1];
    if (gated || idle || kept) return;
    const float* __restrict__ A = jb.A;
    const float* __restrict__ Bm = jb.Bm;
    const int lda = jb.lda, ldb = jb.ldb;
    const int m0 = mt * GEMM_BM, n0 = nt * GEMM_BN;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv >> 1, wn = wv & 1;
    const int ar = tid >> 2, ak = (tid & 3) * 4;     // A: row, 4 consecutive k
    const int br = tid >> 4, bc = (tid & 15) * 4;    // B: k row, 4 consecutive columns
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    const bool kb_lds = stage_kblocks(jb.kb_idx, p0, p1, s_kb, tid);
    auto kbi = [&](int p) { return kb_lds ? s_kb[p - p0] : jb.kb_idx[p]; };
    float4 ra = make_float4(0.f, 0.f, 0.f, 0.f), rb = ra;
    auto gload = [&](int kb) {
        const int k0 = kb * GEMM_BK;
        ra = *reinterpret_cast<const float4*>(A + (int64_t)(m0 + ar) * lda + k0 + ak);
        rb = *reinterpret_cast<const float4*>(Bm + (int64_t)(k0 + br) * ldb + n0 + bc);
    };
    if (p0 < p1) gload(kbi(p0));
    for (int p = p0; p < p1; ++p) {
        __syncthreads();
        s.As[ak + 0][ar] = ra.x;
        s.As[ak + 1][ar] = ra.y;
        s.As[ak + 2][ar] = ra.z;
        s.As[ak + 3][ar] = ra.w;
        *reinterpret_cast<float4*>(&s.Bs[br][bc]) = rb;
        __syncthreads();
        if (p + 1 < p1) gload(kbi(p + 1));
        acc = mfma_tile(s, acc, lane, wm, wn);
    }
    if (jb.act && jb.n_rescored) {
        // threshold decisions of single scores that the FP32 values change (the BF16
        // scores are still in C: each lane reads the elements it overwrites next)
        int flips = 0;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = (r & 3) + 8 * (r >> 2) + 4 * (lane >> 5);
            const int64_t o = (int64_t)(m0 + wm * 32 + i) * jb.ldc + n0 + wn * 32 + (lane & 31);
            const float rn = jb.rnm[m0 + wm * 32 + i];
            const float b16 = fmaxf(jb.C[o], 0.f), b32 = fmaxf(acc[r], 0.f);
            flips += ((b16 * b16) * rn > jb.Th2) != ((b32 * b32) * rn > jb.Th2);
        }
        for (int o = 32; o > 0; o >>= 1) flips += __shfl_xor(flips, o);
        if (lane == 0 && flips) atomicAdd(&jb.n_rescored[1], (unsigned long long)flips);
    }
    store_tile(jb.C, jb.ldc, m0, n0, acc, lane, wm, wn);
}

// ---------------------------------------------------------------------------
// 6-PC temporal convolution, 'same' zero padding, cross-correlation (no flip):
//   Bpc[k*CP + c][w] = sum_m Xw[c][w + m - M/2] * wPCA[k][m]
// One workgroup stages 256 + 64 samples of one whitened channel row in LDS
// (coalesced), each thread produces the K outputs of one column.
// ---------------------------------------------------------------------------
// Bpc[k][c][w] = sum_m wpca[k][m] * Xw[c][w + m - M/2].  A workgroup takes four
// channel rows of one 256-column block; a thread computes 4 consecutive columns
// of its row for every PC, reading the staged row 4 taps at a time as two
// aligned 16-byte words (the row is staged from column w0 - M/2).  Each output
// accumulates its taps in order with one fused multiply-add per tap.
constexpr int PC_ROWS = PC_ROWS_PER_BLOCK;

template <int KK>
__global__ __launch_bounds__(256) void k_pcconv(PcconvBatch bt) {
    constexpr int XW = 256 + 2 * XOFF;           // staged columns per row (>= 256 + M - 1 + 4)
    __shared__ __align__(16) float xs[PC_ROWS][XW];
    __shared__ __align__(16) float wt[MAX_M + 1][8];   // wt[m][k] = wpca[k][m]: a tap's weights, two 16-byte reads
    // interleaved launches: job j = block mod n, on the XCDs where the job's whitening ran
    const int j = bt.ilv ? (int)blockIdx.x % bt.n : job_of(bt.start, bt.n, blockIdx.x);
    const int k = bt.ilv ? (int)blockIdx.x / bt.n : (int)blockIdx.x - bt.start[j];
    const PcconvJob& jb = bt.job[j];
    if (bt.ilv && k >= pc_job_blocks(jb)) return;
    if (jb.active && !*jb.active) return;                      // stream idle this launch
    const int cg = k / jb.wblocks, w0 = (k % jb.wblocks) * 256, tid = threadIdx.x;
    const int M = jb.M, h = M / 2, C = jb.C;
    for (int i = tid; i < PC_ROWS * XW; i += 256) {
        const int r = i / XW, o = i - r * XW, cc = cg * PC_ROWS + r;
        // Xw rows are valid from column -XOFF and past BS + XOFF
        xs[r][o] = (cc < C && o < 256 + M - 1) ? jb.Xw[(int64_t)cc * jb.XS + XOFF + w0 - h + o] : 0.f;
    }
    for (int i = tid; i < M * 8; i += 256) {
        const int m = i >> 3, q = i & 7;
        wt[m][q] = q < KK ? jb.wpca[q * M + m] : 0.f;
    }
    __syncthreads();
    // rows past C store zeros (their staged samples are zero), so the padded rows of
    // the projection's operand never hold stale values
    const int rr = tid >> 6, l = tid & 63, c = cg * PC_ROWS + rr;
    float acc[4][KK];
#pragma unroll
    for (int jj = 0; jj < 4; ++jj)
#pragma unroll
        for (int q = 0; q < KK; ++q) acc[jj][q] = 0.f;
    const float* xp = &xs[rr][4 * l];
    for (int m0 = 0; m0 < (c < C ? M : 0); m0 += 4) {
        const float4 a = *reinterpret_cast<const float4*>(xp + m0);
        const float4 b = *reinterpret_cast<const float4*>(xp + m0 + 4);
        const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int mm = 0; mm < 4; ++mm) {
            if (m0 + mm < M) {
                const float4 wa = *reinterpret_cast<const float4*>(&wt[m0 + mm][0]);
                const float4 wb = *reinterpret_cast<const float4*>(&wt[m0 + mm][4]);
                const float wv8[8] = {wa.x, wa.y, wa.z, wa.w, wb.x, wb.y, wb.z, wb.w};
#pragma unroll
                for (int q = 0; q < KK; ++q) {
#pragma unroll
                    for (int jj = 0; jj < 4; ++jj) acc[jj][q] = fmaf(wv8[q], v[jj + mm], acc[jj][q]);
                }
            }
        }
    }
    const int64_t o0 = (int64_t)c * jb.BS + w0 + 4 * l;
#pragma unroll
    for (int q = 0; q < KK; ++q) {
        const int64_t o = (int64_t)q * jb.CP * jb.BS + o0;
        *reinterpret_cast<float4*>(jb.Bpc + o) = make_float4(acc[0][q], acc[1][q], acc[2][q], acc[3][q]);
        if (jb.Bpcb) {
            // the four BF16 values as one 8-byte store
            const __bf16 v[4] = {(__bf16)acc[0][q], (__bf16)acc[1][q], (__bf16)acc[2][q], (__bf16)acc[3][q]};
            uint2 u;
            __builtin_memcpy(&u, v, 8);
            *reinterpret_cast<uint2*>(jb.Bpcb + o) = u;
        }
    }
}

// ---------------------------------------------------------------------------
// The PC conv as split-BF16 products on MFMA (precision LSS_PREC_BF16X3).  Over
// 16 output columns w0 .. w0 + 15 and the 80 samples j = w0 - 32 + j' the conv is
// the GEMM
//   Bpc[k][c][w0 + w'] = sum_j' Xw[c][w0 - 32 + j'] * Q[j'][k * 16 + w'],
//   Q[j'][k * 16 + w'] = wpca[k][j' - w' + M/2 - 32]   (zero outside 0 .. M - 1),
// whose B operand Q depends on neither the data nor w0: it is split into BF16 hi
// and lo once per handle and staged in LDS as [n][j'] by each workgroup.  The A operand is the
// whitened rows themselves (eight consecutive samples of one channel per lane),
// split while loaded.  Each product is lo_A*hi_B + hi_A*lo_B + hi_A*hi_B with FP32
// accumulation, so an output is within 3 * 2^-16 * sum |wpca||Xw| of the exact
// sum plus the FP32 accumulation error.  A wave owns 32 channels and 64 columns
// (four 16-column sub-tiles); consecutive sub-tiles share four of their five
// 16-sample blocks, which stay in registers.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int PCX_SUB = 4;         // 16-column sub-tiles of a wave

template <int KK>
__global__ __launch_bounds__(256) void k_pcconv_x3(PcconvBatch bt) {
    constexpr int NB = (KK + 1) / 2;             // 32-wide blocks of n = k * 16 + w'
    constexpr int QP = PCX_K + 8;                // LDS pitch of a Q row (conflict-free 16-byte reads)
    __shared__ __attribute__((aligned(16))) __bf16 Qh[NB * 32][QP];
    __shared__ __attribute__((aligned(16))) __bf16 Ql[NB * 32][QP];
    const int j = bt.ilv ? (int)blockIdx.x % bt.n : job_of(bt.start, bt.n, blockIdx.x);
    const int k = bt.ilv ? (int)blockIdx.x / bt.n : (int)blockIdx.x - bt.start[j];
    const PcconvJob& jb = bt.job[j];
    if (bt.ilv && k >= pc_job_blocks(jb)) return;
    if (jb.active && !*jb.active) return;                      // stream idle this launch
    const int C = jb.C;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int cb = (k / jb.wblocks) * PCX_ROWS;                 // the workgroup's channel rows
    const int wb0 = (k % jb.wblocks) * 256 + wv * (16 * PCX_SUB);   // the wave's first column
    static_assert(PCX_K % 8 == 0, "Q rows are copied as 16-byte words");
    for (int i = tid; i < 2 * NB * 32 * (PCX_K / 8); i += 256) {
        const int r = i / (PCX_K / 8), c8 = i - r * (PCX_K / 8);      // r < NB * 32: hi, else lo
        const uint4 v = reinterpret_cast<const uint4*>(jb.q)[i];
        __bf16* dst = r < NB * 32 ? &Qh[r][8 * c8] : &Ql[r - NB * 32][8 * c8];
        *reinterpret_cast<uint4*>(dst) = v;
    }
    // A: channel cb + lane % 32, samples wb0 - 32 + 16 b + 8 (lane / 32) + e of block b
    // (Xw rows hold zeros from column -XOFF and past the batch)
    const int c = cb + (lane & 31);
    const bool cv = c < C;
    const float* xr = jb.Xw + (int64_t)(cv ? c : 0) * jb.XS + XOFF + wb0 - 32 + 8 * (lane >> 5);
    bf16x8 ah[PCX_SUB + 4], al[PCX_SUB + 4];
    auto aload = [&](int b) __attribute__((always_inline)) {
        float4 u = make_float4(0.f, 0.f, 0.f, 0.f), v = u;
        if (cv) {
            u = *reinterpret_cast<const float4*>(xr + 16 * b);
            v = *reinterpret_cast<const float4*>(xr + 16 * b + 4);
        }
        const float x[8] = {u.x, u.y, u.z, u.w, v.x, v.y, v.z, v.w};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const __bf16 hi = (__bf16)x[e];
            ah[b][e] = hi;
            al[b][e] = (__bf16)(x[e] - (float)hi);
        }
    };
    // columns past the batch rounded up to 64 are not computed (a wave's 64 columns are
    // all inside or all outside)
    const bool live = wb0 < jb.ncols;
    if (live) {
#pragma unroll
        for (int b = 0; b < 5; ++b) aload(b);
    }
    __syncthreads();                             // Q staged
    if (!live) return;
#pragma unroll
    for (int t = 0; t < PCX_SUB; ++t) {
        if (t + 5 < PCX_SUB + 4) aload(t + 5);   // the next sub-tile's new block
        // Q is re-read from LDS per sub-tile (held in registers over all four it took
        // 210 VGPRs: one wave per SIMD)
        asm volatile("" ::: "memory");
        f32x16 acc[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[nb][r] = 0.f;
#pragma unroll
        for (int s = 0; s < 5; ++s) {
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const bf16x8 bh = *reinterpret_cast<const bf16x8*>(&Qh[nb * 32 + (lane & 31)][16 * s + 8 * (lane >> 5)]);
                const bf16x8 bl = *reinterpret_cast<const bf16x8*>(&Ql[nb * 32 + (lane & 31)][16 * s + 8 * (lane >> 5)]);
                acc[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[t + s], bh, acc[nb], 0, 0, 0);
                acc[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[t + s], bl, acc[nb], 0, 0, 0);
                acc[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[t + s], bh, acc[nb], 0, 0, 0);
            }
        }
        // D[i][n]: channel cb + i, PC k = n / 16, column wb0 + 16 t + n % 16 (rows past C are
        // zero, so the padded rows of the projection's operand never hold stale values)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int n = nb * 32 + (lane & 31), kq = n >> 4;
            if (kq < KK) {
                float* dst = jb.Bpc + ((int64_t)kq * jb.CP + cb) * jb.BS + wb0 + 16 * t + (n & 15);
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int i = (r & 3) + 8 * (r >> 2) + 4 * (lane >> 5);
                    if (cb + i < jb.CP) dst[(int64_t)i * jb.BS] = acc[nb][r];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// BF16 projection with FP32 accumulation (v_mfma_f32_32x32x16_bf16; lane l holds
// A[l % 32][8 * (l / 32) + e] and B[8 * (l / 32) + e][l % 32], e = 0..7).  Each
// pipeline step gathers four listed 16-wide K-blocks into a 64-deep LDS tile
// (rows k-contiguous, padded), so the block-sparse skipping keeps its 16-row
// granularity.  Optionally writes a second copy of the result (the initial
// scores that the peel's FP32 rescoring compares against).
// ---------------------------------------------------------------------------
constexpr int BK16 = 64;           // K depth of one BF16 step (four 16-blocks)
constexpr int BPAD = 8;

__global__ __launch_bounds__(256) void k_project_bf16(ProjectBatch bt) {
    __shared__ __attribute__((aligned(16))) __bf16 As[GEMM_BM][BK16 + BPAD];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[GEMM_BN][BK16 + BPAD];
    __shared__ float s_nq[GEMM_BN];
    __shared__ int s_act;
    __shared__ int s_kb[PROJ_KB_LDS];
    int j, mt, nt;
    if (!gemm_tile(bt, j, mt, nt)) return;
    const ProjectJob& jb = bt.job[j];
    const bool idle = jb.active && !*jb.active;
    const bool gated = jb.gate && !jb.gate[nt * jb.gate_mt + mt];
    const int p0 = jb.kb_ptr[mt], p1 = jb.kb_ptr[mt + 1];
    if (idle) return;                                          // stream idle this launch
    if (gated) {                                               // tile not proposed: not computed
        if (jb.act && threadIdx.x == 0) jb.act[nt * jb.mtiles + mt] = 0;
        return;
    }
    const __bf16* __restrict__ A = jb.Ab;
    const __bf16* __restrict__ Bm = jb.Bmb;
    const int lda = jb.lda, ldb = jb.ldb;
    const int m0 = mt * GEMM_BM, n0 = nt * GEMM_BN;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv >> 1, wn = wv & 1;
    const int ar = tid >> 2, aq = tid & 3;                  // A: row, sub-block
    const int bq = tid >> 6, brow = (tid & 63) >> 2, bc = (tid & 3) * 16;   // B: sub-block, k row, 16 columns
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    // squared norm of each column's BF16 operand over the tile's listed K-blocks
    // (four threads per column, 16 of each step's 64 rows each): it bounds the
    // BF16 error of every score of the tile (used by the peel's FP32 rescoring)
    float nsq = 0.f;
    const int ncol = tid >> 2, npart = tid & 3;
    const bool kb_lds = stage_kblocks(jb.kb_idx, p0, p1, s_kb, tid);
    auto kbi = [&](int p) { return kb_lds ? s_kb[p - p0] : jb.kb_idx[p]; };
    for (int p = p0; p < p1; p += 4) {
        // gather: sub-block q of this step is K-block kb_idx[p + q] (zeros past the list end)
        {
            const int pa = p + aq;
            uint4 v0 = make_uint4(0, 0, 0, 0), v1 = v0;
            if (pa < p1) {
                const __bf16* src = A + (int64_t)(m0 + ar) * lda + kbi(pa) * GEMM_BK;
                v0 = *reinterpret_cast<const uint4*>(src);
                v1 = *reinterpret_cast<const uint4*>(src + 8);
            }
            __syncthreads();
            *reinterpret_cast<uint4*>(&As[ar][aq * 16]) = v0;
            *reinterpret_cast<uint4*>(&As[ar][aq * 16 + 8]) = v1;
            const int pb = p + bq;
            __bf16 tmp[16];
            if (pb < p1) {
                const __bf16* src = Bm + (int64_t)(kbi(pb) * GEMM_BK + brow) * ldb + n0 + bc;
                const uint4 u0 = *reinterpret_cast<const uint4*>(src);
                const uint4 u1 = *reinterpret_cast<const uint4*>(src + 8);
                __builtin_memcpy(tmp, &u0, 16);
                __builtin_memcpy(tmp + 8, &u1, 16);
            } else {
#pragma unroll
                for (int e = 0; e < 16; ++e) tmp[e] = (__bf16)0.f;
            }
#pragma unroll
            for (int e = 0; e < 16; ++e) Bs[bc + e][bq * 16 + brow] = tmp[e];
            __syncthreads();
        }
#pragma unroll
        for (int kk = 0; kk < BK16 / 16; ++kk) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[wm * 32 + (lane & 31)][kk * 16 + 8 * (lane >> 5)]);
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[wn * 32 + (lane & 31)][kk * 16 + 8 * (lane >> 5)]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        if (jb.Nsq) {
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const float v = (float)Bs[ncol][npart * 16 + e];
                nsq += v * v;
            }
        }
    }
    store_tile(jb.C, jb.ldc, m0, n0, acc, lane, wm, wn);
    if (jb.Nsq) {
        nsq += __shfl_xor(nsq, 1);
        nsq += __shfl_xor(nsq, 2);
        if (npart == 0) {
            jb.Nsq[(int64_t)mt * jb.ldc + n0 + ncol] = nsq;
            s_nq[ncol] = nsq;
        }
    }
    if (jb.act) {
        // Does any score of the tile reach lo2 within the BF16 error bound?  In
        // the sqrt(Cf) scale a BF16 score is within BF16_ERR * N of its FP32 value
        // (N: the column's operand norm over the tile), so the tile is marked when
        // r * sqrt(1/nm) + BF16_ERR * N >= sqrt(lo2) for some score r (a 2^-10
        // margin covers the rounding of these FP32 terms).
        if (tid == 0) s_act = 0;
        __syncthreads();
        const float N = sqrtf(s_nq[wn * 32 + (lane & 31)]);
        const float need = sqrtf(jb.lo2) * (1.f - 0x1p-10f) - BF16_ERR * N;
        bool hit = need <= 0.f;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = (r & 3) + 8 * (r >> 2) + 4 * (lane >> 5);
            const float rn = jb.rnm[m0 + wm * 32 + i];            // 0 on padding rows
            hit = hit || (acc[r] > 0.f && acc[r] * sqrtf(rn) >= need);
        }
        if (__ballot(hit) != 0ull && lane == 0) s_act = 1;
        __syncthreads();
        if (tid == 0) {
            jb.act[nt * jb.mtiles + mt] = (uint8_t)s_act;
            if (s_act && jb.n_rescored) atomicAdd(&jb.n_rescored[0], (unsigned long long)(GEMM_BM * GEMM_BN));
        }
    }
    // the BF16 scores the peel's fallback starts from: needed only where the FP32
    // projection does not overwrite them
    if (jb.C2 && !(jb.act && s_act)) store_tile(jb.C2, jb.ldc, m0, n0, acc, lane, wm, wn);
}

// ---------------------------------------------------------------------------
// Split-BF16 projection: every FP32 operand x is carried as a BF16 pair
// x = hi + lo + e (hi = bf16(x), lo = bf16(x - hi), |e| <= 2^-16 |x|), and each
// product as the three BF16 MFMA products lo_A*hi_B + hi_A*lo_B + hi_A*hi_B
// accumulated in FP32 (the dropped lo_A*lo_B is <= 2^-16 |A||B|), so a score is
// within 3 * 2^-16 * sum |U||Bpc| of the exact dot product plus the FP32
// accumulation error: the FP32 projection's accuracy at a sixth of its MFMA time.
// U's pair is split once at load (Ah, Al); Bpc is read as FP32 and split while
// it is staged.  A step stages four listed 16-deep K-blocks (64 of the
// reduction) in LDS as [row][k] for both operands, so each MFMA operand is one
// 16-byte LDS read; the next step's global loads are in flight while the
// current step's 48 MFMAs run.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_project_x3(ProjectBatch bt) {
    __shared__ __attribute__((aligned(16))) __bf16 Ash[GEMM_BM][BK16 + BPAD];
    __shared__ __attribute__((aligned(16))) __bf16 Asl[GEMM_BM][BK16 + BPAD];
    __shared__ __attribute__((aligned(16))) __bf16 Bsh[GEMM_BN][BK16 + BPAD];
    __shared__ __attribute__((aligned(16))) __bf16 Bsl[GEMM_BN][BK16 + BPAD];
    __shared__ int s_kb[PROJ_KB_LDS];
    int j, mt, nt;
    if (!gemm_tile(bt, j, mt, nt)) return;
    const ProjectJob& jb = bt.job[j];
    // the tile's flags and list bounds are loaded together (one memory latency)
    const bool gated = jb.gate && !jb.gate[nt * jb.gate_mt + mt];   // tile not proposed: not computed
    const bool idle = jb.active && !*jb.active;                     // stream idle this launch
    const int p0 = jb.kb_ptr[mt], p1 = jb.kb_ptr[mt + 1];
    if (gated || idle) return;
    const __bf16* __restrict__ Ah = jb.Ab;
    const __bf16* __restrict__ Al = jb.Al;
    const float* __restrict__ Bm = jb.Bm;
    const int lda = jb.lda, ldb = jb.ldb;
    const int m0 = mt * GEMM_BM, n0 = nt * GEMM_BN;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv >> 1, wn = wv & 1;
    const int ar = tid >> 2, aq = tid & 3;               // A: row, listed block of the step
    const int bn = tid & 63, bg = tid >> 6;              // B: column bn, step rows 8 bg .. and 32 + 8 bg ..
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    const bool kb_lds = stage_kblocks(jb.kb_idx, p0, p1, s_kb, tid);
    auto kbi = [&](int p) { return kb_lds ? s_kb[p - p0] : jb.kb_idx[p]; };
    // this thread's share of a step, in registers (no arrays: they would go to scratch)
    uint4 ah0, ah1, al0, al1;       // A row segment of block aq: hi, then lo (two 16-byte words each)
    float4 b0, b1, b2, b3;          // B column bn: step rows 8 bg .. 8 bg + 7 (b0, b1), 32 + 8 bg .. (b2, b3)
    auto gload = [&](int p) __attribute__((always_inline)) {
        const uint4 z = make_uint4(0u, 0u, 0u, 0u);
        ah0 = ah1 = al0 = al1 = z;
        if (p + aq < p1) {
            const int64_t ao = (int64_t)(m0 + ar) * lda + kbi(p + aq) * GEMM_BK;
            const uint4* sh = reinterpret_cast<const uint4*>(Ah + ao);
            const uint4* sl = reinterpret_cast<const uint4*>(Al + ao);
            ah0 = sh[0]; ah1 = sh[1]; al0 = sl[0]; al1 = sl[1];
        }
        // step rows 8 g .. 8 g + 7 are rows 8 (g & 1) .. of listed block g >> 1; a wave
        // reads 64 consecutive columns of one row per load
        const float4 zf = make_float4(0.f, 0.f, 0.f, 0.f);
        b0 = b1 = b2 = b3 = zf;
        const int64_t l = ldb;
        if (p + (bg >> 1) < p1) {
            const float* src = Bm + (int64_t)(kbi(p + (bg >> 1)) * GEMM_BK + 8 * (bg & 1)) * l + n0 + bn;
            b0 = make_float4(src[0], src[l], src[2 * l], src[3 * l]);
            b1 = make_float4(src[4 * l], src[5 * l], src[6 * l], src[7 * l]);
        }
        if (p + 2 + (bg >> 1) < p1) {
            const float* src = Bm + (int64_t)(kbi(p + 2 + (bg >> 1)) * GEMM_BK + 8 * (bg & 1)) * l + n0 + bn;
            b2 = make_float4(src[0], src[l], src[2 * l], src[3 * l]);
            b3 = make_float4(src[4 * l], src[5 * l], src[6 * l], src[7 * l]);
        }
    };
    // eight consecutive k of column bn as hi and lo BF16 parts, one 16-byte store each
    // (the 64 columns of a wave's stores fall on distinct banks in groups of eight)
    auto split8 = [&](float4 u, float4 v, int k0) __attribute__((always_inline)) {
        const float x[8] = {u.x, u.y, u.z, u.w, v.x, v.y, v.z, v.w};
        __bf16 hi[8], lo[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            hi[e] = (__bf16)x[e];
            lo[e] = (__bf16)(x[e] - (float)hi[e]);
        }
        uint4 uh, ul;
        __builtin_memcpy(&uh, hi, 16);
        __builtin_memcpy(&ul, lo, 16);
        *reinterpret_cast<uint4*>(&Bsh[bn][k0]) = uh;
        *reinterpret_cast<uint4*>(&Bsl[bn][k0]) = ul;
    };
    if (p0 < p1) gload(p0);
    for (int p = p0; p < p1; p += 4) {
        __syncthreads();
        *reinterpret_cast<uint4*>(&Ash[ar][aq * 16]) = ah0;
        *reinterpret_cast<uint4*>(&Ash[ar][aq * 16 + 8]) = ah1;
        *reinterpret_cast<uint4*>(&Asl[ar][aq * 16]) = al0;
        *reinterpret_cast<uint4*>(&Asl[ar][aq * 16 + 8]) = al1;
        split8(b0, b1, 8 * bg);
        split8(b2, b3, 32 + 8 * bg);
        __syncthreads();
        if (p + 4 < p1) gload(p + 4);
        const int nk = min(4, p1 - p);
        for (int kk = 0; kk < nk; ++kk) {
            const int ko = kk * 16 + 8 * (lane >> 5);
            const bf16x8 ah = *reinterpret_cast<const bf16x8*>(&Ash[wm * 32 + (lane & 31)][ko]);
            const bf16x8 al = *reinterpret_cast<const bf16x8*>(&Asl[wm * 32 + (lane & 31)][ko]);
            const bf16x8 bh = *reinterpret_cast<const bf16x8*>(&Bsh[wn * 32 + (lane & 31)][ko]);
            const bf16x8 bl = *reinterpret_cast<const bf16x8*>(&Bsl[wn * 32 + (lane & 31)][ko]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
        }
    }
    store_tile(jb.C, jb.ldc, m0, n0, acc, lane, wm, wn);
}

// nm[t] = sum_{k,c} U[t,k,c]^2 accumulated in FP64, stored as FP64 and rounded FP32;
// rnm = FP32 reciprocal (the LSS matcher multiplies by 1/nm, reciprocal_kernel).
__global__ __launch_bounds__(256) void k_norms(const float* __restrict__ U, int KC,
                                               float* __restrict__ nm, float* __restrict__ rnm,
                                               double* __restrict__ nm64) {
    __shared__ double red[256];
    const int t = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < KC; i += 256) {
        const double u = U[(int64_t)t * KC + i];
        s += u * u;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        nm64[t] = red[0];
        nm[t] = (float)red[0];
        rnm[t] = 1.0f / (float)red[0];
    }
}

// The next epoch of a stream's published column maxima; on wrap-around every
// word is cleared (tag 0 matches no epoch's tags, which are never 0).
__device__ void advance_epoch(unsigned* epoch, unsigned long long* ptag, int64_t words) {
    __shared__ int wrap;
    if (threadIdx.x == 0) {
        const unsigned e = (*epoch + 1u) & EPOCH_MASK;
        *epoch = e;
        wrap = e == 0u;
    }
    __syncthreads();
    if (wrap)
        for (int64_t i = threadIdx.x; i < words; i += blockDim.x) ptag[i] = 0ull;
}

// The per-batch words the peel relies on, for stream j of a launch: iteration
// counts and neighbour flags, phase clock, barrier words, error flags cleared,
// the tag epoch advanced.  Run by one whole workgroup of the whitening launch
// (every peel of the batch comes after it on the stream); a stream idle in this
// launch keeps its words.
__device__ void reset_stream(const ResetBatch& b, int j) {
    if (b.active[j] && !*b.active[j]) return;
    for (int i = threadIdx.x; i < b.n_iter[j]; i += blockDim.x) b.iter_cnt[j][i] = 0;
    if ((int)threadIdx.x < PROF_WORDS) b.prof[j][threadIdx.x] = 0ull;
    if ((int)threadIdx.x < 1 + 2 * BAR_MAX_GROUPS) b.bar[j][threadIdx.x * BAR_LINE] = 0u;
    if (threadIdx.x == 0) *b.err[j] = 0;
    if (b.epoch[j]) advance_epoch(b.epoch[j], b.ptag[j], b.ptag_words[j]);
}

// ---------------------------------------------------------------------------
// Cross-workgroup communication of the persistent peel kernel.
//
// Every value one workgroup hands to another (column maxima, accepted spikes,
// counters) is written and read with agent-scope relaxed atomics, which on
// CDNA are plain vector loads/stores with the coherence bit set: they reach
// and read the device-coherent level directly, so no workgroup needs the
// whole-L2 write-back / invalidate that an agent-scope fence costs on a
// multi-XCD part (and the owner-private score tiles stay resident in L2).
// A barrier then only has to order those accesses: each wave waits for its
// outstanding memory operations, the workgroup meets, one lane arrives on a
// device counter and spins (bounded) until all workgroups have arrived.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T ld_c(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_c(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_spike(Spike* d, const Spike& s) {
    int* q = reinterpret_cast<int*>(d);
    st_c(q + 0, s.w);
    st_c(q + 1, s.t);
    st_c(q + 2, __float_as_int(s.amp));
    st_c(q + 3, s.it);
}
__device__ __forceinline__ Spike ld_spike(const Spike* d) {
    const int* q = reinterpret_cast<const int*>(d);
    Spike s;
    s.w = ld_c(q + 0);
    s.t = ld_c(q + 1);
    s.amp = __int_as_float(ld_c(q + 2));
    s.it = ld_c(q + 3);
    return s;
}

// Barrier number `gen` (1, 2, ...) of a grid of G workgroups; g is this one.
__device__ __forceinline__ void grid_sync(unsigned* bar, int G, int g, unsigned& gen, int* err) {
    __builtin_amdgcn_s_waitcnt(0);       // this wave's coherent stores have completed
    __syncthreads();
    if (threadIdx.x == 0) {
        gen += 1;
        const unsigned* wait_on = bar;
        unsigned target = gen * (unsigned)G;
        if (G >= BAR_HIER_MIN) {
            // arrive on the group counter; the group's last arriver arrives on the
            // top counter, and the grid's last arriver releases every group
            const int ng = (G + BAR_GROUP - 1) / BAR_GROUP, grp = g / BAR_GROUP;
            const unsigned gsize = (unsigned)min(BAR_GROUP, G - grp * BAR_GROUP);
            const unsigned old = __hip_atomic_fetch_add(bar + BAR_LINE * (1 + grp), 1u, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
            if (old == gen * gsize - 1u) {
                const unsigned o2 = __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (o2 == gen * (unsigned)ng - 1u)
                    for (int k = 0; k < ng; ++k) st_c(bar + BAR_LINE * (1 + BAR_MAX_GROUPS + k), gen);
            }
            wait_on = bar + BAR_LINE * (1 + BAR_MAX_GROUPS + grp);
            target = gen;
        } else {
            __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        unsigned spins = 0;
        while (ld_c(wait_on) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > (1u << 26)) {
                __hip_atomic_fetch_or(err, ERR_BARRIER, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
    }
    __syncthreads();
}

// Local synchronisation of the peel loop.  A workgroup publishes that it is done
// with a phase by raising its flag to the iteration number + 1 after its coherent
// stores have completed; a workgroup that depends on a contiguous range of flags
// waits (one thread per flag) until all of them reach the target.  A workgroup
// leaving the loop raises its flags to FLAG_DONE so no neighbour waits on it.
constexpr int FLAG_DONE = 0x7fffffff;

__device__ __forceinline__ void raise_flag(int* f, int v) {
    __builtin_amdgcn_s_waitcnt(0);       // this wave's coherent stores have completed
    __syncthreads();                     // ... and every wave's
    if (threadIdx.x == 0) st_c(f, v);
}

__device__ __forceinline__ void spin_flag(const int* f, int target, int* err) {
    unsigned spins = 0;
    while (ld_c(f) < target) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > (1u << 26)) {
            __hip_atomic_fetch_or(err, ERR_BARRIER, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
    }
}

__device__ __forceinline__ void wait_flags(const int* f, int n, int target, int* err) {
    if ((int)threadIdx.x < n) spin_flag(&f[threadIdx.x], target, err);
    __syncthreads();
}

__device__ __forceinline__ bool better(float v, int o, float bv, int bo) {
    return v > bv || (v == bv && o < bo);
}

// ---------------------------------------------------------------------------
// Persistent peel kernel: workgroup g owns CW consecutive columns of the score
// matrix B and one subset of its templates (rows); only it reads or writes
// them.  With R > 0 the owned tile lives in registers for the whole batch
// (thread (col, rg) holds R consecutive rows); R == 0 keeps it in global
// memory (large template counts).  Per iteration:
//   A  apply the previous iteration's spikes that reach the owned columns
//      (B[t][w] -= a * ctc[s][t][w - w_s + M], spikes in sample order), then
//      Cf = relu(B)^2 / nm (zero in the first and last M columns) and the
//      column max / first-index argmax over the owned templates; publish them
//      and raise the workgroup's flag.
//   B  (Kilosort rules) after the flags of the workgroups within 2M columns:
//      the column maxima over all templates there; Cmax = their max over
//      [w-M, w+M]; accept w when Cmax > Th^2 and Cmax == Cf[w] (|diff| <
//      1e-9), amplitude B[t][w] / nm[t].  Every workgroup does this for the
//      columns within M of its own, so it has the next iteration's
//      subtractions without a further exchange.
//      (LSS rules, one subset) the block accepts its own columns from its
//      neighbours' maxima and publishes the spikes; the neighbours read them
//      after its second flag.
// No grid-wide barrier inside the loop: the waits are on neighbours' flags,
// and the stop test reads a per-iteration count that lags behind.  One owner
// per (row, column) and a fixed application order make the result
// deterministic; no data atomics are used.  Afterwards each workgroup sorts
// its spikes by (column, iteration), writes the events, and computes the PC
// features of the final residual.
// ---------------------------------------------------------------------------
constexpr int PRE = 4;   // neighbour spikes fetched speculatively with the counts
constexpr unsigned PEEL_XCDS = 8;   // XCDs of an MI355X (32 CUs and one L2 each)

__device__ __forceinline__ uint64_t rt_now() { return __builtin_amdgcn_s_memrealtime(); }

// RS: the FP32 rescoring of a BF16 projection is compiled in (a separate
// instance, so the FP32 matcher's register allocation does not carry it).
template <int R, bool RS, bool CK, bool LS>
__global__ __launch_bounds__(PEEL_THREADS) void k_peel(PeelBatch bt) {
    // Every product here is rounded before it is added or compared, as in the
    // FP32 reference; a fused multiply-add would move tie decisions by an ulp.
#pragma clang fp contract(off)
    // Workgroups are dispatched to the 8 XCDs round-robin (from some starting XCD),
    // and each XCD places its share on its own 32 CUs.  With bt.xcd, and every stream
    // of the grid active in this launch, XCD x runs the x-th eighth of the logical
    // workgroups (streams, then column blocks), so neighbouring blocks poll each other
    // and read the same spikes' ctc rows through one L2: with 8 streams, one stream
    // per XCD.  With every stream active no other launch of the group has an active
    // stream (a stream has one batch in flight), so no XCD holds more than
    // ceil(G / 8) <= 32 waiting workgroups.  A launch with idle streams keeps the
    // workgroup order, which spreads each stream over every XCD: launches on other
    // lanes may run peels of their own beside it, and two streams placed on one XCD
    // would not fit.  (A single stream's peel was slower on contiguous XCD shares:
    // 264 against 245 us per 512-template, 3122-column batch.)
    bool all_on = bt.xcd != 0;
    for (int j = 0; all_on && j < bt.n; ++j) all_on = !bt.p[j].active || *bt.p[j].active;
    const unsigned gq = (unsigned)bt.g_start[bt.n] / PEEL_XCDS, gr = (unsigned)bt.g_start[bt.n] % PEEL_XCDS;
    const unsigned xcd = blockIdx.x % PEEL_XCDS;
    const int gl = all_on ? (int)(xcd * gq + min(xcd, gr) + blockIdx.x / PEEL_XCDS) : (int)blockIdx.x;
    const int sj = job_of(bt.g_start, bt.n, gl);
    const PeelParams& p = bt.p[sj];
    if (p.active && !*p.active) return;                         // stream idle: every workgroup leaves
    // LDS is kept under 64 KB so every table is reachable with ds_read's 16-bit
    // immediate offset (no per-row address registers); the norm table comes first.
    constexpr int CWX = R > 0 ? PEEL_THREADS / REG_NRG : MAX_CW;   // columns a workgroup owns at most
    constexpr int NWIN = CWX + 4 * MAX_M + 2;    // columns within 2M of the block
    constexpr int NACC = CWX + 2 * MAX_M + 2;    // columns within M of the block
    __shared__ float s_nm[R > 0 ? REG_MAX_T : 1];    // own rows' reciprocal norms (0 past T)
    __shared__ float s_nmf[R > 0 ? REG_MAX_T : 1];   // every row's norm (+inf past T)
    __shared__ int s_or[R > 0 ? REG_MAX_T : 1];      // every row's file index (INT_MAX past T)
    __shared__ int s_band[R > 0 ? REG_MAX_T : 1];    // every template's nonzero ctc rows: tlo | thi << 16
    __shared__ Spike s_sp[MAX_STAGE];          // staged / newly accepted spikes
    __shared__ Spike s_list[MAX_LIST];         // this workgroup's spikes, whole batch
    __shared__ Spike s_pre[MAX_NBR * PRE];
    __shared__ float s_val[PEEL_THREADS];      // also the rank table of the epilogue
    __shared__ int s_idx[PEEL_THREADS];
    __shared__ int s_org[PEEL_THREADS];
    __shared__ float s_bvv[PEEL_THREADS];
    // local acceptance: the best score of each column within 2M of the block and
    // the maxima of its aligned 8-column groups; the LSS-rule exchange keeps its
    // pooling window in the score array
    __shared__ float s_sc[NWIN];
    __shared__ float s_bm[NWIN / 8 + 1];
    float* const s_win = s_sc;
    __shared__ float s_cf[CWX];                // own column maxima, their argmax and its score
    __shared__ int s_im[CWX];
    __shared__ float s_bv[CWX];
    __shared__ float s_cm[CWX];                // LSS pooled maxima
    __shared__ int s_flag[CWX];
    // FP32 rescoring: each listed pair's correction, then (register tiles) the columns'
    // 64-bit selection keys
    __shared__ __align__(8) float s_bcor[RS ? MAX_BAND + (R > 0 ? 2 * (PEEL_THREADS / REG_NRG) : 0) : 1];
    __shared__ float s_nmx[RS && R > 0 ? PEEL_THREADS / REG_NRG : 1];   //   and each column's largest operand norm
    __shared__ int s_hcnt[MAX_NBR + 1];
    __shared__ __align__(16) int s_hcnt_w[PEEL_THREADS / 64];     // per-wave counts (read as int4)
    __shared__ __align__(16) int s_hcnt_o[PEEL_THREADS / 64];
    // staged spikes' row bands (R == 0; the FP32 rescoring's pair list after phase A)
    __shared__ int s_qlo[R == 0 ? MAX_STAGE : 1], s_qhi[R == 0 ? MAX_STAGE : 1];
    // per-wave list of band-hitting spikes (R > 0; the FP32 rescoring's pair list after phase A)
    __shared__ int s_whit[R > 0 ? PEEL_THREADS : 1];
    __shared__ int s_nsp, s_listn, s_off, s_n, s_tot, s_err, s_ncand, s_kchk, s_nband, s_nnan;
    __shared__ int s_wt[NACC];                 // local acceptance: winning template per column
    __shared__ float s_wb[NACC];               //   and its B value
    int* const s_rank = reinterpret_cast<int*>(s_val);
    static_assert(NWIN >= CWX + 2 * MAX_M + 2, "the LSS pooling window fits the score array");
    static_assert(MAX_LIST <= PEEL_THREADS, "rank table reuses the reduction buffer");
    static_assert(2 * MAX_BAND <= PEEL_THREADS && MAX_BAND <= MAX_STAGE, "rescoring pairs fit the reused arrays");

    const int g = gl - bt.g_start[sj], tid = threadIdx.x;
    // Workgroup g owns column block cb = g / S and template subset sb = g % S:
    // rows [sb * REG_NRG * R, (sb + 1) * REG_NRG * R) of the block's columns.
    // Subset 0 is the block's owner: it accepts spikes, publishes them and
    // writes the events; the other subsets publish per-column partial maxima.
    const int S = R == 64 ? 1 : p.S;        // a 64-row tile holds every template (T <= 1024): one subset
    const int cb = g / S, sb = g - cb * S;
    const int rbase = R > 0 ? sb * REG_NRG * R : 0;
    const bool owner = sb == 0;
    // register tiles use a fixed 64-column width, so column and row strides are
    // compile-time constants (LDS offsets fold into the instructions)
    const int CW = R > 0 ? PEEL_THREADS / REG_NRG : p.CW;
    const int M = p.M, T = p.T, W = p.W;
    const int c0 = cb * CW, c1 = min(W, c0 + CW);
    const int nrg = PEEL_THREADS / CW;
    const int col = tid % CW, rg = tid / CW;
    const int hlo = max(0, (c0 - M) / CW), hhi = min(p.NB - 1, (c1 - 1 + M) / CW);
    const int nh = hhi - hlo + 1;             // <= MAX_NBR because CW >= M
    const int w = c0 + col;
    const bool own = w < c1;
    const bool edge = (w < M) || (w >= W - M);
    // Kilosort rules: every workgroup decides acceptance itself for the columns
    // within M of its block, from the published maxima of the columns within 2M
    // (blocks hlo2..hhi2).  LSS rules (one subset) exchange accepted spikes instead.
    const bool local = !LS;                 // LS: the instance for LSS rules
    unsigned long long* const Pa = p.Pa;                   // scores
    unsigned long long* const Pb = p.Pa + (int64_t)2 * S * W;   // B values
    unsigned long long* const Pt = p.Pa + (int64_t)4 * S * W;   // templates
    unsigned bar_gen = 0;
    if (tid == 0) s_listn = 0;

    // Buffer descriptors: one scalar base per table plus a per-thread offset, so
    // the R rows of a tile need no address registers (row strides go in SGPRs).
    const __amdgpu_buffer_rsrc_t ctc_rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p.ctc), 0, p.ctc_bytes, BUF_DWORD3);
    float breg[R > 0 ? R : 1];
    // the BF16 operand norm of this thread's tile at its column, and (column threads)
    // the largest over the tiles: fixed for the batch
    float nown = 0.f;
    // BF16 mode: this wave's 64-row tile of the column block was recomputed in FP32
    const bool wexact = RS && R > 0 && p.rescore && rbase + rg * (R > 0 ? R : 1) < T &&
                        p.act[cb * p.act_mt + (rbase + rg * (R > 0 ? R : 1)) / GEMM_BM] != 0;
    if (RS && R > 0 && p.rescore) {
        if (w < W && rbase + rg * (R > 0 ? R : 1) < T) nown = sqrtf(p.Nsq[(int64_t)((rbase + rg * R) / GEMM_BM) * p.BS + w]);
        if (tid < CW) {
            float mx = 0.f;
            for (int m = 0; m * GEMM_BM < T && c0 + tid < W; ++m) mx = fmaxf(mx, p.Nsq[(int64_t)m * p.BS + c0 + tid]);
            s_nmx[R > 0 ? tid : 0] = sqrtf(mx);
        }
    }
    if (R > 0) {
        for (int t = tid; t < REG_NRG * R; t += PEEL_THREADS) s_nm[t] = p.rnm[rbase + t];
        for (int t = tid; t < min(p.ctc_rows, REG_MAX_T); t += PEEL_THREADS) {
            s_nmf[t] = p.nm[t];
            s_or[t] = p.orig[t];
            s_band[t] = t < T ? (p.tlo[t] | (p.thi[t] << 16)) : 0;
        }
        // rows past the score matrix read as zero (buffer loads beyond the range)
        const __amdgpu_buffer_rsrc_t b_rs = __builtin_amdgcn_make_buffer_rsrc(p.B, 0, p.B_bytes, BUF_DWORD3);
        // wave rg holds the contiguous row block rbase + [rg*R, rg*R + R) of its 64 columns
        const int voff = ((rbase + rg * R) * p.BS + w) * 4;
        const int rstride = p.BS * 4;
#pragma unroll
        for (int i = 0; i < R; ++i)
            breg[i] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(b_rs, voff, i * rstride, 0));
        __syncthreads();
    }

    // Publish own column cl's maximum.  LSS rules: the column maximum itself.
    // Kilosort rules: this subset's (score, B value, template) of the column, each
    // a 64-bit word carrying the tag, into the buffer of the iteration's parity.
    // A reader of iteration i polls the words until they carry i's tag (or the
    // final tag).  The writer overwrites them only in iteration i + 2, after it
    // has read the iteration i + 1 words of every workgroup within 2M columns
    // (its readers), which they publish after reading iteration i's.
    auto publish = [&](int cl, int buf, unsigned tag) {
        const int ww = c0 + cl;
        if (!local) {
            st_c(&p.Cfmax[ww], s_cf[cl]);
        } else {
            const int64_t k = ((int64_t)buf * S + sb) * W + ww;
            st_c(&Pa[k], ((unsigned long long)__float_as_uint(s_cf[cl]) << 32) | tag);
            st_c(&Pb[k], ((unsigned long long)__float_as_uint(s_bv[cl]) << 32) | tag);
            st_c(&Pt[k], ((unsigned long long)(unsigned)s_im[cl] << 32) | tag);
        }
    };

    __shared__ unsigned s_ep;      // this batch's epoch, shifted into tag position
    __shared__ int s_abort;        // a gather timed out
    if (tid == 0) {
        s_kchk = 0;
        s_abort = 0;
        s_ep = local ? ld_c(p.epoch) << TAG_ITER_BITS : 0u;
    }
    // per iteration: owners done with phase B (high word) and spikes accepted (low word)
    unsigned long long* const itc = reinterpret_cast<unsigned long long*>(p.iter_cnt);
    int* const flagB = p.iter_cnt + 2 * p.max_iter;
    int* const flagA = flagB + p.NB;
    // phase clock: thread 0 accumulates in LDS (no registers held across the loop)
    const bool timing = CK && tid == 0;     // profiling instance only (the clock cost ~5% of the peel)
    __shared__ uint64_t s_ph[NPHASE + 2];      // phases, then the previous mark and the pass count
    if (tid < NPHASE + 2) s_ph[tid] = 0;
    __syncthreads();
    const unsigned ep = s_ep;
    if (timing) s_ph[NPHASE] = rt_now();
    auto mark = [&](int k) {
        if (timing) {
            const uint64_t t = rt_now();
            s_ph[k] += t - s_ph[NPHASE];
            s_ph[NPHASE] = t;
        }
    };
    // Stop once some iteration is known to have accepted no spike anywhere: every
    // owner is done with it (high word of its packed count) and its total (low
    // word) is zero.  Every later iteration then changes nothing, so when a
    // workgroup notices does not change any result.  Run by one thread.
    auto stop_test = [&](int it) {
        int stop = 0;
        while (s_kchk < it) {
            const unsigned long long v = ld_c(&itc[s_kchk]);
            if ((int)(v >> 32) < p.NB) break;
            if ((unsigned)v == 0u) { stop = 1; break; }
            ++s_kchk;
        }
        s_tot = !stop;
        s_err = ld_c(p.err);
    };
    // a wave whose 64-template tile the gate did not propose for this column block
    // scores zero everywhere (its stale rows are neither updated nor read); the mask
    // is fixed for the batch, so it is read once here, not in every iteration
    const bool woff = R > 0 && p.gate && !p.gate[cb * p.gate_mt + (rbase + rg * R) / GEMM_BM];
    for (int it = 0; it <= p.max_iter; ++it) {
        int nsp = 0;
        if (local && it == p.max_iter) break;
        if (it > 0) {
            if (!local) {
                // one coherent round trip: the stop test, the neighbours' counts, and
                // their first PRE spikes
                if (tid == 0) stop_test(it);
                if (tid >= 64 && tid < 64 + nh) s_hcnt[tid - 64] = min(ld_c(&p.slot_cnt[hlo + tid - 64]), p.slot_cap);
                if (tid >= 128 && tid < 128 + nh * PRE) {
                    const int hh = (tid - 128) / PRE, k = (tid - 128) % PRE;
                    s_pre[hh * PRE + k] = ld_spike(&p.slots[(int64_t)(hlo + hh) * p.slot_cap + k]);
                }
                __syncthreads();
                if (s_tot == 0 || (s_err & ERR_BARRIER)) break;
                if (it == p.max_iter) break;
                if (tid == 0) {
                    int n = 0;
                    for (int hh = 0; hh < nh; ++hh) {
                        const int cnt = s_hcnt[hh];
                        for (int k = 0; k < cnt; ++k) {
                            const Spike sp = k < PRE ? s_pre[hh * PRE + k]
                                                     : ld_spike(&p.slots[(int64_t)(hlo + hh) * p.slot_cap + k]);
                            if (sp.w + M >= c0 && sp.w - M < c1) {
                                if (n < MAX_STAGE) s_sp[n++] = sp;
                                else atomicOr(p.err, ERR_STAGE);
                            }
                        }
                    }
                    s_nsp = n;
                }
                __syncthreads();
            }
            // local acceptance has left the spikes within M of this block in s_sp
            nsp = s_nsp;
            if (R == 0) {
                // the global tile tests every (row, spike) pair: the spikes' nonzero
                // ctc row bands go to LDS once
                for (int q = tid; q < nsp; q += PEEL_THREADS) {
                    const int st = s_sp[q].t;
                    s_qlo[q] = p.tlo[st];
                    s_qhi[q] = p.thi[st];
                }
                __syncthreads();
            }
        }
        mark(0);

        // ---------------- phase A ----------------
        if (it == 0 || nsp > 0) {
            if (timing) s_ph[NPHASE + 1] += 1;
            float best = -INFINITY, bval = 0.f;
            int bi = 0, bo = 0x7fffffff;
            // A wave whose rows no spike touched keeps last pass's per-thread
            // maxima in LDS (its registers did not change): it skips the scan.
            bool wdirty = it == 0;
            if (R > 0) {
                // Runs with the whole wave active (every lane tests spikes for the ballot);
                // edge columns and columns past W are updated too but never read.
                // The tables are padded to the tile height (zero ctc rows, +inf norms,
                // INT_MAX file indices), so no row needs a bounds check.  Templates are
                // ordered by main channel, so a spike's nonzero ctc rows form a band
                // [tlo, thi): a wave whose row block misses it skips the spike (a
                // wave-uniform branch).  Otherwise the R loads of the spike are all in
                // flight before any is consumed.  A spike whose window misses this
                // column contributes with amplitude 0 (b - 0*c is b), so lanes never
                // diverge.  Spikes are applied in sample order.
                const int r0 = rbase + rg * R;
                const int lane = tid & 63;
                int* whit = s_whit + rg * 64;
                for (int q0 = 0; q0 < nsp; q0 += 64) {
                    // this wave's spikes whose band reaches its row block, compacted with a ballot
                    const int qq = q0 + lane;
                    bool bh = false;
                    if (qq < nsp && !woff) {
                        const int bd = s_band[s_sp[qq].t];
                        bh = r0 < (bd >> 16) && r0 + R > (bd & 0xffff);
                    }
                    const uint64_t bal = __ballot(bh);
                    if (bh) whit[__popcll(bal & ((1ull << lane) - 1ull))] = qq;
                    const int nhit = __popcll(bal);
                    wdirty = wdirty || nhit > 0;
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_s_waitcnt(0xc07f);     // lgkmcnt(0): the list is in LDS
                    // A load batch holds SPB spikes x RC rows (32 values): a wave usually
                    // has one or two spikes in its band, so taller tiles take more rows of
                    // fewer spikes per batch (each batch costs one memory latency).
                    constexpr int SPB = R <= 8 ? 4 : (R == 16 ? 2 : 1);    // spikes per load batch
                    constexpr int RC = R < 32 ? R : 32;                     // rows per load batch
                    for (int h0 = 0; h0 < nhit; h0 += SPB) {
                        float am[SPB];
                        int vo[SPB];
                        int blo = 1 << 30, bhi = 0;          // rows of the batch's nonzero ctc bands
    #pragma unroll
                        for (int j = 0; j < SPB; ++j) {
                            const bool in = h0 + j < nhit;
                            const Spike sp = s_sp[whit[in ? h0 + j : h0]];
                            const int l = w - sp.w + M;
                            const bool hit = in && l >= 0 && l <= 2 * M;
                            am[j] = hit ? sp.amp : 0.f;
                            vo[j] = ((sp.t * p.ctc_rows + r0) * LP + (hit ? l : 0)) * 4;
                            const int bd = s_band[sp.t];
                            blo = min(blo, bd & 0xffff);
                            bhi = max(bhi, bd >> 16);
                        }
    #pragma unroll
                        for (int i0 = 0; i0 < R; i0 += RC) {
                            // a load batch wholly outside the spikes' nonzero ctc rows changes
                            // nothing (b - a * 0 is b): skipped (wave-uniform)
                            if (RC < R && (r0 + i0 >= bhi || r0 + i0 + RC <= blo)) continue;
                            float c[SPB][RC];
    #pragma unroll
                            for (int j = 0; j < SPB; ++j)
    #pragma unroll
                                for (int i = 0; i < RC; ++i)
                                    c[j][i] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                                        ctc_rs, vo[j], (i0 + i) * LP * 4, 0));
    #pragma unroll
                            for (int i = 0; i < RC; ++i)
    #pragma unroll
                                for (int j = 0; j < SPB; ++j) breg[i0 + i] -= am[j] * c[j][i];
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                }
                mark(6);
            }
            // Global tile: the staged spikes are in sample order, so the ones whose
            // window covers this column form one range [qa, qb).  A column no new
            // spike reaches keeps last pass's maxima (its rows did not change); the
            // others apply only that range, in the same order.
            int qa = 0, qb = 0;
            if (R == 0) {
                int lo = 0, hi = nsp;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (s_sp[mid].w < w - M) lo = mid + 1; else hi = mid;
                }
                qa = lo;
                hi = nsp;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (s_sp[mid].w <= w + M) lo = mid + 1; else hi = mid;
                }
                qb = lo;
                if (qb > qa) wdirty = true;
            }
            if (!wdirty) {
                // nothing to rescan
            } else if (own && woff) {
                best = 0.f; bi = rbase + rg * R; bo = 0x7fffffff; bval = 0.f;
            } else if (own && edge) {
                // Cf is zero for every template: first file index wins the tie
                best = 0.f; bi = p.t_first; bo = 0;
            } else if (own) {
                if (R > 0) {
                    int vb = rg * R;
                    asm volatile("" : "+v"(vb));
                    const float* rnp = s_nm + vb;
                    const int* orp = s_or + rbase + vb;
                    const float* nmp = s_nmf + rbase + vb;
                    // Argmax in two steps.  a = r^2 * (1/nm) is within 3 * 2^-24 (relative)
                    // of Cf = r^2 / nm (Kilosort; LSS multiplies, so a is Cf exactly).  One
                    // branch-free pass keeps the largest a, its row, and the second largest;
                    // when the runner-up is within 2^-20 of the winner (or ties it, LSS) the
                    // rows in that band are decided on exact Cf, smaller file index first on
                    // ties, as torch.max.  Rows go in groups of 8 with a compiler barrier
                    // between groups so the table loads are not all hoisted at once.
                    float amax = -1.f, a2 = -1.f;
                    int bslot = 0;          // winning row slot; its row is vb + bslot
#pragma unroll
                    for (int i0 = 0; i0 < R; i0 += 8) {
#pragma unroll
                        for (int i = i0; i < i0 + 8; ++i) {
                            // selects rather than fmaxf: no NaN canonicalisation (scores are finite)
                            const float r = breg[i] > 0.f ? breg[i] : 0.f;
                            const float a = (r * r) * rnp[i];
                            const bool gt = a > amax;
                            a2 = gt ? amax : (a > a2 ? a : a2);
                            amax = gt ? a : amax;
                            bslot = gt ? i : bslot;
                            bval = gt ? breg[i] : bval;
                        }
                        asm volatile("" ::: "memory");
                    }
                    const float thr = LS ? amax : amax * (1.f - 0x1p-20f);
                    if (a2 < thr) {
                        const float r = fmaxf(bval, 0.f);
                        best = LS ? amax : (r * r) / nmp[bslot];
                        bo = orp[bslot];
                    } else {
                        best = -1.f;
                        bo = 0x7fffffff;
                        // recompute the squares here instead of keeping 64 of them live
#pragma unroll
                        for (int i = 0; i < R; ++i) asm volatile("" : "+v"(breg[i]));
#pragma unroll
                        for (int i0 = 0; i0 < R; i0 += 8) {
#pragma unroll
                            for (int i = i0; i < i0 + 8; ++i) {
                                const float r = fmaxf(breg[i], 0.f);
                                const float a = (r * r) * rnp[i];
                                const float cf = LS ? a : (r * r) / nmp[i];
                                const int o = orp[i];
                                const bool gt = a >= thr && (cf > best || (cf == best && o < bo));
                                best = gt ? cf : best;
                                bo = gt ? o : bo;
                                bslot = gt ? i : bslot;
                                bval = gt ? breg[i] : bval;
                            }
                            asm volatile("" ::: "memory");
                        }
                    }
                    bi = rbase + vb + bslot;
                } else {
                    // each covering spike, in order, updates this thread's rows of its band
                    // (every element sees its updates in spike order, as a row-outer loop
                    // would apply them), then one pass takes the column maximum
                    // Eight rows at a time, every load of a batch in flight before its
                    // stores (the rows of a batch are distinct; the B and ctc pointers
                    // would otherwise keep each row's loads behind the previous store).
                    auto gated = [&](int t) { return p.gate && !p.gate[(w / GEMM_BN) * p.gate_mt + t / GEMM_BM]; };
                    for (int q = qa; q < qb; ++q) {
                        const Spike sp = s_sp[q];
                        const int l = w - sp.w + M;                 // in [0, 2M] for q in [qa, qb)
                        const int t0 = s_qlo[q], t1 = s_qhi[q];
                        const float* cp = p.ctc + (int64_t)sp.t * p.ctc_rows * LP + l;
                        for (int ta = t0 + (rg - t0 % nrg + nrg) % nrg; ta < t1; ta += 8 * nrg) {
                            float bv[8], cv[8];
                            bool in[8];
#pragma unroll
                            for (int k = 0; k < 8; ++k) {
                                const int t = ta + k * nrg;
                                in[k] = t < t1 && !gated(t);
                                bv[k] = in[k] ? p.B[(int64_t)t * p.BS + w] : 0.f;
                                cv[k] = in[k] ? cp[(int64_t)t * LP] : 0.f;
                            }
#pragma unroll
                            for (int k = 0; k < 8; ++k)
                                if (in[k]) p.B[(int64_t)(ta + k * nrg) * p.BS + w] = bv[k] - sp.amp * cv[k];
                        }
                    }
                    for (int ta = rg; ta < T; ta += 8 * nrg) {
                        float bv[8], nv[8];
                        int ov[8];
                        bool in[8], on[8];
#pragma unroll
                        for (int k = 0; k < 8; ++k) {
                            const int t = ta + k * nrg;
                            in[k] = t < T;
                            on[k] = in[k] && !gated(t);
                            bv[k] = on[k] ? p.B[(int64_t)t * p.BS + w] : 0.f;
                            nv[k] = on[k] ? (LS ? p.rnm[t] : p.nm[t]) : 1.f;
                            ov[k] = in[k] ? p.orig[t] : 0;
                        }
#pragma unroll
                        for (int k = 0; k < 8; ++k) {
                            if (!in[k]) continue;
                            const int t = ta + k * nrg;
                            if (!on[k]) {
                                if (better(0.f, ov[k], best, bo)) { best = 0.f; bi = t; bo = ov[k]; bval = 0.f; }
                                continue;
                            }
                            float cf = 0.f;
                            if (!edge) {
                                const float r = fmaxf(bv[k], 0.f);
                                cf = LS ? (r * r) * nv[k] : (r * r) / nv[k];
                            }
                            if (better(cf, ov[k], best, bo)) { best = cf; bi = t; bo = ov[k]; bval = bv[k]; }
                        }
                    }
                }
            }
            mark(1);
            // BF16 mode, register tile: whether some thread of a BF16 tile has a best
            // score that could reach lo2 (its BF16 best plus the bound; the best bounds
            // every row's a = r^2 * (1/nm) from above to within 2^-22), taken with the
            // barrier below
            int need = 0;
            if (RS && R > 0 && p.rescore && own && !edge && !woff && !wexact && rbase + rg * (R > 0 ? R : 1) < T) {
                const float e = sqrtf(p.lo2) - BF16_ERR * nown;
                const float thr = e > 0.f ? e * e * (1.f - 0x1p-18f) : -1.f;
                need = (wdirty ? best : s_val[tid]) >= thr;
            }
            if (wdirty) { s_val[tid] = best; s_idx[tid] = bi; s_org[tid] = bo; s_bvv[tid] = bval; }
            const int any_need = (RS && R > 0) ? __syncthreads_or(need) : (__syncthreads(), 0);
            mark(7);
            if (tid < CW) {
                // per-thread base pointers: the row-group strides become immediate
                // offsets.  The empty asm keeps the base from being treated as loop
                // invariant, which would hoist one address register per row group.
                int vb = tid;
                asm volatile("" : "+v"(vb));
                const float* pv = s_val + vb;
                const int* pi = s_idx + vb;
                const int* po = s_org + vb;
                const float* pb = s_bvv + vb;
                float bv, bb;
                int bI, bO;
                if constexpr (R > 0) {
                    // every row group's (score, file index) read at once, the winner chosen
                    // in registers, then its template and B value: two LDS round trips
                    float v[REG_NRG];
                    int o[REG_NRG];
#pragma unroll
                    for (int r = 0; r < REG_NRG; ++r) { v[r] = pv[r * CW]; o[r] = po[r * CW]; }
                    bv = v[0];
                    bO = o[0];
                    int br = 0;
#pragma unroll
                    for (int r = 1; r < REG_NRG; ++r) {
                        const bool gt = better(v[r], o[r], bv, bO);
                        bv = gt ? v[r] : bv;
                        bO = gt ? o[r] : bO;
                        br = gt ? r : br;
                    }
                    bI = pi[br * CW];
                    bb = pb[br * CW];
                } else {
                    bv = pv[0]; bb = pb[0];
                    bI = pi[0]; bO = po[0];
                    for (int r = 1; r < nrg; ++r) {
                        const int j = r * CW;
                        if (better(pv[j], po[j], bv, bO)) { bv = pv[j]; bI = pi[j]; bO = po[j]; bb = pb[j]; }
                    }
                }
                const int ww = c0 + tid;
                if (ww < c1) {
                    s_cf[tid] = bv;
                    s_im[tid] = bI;
                    s_bv[tid] = bb;
                    if (!local) publish(tid, 0, 0u);
                    if (it == 0 && p.dbg_cfmax && S == 1) { p.dbg_cfmax[ww] = bv; p.dbg_imax[ww] = bI; }
                }
            }
            if (RS && p.rescore) {
                // BF16 projection: the scores of every tile that could reach lo2 (10%
                // under threshold) are the FP32 projection's already.  A score of the
                // other tiles matters only if it could reach lo2 in FP32 (a column under
                // lo2 cannot beat an accepted one, nor be accepted); when one could, the
                // column's candidates are rescored in FP32 before it is published, so
                // pooling, acceptance, the template choice and amplitudes see FP32 scores
                // wherever they can matter.  The correction of (t, w) is fixed
                // (subtractions are applied to both alike).
                const int lane = tid & 63, wv = tid >> 6;
                // FP32 dot of U2 row t with Bpc column ww over the K-blocks listed for t's
                // 64-row tile (the others are exact zeros): four lanes per 16-wide block,
                // 32 blocks in flight per round; every lane of the wave ends with the
                // same, order-fixed sum.
                auto dot32 = [&](int t, int ww) __attribute__((always_inline)) {
                    const int mt = t / GEMM_BM;
                    const int j0 = p.kb_ptr[mt], j1 = p.kb_ptr[mt + 1];
                    const float* urow = p.U2 + (int64_t)t * p.ldu;
                    const float* bcol = p.Bpc + ww;               // element (k, ww) at k * BS
                    const int q4 = (lane & 3) * 4;
                    float acc = 0.f;
                    for (int jr = j0; jr < j1; jr += 32) {
                        int kb[2];
#pragma unroll
                        for (int r = 0; r < 2; ++r) {
                            const int j = jr + 16 * r + (lane >> 2);
                            kb[r] = j < j1 ? p.kb_idx[j] * GEMM_BK + q4 : -1;
                        }
                        float4 u[2], b[2];
#pragma unroll
                        for (int r = 0; r < 2; ++r) {
                            u[r] = kb[r] >= 0 ? *reinterpret_cast<const float4*>(urow + kb[r]) : make_float4(0.f, 0.f, 0.f, 0.f);
                            b[r] = kb[r] >= 0 ? make_float4(bcol[(int64_t)kb[r] * p.BS], bcol[(int64_t)(kb[r] + 1) * p.BS],
                                                            bcol[(int64_t)(kb[r] + 2) * p.BS], bcol[(int64_t)(kb[r] + 3) * p.BS])
                                              : make_float4(0.f, 0.f, 0.f, 0.f);
                        }
#pragma unroll
                        for (int r = 0; r < 2; ++r) acc += ((u[r].x * b[r].x + u[r].y * b[r].y) + u[r].z * b[r].z) + u[r].w * b[r].w;
                    }
                    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
                    return acc;
                };
                // Every template whose score could be the column's FP32 maximum is
                // rescored: a BF16 score b16 has |b16 - b32| <= BF16_ERR * sqrt(nm) * N (N:
                // the norm of the column's BF16 operand over the template's tile), so in
                // the sqrt(Cf) scale each template is within BF16_ERR * N of its FP32 value,
                // and the FP32 maximum lies among the templates with sqrt(Cf16) +
                // BF16_ERR * N >= sqrt(Cf16 of the BF16 best) - BF16_ERR * N(best).  Those
                // (column, template) pairs are listed with their current BF16 score (at most
                // MAX_BAND per workgroup and iteration; more is reported as an error),
                // rescored one wave each (the correction of the column's last winner is
                // cached), and the column keeps the FP32 maximum (larger score, then the
                // smaller file index).
                // a score in an FP32 tile needs no correction
                auto exact_t = [&](int t, int ww) { return p.act[(ww / GEMM_BN) * p.act_mt + t / GEMM_BM] != 0; };
                // register tile: rescoring runs in a pass only when a thread asked for it
                const bool go = R == 0 || any_need != 0;
                float cf16 = 0.f;                        // the column's best BF16 score
                if (go && tid < CW && c0 + tid < c1) {
                    cf16 = s_cf[tid];
                    float sl = -1.f;
                    if (cf16 >= p.lo2) {
                        const float nb = R > 0 ? s_nmx[R > 0 ? tid : 0]
                                               : sqrtf(p.Nsq[(int64_t)(s_im[tid] / GEMM_BM) * p.BS + c0 + tid]);
                        sl = fmaxf(sqrtf(cf16) - BF16_ERR * nb, 0x1p-60f);
                    }
                    s_cm[tid] = sl;                      // the band's floor in the sqrt(Cf) scale
                }
                int listed = 0;
                if (!go) {
                    // no candidate outside the FP32 tiles
                } else if constexpr (R > 0) {
                    // Register tile.  Each thread lists its candidate rows (score plus the
                    // BF16 bound reaching the floor) with their current BF16 score; their
                    // corrections come from the batch's cache or, the first time, from an
                    // FP32 dot (one wave per pair); then each listed pair offers its FP32
                    // score to its column through one 64-bit LDS maximum of (score,
                    // inverted file index, list position), and the column takes the winner.
                    unsigned long long* const s_key = reinterpret_cast<unsigned long long*>(s_bcor + MAX_BAND);
                    int* const s_pl = s_whit;                                        // column | row << 16
                    float* const s_pv = reinterpret_cast<float*>(s_whit) + MAX_BAND;  // BF16 score, then FP32
                    if (tid < CW) s_key[tid] = 0ull;
                    if (tid == 0) { s_nband = 0; s_nnan = 0; }
                    __syncthreads();
                    const float sl = own ? s_cm[col] : -1.f;
                    const int r0 = rbase + rg * R;
                    // A thread's best score over its rows (this pass's scan, or the last
                    // one's when its rows did not change) bounds every row's a = r^2 * (1/nm)
                    // from above to within 2^-22, so a thread whose best is under the
                    // threshold has no row in the band and skips the row tests.
                    const float e = sl - BF16_ERR * nown;
                    const float thr = e > 0.f ? e * e * (1.f - 0x1p-20f) : -1.f;
                    uint64_t mask = 0;
                    if (own && !edge && !woff && sl > 0.f && r0 < T && s_val[tid] >= thr * (1.f - 0x1p-18f)) {
                        const float* rnp = s_nm + rg * R;
#pragma unroll
                        for (int i0 = 0; i0 < R; i0 += 8) {
#pragma unroll
                            for (int i = i0; i < i0 + 8 && i < R; ++i) {
                                const float r = breg[i] > 0.f ? breg[i] : 0.f;
                                const float a = (r * r) * rnp[i];
                                mask |= (uint64_t)(a >= thr && r0 + i < T) << i;
                            }
                            asm volatile("" ::: "memory");
                        }
                    }
                    // the wave's listed rows get consecutive slots (a prefix sum over the
                    // lanes, one LDS atomic per wave); each lane then writes its rows in
                    // row order, the register picked by the unrolled row index
                    {
                        const int cnt = __popcll(mask);
                        int incl = cnt;
#pragma unroll
                        for (int o = 1; o < 64; o <<= 1) {
                            const int y = __shfl_up(incl, o);
                            incl += lane >= o ? y : 0;
                        }
                        const int wtot = __shfl(incl, 63);
                        if (wtot > 0) {
                            int base = 0;
                            if (lane == 63) base = atomicAdd(&s_nband, wtot);
                            int k = __shfl(base, 63) + incl - cnt;
#pragma unroll
                            for (int i = 0; i < R; ++i) {
                                if ((mask >> i) & 1ull) {
                                    if (k < MAX_BAND) {
                                        s_pl[k] = col | ((r0 + i) << 16);
                                        s_pv[k] = breg[i];
                                    }
                                    ++k;
                                }
                            }
                        }
                    }
                    __syncthreads();
                    const int nb = min(s_nband, MAX_BAND);
                    if (tid == 0 && s_nband > MAX_BAND) atomicOr(p.err, ERR_BAND);
                    // cached corrections (NaN: not computed in this batch yet); the missing
                    // ones are compacted (into the staged-spike array, dead until acceptance)
                    // so the waves share them evenly
                    int* const s_nanq = reinterpret_cast<int*>(s_sp);
                    for (int q = tid; q < nb; q += PEEL_THREADS) {
                        const int e2 = s_pl[q];
                        const bool ex = exact_t(e2 >> 16, c0 + (e2 & 0xffff));
                        s_bcor[q] = 0.f;
                        if (!ex) s_nanq[atomicAdd(&s_nnan, 1)] = q;
                    }
                    __syncthreads();
                    const int nnan = s_nnan;
                    for (int j = wv; j < nnan; j += PEEL_THREADS / 64) {
                        const int q = s_nanq[j];
                        const int e2 = s_pl[q];
                        const int cl = e2 & 0xffff, t = e2 >> 16, ww = c0 + cl;
                        const float corr = dot32(t, ww) - p.Binit[(int64_t)t * p.BS + ww];
                        if (lane == 0) s_bcor[q] = corr;
                    }
                    __syncthreads();
                    for (int q = tid; q < nb; q += PEEL_THREADS) {
                        const int e2 = s_pl[q];
                        const int t = e2 >> 16;
                        const float b = s_pv[q] + s_bcor[q];
                        const float r = fmaxf(b, 0.f);
                        const float cf = (r * r) / s_nmf[t];
                        s_pv[q] = b;
                        atomicMax(&s_key[e2 & 0xffff], ((unsigned long long)__float_as_uint(cf) << 32) |
                                                           ((unsigned long long)(0xffffu - (unsigned)s_or[t]) << 16) |
                                                           (unsigned)q);
                    }
                    __syncthreads();
                    if (tid < CW && c0 + tid < c1 && s_cm[tid] > 0.f && s_key[tid] != 0ull) {
                        const unsigned long long kb = s_key[tid];
                        const int q = (int)(kb & 0xffffull);
                        s_cf[tid] = __uint_as_float((unsigned)(kb >> 32));
                        s_im[tid] = s_pl[q] >> 16;
                        s_bv[tid] = s_pv[q];
                    }
                    listed = nb;
                } else {
                    // Global tile: the same list / cache / dot steps over the rows of B in
                    // global memory (each row's BF16 bound from its tile's operand norm);
                    // every (template, column) pair belongs to one workgroup, and each
                    // column scans the list for its FP32 maximum.
                    int* const s_bi = s_qlo;                                   // column | template << 16
                    float* const s_bb = reinterpret_cast<float*>(s_qhi);       // BF16 score, then FP32
                    if (tid == 0) s_nband = 0;
                    __syncthreads();
                    const float sl = own ? s_cm[col] : -1.f;
                    if (own && !edge && !woff && sl > 0.f) {
                        for (int t = rg; t < T; t += nrg) {
                            if (p.gate && !p.gate[(w / GEMM_BN) * p.gate_mt + t / GEMM_BM]) continue;
                            const float b = p.B[(int64_t)t * p.BS + w];
                            const float r = fmaxf(b, 0.f);
                            const float a = (r * r) * p.rnm[t];
                            const float nn = sqrtf(p.Nsq[(int64_t)(t / GEMM_BM) * p.BS + w]);
                            const float e = sl - BF16_ERR * nn;
                            const float thr = e > 0.f ? e * e * (1.f - 0x1p-20f) : -1.f;
                            if (a >= thr) {
                                const int k = atomicAdd(&s_nband, 1);
                                if (k >= MAX_BAND) break;
                                s_bi[k] = col | (t << 16);
                                s_bb[k] = b;
                            }
                        }
                    }
                    __syncthreads();
                    const int nb = min(s_nband, MAX_BAND);
                    if (tid == 0 && s_nband > MAX_BAND) atomicOr(p.err, ERR_BAND);
                    // pairs of BF16 tiles get their correction (compacted so the waves share
                    // them evenly, in the staged-spike array, dead until acceptance)
                    int* const s_nanq = reinterpret_cast<int*>(s_sp);
                    if (tid == 0) s_nnan = 0;
                    __syncthreads();
                    for (int q = tid; q < nb; q += PEEL_THREADS) {
                        const int e2 = s_bi[q];
                        s_bcor[q] = 0.f;
                        if (!exact_t(e2 >> 16, c0 + (e2 & 0xffff))) s_nanq[atomicAdd(&s_nnan, 1)] = q;
                    }
                    __syncthreads();
                    const int nnan = s_nnan;
                    for (int j = wv; j < nnan; j += PEEL_THREADS / 64) {
                        const int q = s_nanq[j];
                        const int e2 = s_bi[q];
                        const int cl = e2 & 0xffff, t = e2 >> 16, ww = c0 + cl;
                        const float corr = dot32(t, ww) - p.Binit[(int64_t)t * p.BS + ww];
                        if (lane == 0) s_bcor[q] = corr;
                    }
                    __syncthreads();
                    if (tid < CW && c0 + tid < c1 && s_cm[tid] > 0.f) {
                        float bv = -1.f, bb = 0.f;
                        int bt = 0, bo = 0x7fffffff;
                        for (int q = 0; q < nb; ++q) {
                            const int e2 = s_bi[q];
                            if ((e2 & 0xffff) != tid) continue;
                            const int t = e2 >> 16;
                            const float b = s_bb[q] + s_bcor[q];
                            const float r = fmaxf(b, 0.f);
                            const float cf = (r * r) / p.nm[t];
                            const int o = p.orig[t];
                            if (better(cf, o, bv, bo)) { bv = cf; bo = o; bt = t; bb = b; }
                        }
                        if (bo != 0x7fffffff) {
                            s_cf[tid] = bv;
                            s_im[tid] = bt;
                            s_bv[tid] = bb;
                        }
                    }
                    listed = nb;
                }
                int flip = 0;
                if (tid < CW && c0 + tid < c1 && cf16 >= p.lo2) flip = (cf16 > p.Th2) != (s_cf[tid] > p.Th2);
                const uint64_t fb = __ballot(flip);
                if (tid == 0 && listed > 0) atomicAdd(&p.n_rescored[0], (unsigned long long)listed);
                if (lane == 0 && fb) atomicAdd(&p.n_rescored[1], (unsigned long long)__popcll(fb));
            }
        }
        mark(8);
        if (local) {
            // ---------------- local acceptance (Kilosort rules) ----------------
            if (tid < CW && c0 + tid < c1) publish(tid, it & 1, ep | (unsigned)(it + 1));
            mark(2);
            const int x0 = c0 - 2 * M;                 // window column 0
            const int nw = (c1 - c0) + 4 * M;          // window: columns [c0 - 2M, c1 + 2M)
            const int na = (c1 - c0) + 2 * M;          // columns [c0 - M, c1 + M): window M .. M + na - 1
            // the stop test, on a thread of the last wave (no window column there)
            if (tid == PEEL_THREADS - 64) stop_test(it);
            // Thread j < nw reads window column j of every subset: the words of this
            // iteration (score; for columns within M also B value and template), and
            // keeps the best (largest score, then the smaller file index: the argmax
            // over all templates).  (The opaque index keeps per-thread addresses out
            // of the peel loop.)
            int pid = tid;
            asm volatile("" : "+v"(pid));
            const int ww = x0 + pid;
            const bool jin = pid < nw && ww >= 0 && ww < W;
            const bool jacc = jin && pid >= M && pid < M + na;
            float bv = -1.f, bb = 0.f;              // scores are >= 0: the first read one wins
            int bt = 0, bo = 0x7fffffff;
            {
                const unsigned tg = ep | (unsigned)(it + 1), tdone = ep | TAG_DONE;
                auto ok = [&](unsigned long long x) { return (unsigned)x == tg || (unsigned)x == tdone; };
                const int64_t base = (int64_t)(it & 1) * S * W + ww;
                // subsets per round, every load of a round in flight at once (tiles of
                // 64 rows and the global tile are never split)
                constexpr int QS = (R == 0 || R == 64) ? 1 : 4;
                unsigned spins = 0;
                for (int s0 = 0; s0 < S; s0 += QS) {
                    unsigned pend = jin ? ((1u << min(QS, S - s0)) - 1u) : 0u;
                    while (pend) {
                        unsigned long long wa[QS], wb[QS], wt[QS];
#pragma unroll
                        for (int q = 0; q < QS; ++q) {
                            wa[q] = wb[q] = wt[q] = 0ull;
                            if ((pend >> q) & 1u) {
                                const int64_t k = base + (int64_t)(s0 + q) * W;
                                wa[q] = ld_c(&Pa[k]);
                                if (jacc) {
                                    wb[q] = ld_c(&Pb[k]);
                                    wt[q] = ld_c(&Pt[k]);
                                }
                            }
                        }
#pragma unroll
                        for (int q = 0; q < QS; ++q) {
                            if (((pend >> q) & 1u) && ok(wa[q]) && (!jacc || (ok(wb[q]) && ok(wt[q])))) {
                                pend &= ~(1u << q);
                                const float v = __uint_as_float((unsigned)(wa[q] >> 32));
                                const int t = (int)(wt[q] >> 32);
                                const int o = (R > 0 && S > 1 && jacc) ? s_or[t] : 0;
                                // subsets come in the order their words arrive; equal
                                // (score, file index) keys only occur on the edge
                                // columns, where every subset reports the same winner
                                if (v > bv || (v == bv && o < bo)) {
                                    bv = v; bo = o; bt = t;
                                    bb = __uint_as_float((unsigned)(wb[q] >> 32));
                                }
                            }
                        }
                        if (pend) {
                            __builtin_amdgcn_s_sleep(1);
                            if (++spins > (1u << 26)) {
                                __hip_atomic_fetch_or(p.err, ERR_BARRIER, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                s_abort = 1;
                                break;
                            }
                        }
                    }
                }
            }
            if (!jin) bv = 0.f;                     // outside the batch
            const int lane = tid & 63, wv = tid >> 6;
            if (wv * 64 < nw) {
                // scores of the window (0 outside the batch), their maxima over aligned
                // 8-column groups (lane butterflies), the winners within M
                if (pid < nw) s_sc[pid] = bv;
                float m = bv;
#pragma unroll
                for (int o = 1; o < 8; o <<= 1) {
                    const float x = __shfl_xor(m, o);
                    m = x > m ? x : m;
                }
                if ((lane & 7) == 0) s_bm[pid >> 3] = m;
                if (jacc) { s_wt[pid - M] = bt; s_wb[pid - M] = bb; }
            }
            const int ln0 = s_listn;
            __syncthreads();
            mark(4);
            if (s_tot == 0 || (s_err & ERR_BARRIER) || s_abort) break;
            // accept column ww when Cmax = max of the scores over [ww - M, ww + M]
            // exceeds Th^2 and equals the column's own score (scores are finite and
            // >= 0: selects rather than fmaxf)
            bool f = false, mine = false;
            if (pid < na) {
                const int j = M + pid, wc = c0 - M + pid;
                const float cfw = s_sc[j];
                if (wc >= 0 && wc < W && cfw > p.Th2) {
                    // window [lo, hi] = the columns before the first whole 8-column
                    // group (< 8), the whole groups (< 16, since 2M + 1 < 128), the
                    // columns after the last (< 8).  Fixed trip counts keep every
                    // read in flight at once; a slot outside the window reads the
                    // column itself, which is in it.
                    const int lo = j - M, hi = j + M;
                    const int blo = (lo + 7) >> 3, bhi = (hi + 1) >> 3;
                    const int hend = min(8 * blo, hi + 1), t0 = 8 * max(blo, bhi);
                    float v[7 + 16 + 7];
#pragma unroll
                    for (int e = 0; e < 7; ++e) v[e] = s_sc[lo + e < hend ? lo + e : j];
#pragma unroll
                    for (int b = 0; b < 16; ++b) v[7 + b] = blo + b < bhi ? s_bm[blo + b] : cfw;
#pragma unroll
                    for (int e = 0; e < 7; ++e) v[23 + e] = s_sc[t0 + e <= hi ? t0 + e : j];
                    float cm = cfw;
#pragma unroll
                    for (int e = 0; e < 30; ++e) cm = v[e] > cm ? v[e] : cm;
                    f = cm > p.Th2 && fabsf(cm - cfw) < 1e-9f;
                }
                mine = f && wc >= c0 && wc < c1;
            }
            // the accepted columns in sample order: all of them are the next
            // iteration's subtractions, those of this block go to its list
            const uint64_t bf = __ballot(f), bm = __ballot(mine);
            if (lane == 0) { s_hcnt_w[wv] = __popcll(bf); s_hcnt_o[wv] = __popcll(bm); }
            __syncthreads();
            if (f) {
                int nb = 0, nbo = 0;
#pragma unroll
                for (int k4 = 0; k4 < PEEL_THREADS / 256; ++k4) {
                    const int4 a = reinterpret_cast<const int4*>(s_hcnt_w)[k4];
                    const int4 o = reinterpret_cast<const int4*>(s_hcnt_o)[k4];
                    nb += (4 * k4 < wv ? a.x : 0) + (4 * k4 + 1 < wv ? a.y : 0) +
                          (4 * k4 + 2 < wv ? a.z : 0) + (4 * k4 + 3 < wv ? a.w : 0);
                    nbo += (4 * k4 < wv ? o.x : 0) + (4 * k4 + 1 < wv ? o.y : 0) +
                           (4 * k4 + 2 < wv ? o.z : 0) + (4 * k4 + 3 < wv ? o.w : 0);
                }
                const uint64_t below = (1ull << lane) - 1ull;
                const int idx = nb + __popcll(bf & below);
                const int t = s_wt[pid];
                const float nmt = R > 0 ? s_nmf[t] : p.nm[t];
                const Spike sp{c0 - M + pid, t, s_wb[pid] / nmt, it};
                if (idx < MAX_STAGE) s_sp[idx] = sp;
                if (mine) {
                    const int ln = ln0 + nbo + __popcll(bm & below);
                    if (ln < MAX_LIST) s_list[ln] = sp;
                }
            }
            if (S > 1 && it == 0 && p.dbg_cfmax && owner && pid >= M && pid < M + (c1 - c0)) {
                p.dbg_cfmax[c0 - M + pid] = s_sc[M + pid];
                p.dbg_imax[c0 - M + pid] = s_wt[pid];
            }
            if (tid == 0) {
                int n = 0, no = 0;
                for (int k = 0; k < PEEL_THREADS / 64; ++k) { n += s_hcnt_w[k]; no += s_hcnt_o[k]; }   // 0 past na
                if (n > MAX_STAGE) atomicOr(p.err, ERR_STAGE);
                if (ln0 + no > MAX_LIST) atomicOr(p.err, ERR_LIST);
                s_nsp = min(n, MAX_STAGE);
                s_listn = min(ln0 + no, MAX_LIST);
                // one atomic counts the owner done and adds its block's spikes, so a
                // reader that sees every owner done also sees the complete total
                if (owner)
                    __hip_atomic_fetch_add(&itc[it], (1ull << 32) | (unsigned)no, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
            }
            __syncthreads();
            mark(3);
        } else {
            // ---------------- LSS rules (one subset): spike exchange ----------------
            // phase A done: the neighbouring blocks may read this block's column maxima
            raise_flag(&flagA[g], it + 1);
            wait_flags(flagA + hlo, nh, it + 1, p.err);
            mark(2);
            const int nwin = (c1 - c0) + 2 * M;
            for (int i = tid; i < nwin; i += PEEL_THREADS) {
                const int ww = c0 - M + i;
                float v = -INFINITY;
                if (ww >= c0 && ww < c1) v = s_cf[ww - c0];
                else if (ww >= 0 && ww < W) v = ld_c(&p.Cfmax[ww]);
                s_win[i] = v;
            }
            if (tid == 0) s_ncand = 0;
            __syncthreads();
            float* const s_part = reinterpret_cast<float*>(s_whit);     // free here (R > 0)
            if (R > 0) {
                // pooled maximum over each own column's 2M+1 window: the nrg row groups take
                // interleaved slices, the column's thread then combines the nrg slices
                float m = -INFINITY;
                for (int i = rg; i <= 2 * M; i += nrg) m = fmaxf(m, s_win[col + i]);
                s_part[tid] = m;
                __syncthreads();
            }
            if (tid < CW) {
                int f = 0;
                if (c0 + tid < c1) {
                    float cm = -INFINITY;
                    if (R > 0) {
                        for (int q = 0; q < nrg; ++q) cm = fmaxf(cm, s_part[q * CW + tid]);
                    } else {
                        for (int i = 0; i <= 2 * M; ++i) cm = fmaxf(cm, s_win[tid + i]);
                    }
                    const float cfw = s_win[tid + M];
                    f = cm >= p.Th2 && (cm - cfw) < 1e-4f;
                    s_cm[tid] = cm;
                }
                s_flag[tid] = f;
            }
            __syncthreads();
            // every template within 1e-4 of the pooled max (myGPUhelpers.cu:465); the
            // candidates are collected, then ordered by (column, template)
            if (own && s_flag[col]) {
                const float cm = s_cm[col];
                if (R > 0) {
                    int vb = rg * R;
                    asm volatile("" : "+v"(vb));
#pragma unroll
                    for (int i = 0; i < (R > 0 ? R : 1); ++i) {
                        const int t = vb + i;
                        const float b = breg[i];
                        const float r = fmaxf(b, 0.f);
                        const float cf = (r * r) * s_nm[t];      // padded rows: rnm = 0
                        if (fabsf(cm - cf) < 1e-4f) {
                            const int k = atomicAdd(&s_ncand, 1);
                            if (k < MAX_STAGE) s_sp[k] = Spike{w, t, b * s_nm[t], it};
                        }
                    }
                } else {
                    for (int t = rg; t < T; t += nrg) {
                        const float b = p.B[(int64_t)t * p.BS + w];
                        const float r = fmaxf(b, 0.f);
                        const float cf = (r * r) * p.rnm[t];
                        if (fabsf(cm - cf) < 1e-4f) {
                            const int k = atomicAdd(&s_ncand, 1);
                            if (k < MAX_STAGE) s_sp[k] = Spike{w, t, b * p.rnm[t], it};
                        }
                    }
                }
            }
            __syncthreads();
            if (tid == 0) {
                int n = min(s_ncand, MAX_STAGE), ln = s_listn;
                if (s_ncand > MAX_STAGE) atomicOr(p.err, ERR_SLOT);
                for (int i = 1; i < n; ++i) {          // insertion sort by (column, template)
                    const Spike x = s_sp[i];
                    int j = i - 1;
                    while (j >= 0 && (s_sp[j].w > x.w || (s_sp[j].w == x.w && s_sp[j].t > x.t))) {
                        s_sp[j + 1] = s_sp[j];
                        --j;
                    }
                    s_sp[j + 1] = x;
                }
                if (n > p.slot_cap) { atomicOr(p.err, ERR_SLOT); n = p.slot_cap; }
                for (int i = 0; i < n; ++i) {
                    if (ln < MAX_LIST) s_list[ln] = s_sp[i];
                    else atomicOr(p.err, ERR_LIST);
                    ++ln;
                }
                s_n = n;
                s_listn = min(ln, MAX_LIST);
            }
            __syncthreads();
            const int n = s_n;
            for (int i = tid; i < n; i += PEEL_THREADS) st_spike(&p.slots[(int64_t)cb * p.slot_cap + i], s_sp[i]);
            if (tid == 0) st_c(&p.slot_cnt[cb], n);
            mark(3);
            // one atomic counts the block done and adds its spikes, so a reader that
            // sees every block done also sees the complete total
            if (tid == 0)
                __hip_atomic_fetch_add(&itc[it], (1ull << 32) | (unsigned)n, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            raise_flag(&flagB[cb], it + 1);
            // the next exchange reads the slots of blocks hlo..hhi; waiting for them also
            // keeps this workgroup's next phase A from overwriting column maxima that a
            // neighbour is still pooling
            wait_flags(flagB + hlo, nh, it + 1, p.err);
            mark(4);
        }
    }
    mark(0);
    if (local && tid < CW && c0 + tid < c1) {
        // final maxima for the neighbours still iterating: every later iteration's
        // (no score changes after an iteration without spikes, or none follows)
        publish(tid, 0, ep | TAG_DONE);
        publish(tid, 1, ep | TAG_DONE);
    }
    if (!local) {
        raise_flag(&flagA[g], FLAG_DONE);
        if (tid == 0) st_c(&flagB[cb], FLAG_DONE);
    }

    // ---------------- events and features ----------------
    // The owners publish their lists for the neighbours' features.  Every
    // workgroup of a column block holds the block's list (all of them accepted
    // the block's spikes alike) and takes a share of its spikes' features (spike
    // i goes to subset i % S), so a block dense in spikes is spread over its S
    // workgroups; the owner writes the events.
    if (owner) {
        const int nl = min(s_listn, MAX_LIST);
        for (int i = tid; i < nl; i += PEEL_THREADS) st_spike(&p.list[(int64_t)cb * p.cap_g + i], s_list[i]);
        if (tid == 0) st_c(&p.list_cnt[cb], nl);
    }
    grid_sync(p.bar, p.G, g, bar_gen, p.err);
    // Blocks whose spikes lie within M - 1 columns (the features' overlap terms).
    const int flo = max(0, (c0 - (M - 1)) / CW), fhi = min(p.NB - 1, (c1 - 1 + M - 1) / CW);
    const int nfb = fhi - flo + 1;               // <= MAX_NBR because CW >= M
    __shared__ int s_nbn[MAX_NBR], s_nbo[MAX_NBR], s_nbt;
    {   // one round trip: the spike counts of the blocks before this one (its offset in
        // the event list: wave 0 sums them) and of the neighbour blocks
        if (tid < 64) {
            int part = 0;
            for (int hh = tid; hh < cb; hh += 64) part += ld_c(&p.list_cnt[hh]);
            for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
            if (tid == 0) s_off = part;
        } else if (tid < 64 + nfb) {
            const int hh = flo + tid - 64;
            s_nbn[tid - 64] = hh == cb ? 0 : min(ld_c(&p.list_cnt[hh]), MAX_LIST);
        }
        __syncthreads();
    }
    // the neighbour blocks' lists, staged once in the (now free) spike array
    if (tid == 0) {
        int o = 0;
        for (int i = 0; i < nfb; ++i) { s_nbo[i] = o; o += s_nbn[i]; }
        s_nbt = o;
    }
    __syncthreads();
    const bool nb_lds = p.do_feat && s_nbt <= MAX_STAGE;
    if (nb_lds)
        for (int i = 0; i < nfb; ++i)
            for (int q = tid; q < s_nbn[i]; q += PEEL_THREADS)
                s_sp[s_nbo[i] + q] = ld_spike(&p.list[(int64_t)(flo + i) * p.cap_g + q]);
    const int nmine = min(s_listn, MAX_LIST);
    const int off = s_off;
    for (int i = tid; i < nmine; i += PEEL_THREADS) {
        const Spike si = s_list[i];
        int r = 0;
        for (int j = 0; j < nmine; ++j) {
            const Spike sj = s_list[j];
            r += (sj.w < si.w) || (sj.w == si.w && (sj.it < si.it || (sj.it == si.it && j < i)));
        }
        s_rank[i] = r;
        const int d = off + r;
        if (owner && d < p.ev_cap && !p.cent) {     // with relabelling the event is written below
            lss_event e;
            e.sample = si.w - p.half + p.nt0min;     // the host adds the batch's first sample
            e.template_id = p.orig[si.t];
            e.amplitude = si.amp;
            p.ev[d] = e;
        }
    }
    __syncthreads();
    if (p.do_feat) {
        const int lane = tid & 63, wv = tid >> 6;
        const int K = p.K, nC = p.nC, C = p.C, L2 = 2 * M - 1;
        float* stage = reinterpret_cast<float*>(s_idx) + wv * 64;    // s_idx is free after the offset sum
        for (int i = sb + S * wv; i < nmine; i += S * (PEEL_THREADS / 64)) {
            const int d = off + s_rank[i];
            if (d >= p.ev_cap) continue;                             // uniform per wave
            const Spike si = s_list[i];
            const bool mine = lane < nC * K;
            const int j = mine ? lane / K : 0, k = mine ? lane % K : 0;
            const int ch = p.iCC[j * C + p.iU[si.t]];
            float f = 0.f;
            if (mine)
                f = p.Bpc[(int64_t)(k * p.CP + ch) * p.BS + si.w] +
                    si.amp * p.U[((int64_t)si.t * K + k) * C + ch];
            // the spikes within +-(M-1) columns, in list order: a wave reads 64 at a time
            // and takes the near ones from the ballot, lowest index first
            for (int hh = flo; hh <= fhi; ++hh) {
                const int cnt = hh == cb ? nmine : s_nbn[hh - flo];
                const Spike* lst = p.list + (int64_t)hh * p.cap_g;
                const Spike* lsl = hh == cb ? s_list : s_sp + s_nbo[hh - flo];
                for (int q0 = 0; q0 < cnt; q0 += 64) {
                    const int q = q0 + lane;
                    Spike sq{};
                    bool near = false;
                    if (q < cnt) {
                        sq = (hh == cb || nb_lds) ? lsl[q] : ld_spike(&lst[q]);
                        const int dd = si.w - sq.w;
                        near = dd >= -(M - 1) && dd <= M - 1;
                    }
                    unsigned long long mask = __ballot(near);
                    while (mask) {
                        const int b = __builtin_ctzll(mask);
                        mask &= mask - 1;
                        const int wq = __shfl(sq.w, b), tq = __shfl(sq.t, b);
                        const float aq = __shfl(sq.amp, b);
                        if (mine) {
                            const int dd = si.w - wq;
                            // all 2K loads in flight before the sum (in k2 order)
                            float uu[MAX_K], gg[MAX_K];
#pragma unroll
                            for (int k2 = 0; k2 < MAX_K; ++k2) {
                                uu[k2] = k2 < K ? p.U[((int64_t)tq * K + k2) * C + ch] : 0.f;
                                gg[k2] = k2 < K ? p.Gt[(k2 * K + k) * L2 + dd + M - 1] : 0.f;
                            }
                            float gs = 0.f;
#pragma unroll
                            for (int k2 = 0; k2 < MAX_K; ++k2)
                                if (k2 < K) gs += uu[k2] * gg[k2];
                            f -= aq * gs;
                        }
                    }
                }
            }
            if (mine) p.feat[(int64_t)d * nC * K + lane] = f;
            if (p.pos) {
                // position (compute_spike_positions_kernel, OnlineSpikesV2.cu:190-245): the
                // mass of nearest channel jj is the sum of its squared features; the
                // channel coordinates are averaged with weights mass / total, in order
                stage[lane] = f;
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_s_waitcnt(0xc07f);               // lgkmcnt(0): the features are in LDS
                // lane jj < nC loads nearest channel jj's mass and coordinates; the sums
                // then run over jj in order on every lane (broadcasts), as a serial loop would
                const int iu = p.iU[si.t];
                float mj = 0.f, cxj = 0.f, cyj = 0.f;
                if (lane < nC) {
                    for (int kk = 0; kk < K; ++kk) mj += stage[lane * K + kk] * stage[lane * K + kk];
                    const int c2 = p.iCC[lane * C + iu];
                    cxj = p.xc[c2];
                    cyj = p.yc[c2];
                }
                float total = 0.f;
                for (int jj = 0; jj < nC; ++jj) total += __shfl(mj, jj);
                float x = 0.f, y = 0.f;
                for (int jj = 0; jj < nC; ++jj) {
                    const float m = __shfl(mj, jj);
                    const float wgt = total > 0.f ? m / total : 0.f;
                    x += __shfl(cxj, jj) * wgt;
                    y += __shfl(cyj, jj) * wgt;
                }
                if (lane == 0) {
                    p.pos[2 * (int64_t)d] = x;
                    p.pos[2 * (int64_t)d + 1] = y;
                }
                __builtin_amdgcn_wave_barrier();
                if (p.cent) {
                    // nearest final cluster (closestCluster, OnlineSpikesV2.cu:1058-1072):
                    // Euclidean distance, the first index among equals
                    float bd = INFINITY;
                    int bc = 0x7fffffff;
                    for (int c2 = lane; c2 < p.n_cent; c2 += 64) {
                        const float dx = x - p.cent[2 * c2], dy = y - p.cent[2 * c2 + 1];
                        const float dist = sqrtf(dx * dx + dy * dy);
                        if (dist < bd) { bd = dist; bc = c2; }
                    }
                    for (int m = 32; m > 0; m >>= 1) {
                        const float od = __shfl_xor(bd, m);
                        const int oc = __shfl_xor(bc, m);
                        if (od < bd || (od == bd && oc < bc)) { bd = od; bc = oc; }
                    }
                    if (lane == 0) {
                        lss_event e;
                        e.sample = si.w - p.half + p.nt0min;
                        e.template_id = bc != 0x7fffffff ? bc : p.orig[si.t];
                        e.amplitude = si.amp;
                        p.ev[d] = e;
                    }
                }
            }
        }
    }
    mark(5);
    if (timing) {
        for (int k = 0; k < NPHASE; ++k) {
            atomicAdd(&p.prof[k], (unsigned long long)s_ph[k]);
            atomicMax(&p.prof[NPHASE + k], (unsigned long long)s_ph[k]);
        }
        atomicAdd(&p.prof[2 * NPHASE], (unsigned long long)s_ph[NPHASE + 1]);
        atomicAdd(&p.prof[2 * NPHASE + 1], 1ull);
    }
    if (cb == p.NB - 1 && owner && tid == 0) {
        // iterations run: the first that accepted nothing, or the cap (the grid
        // barrier of the epilogue has made every count final)
        int n_iter = p.max_iter;
        for (int k = 0; k < p.max_iter; ++k)
            if ((unsigned)ld_c(&itc[k]) == 0u) { n_iter = k; break; }
        p.hdr->count = off + nmine;
        p.hdr->n_iter = n_iter;
        p.hdr->err = ld_c(p.err);
    }
}


// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_whiten(const WhitenBatch& b, const ResetBatch* rb, hipStream_t s) {
    WhitenLaunch wl{};
    wl.w = b;
    if (rb) wl.r = *rb;                          // else n = 0: nothing to reset
    // a launch's streams share one precision mode
    if (b.n > 0 && b.job[0].Wrh)
        hipLaunchKernelGGL(k_whiten_x3, dim3(b.start[b.n]), dim3(256), 0, s, wl);
    else
        hipLaunchKernelGGL(k_whiten, dim3(b.start[b.n]), dim3(256), 0, s, wl);
}

void launch_pcconv(const PcconvBatch& b, int K, hipStream_t s) {
    const bool x3 = b.job[0].x3 != 0;           // the jobs of a launch share the precision
    switch (K) {
#define LSS_PC_CASE(KK) \
    case KK: \
        if (x3) hipLaunchKernelGGL(k_pcconv_x3<KK>, dim3(b.start[b.n]), dim3(256), 0, s, b); \
        else hipLaunchKernelGGL(k_pcconv<KK>, dim3(b.start[b.n]), dim3(256), 0, s, b); \
        break;
        LSS_PC_CASE(1) LSS_PC_CASE(2) LSS_PC_CASE(3) LSS_PC_CASE(4)
        LSS_PC_CASE(5) LSS_PC_CASE(6) LSS_PC_CASE(7) LSS_PC_CASE(8)
#undef LSS_PC_CASE
        default: break;
    }
}

// BF16 mode: the BF16 projection of every tile, then the FP32 projection of the
// tiles it marked (the same jobs; act selects the tiles).  Split-BF16: one launch.
void launch_project(const ProjectBatch& b, int precision, hipStream_t s) {
    if (precision == LSS_PREC_BF16X3) {
        hipLaunchKernelGGL(k_project_x3, dim3(b.start[b.n]), dim3(256), 0, s, b);
        return;
    }
    if (precision == LSS_PREC_BF16_RESCORE) hipLaunchKernelGGL(k_project_bf16, dim3(b.start[b.n]), dim3(256), 0, s, b);
    hipLaunchKernelGGL(k_project, dim3(b.start[b.n]), dim3(256), 0, s, b);
}

void launch_norms(const float* U, int KC, int T, float* nm, float* rnm, double* nm64, hipStream_t s) {
    hipLaunchKernelGGL(k_norms, dim3(T), dim3(256), 0, s, U, KC, nm, rnm, nm64);
}

// The peel instance: tile height, FP32 rescoring compiled in (Kilosort rules only),
// the phase clock (profiling), LSS rules.  Each leaves the code of the others out.
static const void* peel_fn(int R, bool rescore, bool clock, bool lss) {
#define LSS_PEEL_FN(RR)                                                                                         \
    if (lss) return clock ? reinterpret_cast<const void*>(k_peel<RR, false, true, true>)                        \
                          : reinterpret_cast<const void*>(k_peel<RR, false, false, true>);                      \
    return clock ? (rescore ? reinterpret_cast<const void*>(k_peel<RR, true, true, false>)                      \
                            : reinterpret_cast<const void*>(k_peel<RR, false, true, false>))                    \
                 : (rescore ? reinterpret_cast<const void*>(k_peel<RR, true, false, false>)                     \
                            : reinterpret_cast<const void*>(k_peel<RR, false, false, false>))
    switch (R) {
        case 8: LSS_PEEL_FN(8);
        case 16: LSS_PEEL_FN(16);
        case 32: LSS_PEEL_FN(32);
        case 64: LSS_PEEL_FN(64);
        default: LSS_PEEL_FN(0);
    }
#undef LSS_PEEL_FN
}

int peel_rows(int T, int CW) {
    if (CW != PEEL_THREADS / REG_NRG || T > REG_MAX_T) return 0;
    const int per = (T + REG_NRG - 1) / REG_NRG;
    for (int R : {8, 16, 32, 64})
        if (per <= R) return R;
    return 0;
}

hipError_t launch_peel(const PeelBatch& b, int R, bool cooperative, hipStream_t s) {
    if (R == 64)                            // the 64-row instance assumes one template subset
        for (int j = 0; j < b.n; ++j)
            if (b.p[j].S != 1) return hipErrorInvalidValue;
    PeelBatch copy = b;
    void* args[] = {&copy};
    const int G = b.g_start[b.n];
    const void* fn = peel_fn(R, b.p[0].rescore != 0, b.p[0].clock != 0, b.p[0].lss != 0);
    if (cooperative) return hipLaunchCooperativeKernel(fn, dim3(G), dim3(PEEL_THREADS), args, 0, s);
    return hipLaunchKernel(fn, dim3(G), dim3(PEEL_THREADS), args, 0, s);
}

int peel_max_blocks(int device, int R) {
    int per_cu = 1 << 30, cus = 0;
    for (int v = 0; v < 6; ++v) {          // rescore x clock (Kilosort rules), clock (LSS rules)
        int k = 0;
        const bool lss = v >= 4, rescore = !lss && (v & 1), clock = lss ? v == 5 : (v >> 1) != 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, peel_fn(R, rescore, clock, lss), PEEL_THREADS, 0) !=
            hipSuccess)
            return 0;
        per_cu = k < per_cu ? k : per_cu;
    }
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    return per_cu * cus;
}

}  // namespace lss
