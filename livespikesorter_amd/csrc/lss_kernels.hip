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

// Xw[c][XOFF + w] = sum_c' Wr[c][c'] * x[w][cmap[c']]   (rows c < CP)
// Columns w >= W read zero input and so store zero: they are the right-hand
// zero padding the PC convolution reads.
__global__ __launch_bounds__(256) void k_whiten(WhitenBatch bt) {
    __shared__ GemmTile s;
    const int j = job_of(bt.start, bt.n, blockIdx.x);
    const WhitenJob& jb = bt.job[j];
    int mt, nt;
    if (!xcd_tile(blockIdx.x - bt.start[j], jb.mtiles, jb.ntiles, mt, nt)) return;
    const float* __restrict__ Wr = jb.Wr;
    const int16_t* __restrict__ x = jb.xptr ? *jb.xptr : jb.x;
    if (!x) return;                                             // stream idle this launch
    const int* __restrict__ cmap = jb.cmap;
    const int CP = jb.CP, W = jb.W, C = jb.C;
    const int64_t ld = jb.ld;
    const int m0 = mt * GEMM_BM, n0 = nt * GEMM_BN;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv >> 1, wn = wv & 1;
    const int ar = tid >> 2, ak = (tid & 3) * 4;   // A: row, 4 consecutive k
    const int bn = tid >> 2, bk = (tid & 3) * 4;   // B: column (sample), 4 consecutive channels
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    const int p0 = jb.kb_ptr[mt], p1 = jb.kb_ptr[mt + 1];
    float4 ra = make_float4(0.f, 0.f, 0.f, 0.f);
    float rb[4] = {0.f, 0.f, 0.f, 0.f};
    auto gload = [&](int kb) {
        const int k0 = kb * GEMM_BK;
        ra = *reinterpret_cast<const float4*>(Wr + (int64_t)(m0 + ar) * CP + k0 + ak);
        const int n = n0 + bn;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int c = k0 + bk + i;
            rb[i] = (n < W && c < C) ? (jb.xf ? jb.xf[(int64_t)c * jb.xfs + n] : (float)x[(int64_t)n * ld + cmap[c]])
                                     : 0.f;
        }
    };
    if (p0 < p1) gload(jb.kb_idx[p0]);
    for (int p = p0; p < p1; ++p) {
        __syncthreads();
        s.As[ak + 0][ar] = ra.x;
        s.As[ak + 1][ar] = ra.y;
        s.As[ak + 2][ar] = ra.z;
        s.As[ak + 3][ar] = ra.w;
#pragma unroll
        for (int i = 0; i < 4; ++i) s.Bs[bk + i][bn] = rb[i];
        __syncthreads();
        if (p + 1 < p1) gload(jb.kb_idx[p + 1]);
        acc = mfma_tile(s, acc, lane, wm, wn);
    }
    store_tile(jb.Xw + XOFF, jb.XS, m0, n0, acc, lane, wm, wn);
}

// B[t][w] = sum_kk A[t][kk] * Bm[kk][w]   (A = spatial templates [TP][K*CP], Bm = PC conv [K*CP][BS])
__global__ __launch_bounds__(256) void k_project(ProjectBatch bt) {
    __shared__ GemmTile s;
    const int j = job_of(bt.start, bt.n, blockIdx.x);
    const ProjectJob& jb = bt.job[j];
    int mt, nt;
    if (!xcd_tile(blockIdx.x - bt.start[j], jb.mtiles, jb.ntiles, mt, nt)) return;
    if (jb.gate && !jb.gate[nt * jb.gate_mt + mt]) return;     // tile not proposed: not computed
    if (jb.active && !*jb.active) return;                      // stream idle this launch
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
    const int p0 = jb.kb_ptr[mt], p1 = jb.kb_ptr[mt + 1];
    float4 ra = make_float4(0.f, 0.f, 0.f, 0.f), rb = ra;
    auto gload = [&](int kb) {
        const int k0 = kb * GEMM_BK;
        ra = *reinterpret_cast<const float4*>(A + (int64_t)(m0 + ar) * lda + k0 + ak);
        rb = *reinterpret_cast<const float4*>(Bm + (int64_t)(k0 + br) * ldb + n0 + bc);
    };
    if (p0 < p1) gload(jb.kb_idx[p0]);
    for (int p = p0; p < p1; ++p) {
        __syncthreads();
        s.As[ak + 0][ar] = ra.x;
        s.As[ak + 1][ar] = ra.y;
        s.As[ak + 2][ar] = ra.z;
        s.As[ak + 3][ar] = ra.w;
        *reinterpret_cast<float4*>(&s.Bs[br][bc]) = rb;
        __syncthreads();
        if (p + 1 < p1) gload(jb.kb_idx[p + 1]);
        acc = mfma_tile(s, acc, lane, wm, wn);
    }
    store_tile(jb.C, jb.ldc, m0, n0, acc, lane, wm, wn);
}

// ---------------------------------------------------------------------------
// 6-PC temporal convolution, 'same' zero padding, cross-correlation (no flip):
//   Bpc[k*CP + c][w] = sum_m Xw[c][w + m - M/2] * wPCA[k][m]
// One workgroup stages 256 + 64 samples of one whitened channel row in LDS
// (coalesced), each thread produces the K outputs of one column.
// ---------------------------------------------------------------------------
template <int KK>
__global__ __launch_bounds__(256) void k_pcconv(PcconvBatch bt) {
    __shared__ float xs[256 + 2 * XOFF];
    const int j = job_of(bt.start, bt.n, blockIdx.x);
    const PcconvJob& jb = bt.job[j];
    if (jb.active && !*jb.active) return;                      // stream idle this launch
    const int k = blockIdx.x - bt.start[j];
    const int c = k / jb.wblocks, w0 = (k % jb.wblocks) * 256, tid = threadIdx.x;
    const int M = jb.M;
    const float* __restrict__ wpca = jb.wpca;
    const float* row = jb.Xw + (int64_t)c * jb.XS + XOFF;      // row[w] valid for w >= -XOFF
    for (int i = tid; i < 256 + 2 * XOFF; i += 256) xs[i] = row[w0 - XOFF + i];
    __syncthreads();
    const int h = M / 2;
    float acc[KK];
#pragma unroll
    for (int q = 0; q < KK; ++q) acc[q] = 0.f;
    const float* xp = xs + tid + XOFF - h;
    for (int m = 0; m < M; ++m) {
        const float v = xp[m];
#pragma unroll
        for (int q = 0; q < KK; ++q) acc[q] = fmaf(wpca[q * M + m], v, acc[q]);
    }
#pragma unroll
    for (int q = 0; q < KK; ++q) jb.Bpc[(int64_t)(q * jb.CP + c) * jb.BS + w0 + tid] = acc[q];
    if (jb.Bpcb) {
#pragma unroll
        for (int q = 0; q < KK; ++q) jb.Bpcb[(int64_t)(q * jb.CP + c) * jb.BS + w0 + tid] = (__bf16)acc[q];
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
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int BK16 = 64;           // K depth of one BF16 step (four 16-blocks)
constexpr int BPAD = 8;

__global__ __launch_bounds__(256) void k_project_bf16(ProjectBatch bt) {
    __shared__ __attribute__((aligned(16))) __bf16 As[GEMM_BM][BK16 + BPAD];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[GEMM_BN][BK16 + BPAD];
    const int j = job_of(bt.start, bt.n, blockIdx.x);
    const ProjectJob& jb = bt.job[j];
    int mt, nt;
    if (!xcd_tile(blockIdx.x - bt.start[j], jb.mtiles, jb.ntiles, mt, nt)) return;
    if (jb.gate && !jb.gate[nt * jb.gate_mt + mt]) return;     // tile not proposed: not computed
    if (jb.active && !*jb.active) return;                      // stream idle this launch
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
    const int p0 = jb.kb_ptr[mt], p1 = jb.kb_ptr[mt + 1];
    for (int p = p0; p < p1; p += 4) {
        // gather: sub-block q of this step is K-block kb_idx[p + q] (zeros past the list end)
        {
            const int pa = p + aq;
            uint4 v0 = make_uint4(0, 0, 0, 0), v1 = v0;
            if (pa < p1) {
                const __bf16* src = A + (int64_t)(m0 + ar) * lda + jb.kb_idx[pa] * GEMM_BK;
                v0 = *reinterpret_cast<const uint4*>(src);
                v1 = *reinterpret_cast<const uint4*>(src + 8);
            }
            __syncthreads();
            *reinterpret_cast<uint4*>(&As[ar][aq * 16]) = v0;
            *reinterpret_cast<uint4*>(&As[ar][aq * 16 + 8]) = v1;
            const int pb = p + bq;
            __bf16 tmp[16];
            if (pb < p1) {
                const __bf16* src = Bm + (int64_t)(jb.kb_idx[pb] * GEMM_BK + brow) * ldb + n0 + bc;
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
    }
    store_tile(jb.C, jb.ldc, m0, n0, acc, lane, wm, wn);
    if (jb.C2) store_tile(jb.C2, jb.ldc, m0, n0, acc, lane, wm, wn);
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

__global__ void k_reset(unsigned* bar, int* iter_cnt, int n_iter, int* err, unsigned long long* prof) {
    for (int i = threadIdx.x; i < n_iter; i += blockDim.x) iter_cnt[i] = 0;
    if ((int)threadIdx.x < PROF_WORDS) prof[threadIdx.x] = 0ull;
    if ((int)threadIdx.x < 1 + 2 * BAR_MAX_GROUPS) bar[threadIdx.x * BAR_LINE] = 0u;
    if (threadIdx.x == 0) *err = 0;
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

__device__ __forceinline__ void wait_flags(const int* f, int n, int target, int* err) {
    if ((int)threadIdx.x < n) {
        unsigned spins = 0;
        while (ld_c(&f[threadIdx.x]) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > (1u << 26)) {
                __hip_atomic_fetch_or(err, ERR_BARRIER, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
    }
    __syncthreads();
}

__device__ __forceinline__ bool better(float v, int o, float bv, int bo) {
    return v > bv || (v == bv && o < bo);
}

// ---------------------------------------------------------------------------
// Persistent peel kernel: one workgroup owns CW consecutive columns of the
// score matrix B (all templates); only the owner reads or writes them.  With
// R > 0 the owned tile lives in registers for the whole batch (thread (col,
// rg) holds rows rg, rg + nrg, ...); R == 0 keeps it in global memory (large
// template counts).  Per iteration:
//   A  apply the previous iteration's spikes that reach the owned columns
//      (B[t][w] -= a * ctc[s][t][w - w_s + M], spikes in sample order), then
//      Cf = relu(B)^2 / nm (zero in the first and last M columns) and the
//      column max / first-index argmax over templates;        -- barrier --
//   B  Cmax = max of Cfmax over [w-M, w+M]; accept w when Cmax > Th^2 and
//      Cmax == Cfmax[w] (|diff| < 1e-9); amplitude B[t][w] / nm[t];
//                                                              -- barrier --
//   the next iteration starts with one coherent round trip that reads the
//   spike total (stop when zero) and the neighbours' new spikes.
// One owner per column and a fixed application order make the result
// deterministic; no data atomics are used.  Afterwards each workgroup sorts
// its spikes by (column, iteration), writes the events, and computes the PC
// features of the final residual.
// ---------------------------------------------------------------------------
constexpr int PRE = 4;   // neighbour spikes fetched speculatively with the counts

__device__ __forceinline__ uint64_t rt_now() { return __builtin_amdgcn_s_memrealtime(); }

template <int R>
__global__ __launch_bounds__(PEEL_THREADS) void k_peel(PeelBatch bt) {
    // Every product here is rounded before it is added or compared, as in the
    // FP32 reference; a fused multiply-add would move tie decisions by an ulp.
#pragma clang fp contract(off)
    const int sj = job_of(bt.g_start, bt.n, blockIdx.x);
    const PeelParams& p = bt.p[sj];
    if (p.active && !*p.active) return;                         // stream idle: every workgroup leaves
    // LDS is kept under 64 KB so every table is reachable with ds_read's 16-bit
    // immediate offset (no per-row address registers); the norm table comes first.
    __shared__ float s_nm[R > 0 ? REG_MAX_T : 1];    // per-template reciprocal norm (0 past T)
    __shared__ int s_or[R > 0 ? REG_MAX_T : 1];      // per-template file index (INT_MAX past T)
    __shared__ Spike s_sp[MAX_STAGE];          // staged / newly accepted spikes
    __shared__ Spike s_list[MAX_LIST];         // this workgroup's spikes, whole batch
    __shared__ Spike s_pre[MAX_NBR * PRE];
    __shared__ float s_val[PEEL_THREADS];      // also the rank table of the epilogue
    __shared__ int s_idx[PEEL_THREADS];
    __shared__ int s_org[PEEL_THREADS];
    __shared__ float s_bvv[PEEL_THREADS];
    __shared__ float s_win[MAX_CW + 2 * MAX_M + 2];
    __shared__ float s_cf[MAX_CW];             // own column maxima, their argmax and its score
    __shared__ int s_im[MAX_CW];
    __shared__ float s_bv[MAX_CW];
    __shared__ float s_cm[MAX_CW];
    __shared__ int s_flag[MAX_CW];
    __shared__ int s_rt[MAX_CW];               // FP32 rescoring: template whose correction is cached
    __shared__ float s_rd[MAX_CW];             //   and the correction (FP32 score - BF16 score)
    __shared__ int s_hcnt[MAX_NBR + 1];
    __shared__ int s_hcnt_w[MAX_CW / 64];
    __shared__ int s_qlo[R == 0 ? MAX_STAGE : 1], s_qhi[R == 0 ? MAX_STAGE : 1];   // staged spikes' row bands
    __shared__ int s_whit[R > 0 ? PEEL_THREADS : 1];   // per-wave list of band-hitting spikes
    __shared__ int s_nsp, s_listn, s_off, s_n, s_tot, s_err, s_ncand, s_kchk, s_cntadd;
    __shared__ unsigned long long s_wkey[R > 0 ? PEEL_THREADS / REG_NRG : 1];   // split: winner key per column
    __shared__ int s_wt[R > 0 ? PEEL_THREADS / REG_NRG : 1];                   //   its template
    __shared__ float s_wb[R > 0 ? PEEL_THREADS / REG_NRG : 1];                 //   and its B value
    int* const s_rank = reinterpret_cast<int*>(s_val);
    static_assert(MAX_LIST <= PEEL_THREADS, "rank table reuses the reduction buffer");

    const int g = blockIdx.x - bt.g_start[sj], tid = threadIdx.x;
    // Workgroup g owns column block cb = g / S and template subset sb = g % S:
    // rows [sb * REG_NRG * R, (sb + 1) * REG_NRG * R) of the block's columns.
    // Subset 0 is the block's owner: it accepts spikes, publishes them and
    // writes the events; the other subsets publish per-column partial maxima.
    const int S = p.S;
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
    unsigned bar_gen = 0;
    if (tid == 0) s_listn = 0;
    for (int i = tid; i < MAX_CW; i += PEEL_THREADS) s_rt[i] = -1;

    // Buffer descriptors: one scalar base per table plus a per-thread offset, so
    // the R rows of a tile need no address registers (row strides go in SGPRs).
    const __amdgpu_buffer_rsrc_t ctc_rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p.ctc), 0, p.ctc_bytes, BUF_DWORD3);
    float breg[R > 0 ? R : 1];
    if (R > 0) {
        for (int t = tid; t < REG_NRG * R; t += PEEL_THREADS) {
            s_nm[t] = p.rnm[rbase + t];
            s_or[t] = p.orig[rbase + t];
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

    // Publish own column cl's maximum: the column maximum itself (one subset) or
    // this subset's partial (score, template, B value).
    auto publish = [&](int cl) {
        const int ww = c0 + cl;
        if (S == 1) {
            st_c(&p.Cfmax[ww], s_cf[cl]);
        } else {
            const int64_t k = (int64_t)sb * W + ww;
            st_c(&p.Pcf[k], s_cf[cl]);
            st_c(&p.Pim[k], s_im[cl]);
            st_c(&p.Pbv[k], s_bv[cl]);
        }
    };

    // The winning partial of column ww over the subsets: largest score, smaller
    // file index on ties (the argmax over all templates).
    auto subset_winner = [&](int ww, float& v, int& t, float& b) {
        v = -1.f;
        int bo2 = 0x7fffffff;
        for (int s2 = 0; s2 < S; ++s2) {
            const int64_t k = (int64_t)s2 * W + ww;
            const float c = ld_c(&p.Pcf[k]);
            const int tt = ld_c(&p.Pim[k]);
            const int o = p.orig[tt];
            if (c > v || (c == v && o < bo2)) { v = c; bo2 = o; t = tt; b = ld_c(&p.Pbv[k]); }
        }
    };

    if (tid == 0) s_kchk = 0;
    // per iteration: owners done with phase B (high word) and spikes accepted (low word)
    unsigned long long* const itc = reinterpret_cast<unsigned long long*>(p.iter_cnt);
    int* const flagB = p.iter_cnt + 2 * p.max_iter;
    int* const flagA = flagB + p.NB;
    // phase clock: thread 0 accumulates in LDS (no registers held across the loop)
    const bool timing = (tid == 0);
    __shared__ uint64_t s_ph[NPHASE + 2];      // phases, then the previous mark and the pass count
    if (tid < NPHASE + 2) s_ph[tid] = 0;
    __syncthreads();
    if (timing) s_ph[NPHASE] = rt_now();
    auto mark = [&](int k) {
        if (timing) {
            const uint64_t t = rt_now();
            s_ph[k] += t - s_ph[NPHASE];
            s_ph[NPHASE] = t;
        }
    };
    for (int it = 0; it <= p.max_iter; ++it) {
        int nsp = 0;
        if (it > 0) {
            // one coherent round trip: spike total of the last iteration, the
            // neighbours' counts, and their first PRE spikes
            if (tid == 0) {
                // Stop once some iteration is known to have accepted no spike anywhere:
                // all owners are done with it (done_cnt) and its total is zero.  Every
                // later iteration then changes nothing, so when a workgroup notices
                // does not change any result.
                int stop = 0;
                while (s_kchk < it) {
                    const unsigned long long v = ld_c(&itc[s_kchk]);
                    if ((int)(v >> 32) < p.NB) break;
                    if ((unsigned)v == 0u) { stop = 1; break; }
                    ++s_kchk;
                }
                s_tot = !stop;
                s_err = ld_c(p.err);
            }
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
            // a wave whose 64-template tile the gate did not propose for this column
            // block scores zero everywhere (its stale rows are neither updated nor read)
            const bool woff = R > 0 && p.gate && !p.gate[cb * p.gate_mt + (rbase + rg * R) / GEMM_BM];
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
                        const int st = s_sp[qq].t;
                        bh = r0 < p.thi[st] && r0 + R > p.tlo[st];
                    }
                    const uint64_t bal = __ballot(bh);
                    if (bh) whit[__popcll(bal & ((1ull << lane) - 1ull))] = qq;
                    const int nhit = __popcll(bal);
                    wdirty = wdirty || nhit > 0;
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_s_waitcnt(0xc07f);     // lgkmcnt(0): the list is in LDS
                    for (int h0 = 0; h0 < nhit; h0 += 4) {
                        float am[4];
                        int vo[4];
    #pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const bool in = h0 + j < nhit;
                            const Spike sp = s_sp[whit[in ? h0 + j : h0]];
                            const int l = w - sp.w + M;
                            const bool hit = in && l >= 0 && l <= 2 * M;
                            am[j] = hit ? sp.amp : 0.f;
                            vo[j] = ((sp.t * p.ctc_rows + r0) * LP + (hit ? l : 0)) * 4;
                        }
                        constexpr int RC = R < 8 ? R : (R >= 64 ? 4 : 8);    // rows per load batch
    #pragma unroll
                        for (int i0 = 0; i0 < R; i0 += RC) {
                            float c[4][RC];
    #pragma unroll
                            for (int j = 0; j < 4; ++j)
    #pragma unroll
                                for (int i = 0; i < RC; ++i)
                                    c[j][i] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                                        ctc_rs, vo[j], (i0 + i) * LP * 4, 0));
    #pragma unroll
                            for (int i = 0; i < RC; ++i)
    #pragma unroll
                                for (int j = 0; j < 4; ++j) breg[i0 + i] -= am[j] * c[j][i];
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
                    const int* orp = s_or + vb;
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
                    const float thr = p.lss ? amax : amax * (1.f - 0x1p-20f);
                    if (a2 < thr) {
                        const float r = fmaxf(bval, 0.f);
                        best = p.lss ? amax : (r * r) / p.nm[rbase + vb + bslot];
                        bo = orp[bslot];
                    } else {
                        best = -1.f;
                        bo = 0x7fffffff;
                        const float* nmp = p.nm + rbase + vb;
                        // recompute the squares here instead of keeping 64 of them live
#pragma unroll
                        for (int i = 0; i < R; ++i) asm volatile("" : "+v"(breg[i]));
#pragma unroll
                        for (int i0 = 0; i0 < R; i0 += 8) {
#pragma unroll
                            for (int i = i0; i < i0 + 8; ++i) {
                                const float r = fmaxf(breg[i], 0.f);
                                const float a = (r * r) * rnp[i];
                                const float cf = p.lss ? a : (r * r) / nmp[i];
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
                    for (int q = qa; q < qb; ++q) {
                        const Spike sp = s_sp[q];
                        const int l = w - sp.w + M;                 // in [0, 2M] for q in [qa, qb)
                        const int t0 = s_qlo[q], t1 = s_qhi[q];
                        const float* cp = p.ctc + (int64_t)sp.t * p.ctc_rows * LP + l;
                        for (int t = t0 + (rg - t0 % nrg + nrg) % nrg; t < t1; t += nrg) {
                            if (p.gate && !p.gate[(w / GEMM_BN) * p.gate_mt + t / GEMM_BM]) continue;
                            float* bp = p.B + (int64_t)t * p.BS + w;
                            *bp -= sp.amp * cp[(int64_t)t * LP];
                        }
                    }
                    for (int t = rg; t < T; t += nrg) {
                        if (p.gate && !p.gate[(w / GEMM_BN) * p.gate_mt + t / GEMM_BM]) {
                            if (better(0.f, p.orig[t], best, bo)) { best = 0.f; bi = t; bo = p.orig[t]; bval = 0.f; }
                            continue;
                        }
                        const float b = p.B[(int64_t)t * p.BS + w];
                        float cf = 0.f;
                        if (!edge) {
                            const float r = fmaxf(b, 0.f);
                            cf = p.lss ? (r * r) * p.rnm[t] : (r * r) / p.nm[t];
                        }
                        const int o = p.orig[t];
                        if (better(cf, o, best, bo)) { best = cf; bi = t; bo = o; bval = b; }
                    }
                }
            }
            if (wdirty) { s_val[tid] = best; s_idx[tid] = bi; s_org[tid] = bo; s_bvv[tid] = bval; }
            __syncthreads();
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
                float bv = pv[0], bb = pb[0];
                int bI = pi[0], bO = po[0];
                for (int r = 1; r < nrg; ++r) {
                    const int j = r * CW;
                    if (better(pv[j], po[j], bv, bO)) { bv = pv[j]; bI = pi[j]; bO = po[j]; bb = pb[j]; }
                }
                const int ww = c0 + tid;
                if (ww < c1) {
                    s_cf[tid] = bv;
                    s_im[tid] = bI;
                    s_bv[tid] = bb;
                    if (!p.rescore) publish(tid);
                    if (it == 0 && p.dbg_cfmax && S == 1) { p.dbg_cfmax[ww] = bv; p.dbg_imax[ww] = bI; }
                }
            }
            if (p.rescore) {
                // BF16 projection: every column whose best score reaches lo2 (10% under
                // threshold) is rescored in FP32 before it is published, so pooling,
                // acceptance and amplitudes see FP32 scores wherever they can matter
                // (a column under lo2 in BF16 cannot beat an accepted one in FP32).
                // The correction of (t, w) is fixed (subtractions are applied to both
                // alike), so it is computed once per (column, template) and cached.
                // The columns' own threads (tid < CW, written just above) list the
                // columns that need a dot product; one barrier publishes the list.
                const int lane = tid & 63, wv = tid >> 6;
                bool need = false;
                if (tid < CW && c0 + tid < c1) need = s_cf[tid] >= p.lo2 && s_rt[tid] != s_im[tid];
                const uint64_t bal = __ballot(need);
                if (tid < CW && lane == 0) s_hcnt_w[wv] = __popcll(bal);
                if (CW > 64) __syncthreads();            // wave counts of other waves
                int nres = 0, base = 0;
                if (tid < CW) {
                    for (int k = 0; k < (CW + 63) / 64; ++k) {
                        if (k == wv) base = nres;
                        nres += s_hcnt_w[k];
                    }
                    if (need) s_flag[base + __popcll(bal & ((1ull << lane) - 1ull))] = tid;
                    if (tid == 0) s_ncand = nres;
                }
                __syncthreads();
                nres = s_ncand;
                for (int q = wv; q < nres; q += PEEL_THREADS / 64) {
                    const int cl = s_flag[q];
                    const int t = s_im[cl], ww = c0 + cl;
                    // FP32 dot of U2 row t with Bpc column ww over the K-blocks listed for
                    // t's 64-row tile (the others are exact zeros)
                    const int mt = t / GEMM_BM;
                    const int e0 = p.kb_ptr[mt] * GEMM_BK, e1 = p.kb_ptr[mt + 1] * GEMM_BK;
                    const float* urow = p.U2 + (int64_t)t * p.ldu;
                    // 8 elements per lane in flight at once: the block index, then the
                    // template weight and the (scattered) Bpc column entry
                    float acc = 0.f;
                    for (int e = e0 + lane; e < e1; e += 64 * 8) {
                        float uq[8], bq[8];
#pragma unroll
                        for (int q = 0; q < 8; ++q) {
                            const int ee = e + 64 * q;
                            const int kc = ee < e1 ? p.kb_idx[ee / GEMM_BK] * GEMM_BK + ee % GEMM_BK : -1;
                            uq[q] = kc >= 0 ? urow[kc] : 0.f;
                            bq[q] = kc >= 0 ? p.Bpc[(int64_t)kc * p.BS + ww] : 0.f;
                        }
#pragma unroll
                        for (int q = 0; q < 8; ++q) acc += uq[q] * bq[q];
                    }
                    // butterfly: every lane ends with the same, order-fixed sum
                    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
                    if (lane == 0) {
                        s_rd[cl] = acc - p.Binit[(int64_t)t * p.BS + ww];
                        s_rt[cl] = t;
                    }
                }
                if (nres > 0) __syncthreads();
                int flip = 0;
                if (tid < CW && c0 + tid < c1) {
                    float cf = s_cf[tid];
                    if (cf >= p.lo2) {
                        const int t = s_im[tid];
                        const float b = s_bv[tid] + s_rd[tid];
                        const float r = fmaxf(b, 0.f);
                        const float c32 = (r * r) / p.nm[t];
                        flip = (cf > p.Th2) != (c32 > p.Th2);
                        s_bv[tid] = b;
                        s_cf[tid] = c32;
                    }
                    publish(tid);
                }
                const uint64_t fb = __ballot(flip);
                if (lane == 0 && wv == 0) {
                    if (nres > 0) atomicAdd(&p.n_rescored[0], (unsigned long long)nres);
                }
                if (lane == 0 && fb) atomicAdd(&p.n_rescored[1], (unsigned long long)__popcll(fb));
            }
        }
        mark(1);
        // phase A done: the owners of this and the neighbouring blocks may read this
        // workgroup's column maxima (or partials); an owner waits for every workgroup
        // of blocks hlo..hhi (contiguous ids) before pooling over +-M columns
        raise_flag(&flagA[g], it + 1);
        if (owner) wait_flags(flagA + hlo * S, (hhi - hlo + 1) * S, it + 1, p.err);
        mark(2);

        // ---------------- phase B (the column block's owner) ----------------
        const int nwin = (c1 - c0) + 2 * M;
        if (S == 1) {
            for (int i = tid; i < nwin; i += PEEL_THREADS) {
                const int ww = c0 - M + i;
                float v = -INFINITY;
                if (ww >= c0 && ww < c1) v = s_cf[ww - c0];
                else if (ww >= 0 && ww < W) v = ld_c(&p.Cfmax[ww]);
                s_win[i] = v;
            }
        } else if (owner) {
            // column maxima = max over the subsets' partials; scores are >= 0, so their
            // bit patterns order as integers (columns outside the batch stay at 0)
            int* const s_wi = reinterpret_cast<int*>(s_win);
            for (int i = tid; i < nwin; i += PEEL_THREADS) s_wi[i] = 0;
            __syncthreads();
            if (R > 0 && tid < CW) s_wkey[tid] = 0ull;
            __syncthreads();
            for (int i = tid; i < nwin * S; i += PEEL_THREADS) {
                const int s2 = i / nwin, iw = i - s2 * nwin, ww = c0 - M + iw;
                if (ww >= 0 && ww < W) atomicMax(&s_wi[iw], __float_as_int(ld_c(&p.Pcf[(int64_t)s2 * W + ww])));
            }
            if (R > 0) {
                // the winning partial of every own column, from the same round trip:
                // key = (score bits, ~file index), so the largest score wins and the
                // smaller file index breaks ties (scores are >= 0)
                const int s2 = tid / CW, cl = tid - s2 * CW;
                float v = 0.f, b = 0.f;
                int t = 0;
                const bool in = s2 < S && c0 + cl < c1;
                if (in) {
                    const int64_t k = (int64_t)s2 * W + c0 + cl;
                    v = ld_c(&p.Pcf[k]);
                    t = ld_c(&p.Pim[k]);
                    b = ld_c(&p.Pbv[k]);
                }
                const unsigned long long key =
                    ((unsigned long long)__float_as_uint(v) << 32) | (unsigned)(0x7fffffff - s_or[in ? t : 0]);
                if (in) atomicMax(&s_wkey[cl], key);
                __syncthreads();
                if (in && s_wkey[cl] == key) {
                    s_wt[cl] = t;
                    s_wb[cl] = b;
                }
            }
        }
        if (!owner) {
            // the owner accepts and publishes; nothing to do until the barrier
        } else {
        if (tid == 0) s_ncand = 0;
        __syncthreads();
        float* const s_part = reinterpret_cast<float*>(s_whit);     // free in phase B (R > 0)
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
                f = p.lss ? (cm >= p.Th2 && (cm - cfw) < 1e-4f) : (cm > p.Th2 && fabsf(cm - cfw) < 1e-9f);
                s_cm[tid] = cm;
                if (S > 1 && it == 0 && p.dbg_cfmax) {
                    float v; int t; float b;
                    subset_winner(c0 + tid, v, t, b);
                    p.dbg_cfmax[c0 + tid] = v;
                    p.dbg_imax[c0 + tid] = t;
                }
            }
            s_flag[tid] = f;
        }
        __syncthreads();
        if (p.lss) {
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
        }
        if (!p.lss) {
            // one spike per accepted column: compact the flags with per-wave ballots
            const int lane = tid & 63, wv = tid >> 6;
            const int nwv = (CW + 63) / 64;
            const bool f = tid < CW && s_flag[tid];
            const uint64_t bal = __ballot(f);
            if (tid < CW && lane == 0) s_hcnt_w[wv] = __popcll(bal);
            __syncthreads();
            if (tid < CW) {
                int base = 0;
                for (int k = 0; k < wv; ++k) base += s_hcnt_w[k];
                if (f) {
                    const int idx = base + __popcll(bal & ((1ull << lane) - 1ull));
                    if (S == 1) {
                        const int t = s_im[tid];
                        const Spike sp{c0 + tid, t, s_bv[tid] / p.nm[t], it};
                        if (idx < p.slot_cap) s_sp[idx] = sp;
                        const int ln = s_listn + idx;
                        if (ln < MAX_LIST) s_list[ln] = sp;
                    } else {
                        s_sp[idx] = Spike{c0 + tid, 0, 0.f, it};   // idx < CW <= MAX_STAGE
                    }
                }
            }
            __syncthreads();
            if (S > 1) {
                // template and amplitude of each accepted column from the subsets'
                // partials: one wave per column, one lane per subset, then a
                // butterfly argmax (largest score, smaller file index on ties)
                int n = 0;
                for (int k = 0; k < nwv; ++k) n += s_hcnt_w[k];
                for (int q = tid; q < n; q += PEEL_THREADS) {
                    const int cl = s_sp[q].w - c0;
                    const int t = s_wt[cl];
                    s_sp[q].t = t;
                    s_sp[q].amp = s_wb[cl] / p.nm[t];
                }
                __syncthreads();
                for (int i = tid; i < min(n, p.slot_cap); i += PEEL_THREADS)
                    if (s_listn + i < MAX_LIST) s_list[s_listn + i] = s_sp[i];
                __syncthreads();        // s_listn is advanced below
            }
            if (tid == 0) {
                int n = 0;
                for (int k = 0; k < nwv; ++k) n += s_hcnt_w[k];
                if (n > p.slot_cap) atomicOr(p.err, ERR_SLOT);
                if (s_listn + n > MAX_LIST) atomicOr(p.err, ERR_LIST);
                s_n = min(n, p.slot_cap);
                s_listn = min(s_listn + n, MAX_LIST);
                s_cntadd = n;
            }
        } else if (tid == 0) {
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
            s_cntadd = n;
        }
        __syncthreads();
        {
            const int n = s_n;
            for (int i = tid; i < n; i += PEEL_THREADS) st_spike(&p.slots[(int64_t)cb * p.slot_cap + i], s_sp[i]);
            if (tid == 0) st_c(&p.slot_cnt[cb], n);
        }
        }   // owner
        mark(3);
        if (owner) {
            // one atomic counts the owner done and adds its spikes, so a reader that
            // sees every owner done also sees the complete total
            if (tid == 0)
                __hip_atomic_fetch_add(&itc[it], (1ull << 32) | (unsigned)s_cntadd, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            raise_flag(&flagB[cb], it + 1);
        }
        // the next exchange reads the slots of blocks hlo..hhi; waiting for their owners
        // also keeps this workgroup's next phase A from overwriting column maxima that
        // a neighbouring owner is still pooling
        wait_flags(flagB + hlo, hhi - hlo + 1, it + 1, p.err);
        mark(4);
    }
    mark(0);
    raise_flag(&flagA[g], FLAG_DONE);
    if (owner && tid == 0) st_c(&flagB[cb], FLAG_DONE);

    // ---------------- events and features ----------------
    // The owners publish their lists.  After the barrier every workgroup of a
    // column block holds the block's list and takes a share of its spikes'
    // features (spike i goes to subset i % S), so a block dense in spikes is
    // spread over its S workgroups; the owner writes the events.
    if (owner) {
        const int nl = min(s_listn, MAX_LIST);
        for (int i = tid; i < nl; i += PEEL_THREADS) st_spike(&p.list[(int64_t)cb * p.cap_g + i], s_list[i]);
        if (tid == 0) st_c(&p.list_cnt[cb], nl);
    }
    grid_sync(p.bar, p.G, g, bar_gen, p.err);
    if (!owner) {
        if (tid == 0) s_listn = min(ld_c(&p.list_cnt[cb]), MAX_LIST);
        __syncthreads();
        for (int i = tid; i < s_listn; i += PEEL_THREADS) s_list[i] = ld_spike(&p.list[(int64_t)cb * p.cap_g + i]);
    }
    {   // spikes of the blocks before this one: the block's offset in the event list
        int part = 0;
        for (int hh = tid; hh < cb; hh += PEEL_THREADS) part += ld_c(&p.list_cnt[hh]);
        s_idx[tid] = part;
        __syncthreads();
        for (int o = PEEL_THREADS / 2; o > 0; o >>= 1) {
            if (tid < o) s_idx[tid] += s_idx[tid + o];
            __syncthreads();
        }
        if (tid == 0) s_off = s_idx[0];
        __syncthreads();
    }
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
        const int flo = max(0, (c0 - (M - 1)) / CW), fhi = min(p.NB - 1, (c1 - 1 + M - 1) / CW);
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
                f = p.Bpc[(int64_t)(k * p.CP + ch) * p.BS + si.w] + si.amp * p.U[((int64_t)si.t * K + k) * C + ch];
            // the spikes within +-(M-1) columns, in list order: a wave reads 64 at a time
            // and takes the near ones from the ballot, lowest index first
            for (int hh = flo; hh <= fhi; ++hh) {
                const int cnt = hh == cb ? nmine : ld_c(&p.list_cnt[hh]);
                const Spike* lst = p.list + (int64_t)hh * p.cap_g;
                for (int q0 = 0; q0 < cnt; q0 += 64) {
                    const int q = q0 + lane;
                    Spike sq{};
                    bool near = false;
                    if (q < cnt) {
                        sq = hh == cb ? s_list[q] : ld_spike(&lst[q]);
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
                            float gs = 0.f;
                            for (int k2 = 0; k2 < K; ++k2)
                                gs += p.U[((int64_t)tq * K + k2) * C + ch] * p.Gt[(k2 * K + k) * L2 + dd + M - 1];
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

// One block per stream: the per-batch counters of every stream of a group.
__global__ void k_reset_batch(ResetBatch b) {
    const int j = blockIdx.x;
    if (j >= b.n) return;
    if (b.active[j] && !*b.active[j]) return;     // idle stream (maybe running in another launch)
    for (int i = threadIdx.x; i < b.n_iter[j]; i += blockDim.x) b.iter_cnt[j][i] = 0;
    if ((int)threadIdx.x < PROF_WORDS) b.prof[j][threadIdx.x] = 0ull;
    if ((int)threadIdx.x < 1 + 2 * BAR_MAX_GROUPS) b.bar[j][threadIdx.x * BAR_LINE] = 0u;
    if (threadIdx.x == 0) *b.err[j] = 0;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_reset(unsigned* bar, int* iter_cnt, int n_iter, int* err, unsigned long long* prof, hipStream_t s) {
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(256), 0, s, bar, iter_cnt, n_iter, err, prof);
}

void launch_reset_batch(const ResetBatch& b, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_batch, dim3(b.n), dim3(256), 0, s, b);
}

void launch_whiten(const WhitenBatch& b, hipStream_t s) {
    hipLaunchKernelGGL(k_whiten, dim3(b.start[b.n]), dim3(256), 0, s, b);
}

void launch_pcconv(const PcconvBatch& b, int K, hipStream_t s) {
    switch (K) {
#define LSS_PC_CASE(KK) \
    case KK: hipLaunchKernelGGL(k_pcconv<KK>, dim3(b.start[b.n]), dim3(256), 0, s, b); break;
        LSS_PC_CASE(1) LSS_PC_CASE(2) LSS_PC_CASE(3) LSS_PC_CASE(4)
        LSS_PC_CASE(5) LSS_PC_CASE(6) LSS_PC_CASE(7) LSS_PC_CASE(8)
#undef LSS_PC_CASE
        default: break;
    }
}

void launch_project(const ProjectBatch& b, bool bf16, hipStream_t s) {
    if (bf16) hipLaunchKernelGGL(k_project_bf16, dim3(b.start[b.n]), dim3(256), 0, s, b);
    else hipLaunchKernelGGL(k_project, dim3(b.start[b.n]), dim3(256), 0, s, b);
}

void launch_norms(const float* U, int KC, int T, float* nm, float* rnm, double* nm64, hipStream_t s) {
    hipLaunchKernelGGL(k_norms, dim3(T), dim3(256), 0, s, U, KC, nm, rnm, nm64);
}

static const void* peel_fn(int R) {
    switch (R) {
        case 8: return reinterpret_cast<const void*>(k_peel<8>);
        case 16: return reinterpret_cast<const void*>(k_peel<16>);
        case 32: return reinterpret_cast<const void*>(k_peel<32>);
        case 64: return reinterpret_cast<const void*>(k_peel<64>);
        default: return reinterpret_cast<const void*>(k_peel<0>);
    }
}

int peel_rows(int T, int CW) {
    if (CW != PEEL_THREADS / REG_NRG || T > REG_MAX_T) return 0;
    const int per = (T + REG_NRG - 1) / REG_NRG;
    for (int R : {8, 16, 32, 64})
        if (per <= R) return R;
    return 0;
}

hipError_t launch_peel(const PeelBatch& b, int R, bool cooperative, hipStream_t s) {
    PeelBatch copy = b;
    void* args[] = {&copy};
    const int G = b.g_start[b.n];
    if (cooperative)
        return hipLaunchCooperativeKernel(peel_fn(R), dim3(G), dim3(PEEL_THREADS), args, 0, s);
    return hipLaunchKernel(peel_fn(R), dim3(G), dim3(PEEL_THREADS), args, 0, s);
}

int peel_max_blocks(int device, int R) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, peel_fn(R), PEEL_THREADS, 0) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    return per_cu * cus;
}

}  // namespace lss
