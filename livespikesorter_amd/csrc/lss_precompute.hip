// Load-time precompute from the raw Kilosort outputs (Wall3, wPCA):
//   ctc[t][s][l] = sum_{k1,k2} (sum_c U[t,k1,c] U[s,k2,c]) WtW[k1][k2][l]
// (KS/template_matching.py:110-121, read by OnlineSpikesV2.cu:604-620 from ctc.npy),
// accumulated in FP64 and rounded once to FP32, written straight into the peel
// kernel's layout ctc_i[s'][t'][l] (internal template order, LP-padded lags).
// (KS/ = src/Python/run_online_spikes/Lib/site-packages/kilosort/)

#include "lss_kernels.h"

namespace lss {

constexpr int CTC_WAVES = 4;

// One workgroup per s'; each wave takes t' = wave, wave + 4, ...  The K x K
// channel dot products of a pair run over the overlap of the two templates'
// nonzero channel spans only (disjoint spans give an exact-zero row, left as
// the zero the output was cleared to), lanes striding over channels, then a
// butterfly sum leaves every product in every lane; lanes then own lags
// l = lane and lane + 64.  tlo/thi collect the t' range of each s' whose rows
// matter: nonzero, or (tau > 0) with an entry above tau * sqrt(nm[s'] nm[t']).
template <int K>
__global__ __launch_bounds__(64 * CTC_WAVES) void k_ctc(const float* __restrict__ U, const double* __restrict__ wtw,
                                                        const int2* __restrict__ span, int T, int C, int L, int TS,
                                                        float* __restrict__ ctc, int* __restrict__ tlo,
                                                        int* __restrict__ thi, const double* __restrict__ nm,
                                                        double tau) {
    extern __shared__ double s_wtw[];   // [K*K][L]
    const int sp = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < K * K * L; i += 64 * CTC_WAVES) s_wtw[i] = wtw[i];
    __syncthreads();
    const int2 ss = span[sp];
    const float* us = U + (int64_t)sp * K * C;
    int lo = T, hi = 0;
    for (int tp = wave; tp < T; tp += CTC_WAVES) {
        const int2 st = span[tp];
        const int c0 = max(ss.x, st.x), c1 = min(ss.y, st.y);
        if (c0 >= c1) continue;
        const float* ut = U + (int64_t)tp * K * C;
        double acc[K][K];
#pragma unroll
        for (int a = 0; a < K; ++a)
#pragma unroll
            for (int b = 0; b < K; ++b) acc[a][b] = 0.0;
        for (int c = c0 + lane; c < c1; c += 64) {
            double vt[K], vs[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                vt[k] = ut[(int64_t)k * C + c];
                vs[k] = us[(int64_t)k * C + c];
            }
#pragma unroll
            for (int a = 0; a < K; ++a)
#pragma unroll
                for (int b = 0; b < K; ++b) acc[a][b] = fma(vt[a], vs[b], acc[a][b]);
        }
#pragma unroll
        for (int a = 0; a < K; ++a)
#pragma unroll
            for (int b = 0; b < K; ++b)
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) acc[a][b] += __shfl_xor(acc[a][b], o, 64);
        bool nz = false;
        const double cut = tau > 0.0 ? tau * sqrt(nm[sp] * nm[tp]) : 0.0;
        float* out = ctc + ((int64_t)sp * TS + tp) * LP;
        for (int l = lane; l < L; l += 64) {
            double v = 0.0;
#pragma unroll
            for (int a = 0; a < K; ++a)
#pragma unroll
                for (int b = 0; b < K; ++b) v = fma(acc[a][b], s_wtw[(a * K + b) * L + l], v);
            const float f = (float)v;
            out[l] = f;
            nz |= fabs((double)f) > cut;
        }
        if (__any(nz)) {
            lo = min(lo, tp);
            hi = tp + 1;
        }
    }
    __shared__ int s_lo[CTC_WAVES], s_hi[CTC_WAVES];
    if (lane == 0) {
        s_lo[wave] = lo;
        s_hi[wave] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < CTC_WAVES; ++w) {
            lo = min(lo, s_lo[w]);
            hi = max(hi, s_hi[w]);
        }
        tlo[sp] = lo;
        thi[sp] = hi;
    }
}

// U [T][K][C] FP32 (internal order), wtw [K*K][L] FP64, span [T] nonzero channel
// range of each template; ctc [T][TS][LP] must be zero on entry.
void launch_ctc(const float* U, const double* wtw, const int2* span, int T, int K, int C, int M, int TS, float* ctc,
                int* tlo, int* thi, const double* nm, double tau, hipStream_t s) {
    const int L = 2 * M + 1;
    const size_t lds = sizeof(double) * K * K * L;
    const dim3 grid(T), block(64 * CTC_WAVES);
    switch (K) {
#define LSS_CTC_CASE(k) \
    case k:                                                                                    \
        (void)hipFuncSetAttribute((const void*)k_ctc<k>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
        k_ctc<k><<<grid, block, lds, s>>>(U, wtw, span, T, C, L, TS, ctc, tlo, thi, nm, tau);  \
        break;
        LSS_CTC_CASE(1) LSS_CTC_CASE(2) LSS_CTC_CASE(3) LSS_CTC_CASE(4)
        LSS_CTC_CASE(5) LSS_CTC_CASE(6) LSS_CTC_CASE(7) LSS_CTC_CASE(8)
#undef LSS_CTC_CASE
        default: break;
    }
}

}  // namespace lss
