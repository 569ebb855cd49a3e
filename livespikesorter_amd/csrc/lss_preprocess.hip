// Preprocessing in front of whitening, per batch of W samples x C channels:
//   channel means   OnlineSpikesV2.cu:774-779, myGPUhelpers.cu:99-119,704-712
//                   (src/Python/OnlineKilosort.py:118)               -> k_pre_stats
//   CAR (lower median over channels per sample)
//                   OnlineSpikesV2.cu:782-786, myGPUhelpers.cu:480-543
//                   (OnlineKilosort.py:119-121)                      -> k_pre_car
//   high-pass       OnlineSpikesV2.cu:789-794,1006-1056 (cuFFT circular filter +
//                   shift_kernel, myGPUhelpers.cu:72-83); OnlineKilosort.py:124-128
//                   with fft_highpass (KS/preprocessing.py:139-154)  -> k_pre_fir
//   peak-to-peak    OnlineSpikesV2.cu:771,1137-1146 (NPP max/min)    -> k_pre_stats/k_pre_car
//
// The FFT filter of the references is the circular correlation
//   out[k] = sum_m x[(k + s + m) mod W] h[m],   h = impulse response cropped to W,
// s = -(W/2) for Kilosort's fftshift and +(W/2) for LSS's shift_kernel.  Here it is
// computed directly over the taps whose magnitude matters (the host drops the
// response's tails, whose summed magnitude is below 2^-30 of the whole), in FP32
// with a register-resident sliding window; no transform, no complex buffers.
// (KS/ = src/Python/run_online_spikes/Lib/site-packages/kilosort/)

#include "lss_kernels.h"

#include <mutex>
#include <set>
#include <utility>

namespace lss {

// Stats: per (slab of PRE_SLAB samples, 64 channels): integer sum, min and max.
// Lanes own channels (row reads are contiguous), the four waves interleave samples.
__global__ __launch_bounds__(256) void k_pre_stats(PreBatch bt) {
    const int j = [&] {
        int q = 0;
        while (q + 1 < bt.n && (int)blockIdx.x >= bt.start_stats[q + 1]) ++q;
        return q;
    }();
    const PreJob& jb = bt.job[j];
    const int16_t* __restrict__ x = jb.xptr ? *jb.xptr : jb.x;
    if (!x) return;
    const int b = blockIdx.x - bt.start_stats[j];
    const int ctiles = (jb.C + 63) / 64;
    const int slab = b / ctiles, ct = b % ctiles;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c = ct * 64 + lane;
    const int w0 = slab * PRE_SLAB, w1 = min(jb.W, w0 + PRE_SLAB);
    int sum = 0, mn = 32767, mx = -32768;
    if (c < jb.C) {
        const int16_t* __restrict__ xc = x + jb.cmap[c];
#pragma unroll 8
        for (int w = w0 + wv; w < w1; w += 4) {
            const int v = xc[(int64_t)w * jb.ld];
            sum += v;
            mn = min(mn, v);
            mx = max(mx, v);
        }
    }
    __shared__ int s_sum[4][64], s_mn[4][64], s_mx[4][64];
    s_sum[wv][lane] = sum;
    s_mn[wv][lane] = mn;
    s_mx[wv][lane] = mx;
    __syncthreads();
    if (wv == 0 && c < jb.C) {
        for (int q = 1; q < 4; ++q) {
            sum += s_sum[q][lane];
            mn = min(mn, s_mn[q][lane]);
            mx = max(mx, s_mx[q][lane]);
        }
        const int64_t o = (int64_t)slab * jb.C + c;
        jb.psum[o] = sum;
        jb.pmin[o] = mn;
        jb.pmax[o] = mx;
    }
}

// Means: one workgroup per job, a thread per channel sums the slab partials in
// int64 and divides once (FP64 division rounded to FP32: the correctly rounded
// quotient); the workgroup reduces min/max to the batch's peak-to-peak.
__global__ __launch_bounds__(PRE_MAX_C) void k_pre_mean(PreBatch bt) {
    const PreJob& jb = bt.job[blockIdx.x];
    if (!(jb.xptr ? *jb.xptr : jb.x)) return;
    const int c = threadIdx.x;
    int mn = 32767, mx = -32768;
    if (c < jb.C) {
        int64_t s = 0;
#pragma unroll 8
        for (int q = 0; q < jb.nslab; ++q) {
            const int64_t o = (int64_t)q * jb.C + c;
            s += jb.psum[o];
            mn = min(mn, jb.pmin[o]);
            mx = max(mx, jb.pmax[o]);
        }
        jb.mean[c] = (jb.flags & PRE_MEAN) ? (float)((double)s / (double)jb.W) : 0.f;
    }
    for (int o = 32; o > 0; o >>= 1) {
        mn = min(mn, __shfl_xor(mn, o, 64));
        mx = max(mx, __shfl_xor(mx, o, 64));
    }
    __shared__ int s_mn[PRE_MAX_C / 64], s_mx[PRE_MAX_C / 64];
    if ((c & 63) == 0) {
        s_mn[c >> 6] = mn;
        s_mx[c >> 6] = mx;
    }
    __syncthreads();
    if (c == 0 && jb.p2p) {
        for (int q = 1; q < PRE_MAX_C / 64; ++q) {
            mn = min(mn, s_mn[q]);
            mx = max(mx, s_mx[q]);
        }
        *jb.p2p = (float)(mx - mn);
    }
}

// Order-preserving map of FP32 to uint32 (negative values reversed), and back.
__device__ __forceinline__ uint32_t fkey(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unkey(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// Lower median of one row (C values in LDS) by a wave: a most-significant-bit-
// first radix select of the k-th smallest key, k = (C-1)/2.  At bit b the keys
// that agree with the chosen prefix above b and have a 0 at b are counted by
// ballots; k falls among them (bit 0) or past them (bit 1, k reduced).  The
// selected key is the k-th order statistic itself, so the value is exact.
template <int NJ>
__device__ __forceinline__ float row_median(const float* __restrict__ row, int C, int lane) {
    uint32_t key[NJ];
#pragma unroll
    for (int q = 0; q < NJ; ++q) {
        const int c = lane + 64 * q;
        key[q] = c < C ? fkey(row[c]) : 0xffffffffu;   // padding sorts last
    }
    uint32_t prefix = 0;
    int k = (C - 1) / 2;
    for (int b = 31; b >= 0; --b) {
        const uint32_t want = prefix >> b;
        int cnt0 = 0;
#pragma unroll
        for (int q = 0; q < NJ; ++q) cnt0 += __popcll(__ballot((key[q] >> b) == want));
        if (k >= cnt0) {
            k -= cnt0;
            prefix |= 1u << b;
        }
    }
    return unkey(prefix);
}

// CAR: PRE_ROWS samples per workgroup.  The tile x[w][c] - mean[c] sits in LDS
// (row stride odd, so the transposed write-out reads it without bank
// conflicts); each of the 16 waves takes the median of PRE_ROWS / 16 rows.  The result is
// written channel-major, X1[c][w] = x - mean - median.
template <int NJ>
__global__ __launch_bounds__(1024) void k_pre_car(PreBatch bt) {
    extern __shared__ float s_dyn[];
    const int j = [&] {
        int q = 0;
        while (q + 1 < bt.n && (int)blockIdx.x >= bt.start_car[q + 1]) ++q;
        return q;
    }();
    const PreJob& jb = bt.job[j];
    const int16_t* __restrict__ x = jb.xptr ? *jb.xptr : jb.x;
    if (!x) return;
    const int b = blockIdx.x - bt.start_car[j];
    const int C = jb.C, W = jb.W, SP = C | 1;
    float* s_mean = s_dyn;                       // [C]
    float* s_med = s_dyn + C;                    // [PRE_ROWS]
    float* s_x = s_med + PRE_ROWS;               // [PRE_ROWS][SP]
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int c = tid; c < C; c += 1024) s_mean[c] = jb.mean[c];
    __syncthreads();
    const int w0 = b * PRE_ROWS;
    for (int i = tid; i < PRE_ROWS * C; i += 1024) {
        const int r = i / C, c = i - r * C;
        const int w = w0 + r;
        s_x[r * SP + c] = w < W ? (float)x[(int64_t)w * jb.ld + jb.cmap[c]] - s_mean[c] : 0.f;
    }
    __syncthreads();
    for (int r = wv; r < PRE_ROWS; r += 16) {
        float med = 0.f;
        if ((jb.flags & PRE_CAR) && w0 + r < W) med = row_median<NJ>(s_x + r * SP, C, lane);
        if (lane == 0) {
            s_med[r] = med;
            if (w0 + r < W) jb.med[w0 + r] = med;
        }
    }
    __syncthreads();
    for (int i = tid; i < PRE_ROWS * C; i += 1024) {
        const int r = i & (PRE_ROWS - 1), c = i / PRE_ROWS;
        const int w = w0 + r;
        if (w < W) jb.X1[(int64_t)c * jb.PS + w] = s_x[r * SP + c] - s_med[r];
    }
}

// High-pass: one workgroup per (channel, PRE_OUT outputs); thread t owns outputs
// 4t..4t+3.  The window of the circularly indexed input is staged in LDS PRE_TCH
// taps at a time; per group of four taps a thread reads one new float4 of it and
// the four taps themselves through the scalar path (uniform loads), for 16 FMAs.
__global__ __launch_bounds__(256) void k_pre_fir(PreBatch bt) {
    __shared__ __attribute__((aligned(16))) float s_w[PRE_OUT + PRE_TCH + 8];
    __shared__ __attribute__((aligned(16))) float s_h[PRE_TCH];     // this chunk's taps (broadcast reads)
    const int j = [&] {
        int q = 0;
        while (q + 1 < bt.n && (int)blockIdx.x >= bt.start_fir[q + 1]) ++q;
        return q;
    }();
    const PreJob& jb = bt.job[j];
    if (!(jb.xptr ? *jb.xptr : jb.x)) return;
    const int b = blockIdx.x - bt.start_fir[j];
    const int W = jb.W;
    const int nb = (W + PRE_OUT - 1) / PRE_OUT;
    const int c = b / nb, k0 = (b % nb) * PRE_OUT;
    const float* __restrict__ xr = jb.X1 + (int64_t)c * jb.PS;
    const float* __restrict__ taps = jb.taps;
    const int tid = threadIdx.x;
    float y0 = 0.f, y1 = 0.f, y2 = 0.f, y3 = 0.f;
    for (int base = 0; base < jb.ntap; base += PRE_TCH) {
        const int nt = min(PRE_TCH, jb.ntap - base);
        const int nt4 = nt & ~3;
        __syncthreads();
        // window element e = x[(k0 + shift + base + e) mod W]
        const int i0 = (k0 + jb.shift + base) % W;
        for (int e = tid; e < PRE_OUT + nt4 + 4; e += 256) s_w[e] = xr[(i0 + e) % W];
        for (int e = tid; e < nt; e += 256) s_h[e] = taps[base + e];
        __syncthreads();
        const float* wp = s_w + 4 * tid;
        const float* hp = s_h;
        float4 A = *reinterpret_cast<const float4*>(wp);
        auto step4 = [&](const float4& B, float h0, float h1, float h2, float h3) {
            y0 = fmaf(A.x, h0, y0); y0 = fmaf(A.y, h1, y0); y0 = fmaf(A.z, h2, y0); y0 = fmaf(A.w, h3, y0);
            y1 = fmaf(A.y, h0, y1); y1 = fmaf(A.z, h1, y1); y1 = fmaf(A.w, h2, y1); y1 = fmaf(B.x, h3, y1);
            y2 = fmaf(A.z, h0, y2); y2 = fmaf(A.w, h1, y2); y2 = fmaf(B.x, h2, y2); y2 = fmaf(B.y, h3, y2);
            y3 = fmaf(A.w, h0, y3); y3 = fmaf(B.x, h1, y3); y3 = fmaf(B.y, h2, y3); y3 = fmaf(B.z, h3, y3);
            A = B;
        };
        // sixteen taps per iteration (four broadcast 16-byte reads) and four window reads
        const int nt16 = nt & ~15;
        int q = 0;
        for (; q < nt16; q += 16) {
            float h[16];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float4 t4 = *reinterpret_cast<const float4*>(hp + q + 4 * i);
                h[4 * i] = t4.x; h[4 * i + 1] = t4.y; h[4 * i + 2] = t4.z; h[4 * i + 3] = t4.w;
            }
            const float4 B0 = *reinterpret_cast<const float4*>(wp + q + 4);
            const float4 B1 = *reinterpret_cast<const float4*>(wp + q + 8);
            const float4 B2 = *reinterpret_cast<const float4*>(wp + q + 12);
            const float4 B3 = *reinterpret_cast<const float4*>(wp + q + 16);
            step4(B0, h[0], h[1], h[2], h[3]);
            step4(B1, h[4], h[5], h[6], h[7]);
            step4(B2, h[8], h[9], h[10], h[11]);
            step4(B3, h[12], h[13], h[14], h[15]);
        }
        for (; q < nt4; q += 4) step4(*reinterpret_cast<const float4*>(wp + q + 4), hp[q], hp[q + 1], hp[q + 2], hp[q + 3]);
        // the last nt - nt4 taps one at a time
        for (int q = nt4; q < nt; ++q) {
            const float h = hp[q];
            y0 = fmaf(wp[q], h, y0);
            y1 = fmaf(wp[q + 1], h, y1);
            y2 = fmaf(wp[q + 2], h, y2);
            y3 = fmaf(wp[q + 3], h, y3);
        }
    }
    float* __restrict__ out = jb.Xp + (int64_t)c * jb.PS;
    const int k = k0 + 4 * tid;
    if (k + 0 < W) out[k + 0] = y0;
    if (k + 1 < W) out[k + 1] = y1;
    if (k + 2 < W) out[k + 2] = y2;
    if (k + 3 < W) out[k + 3] = y3;
}

static size_t car_lds(int C) { return sizeof(float) * (C + PRE_ROWS + (size_t)PRE_ROWS * (C | 1)); }

void launch_preprocess(PreBatch& b, hipStream_t s) {
    if (b.n <= 0) return;
    int C = 0;
    b.start_stats[0] = b.start_car[0] = b.start_fir[0] = 0;
    for (int i = 0; i < b.n; ++i) {
        const PreJob& jb = b.job[i];
        C = std::max(C, jb.C);
        b.start_stats[i + 1] = b.start_stats[i] + jb.nslab * ((jb.C + 63) / 64);
        b.start_car[i + 1] = b.start_car[i] + (jb.W + PRE_ROWS - 1) / PRE_ROWS;
        const bool fir = (jb.flags & PRE_HP) && jb.ntap > 0;
        b.start_fir[i + 1] = b.start_fir[i] + (fir ? jb.C * ((jb.W + PRE_OUT - 1) / PRE_OUT) : 0);
    }
    k_pre_stats<<<b.start_stats[b.n], 256, 0, s>>>(b);
    k_pre_mean<<<b.n, PRE_MAX_C, 0, s>>>(b);
    const size_t lds = car_lds(C);
    const int NJ = (C + 63) / 64;
    // the CAR kernels may need more than 64 KB of LDS: raised once per device and channel class
    static std::mutex mu;
    static std::set<std::pair<int, int>> raised;
    int dev = 0;
    (void)hipGetDevice(&dev);
    bool raise = false;
    {
        std::lock_guard<std::mutex> lk(mu);
        raise = raised.insert({dev, NJ}).second;
    }
    switch (NJ) {
#define LSS_CAR_CASE(nj)                                                                                 \
    case nj:                                                                                             \
        if (raise)                                                                                       \
            (void)hipFuncSetAttribute((const void*)k_pre_car<nj>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                      (int)car_lds(PRE_MAX_C));                                          \
        k_pre_car<nj><<<b.start_car[b.n], 1024, lds, s>>>(b);                                            \
        break;
        LSS_CAR_CASE(1) LSS_CAR_CASE(2) LSS_CAR_CASE(3) LSS_CAR_CASE(4)
        LSS_CAR_CASE(5) LSS_CAR_CASE(6) LSS_CAR_CASE(7) LSS_CAR_CASE(8)
        LSS_CAR_CASE(9) LSS_CAR_CASE(10) LSS_CAR_CASE(11) LSS_CAR_CASE(12)
        LSS_CAR_CASE(13) LSS_CAR_CASE(14) LSS_CAR_CASE(15) LSS_CAR_CASE(16)
#undef LSS_CAR_CASE
        default: break;
    }
    if (b.start_fir[b.n] > 0) k_pre_fir<<<b.start_fir[b.n], 256, 0, s>>>(b);
}

}  // namespace lss
