// Long-run z-score baseline of binned spike counts, FP64 on the GPU.
// Follows the Decoder's ZScoreSdmProcessor (src/Decoder/ZScoreSdmProcessor.cpp):
// spikes are counted per bin of bin_samples (onSpikes, :85-94: bin = time /
// binSamples); when a bin completes its count updates a running mean and M2 by
// Welford's method while the baseline is open (computeBinValue, :96-114), and
// is scored z = (count - mean) / sd against a frozen baseline, with direction
// +1 / -1 beyond +-trigger_z (:116-140).  Series 0 is the total count (the
// reference's one series); series 1 + t counts template t.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/lss.h"

namespace lss {

struct ZsState {                 // device pointers
    int* counts;                 // [ring][S] pending bin counts
    double* mean;                // [S]
    double* m2;                  // [S]
    long long* seen;             // [S] bins folded into the baseline
    double* base_mean;           // [S] frozen baseline
    double* base_sd;             // [S]
    unsigned long long* flags;   // [0] late events (bin already closed), [1] events past the ring
};

// Bin a batch of events.  C++ integer division (truncation), as time / m_binSamples.
__global__ __launch_bounds__(256) void k_zs_bin(ZsState st, const lss_event* __restrict__ ev, int64_t n,
                                                int64_t bin_samples, int64_t first_open, int ring, int S) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const lss_event e = ev[i];
    const int64_t b = e.sample / bin_samples;
    if (b < first_open) {
        atomicAdd(&st.flags[0], 1ull);
        return;
    }
    if (b >= first_open + ring) {
        atomicAdd(&st.flags[1], 1ull);
        return;
    }
    int* row = st.counts + (int64_t)(b % ring) * S;
    atomicAdd(&row[0], 1);
    const int s = 1 + e.template_id;
    if (e.template_id >= 0 && s < S) atomicAdd(&row[s], 1);
}

// Close n_close bins from first_open in order: one thread per series, the
// Welford update (or the z-score against the frozen baseline) per bin, then the
// bin's slot is cleared for reuse.  Contraction is off so every operation rounds
// as written: the result is the sequential FP64 recurrence exactly.
__global__ __launch_bounds__(256) void k_zs_close(ZsState st, int64_t first_open, int n_close, int ring, int S,
                                                  int frozen, double trigger_z, float* __restrict__ z_out,
                                                  int8_t* __restrict__ dir_out) {
#pragma clang fp contract(off)
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= S) return;
    double mean = st.mean[s], m2 = st.m2[s];
    long long seen = st.seen[s];
    const double bm = st.base_mean[s], bsd = st.base_sd[s];
    for (int k = 0; k < n_close; ++k) {
        int* slot = st.counts + (int64_t)((first_open + k) % ring) * S + s;
        const double x = (double)*slot;
        *slot = 0;
        float z = 0.f;
        int8_t dir = 0;
        if (!frozen) {
            seen += 1;
            const double delta = x - mean;
            mean += delta / (double)seen;
            const double delta2 = x - mean;
            m2 += delta * delta2;
        } else if (bsd > 0.0) {
            const double zv = (x - bm) / bsd;
            if (zv > trigger_z) dir = 1;
            else if (zv < -trigger_z) dir = -1;
            z = (float)zv;
        }
        if (z_out) {
            z_out[(int64_t)k * S + s] = z;
            dir_out[(int64_t)k * S + s] = dir;
        }
    }
    st.mean[s] = mean;
    st.m2[s] = m2;
    st.seen[s] = seen;
}

void launch_zs_bin(const ZsState& st, const lss_event* ev, int64_t n, int64_t bin_samples, int64_t first_open,
                   int ring, int S, hipStream_t s) {
    if (n <= 0) return;
    k_zs_bin<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(st, ev, n, bin_samples, first_open, ring, S);
}

void launch_zs_close(const ZsState& st, int64_t first_open, int n_close, int ring, int S, int frozen,
                     double trigger_z, float* z_out, int8_t* dir_out, hipStream_t s) {
    if (n_close <= 0) return;
    k_zs_close<<<(S + 255) / 256, 256, 0, s>>>(st, first_open, n_close, ring, S, frozen, trigger_z, z_out, dir_out);
}

}  // namespace lss

// ---------------------------------------------------------------------------
// C ABI (host)
// ---------------------------------------------------------------------------
#include <algorithm>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace lss {
void set_last_error(const std::string& msg);   // lss_capi.cpp (lss_last_error)
}

namespace {

struct ZsError : std::runtime_error {
    int code;
    ZsError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define ZSCHK(x)                                                                                     \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) throw ZsError(LSS_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class F>
int zrun(F&& f) {
    try {
        f();
        lss::set_last_error("");
        return LSS_OK;
    } catch (const ZsError& e) {
        lss::set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        lss::set_last_error(e.what());
        return LSS_ERR_IO;
    }
}

}  // namespace

struct lss_zscore {
    int dev = 0, S = 0, ring = 0;
    int64_t bin = 1, first_open = 0;
    int frozen = 0;
    double trigger_z = 1.0;
    hipStream_t stream = nullptr;
    lss::ZsState st{};
    lss_event* d_ev = nullptr;
    int64_t ev_cap = 0;
    float* d_z = nullptr;
    int8_t* d_dir = nullptr;
    int close_cap = 0;
    ~lss_zscore() {
        (void)hipSetDevice(dev);
        void* ptrs[] = {st.counts, st.mean, st.m2, st.seen, st.base_mean, st.base_sd, st.flags, d_ev, d_z, d_dir};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" {

int lss_zscore_bin_samples(double bin_ms, float samp_rate, int64_t* bin_samples) {
    return zrun([&] {
        if (!bin_samples) throw ZsError(LSS_ERR_ARG, "null argument");
        // m_binSamples = max(1, llround(binMs / 1000 * rate)) (ZScoreSdmProcessor.cpp:27-28)
        *bin_samples = std::max<int64_t>(1, (int64_t)std::llround((bin_ms / 1000.0) * (double)samp_rate));
    });
}

int lss_zscore_create(int32_t device, int32_t n_templates, int64_t bin_samples, int32_t ring_bins,
                      lss_zscore** out) {
    return zrun([&] {
        if (!out) throw ZsError(LSS_ERR_ARG, "null argument");
        *out = nullptr;
        if (n_templates < 0 || bin_samples <= 0 || ring_bins <= 0)
            throw ZsError(LSS_ERR_ARG, "need n_templates >= 0, bin_samples > 0, ring_bins > 0");
        auto z = std::make_unique<lss_zscore>();
        z->dev = device;
        z->S = 1 + n_templates;
        z->ring = ring_bins;
        z->bin = bin_samples;
        ZSCHK(hipSetDevice(device));
        ZSCHK(hipStreamCreateWithFlags(&z->stream, hipStreamNonBlocking));
        const size_t S = (size_t)z->S;
        ZSCHK(hipMalloc(&z->st.counts, sizeof(int) * S * (size_t)ring_bins));
        ZSCHK(hipMalloc(&z->st.mean, sizeof(double) * S));
        ZSCHK(hipMalloc(&z->st.m2, sizeof(double) * S));
        ZSCHK(hipMalloc(&z->st.seen, sizeof(long long) * S));
        ZSCHK(hipMalloc(&z->st.base_mean, sizeof(double) * S));
        ZSCHK(hipMalloc(&z->st.base_sd, sizeof(double) * S));
        ZSCHK(hipMalloc(&z->st.flags, sizeof(unsigned long long) * 2));
        ZSCHK(hipMemset(z->st.counts, 0, sizeof(int) * S * (size_t)ring_bins));
        ZSCHK(hipMemset(z->st.mean, 0, sizeof(double) * S));
        ZSCHK(hipMemset(z->st.m2, 0, sizeof(double) * S));
        ZSCHK(hipMemset(z->st.seen, 0, sizeof(long long) * S));
        ZSCHK(hipMemset(z->st.base_mean, 0, sizeof(double) * S));
        ZSCHK(hipMemset(z->st.base_sd, 0, sizeof(double) * S));
        ZSCHK(hipMemset(z->st.flags, 0, sizeof(unsigned long long) * 2));
        *out = z.release();
    });
}

int lss_zscore_add(lss_zscore* z, const lss_event* ev, int64_t n) {
    return zrun([&] {
        if (!z || (n > 0 && !ev) || n < 0) throw ZsError(LSS_ERR_ARG, "bad argument");
        if (n == 0) return;
        ZSCHK(hipSetDevice(z->dev));
        if (n > z->ev_cap) {
            ZSCHK(hipStreamSynchronize(z->stream));
            if (z->d_ev) ZSCHK(hipFree(z->d_ev));
            z->d_ev = nullptr;
            z->ev_cap = std::max<int64_t>(n, 2 * z->ev_cap);
            ZSCHK(hipMalloc(&z->d_ev, sizeof(lss_event) * (size_t)z->ev_cap));
        }
        ZSCHK(hipMemcpyAsync(z->d_ev, ev, sizeof(lss_event) * (size_t)n, hipMemcpyHostToDevice, z->stream));
        lss::launch_zs_bin(z->st, z->d_ev, n, z->bin, z->first_open, z->ring, z->S, z->stream);
        ZSCHK(hipGetLastError());
        ZSCHK(hipStreamSynchronize(z->stream));
    });
}

int lss_zscore_complete(lss_zscore* z, int64_t end_sample, float* z_out, int8_t* dir_out, int64_t cap_bins,
                        int64_t* n_closed) {
    return zrun([&] {
        if (!z || !n_closed || (z_out && !dir_out)) throw ZsError(LSS_ERR_ARG, "bad argument");
        *n_closed = 0;
        ZSCHK(hipSetDevice(z->dev));
        unsigned long long flags[2];
        ZSCHK(hipMemcpy(flags, z->st.flags, sizeof flags, hipMemcpyDeviceToHost));
        if (flags[1]) throw ZsError(LSS_ERR_CAPACITY, "events fell past the pending-bin ring (raise ring_bins)");
        // bins ending at or before end_sample: [first_open, floor(end_sample / bin))
        const int64_t last = end_sample >= 0 ? end_sample / z->bin : -((-end_sample + z->bin - 1) / z->bin);
        int64_t n = last - z->first_open;
        if (n <= 0) return;
        if (z_out && n > cap_bins) throw ZsError(LSS_ERR_CAPACITY, "z output capacity too small");
        while (n > 0) {
            const int k = (int)std::min<int64_t>(n, z->ring);
            if (z_out && k > z->close_cap) {
                if (z->d_z) ZSCHK(hipFree(z->d_z));
                if (z->d_dir) ZSCHK(hipFree(z->d_dir));
                z->d_z = nullptr;
                z->d_dir = nullptr;
                z->close_cap = k;
                ZSCHK(hipMalloc(&z->d_z, sizeof(float) * (size_t)k * z->S));
                ZSCHK(hipMalloc(&z->d_dir, (size_t)k * z->S));
            }
            lss::launch_zs_close(z->st, z->first_open, k, z->ring, z->S, z->frozen, z->trigger_z,
                                 z_out ? z->d_z : nullptr, z_out ? z->d_dir : nullptr, z->stream);
            ZSCHK(hipGetLastError());
            if (z_out) {
                const int64_t o = *n_closed * z->S;
                ZSCHK(hipMemcpyAsync(z_out + o, z->d_z, sizeof(float) * (size_t)k * z->S, hipMemcpyDeviceToHost,
                                     z->stream));
                ZSCHK(hipMemcpyAsync(dir_out + o, z->d_dir, (size_t)k * z->S, hipMemcpyDeviceToHost, z->stream));
            }
            ZSCHK(hipStreamSynchronize(z->stream));
            z->first_open += k;
            *n_closed += k;
            n -= k;
        }
    });
}

int lss_zscore_freeze(lss_zscore* z, const double* mean, const double* sd, double trigger_z) {
    return zrun([&] {
        if (!z || (!mean) != (!sd)) throw ZsError(LSS_ERR_ARG, "bad argument");
        ZSCHK(hipSetDevice(z->dev));
        const size_t S = (size_t)z->S;
        std::vector<double> m(S), s(S);
        if (mean) {
            m.assign(mean, mean + S);
            s.assign(sd, sd + S);
        } else {
            // the running baseline: mean, and the sample standard deviation sqrt(M2 / (n - 1))
            std::vector<double> m2(S);
            std::vector<long long> seen(S);
            ZSCHK(hipMemcpy(m.data(), z->st.mean, sizeof(double) * S, hipMemcpyDeviceToHost));
            ZSCHK(hipMemcpy(m2.data(), z->st.m2, sizeof(double) * S, hipMemcpyDeviceToHost));
            ZSCHK(hipMemcpy(seen.data(), z->st.seen, sizeof(long long) * S, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < S; ++i) s[i] = seen[i] > 1 ? std::sqrt(m2[i] / (double)(seen[i] - 1)) : 0.0;
        }
        ZSCHK(hipMemcpy(z->st.base_mean, m.data(), sizeof(double) * S, hipMemcpyHostToDevice));
        ZSCHK(hipMemcpy(z->st.base_sd, s.data(), sizeof(double) * S, hipMemcpyHostToDevice));
        z->trigger_z = trigger_z;
        z->frozen = 1;
    });
}

int lss_zscore_stats(lss_zscore* z, int64_t* bins_seen, double* mean, double* m2, int64_t* late, int64_t* first_open) {
    return zrun([&] {
        if (!z) throw ZsError(LSS_ERR_ARG, "null argument");
        ZSCHK(hipSetDevice(z->dev));
        ZSCHK(hipStreamSynchronize(z->stream));
        const size_t S = (size_t)z->S;
        if (bins_seen) {
            std::vector<long long> seen(S);
            ZSCHK(hipMemcpy(seen.data(), z->st.seen, sizeof(long long) * S, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < S; ++i) bins_seen[i] = seen[i];
        }
        if (mean) ZSCHK(hipMemcpy(mean, z->st.mean, sizeof(double) * S, hipMemcpyDeviceToHost));
        if (m2) ZSCHK(hipMemcpy(m2, z->st.m2, sizeof(double) * S, hipMemcpyDeviceToHost));
        if (late) {
            unsigned long long f[2];
            ZSCHK(hipMemcpy(f, z->st.flags, sizeof f, hipMemcpyDeviceToHost));
            *late = (int64_t)f[0];
        }
        if (first_open) *first_open = z->first_open;
    });
}

int lss_zscore_destroy(lss_zscore* z) {
    delete z;
    return 0;
}

}  // extern "C"
