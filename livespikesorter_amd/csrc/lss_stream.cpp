// Host-side streaming pieces around the matcher (no device code):
//   window and skip policy    runSpikeSorting, OnlineSpikesV2.cu:698-745,822,844
//   saveSpikes                 OnlineSpikesV2.cu:1173-1206
//   spikeOutput.txt lines      writeSpikesToFile, OnlineSpikesV2.cu:1148-1166
//   OnlineSpikesPayload and SorterParameters in cereal's binary layout
//                              (src/Networking/onlineSpikesPayload.h:8-48,
//                               sorterParameters.h:6-29; External/cereal)
//   .ap.bin replay source      (interleaved int16, channels fastest; the file the
//                               reference's verify_batch.py:33-39 reads)
#include "../../include/lss.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace lss {
void set_last_error(const std::string& msg);   // lss_capi.cpp (lss_last_error)
}

namespace {

struct StreamError : std::runtime_error {
    int code;
    StreamError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

template <class F>
int run(F&& f) {
    try {
        f();
        lss::set_last_error("");
        return LSS_OK;
    } catch (const StreamError& e) {
        lss::set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        lss::set_last_error(e.what());
        return LSS_ERR_IO;
    }
}

// ---- cereal binary archive: raw little-endian values, sizes as uint64 ----
struct Writer {
    uint8_t* out;
    int64_t cap, n = 0;
    void raw(const void* p, size_t k) {
        if (out && n + (int64_t)k <= cap) std::memcpy(out + n, p, k);
        n += (int64_t)k;
    }
    template <class T> void put(T v) { raw(&v, sizeof(T)); }
    void lng(int64_t v, int lb) {
        if (lb == 8) put<int64_t>(v);
        else put<int32_t>((int32_t)v);
    }
};

struct Reader {
    const uint8_t* p;
    int64_t size, n = 0;
    void raw(void* d, size_t k) {
        if (n + (int64_t)k > size) throw StreamError(LSS_ERR_ARG, "payload truncated");
        if (d) std::memcpy(d, p + n, k);
        n += (int64_t)k;
    }
    template <class T> T get() {
        T v;
        raw(&v, sizeof(T));
        return v;
    }
    int64_t lng(int lb) { return lb == 8 ? get<int64_t>() : (int64_t)get<int32_t>(); }
};

void check_long_bytes(int lb) {
    if (lb != 4 && lb != 8) throw StreamError(LSS_ERR_ARG, "long_bytes must be 4 (LLP64) or 8 (LP64)");
}

}  // namespace

struct lss_replay {
    int fd = -1;
    const int16_t* base = nullptr;
    size_t bytes = 0;
    int n_file = 0;
    std::vector<int> chans;
    int64_t n_samples = 0;
    ~lss_replay() {
        if (base) munmap(const_cast<int16_t*>(base), bytes);
        if (fd >= 0) close(fd);
    }
};

extern "C" {

int lss_window_plan(const lss_window_cfg* cfg, int64_t processed_ct, int64_t stream_ct, lss_window_step* s) {
    return run([&] {
        if (!cfg || !s) throw StreamError(LSS_ERR_ARG, "null argument");
        if (cfg->min_window <= 0 || cfg->max_window <= 0) throw StreamError(LSS_ERR_ARG, "windows must be positive");
        const int64_t minW = cfg->min_window, maxW = cfg->max_window;
        const int64_t latest = stream_ct;
        *s = lss_window_step{};
        // allowedCt = latestCt - timeBehind * samplingRate / 1'000 (:707) evaluates in FP32
        const float allowed_f = (float)latest - ((float)cfg->time_behind_ms * cfg->samp_rate) / 1000.0f;
        const int64_t allowed = (int64_t)allowed_f;
        s->allowed_ct = allowed;
        const int64_t behind = latest - processed_ct;
        // fetchLatest (dataSocket.cpp:103-126): lToGet = latest - start as unsigned, capped at the
        // socket's maximum (the max window), read ending at latest
        auto fetch_latest = [&](int64_t start) {
            uint64_t to_get = (uint64_t)latest - (uint64_t)start;
            if (to_get > (uint64_t)maxW) to_get = (uint64_t)maxW;
            s->fetch_start = latest - (int64_t)to_get;
            s->fetch_n = (int64_t)to_get;
            s->latest_ct = latest;
        };
        // fetchFromPlace (dataSocket.cpp:156-168): min(max window, latest - start) from start
        auto fetch_place = [&](int64_t start) {
            const int64_t to_get = std::min<int64_t>(maxW, latest - start);
            s->fetch_start = start;
            s->fetch_n = std::max<int64_t>(0, to_get);
            s->latest_ct = start + s->fetch_n;
        };
        if (behind <= maxW) {
            s->kind = LSS_WIN_LATEST;
            fetch_latest(processed_ct);
            s->buf_row = 0;
            s->batch_n = std::min<int64_t>(behind * (behind >= 0), maxW) + minW;
            s->skip = 0;
        } else if (cfg->time_behind_ms == 0) {
            s->kind = LSS_WIN_SKIP_LATEST;
            fetch_latest(processed_ct);
            s->buf_row = minW;
            s->batch_n = maxW;
            s->skip = 1;
        } else if (processed_ct >= allowed) {
            s->kind = LSS_WIN_CATCH_UP;
            fetch_place(processed_ct);
            s->buf_row = minW;
            s->batch_n = maxW + minW;
            s->skip = 0;
        } else if (cfg->small_skip) {
            s->kind = LSS_WIN_SMALL_SKIP;
            fetch_place(allowed);
            s->buf_row = minW;
            s->batch_n = maxW;
            s->skip = 1;
        } else {
            s->kind = LSS_WIN_BIG_SKIP;
            fetch_latest(processed_ct);
            s->buf_row = minW;
            s->batch_n = maxW;
            s->skip = 1;
        }
        // the device copy starts at row min_window of the fetch buffer when skipping (:750-757)
        s->batch_row = s->skip ? minW : 0;
        // saveSpikes(numSpikes, latestCt - currBatchNumSamples + 1, currBatchNumSamples - minWindow) (:822)
        s->save_start = s->latest_ct - s->batch_n + 1;
        s->save_end = s->batch_n - minW;
        // processedCt = latestCt - lookback (:844)
        s->next_processed_ct = s->latest_ct - cfg->lookback;
    });
}

int lss_save_spikes(const lss_event* ev, int64_t n, int32_t M, int32_t lookback, int32_t dt, int64_t save_start,
                    int64_t save_end, int64_t* last_spike, int32_t n_templates, lss_event* out, int64_t* src,
                    int64_t cap, int64_t* n_out) {
    return run([&] {
        if ((!ev && n) || !last_spike || !n_out || (!out && cap)) throw StreamError(LSS_ERR_ARG, "null argument");
        *n_out = 0;
        std::vector<lss_event> keep;
        std::vector<int64_t> from;
        for (int64_t i = 0; i < n; ++i) {
            const int32_t t = ev[i].template_id;
            if (t < 0 || t >= n_templates) throw StreamError(LSS_ERR_ARG, "template id outside the last-spike table");
            // the matcher's sample (Kilosort convention, column - M/2 + nt0min with column 0 at
            // batch row 0) minus M, then minus the lookback (:1186-1187)
            const int64_t idx = ev[i].sample - M - lookback;
            const int64_t abs_t = save_start + idx;
            // dedup against the template's last kept time, updated before the end-valid test
            if (abs_t - last_spike[t] < dt) continue;
            last_spike[t] = abs_t;
            if (idx >= save_end) continue;
            lss_event e = ev[i];
            e.sample = abs_t;
            // insert_sorted: after equal times (std::upper_bound)
            auto it = std::upper_bound(keep.begin(), keep.end(), e,
                                       [](const lss_event& a, const lss_event& b) { return a.sample < b.sample; });
            from.insert(from.begin() + (it - keep.begin()), i);
            keep.insert(it, e);
        }
        if ((int64_t)keep.size() > cap) throw StreamError(LSS_ERR_CAPACITY, "output capacity too small");
        std::copy(keep.begin(), keep.end(), out);
        if (src) std::copy(from.begin(), from.end(), src);
        *n_out = (int64_t)keep.size();
    });
}

int lss_format_spikes(int64_t recording_offset, const lss_event* ev, const float* y, int64_t n, char* out,
                      int64_t cap, int64_t* size) {
    return run([&] {
        if ((!ev && n) || !size) throw StreamError(LSS_ERR_ARG, "null argument");
        std::string s;
        char line[128];
        for (int64_t i = 0; i < n; ++i) {
            // std::ostream defaults: integers in decimal, floats as %g with 6 significant digits
            const int k = std::snprintf(line, sizeof line, "%lld,%d,%g,%g\n",
                                        (long long)(recording_offset + ev[i].sample), ev[i].template_id,
                                        (double)ev[i].amplitude, y ? (double)y[i] : 0.0);
            s.append(line, (size_t)k);
        }
        *size = (int64_t)s.size();
        if (out) {
            if (cap < (int64_t)s.size()) throw StreamError(LSS_ERR_CAPACITY, "output capacity too small");
            std::memcpy(out, s.data(), s.size());
        }
    });
}

int lss_payload_encode(const lss_payload_header* h, const int64_t* times, const int64_t* templates,
                       const float* amplitudes, int64_t n, int32_t long_bytes, uint8_t* out, int64_t cap,
                       int64_t* size) {
    return run([&] {
        if (!h || !size || n < 0 || (n && (!times || !templates || !amplitudes)))
            throw StreamError(LSS_ERR_ARG, "null argument");
        check_long_bytes(long_bytes);
        Writer w{out, cap};
        w.lng(h->recording_offset, long_bytes);
        w.lng(h->stream_sample_ct, long_bytes);
        w.put<uint64_t>((uint64_t)n);
        for (int64_t i = 0; i < n; ++i) w.lng(times[i], long_bytes);
        w.put<uint64_t>((uint64_t)n);
        for (int64_t i = 0; i < n; ++i) w.lng(templates[i], long_bytes);
        w.put<uint64_t>((uint64_t)n);
        w.raw(amplitudes, sizeof(float) * (size_t)n);
        w.put<double>(h->vrms);
        w.put<float>(h->p2p);
        w.lng(h->process_time, long_bytes);
        w.lng(h->event_stream_sample_ct, long_bytes);
        w.put<int16_t>(h->predict_label);
        w.put<int16_t>(h->label);
        w.put<int16_t>(h->n_trials);
        w.put<int16_t>(h->n_correct);
        w.put<double>(h->confidence);
        *size = w.n;
        if (out && w.n > cap) throw StreamError(LSS_ERR_CAPACITY, "output capacity too small");
    });
}

int lss_payload_decode(const uint8_t* buf, int64_t size, int32_t long_bytes, lss_payload_header* h, int64_t* times,
                       int64_t* templates, float* amplitudes, int64_t cap, int64_t* n) {
    return run([&] {
        if (!buf || !h || !n) throw StreamError(LSS_ERR_ARG, "null argument");
        check_long_bytes(long_bytes);
        Reader r{buf, size};
        *h = lss_payload_header{};
        h->recording_offset = r.lng(long_bytes);
        h->stream_sample_ct = r.lng(long_bytes);
        const int64_t lb = long_bytes;
        auto count = [&](int64_t elem) {
            const uint64_t k = r.get<uint64_t>();
            if (k > (uint64_t)((size - r.n) / elem)) throw StreamError(LSS_ERR_ARG, "payload vector length exceeds the data");
            return (int64_t)k;
        };
        const int64_t nt = count(lb);
        *n = nt;
        const bool store = times && templates && amplitudes;
        if (store && nt > cap) throw StreamError(LSS_ERR_CAPACITY, "output capacity too small");
        for (int64_t i = 0; i < nt; ++i) {
            const int64_t v = r.lng(long_bytes);
            if (store) times[i] = v;
        }
        if (count(lb) != nt) throw StreamError(LSS_ERR_ARG, "payload vectors of different lengths");
        for (int64_t i = 0; i < nt; ++i) {
            const int64_t v = r.lng(long_bytes);
            if (store) templates[i] = v;
        }
        if (count(4) != nt) throw StreamError(LSS_ERR_ARG, "payload vectors of different lengths");
        r.raw(store ? amplitudes : nullptr, sizeof(float) * (size_t)nt);
        h->vrms = r.get<double>();
        h->p2p = r.get<float>();
        h->process_time = r.lng(long_bytes);
        h->event_stream_sample_ct = r.lng(long_bytes);
        h->predict_label = r.get<int16_t>();
        h->label = r.get<int16_t>();
        h->n_trials = r.get<int16_t>();
        h->n_correct = r.get<int16_t>();
        h->confidence = r.get<double>();
        if (r.n != size) throw StreamError(LSS_ERR_ARG, "trailing bytes after the payload");
    });
}

int lss_sorter_params_encode(const lss_sorter_params* p, const double* chanpos, int64_t n_chanpos,
                             const int32_t* neuron_indices, int64_t n_indices, int32_t long_bytes, uint8_t* out,
                             int64_t cap, int64_t* size) {
    return run([&] {
        if (!p || !size || (n_chanpos && !chanpos) || (n_indices && !neuron_indices) || n_chanpos < 0 || n_indices < 0)
            throw StreamError(LSS_ERR_ARG, "null argument");
        check_long_bytes(long_bytes);
        Writer w{out, cap};
        w.lng(p->T, long_bytes);
        w.lng(p->C, long_bytes);
        w.lng(p->M, long_bytes);
        w.lng(p->N, long_bytes);
        w.put<float>(p->samp_rate);
        w.put<uint64_t>((uint64_t)n_chanpos);
        w.raw(chanpos, sizeof(double) * (size_t)n_chanpos);
        w.put<uint64_t>((uint64_t)n_indices);
        w.raw(neuron_indices, sizeof(int32_t) * (size_t)n_indices);
        *size = w.n;
        if (out && w.n > cap) throw StreamError(LSS_ERR_CAPACITY, "output capacity too small");
    });
}

int lss_sorter_params_decode(const uint8_t* buf, int64_t size, int32_t long_bytes, lss_sorter_params* p,
                             double* chanpos, int64_t chanpos_cap, int64_t* n_chanpos, int32_t* neuron_indices,
                             int64_t indices_cap, int64_t* n_indices) {
    return run([&] {
        if (!buf || !p || !n_chanpos || !n_indices) throw StreamError(LSS_ERR_ARG, "null argument");
        check_long_bytes(long_bytes);
        Reader r{buf, size};
        p->T = r.lng(long_bytes);
        p->C = r.lng(long_bytes);
        p->M = r.lng(long_bytes);
        p->N = r.lng(long_bytes);
        p->samp_rate = r.get<float>();
        const uint64_t nc = r.get<uint64_t>();
        if (nc > (uint64_t)((size - r.n) / 8)) throw StreamError(LSS_ERR_ARG, "vector length exceeds the data");
        *n_chanpos = (int64_t)nc;
        if (chanpos && (int64_t)nc > chanpos_cap) throw StreamError(LSS_ERR_CAPACITY, "output capacity too small");
        r.raw(chanpos, sizeof(double) * nc);
        const uint64_t ni = r.get<uint64_t>();
        if (ni > (uint64_t)((size - r.n) / 4)) throw StreamError(LSS_ERR_ARG, "vector length exceeds the data");
        *n_indices = (int64_t)ni;
        if (neuron_indices && (int64_t)ni > indices_cap) throw StreamError(LSS_ERR_CAPACITY, "output capacity too small");
        r.raw(neuron_indices, sizeof(int32_t) * ni);
        if (r.n != size) throw StreamError(LSS_ERR_ARG, "trailing bytes after the parameters");
    });
}

int lss_replay_open(const char* path, int32_t n_chan_file, const int32_t* channels, int32_t n_channels,
                    lss_replay** out) {
    return run([&] {
        if (!path || !out) throw StreamError(LSS_ERR_ARG, "null argument");
        *out = nullptr;
        if (n_chan_file <= 0) throw StreamError(LSS_ERR_ARG, "channel count of the file must be positive");
        auto r = std::make_unique<lss_replay>();
        r->n_file = n_chan_file;
        if (channels) {
            if (n_channels <= 0) throw StreamError(LSS_ERR_ARG, "empty channel subset");
            r->chans.assign(channels, channels + n_channels);
        } else {
            // all channels but the last (the sync channel of an imec .ap.bin)
            r->chans.resize(std::max(1, n_chan_file - 1));
            for (size_t i = 0; i < r->chans.size(); ++i) r->chans[i] = (int)i;
        }
        for (int c : r->chans)
            if (c < 0 || c >= n_chan_file) throw StreamError(LSS_ERR_ARG, "channel outside the file's channels");
        r->fd = open(path, O_RDONLY);
        if (r->fd < 0) throw StreamError(LSS_ERR_IO, std::string("cannot open ") + path);
        struct stat st {};
        if (fstat(r->fd, &st) != 0) throw StreamError(LSS_ERR_IO, std::string("cannot stat ") + path);
        r->bytes = (size_t)st.st_size;
        const size_t row = sizeof(int16_t) * (size_t)n_chan_file;
        r->n_samples = (int64_t)(r->bytes / row);
        if (r->bytes) {
            void* m = mmap(nullptr, r->bytes, PROT_READ, MAP_PRIVATE, r->fd, 0);
            if (m == MAP_FAILED) throw StreamError(LSS_ERR_IO, std::string("cannot map ") + path);
            r->base = static_cast<const int16_t*>(m);
        }
        *out = r.release();
    });
}

int lss_replay_samples(lss_replay* r, int64_t* n_samples, int32_t* n_channels) {
    return run([&] {
        if (!r || !n_samples) throw StreamError(LSS_ERR_ARG, "null argument");
        *n_samples = r->n_samples;
        if (n_channels) *n_channels = (int32_t)r->chans.size();
    });
}

int lss_replay_fetch(lss_replay* r, int64_t start, int64_t n, int16_t* out, int64_t* n_out) {
    return run([&] {
        if (!r || !n_out || (n > 0 && !out)) throw StreamError(LSS_ERR_ARG, "null argument");
        *n_out = 0;
        if (start < 0 || n < 0) throw StreamError(LSS_ERR_ARG, "negative range");
        const int64_t end = std::min(r->n_samples, start + n);
        const int C = (int)r->chans.size();
        for (int64_t s = start; s < end; ++s) {
            const int16_t* src = r->base + s * r->n_file;
            int16_t* dst = out + (s - start) * C;
            for (int c = 0; c < C; ++c) dst[c] = src[r->chans[c]];
        }
        *n_out = std::max<int64_t>(0, end - start);
    });
}

int lss_replay_close(lss_replay* r) {
    delete r;
    return 0;
}

}  // extern "C"
