// C++ adapter with the shape of LSS's OnlineSpikesV2 sorter (src/SpikeSorter/
// OnlineSpikesV2.h:29-34): constructed from the input parameters, a decoder
// connection and a shared DataSocket, then runSpikeSorting() runs the live loop
// (OnlineSpikesV2.cu:662-850) over the socket.  Header-only, above the C ABI of
// lss.h: the window policy (lss_window_plan), the matcher (lss_create /
// lss_process_batch / lss_query / lss_last_positions), saveSpikes
// (lss_save_spikes), spikeOutput.txt (lss_format_spikes) and the decoder payload
// (lss_payload_encode).
//
// Departures, each forced by this library's boundary:
//  * DataSocket fetches int16 rows [n][lNChans] (SpikeGLX's own sample type);
//    the reference converts them to float before use (dataSocket.cpp:118-121),
//    which changes no value.
//  * The decoder connection (a UDP socket to mainAddr, establishDecoderConnection
//    / sendPayload) is a PayloadSink that receives the encoded bytes; the
//    networking layer is outside this library.
//  * The matcher is a BatchMatcher: LssMatcher runs this library's GPU matcher;
//    a test can supply another (the loop's bookkeeping is then checked alone).
#pragma once

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "lss.h"

namespace lss {

typedef unsigned long long t_ull;
enum streamTypes { NIDQ, OBX, IMEC };

// Per-sorter socket parameters (dataSocket.h:17-21).
struct OSSSpecificParams {
    long lNChans;
    std::vector<int> vImecChannels;
    int substream;
};

// The acquisition socket (dataSocket.h:23-57): the stream's sample count, waits,
// and the two fetch forms runSpikeSorting uses.  fetchLatest(data, os, start)
// copies the newest min(latest - start, max size) samples (the min size when
// start is ULLONG_MAX); fetchFromPlace(data, os, start) copies
// min(max size, latest - start) samples from start.  Both return the stream
// count one past the last sample copied.  (The NIDQ and event calls of the
// interface serve the decoder and the threshold-crossing sorter, not this one.)
class DataSocket {
public:
    virtual ~DataSocket() = default;
    virtual bool connect() { return false; }
    virtual void disconnect() {}
    virtual t_ull getStreamSampleCt(int streamType, OSSSpecificParams osParams) = 0;
    virtual t_ull fetchLatest(int16_t* data, OSSSpecificParams osParams, t_ull lStartCt = ULLONG_MAX) = 0;
    virtual t_ull fetchFromPlace(int16_t* data, OSSSpecificParams osParams, t_ull lStartCt) = 0;
    virtual bool isRunning() { return true; }
    virtual bool startRun() { return true; }
    virtual bool stopRun() { return true; }
    virtual void waitUntil(t_ull lCt, OSSSpecificParams osParams) { (void)lCt; (void)osParams; }
};

inline void check(int rc, const char* what) {
    if (rc != LSS_OK) throw std::runtime_error(std::string(what) + ": " + lss_last_error());
}

// A recorded .ap.bin played as the stream (lss_replay_*), with the stream's
// sample count following a script: count[0] before the loop, and each
// waitUntil moves to the next count (which must reach the awaited one).  The
// run ends (isRunning false) when the script is used up.
class ReplayDataSocket : public DataSocket {
public:
    ReplayDataSocket(const std::string& path, int n_chan_file, const std::vector<int>& channels, long lMaxSize,
                     long lMinSize, std::vector<int64_t> counts)
        : max_size_(lMaxSize), min_size_(lMinSize), counts_(std::move(counts)) {
        if (counts_.empty()) throw std::invalid_argument("the replay clock needs at least its start count");
        check(lss_replay_open(path.c_str(), n_chan_file, channels.empty() ? nullptr : channels.data(),
                              (int32_t)channels.size(), &r_),
              "lss_replay_open");
        int64_t n = 0;
        check(lss_replay_samples(r_, &n, &nch_), "lss_replay_samples");
    }
    ~ReplayDataSocket() override {
        if (r_) lss_replay_close(r_);
    }
    ReplayDataSocket(const ReplayDataSocket&) = delete;
    ReplayDataSocket& operator=(const ReplayDataSocket&) = delete;

    int channels() const { return nch_; }
    t_ull getStreamSampleCt(int, OSSSpecificParams) override { return (t_ull)counts_[pos_]; }
    bool isRunning() override { return running_; }
    void waitUntil(t_ull lCt, OSSSpecificParams) override {
        if (pos_ + 1 >= counts_.size()) {
            running_ = false;
            return;
        }
        ++pos_;
        if ((t_ull)counts_[pos_] < lCt) throw std::runtime_error("replay clock: a count stops short of the awaited one");
    }
    t_ull fetchLatest(int16_t* data, OSSSpecificParams, t_ull lStartCt) override {
        const t_ull latest = (t_ull)counts_[pos_];
        t_ull n = lStartCt == ULLONG_MAX ? (t_ull)min_size_ : latest - lStartCt;
        n = std::min(n, (t_ull)max_size_);
        return copy(data, latest - n, n);
    }
    t_ull fetchFromPlace(int16_t* data, OSSSpecificParams, t_ull lStartCt) override {
        const t_ull latest = (t_ull)counts_[pos_];
        return copy(data, lStartCt, std::min((t_ull)max_size_, latest - lStartCt));
    }

private:
    t_ull copy(int16_t* data, t_ull start, t_ull n) {
        int64_t got = 0;
        check(lss_replay_fetch(r_, (int64_t)start, (int64_t)n, data, &got), "lss_replay_fetch");
        return start + (t_ull)got;
    }
    lss_replay* r_ = nullptr;
    int32_t nch_ = 0;
    long max_size_, min_size_;
    std::vector<int64_t> counts_;
    size_t pos_ = 0;
    bool running_ = true;
};

// The sorter's parameters (the fields of InputParameters that OnlineSpikesV2
// reads, OnlineSpikesV2.cu:248-281,356-372), plus the matcher's own settings.
struct InputParameters {
    std::string sInputFolder;       // oss_input directory
    std::string sOSSOutputFolder;   // spikeOutput.txt goes here ("" = no file)
    unsigned uSelectedDevice = 0;
    int iMinScanWindow = 1500;
    int iMaxScanWindow = 3000;
    int iTimeBehind = 0;            // ms behind tolerated; 0 = skip ahead
    int iDownsampling = 1;
    float fImecSamplingRate = 30000.f;
    bool bSmallskip = false;
    int iSubstream = 0;
    std::vector<int> vImecChannels; // channels the socket fetches
    // matcher settings (lss_config): semantics LSS_SEM_* and precision LSS_PREC_*
    int iSemantics = LSS_SEM_LSS;
    int iPrecision = LSS_PREC_FP32;
    int iPreprocess = LSS_PRE_ALL;
    int iLongBytes = 4;             // `long` width on the wire (4: the reference's Windows build)
    int64_t lRecordingOffset = 0;
};

// One batch matched: events in the matcher's sample convention (first sample 0
// of the batch), each event's y position (0 when the matcher has none), and the
// batch's peak-to-peak.
class BatchMatcher {
public:
    virtual ~BatchMatcher() = default;
    virtual lss_info info() = 0;
    virtual void match(const int16_t* rows, int64_t n, int32_t ld, std::vector<lss_event>& ev,
                       std::vector<float>& y, float& p2p) = 0;
};

// This library's matcher on one HIP device.
class LssMatcher : public BatchMatcher {
public:
    LssMatcher(const InputParameters& p, int max_batch) {
        lss_config cfg;
        check(lss_config_init(&cfg), "lss_config_init");
        cfg.device = (int32_t)p.uSelectedDevice;
        cfg.max_batch = max_batch;
        cfg.semantics = p.iSemantics;
        cfg.precision = p.iPrecision;
        cfg.preprocess = p.iPreprocess;
        check(lss_create(p.sInputFolder.c_str(), &cfg, &h_), "lss_create");
        check(lss_query(h_, &info_), "lss_query");
        cap_ = std::max<int64_t>(info_.max_batch, 1);
    }
    ~LssMatcher() override {
        if (h_) lss_destroy(h_);
    }
    LssMatcher(const LssMatcher&) = delete;
    LssMatcher& operator=(const LssMatcher&) = delete;

    lss_info info() override { return info_; }
    void match(const int16_t* rows, int64_t n, int32_t ld, std::vector<lss_event>& ev, std::vector<float>& y,
               float& p2p) override {
        ev.resize((size_t)cap_);
        int64_t ne = 0;
        check(lss_process_batch(h_, rows, n, ld, nullptr, 0, ev.data(), cap_, &ne, nullptr), "lss_process_batch");
        ev.resize((size_t)ne);
        std::vector<float> xy(2 * (size_t)std::max<int64_t>(ne, 1));
        int64_t np = 0;
        y.assign((size_t)ne, 0.f);
        if (lss_last_positions(h_, xy.data(), ne, &np) == LSS_OK)
            for (int64_t i = 0; i < np && i < ne; ++i) y[(size_t)i] = xy[2 * (size_t)i + 1];
        lss_info q;
        check(lss_query(h_, &q), "lss_query");
        p2p = q.p2p;
    }

private:
    lss_handle* h_ = nullptr;
    lss_info info_{};
    int64_t cap_ = 0;
};

// Where the per-batch payload goes (the decoder connection).
class PayloadSink {
public:
    virtual ~PayloadSink() = default;
    virtual void send(const std::vector<uint8_t>& payload) = 0;
};

class OnlineSpikesV2 {
public:
    // The matcher is this library's GPU matcher on params.uSelectedDevice.
    OnlineSpikesV2(const InputParameters& params, PayloadSink* decoder, DataSocket* socket)
        : OnlineSpikesV2(params, decoder, socket, nullptr) {}
    // The same with a caller's matcher (not owned; it must accept batches of
    // iMinScanWindow + iMaxScanWindow samples).
    OnlineSpikesV2(const InputParameters& params, PayloadSink* decoder, DataSocket* socket, BatchMatcher* matcher)
        : p_(params), sink_(decoder), sock_(socket) {
        if (!sock_) throw std::invalid_argument("OnlineSpikesV2 needs a data socket");
        if (p_.iDownsampling < 1) throw std::invalid_argument("downsampling must be >= 1");
        if (matcher) {
            m_ = matcher;
        } else {
            own_.reset(new LssMatcher(p_, p_.iMinScanWindow + p_.iMaxScanWindow));
            m_ = own_.get();
        }
        const lss_info in = m_->info();
        C_ = p_.vImecChannels.empty() ? in.C : (int)p_.vImecChannels.size();
        if (C_ != in.C) throw std::invalid_argument("the socket's channel subset must match the matcher's channels");
        cfg_.min_window = p_.iMinScanWindow;
        cfg_.max_window = p_.iMaxScanWindow;
        cfg_.time_behind_ms = p_.iTimeBehind;
        cfg_.samp_rate = p_.fImecSamplingRate / (float)p_.iDownsampling;
        cfg_.small_skip = p_.bSmallskip ? 1 : 0;
        cfg_.lookback = 2 * in.M;                  // OnlineSpikesV2.cu:461
        M_ = in.M;
        dt_ = in.dt;
        // lastSpikeTime per template (or final cluster when relabelling), zero-initialised
        last_spike_.assign((size_t)std::max(in.T, 1 << 16), 0);
        buf_.assign((size_t)(p_.iMinScanWindow + p_.iMaxScanWindow) * C_, 0);
        if (!p_.sOSSOutputFolder.empty()) spike_path_ = p_.sOSSOutputFolder + "spikeOutput.txt";
    }

    // The live loop (OnlineSpikesV2.cu:662-850): wait for min_window new
    // samples, plan the fetch, fetch into the persistent buffer, match the batch
    // cut from it, keep the spikes saveSpikes keeps, send the payload and append
    // the spike lines; until the socket stops running.
    void runSpikeSorting() {
        OSSSpecificParams os{(long)C_, p_.vImecChannels, p_.iSubstream};
        int64_t latest = (int64_t)sock_->getStreamSampleCt(IMEC, os);
        int64_t processed = latest;
        std::vector<lss_event> ev, saved;
        std::vector<int64_t> src;
        std::vector<float> y;
        while (sock_->isRunning()) {
            sock_->waitUntil((t_ull)(latest + cfg_.min_window), os);
            if (!sock_->isRunning()) break;
            const auto t0 = std::chrono::steady_clock::now();
            const int64_t ct_now = (int64_t)sock_->getStreamSampleCt(IMEC, os);
            lss_window_step st;
            check(lss_window_plan(&cfg_, processed, ct_now, &st), "lss_window_plan");
            int16_t* dst = buf_.data() + (size_t)st.buf_row * C_;
            t_ull got;
            switch (st.kind) {                     // :709-745
                case LSS_WIN_LATEST:
                case LSS_WIN_SKIP_LATEST:
                case LSS_WIN_BIG_SKIP: got = sock_->fetchLatest(dst, os, (t_ull)processed); break;
                case LSS_WIN_CATCH_UP: got = sock_->fetchFromPlace(dst, os, (t_ull)processed); break;
                default: got = sock_->fetchFromPlace(dst, os, (t_ull)st.allowed_ct); break;
            }
            follow_fetch(st, (int64_t)got, processed);
            float p2p = 0.f;
            m_->match(buf_.data() + (size_t)st.batch_row * C_, st.batch_n, C_, ev, y, p2p);
            saved.resize(std::max<size_t>(ev.size(), 1));
            src.resize(saved.size());
            int64_t ns = 0;
            check(lss_save_spikes(ev.data(), (int64_t)ev.size(), M_, cfg_.lookback, dt_, st.save_start, st.save_end,
                                  last_spike_.data(), (int32_t)last_spike_.size(), saved.data(), src.data(),
                                  (int64_t)saved.size(), &ns),
                  "lss_save_spikes");
            saved.resize((size_t)ns);
            std::vector<float> ys((size_t)ns);
            for (int64_t i = 0; i < ns; ++i) ys[(size_t)i] = y.empty() ? 0.f : y[(size_t)src[(size_t)i]];
            const auto t1 = std::chrono::steady_clock::now();
            send(st, saved, p2p, std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count());
            write_spikes(saved, ys);
            latest = st.latest_ct;
            processed = st.next_processed_ct;      // :844
            ++batches_;
        }
    }

    int64_t batches() const { return batches_; }
    const lss_window_cfg& window_config() const { return cfg_; }

private:
    // The batch follows the count the fetch returned (:710-715): unchanged when
    // the stream did not advance between the plan and the fetch (always so for
    // a replay).  A fetchLatest took the newest min(count - processedCt,
    // max_window) samples; a fetchFromPlace started where it was asked to.
    void follow_fetch(lss_window_step& st, int64_t got, int64_t processed) const {
        if (got == st.latest_ct) return;
        const bool latest_form = st.kind == LSS_WIN_LATEST || st.kind == LSS_WIN_SKIP_LATEST ||
                                 st.kind == LSS_WIN_BIG_SKIP;
        if (latest_form) st.fetch_start = got - std::min<int64_t>(std::max<int64_t>(got - processed, 0), cfg_.max_window);
        st.fetch_n = got - st.fetch_start;
        st.latest_ct = got;
        if (st.kind == LSS_WIN_LATEST)
            st.batch_n = std::min<int64_t>(std::max<int64_t>(got - processed, 0), cfg_.max_window) + cfg_.min_window;
        st.save_start = got - st.batch_n + 1;
        st.save_end = st.batch_n - cfg_.min_window;
        st.next_processed_ct = got - cfg_.lookback;
    }

    void send(const lss_window_step& st, const std::vector<lss_event>& saved, float p2p, int64_t ms) {
        if (!sink_) return;
        const size_t n = saved.size();
        std::vector<int64_t> times(n), templates(n);
        std::vector<float> amps(n);
        for (size_t i = 0; i < n; ++i) {
            times[i] = saved[i].sample;
            templates[i] = saved[i].template_id;
            amps[i] = saved[i].amplitude;
        }
        lss_payload_header h{};
        h.recording_offset = p_.lRecordingOffset;
        h.stream_sample_ct = st.latest_ct;
        h.p2p = p2p;
        h.process_time = ms;
        int64_t size = 0;
        check(lss_payload_encode(&h, times.data(), templates.data(), amps.data(), (int64_t)n, p_.iLongBytes,
                                 nullptr, 0, &size),
              "lss_payload_encode");
        std::vector<uint8_t> out((size_t)size);
        check(lss_payload_encode(&h, times.data(), templates.data(), amps.data(), (int64_t)n, p_.iLongBytes,
                                 out.data(), size, &size),
              "lss_payload_encode");
        sink_->send(out);
    }

    void write_spikes(const std::vector<lss_event>& saved, const std::vector<float>& y) {
        if (spike_path_.empty()) return;
        int64_t size = 0;
        check(lss_format_spikes(p_.lRecordingOffset, saved.data(), y.data(), (int64_t)saved.size(), nullptr, 0,
                                &size),
              "lss_format_spikes");
        std::string text((size_t)size, '\0');
        check(lss_format_spikes(p_.lRecordingOffset, saved.data(), y.data(), (int64_t)saved.size(), &text[0], size,
                                &size),
              "lss_format_spikes");
        std::ofstream f(spike_path_, std::ios::app);   // appended, as the reference reopens it (:1154)
        if (!f) throw std::runtime_error("cannot open " + spike_path_);
        f.write(text.data(), (std::streamsize)text.size());
    }

    InputParameters p_;
    PayloadSink* sink_;
    DataSocket* sock_;
    std::unique_ptr<LssMatcher> own_;
    BatchMatcher* m_ = nullptr;
    lss_window_cfg cfg_{};
    int C_ = 0, M_ = 0, dt_ = 0;
    std::vector<int64_t> last_spike_;
    std::vector<int16_t> buf_;
    std::string spike_path_;
    int64_t batches_ = 0;
};

}  // namespace lss
