// Drives lss::OnlineSpikesV2 (include/lss_online_spikes.hpp) over a replayed
// .ap.bin and writes what the loop produced: every payload the decoder would
// receive (payloads.bin: uint64 length + bytes each) and spikeOutput.txt.
//
//   adapter_replay <stub|gpu> <oss_input_dir> <ap.bin> <n_chan_file> <min_window> <max_window>
//                  <time_behind_ms> <small_skip> <semantics> <preprocess> <counts.txt> <out_dir>
//
// counts.txt: the stream's sample count before the loop, then one per batch.
// "stub" matches with a deterministic host rule (no GPU), so the loop's own
// bookkeeping can be compared on any machine; "gpu" runs the library's matcher.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "lss_online_spikes.hpp"

namespace {

// The host rule (mirrored by tests/test_adapter.py): row t is an event when its
// channel-0 value exceeds 300 and is the strict maximum of rows t-2 .. t+2;
// template = channel-1 value & 7, amplitude = channel-0 value / 8, sample =
// t - M/2 + nt0min.  p2p = max - min of channel 0 over the batch.
class StubMatcher : public lss::BatchMatcher {
public:
    StubMatcher(int C, int M, int nt0min, int dt) {
        info_ = lss_info{};
        info_.C = C;
        info_.T = 8;
        info_.M = M;
        info_.nt0min = nt0min;
        info_.dt = dt;
    }
    lss_info info() override { return info_; }
    void match(const int16_t* rows, int64_t n, int32_t ld, std::vector<lss_event>& ev, std::vector<float>& y,
               float& p2p) override {
        ev.clear();
        y.clear();
        int lo = 32767, hi = -32768;
        for (int64_t t = 0; t < n; ++t) {
            const int v = rows[t * ld];
            lo = std::min(lo, v);
            hi = std::max(hi, v);
            if (v <= 300) continue;
            bool peak = true;
            for (int64_t u = t - 2; u <= t + 2 && peak; ++u)
                if (u != t && u >= 0 && u < n && rows[u * ld] >= v) peak = false;
            if (!peak) continue;
            ev.push_back(lss_event{t - info_.M / 2 + info_.nt0min, rows[t * ld + 1] & 7, (float)v / 8.f});
        }
        p2p = n > 0 ? (float)(hi - lo) : 0.f;
    }

private:
    lss_info info_;
};

class FileSink : public lss::PayloadSink {
public:
    explicit FileSink(const std::string& path) : f_(path, std::ios::binary) {
        if (!f_) throw std::runtime_error("cannot write " + path);
    }
    void send(const std::vector<uint8_t>& p) override {
        const uint64_t n = p.size();
        f_.write(reinterpret_cast<const char*>(&n), sizeof n);
        f_.write(reinterpret_cast<const char*>(p.data()), (std::streamsize)p.size());
    }

private:
    std::ofstream f_;
};

}  // namespace

int main(int argc, char** argv) {
    if (argc != 13) {
        std::fprintf(stderr, "usage: adapter_replay <stub|gpu> <oss_input_dir> <ap.bin> <n_chan_file> <min_window> "
                             "<max_window> <time_behind_ms> <small_skip> <semantics> <preprocess> <counts.txt> <out_dir>\n");
        return 2;
    }
    try {
        const std::string mode = argv[1];
        lss::InputParameters p;
        p.sInputFolder = argv[2];
        const std::string bin = argv[3];
        const int n_chan_file = std::atoi(argv[4]);
        p.iMinScanWindow = std::atoi(argv[5]);
        p.iMaxScanWindow = std::atoi(argv[6]);
        p.iTimeBehind = std::atoi(argv[7]);
        p.bSmallskip = std::atoi(argv[8]) != 0;
        p.iSemantics = std::atoi(argv[9]);
        p.iPreprocess = std::atoi(argv[10]);
        std::vector<int64_t> counts;
        {
            std::ifstream f(argv[11]);
            long long c;
            while (f >> c) counts.push_back(c);
        }
        std::string out = argv[12];
        if (!out.empty() && out.back() != '/') out += '/';
        p.sOSSOutputFolder = out;
        lss::ReplayDataSocket sock(bin, n_chan_file, {}, p.iMaxScanWindow, p.iMinScanWindow, counts);
        FileSink sink(out + "payloads.bin");
        StubMatcher stub(sock.channels(), 61, 20, 15);
        lss::OnlineSpikesV2 sorter(p, &sink, &sock, mode == "stub" ? &stub : nullptr);
        sorter.runSpikeSorting();
        std::printf("batches %lld\n", (long long)sorter.batches());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "adapter_replay: %s\n", e.what());
        return 1;
    }
    return 0;
}
