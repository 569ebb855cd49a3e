"""Host-side streaming pieces (no GPU): the cereal wire formats against bytes
made by the reference's own headers (tests/golden/wire, written by
tests/golden/make_wire_vectors.py), spikeOutput.txt lines, the .ap.bin replay
source, the window policy and saveSpikes on hand-worked cases."""
from pathlib import Path

import numpy as np
import pytest

from livespikesorter_amd import EVENT_DTYPE, LssError
from livespikesorter_amd import stream as S
from livespikesorter_amd._capi import WIN_BIG_SKIP, WIN_CATCH_UP, WIN_LATEST, WIN_SKIP_LATEST, WIN_SMALL_SKIP

GOLD = Path(__file__).resolve().parent / "golden" / "wire"


def golden_payload():
    return S.Payload(recording_offset=123456789, stream_sample_ct=9876543210,
                     times=np.array([1000, 1003, 2047, 3000000001], np.int64),
                     templates=np.array([5, 17, 5, 1023], np.int64),
                     amplitudes=np.array([12.5, -3.25, 1e-3, 65504.0], np.float32),
                     vrms=7.125, p2p=812.0, process_time=431, event_stream_sample_ct=-7,
                     predict_label=3, label=-2, n_trials=40, n_correct=31, confidence=0.8125)


def test_payload_matches_reference_cereal_bytes():
    want = (GOLD / "payload_lp64.bin").read_bytes()
    assert S.encode_payload(golden_payload(), S.LONG_LP64) == want
    p = S.decode_payload(want, S.LONG_LP64)
    g = golden_payload()
    for f in ("recording_offset", "stream_sample_ct", "vrms", "p2p", "process_time", "event_stream_sample_ct",
              "predict_label", "label", "n_trials", "n_correct", "confidence"):
        assert getattr(p, f) == getattr(g, f), f
    assert np.array_equal(p.times, g.times) and np.array_equal(p.templates, g.templates)
    assert np.array_equal(p.amplitudes, g.amplitudes)
    assert S.encode_payload(S.Payload(), S.LONG_LP64) == (GOLD / "payload_empty_lp64.bin").read_bytes()


def test_payload_llp64_narrows_only_the_long_fields():
    """The Windows build's 4-byte `long`: the LP64 golden with each long field
    narrowed, everything else byte-identical."""
    g = (GOLD / "payload_lp64.bin").read_bytes()
    o, parts = 0, []

    def take(k):
        nonlocal o
        b = g[o:o + k]
        o += k
        return b

    def long_():
        parts.append((int.from_bytes(take(8), "little", signed=True) & 0xFFFFFFFF).to_bytes(4, "little"))

    long_(); long_()                                   # recordingOffset, streamSampleCt
    for _ in range(2):                                 # Times, Templates
        n = int.from_bytes(g[o:o + 8], "little")
        parts.append(take(8))
        for _ in range(n):
            long_()
    n = int.from_bytes(g[o:o + 8], "little")
    parts.append(take(8 + 4 * n))                      # Amplitudes
    parts.append(take(8 + 4))                          # VRMS, P2P
    long_(); long_()                                   # processTime, eventStreamSampleCt
    parts.append(take(8 + 8))                          # 4 x int16, confidence
    assert o == len(g)
    gp = golden_payload()
    gp.times = np.array([1000, 1003, 2047, -1294967295], np.int64)   # 3000000001 wraps in 32 bits
    gp.stream_sample_ct = int(np.int64(9876543210).astype(np.int32))
    assert S.encode_payload(golden_payload(), S.LONG_LLP64) == b"".join(parts)
    back = S.decode_payload(b"".join(parts), S.LONG_LLP64)
    assert np.array_equal(back.times, gp.times) and back.stream_sample_ct == gp.stream_sample_ct


def test_sorter_params_match_reference_cereal_bytes():
    p = S.SorterParameters(1024, 384, 61, 3122, 30000.0, np.array([0.0, 20.0, 43.5, -16.25]),
                           np.array([0, 1, 2, 7, 1023], np.int32))
    want = (GOLD / "sorter_params_lp64.bin").read_bytes()
    assert S.encode_sorter_params(p, S.LONG_LP64) == want
    q = S.decode_sorter_params(want, S.LONG_LP64)
    assert (q.T, q.C, q.M, q.N, q.samp_rate) == (1024, 384, 61, 3122, 30000.0)
    assert np.array_equal(q.chanpos, p.chanpos) and np.array_equal(q.neuron_indices, p.neuron_indices)
    assert len(S.encode_sorter_params(p, S.LONG_LLP64)) == len(want) - 16


def test_decode_rejects_malformed_bytes():
    g = (GOLD / "payload_lp64.bin").read_bytes()
    with pytest.raises(LssError):
        S.decode_payload(g[:-1], S.LONG_LP64)
    with pytest.raises(LssError):
        S.decode_payload(g + b"\0", S.LONG_LP64)
    bad = bytearray(g)
    bad[16:24] = (1 << 40).to_bytes(8, "little")      # absurd vector length
    with pytest.raises(LssError):
        S.decode_payload(bytes(bad), S.LONG_LP64)
    with pytest.raises(LssError):
        S.decode_payload(g, 2)


def test_spike_lines_use_ostream_defaults():
    ev = np.zeros(4, EVENT_DTYPE)
    ev["sample"] = [10, 11, 12, 3000000000]
    ev["template"] = [0, 7, 1023, 5]
    ev["amplitude"] = [12.5, 1e-3, 1234567.0, -3.25]
    y = np.array([0.0, 20.0, 43.5, 7654321.0], np.float32)
    text = S.format_spike_lines(ev, y, recording_offset=100)
    assert text == ("110,0,12.5,0\n111,7,0.001,20\n112,1023,1.23457e+06,43.5\n"
                    "3000000100,5,-3.25,7.65432e+06\n")
    assert text.splitlines()[0] == "%d,%d,%g,%g" % (110, 0, 12.5, 0.0)
    assert S.format_spike_lines(ev[:0]) == ""


def test_replay_source_reads_interleaved_channels(tmp_path):
    r = np.random.default_rng(1)
    x = r.integers(-2000, 2000, size=(5000, 384)).astype(np.int16)
    sync = (np.arange(5000) % 7).astype(np.int16)
    f = S.write_ap_bin(tmp_path / "rec.imec0.ap.bin", x, 385, sync)
    assert f.stat().st_size == 5000 * 385 * 2
    with S.ReplaySource(f) as src:
        assert (src.n_samples, src.n_channels) == (5000, 384)
        assert np.array_equal(src.fetch(0, 100), x[:100])
        assert np.array_equal(src.fetch(1234, 3000), x[1234:4234])
        assert np.array_equal(src.fetch(4900, 500), x[4900:])       # short at the end
        assert src.fetch(6000, 10).shape == (0, 384)
    with S.ReplaySource(f, 385, channels=[384, 3, 0]) as src:
        got = src.fetch(10, 20)
        assert np.array_equal(got[:, 0], sync[10:30])
        assert np.array_equal(got[:, 1:], x[10:30][:, [3, 0]])
    with pytest.raises(LssError):
        S.ReplaySource(f, 385, channels=[385])
    with pytest.raises(LssError):
        S.ReplaySource(tmp_path / "missing.bin")


CFG = dict(min_window=100, max_window=3000, samp_rate=30000.0, lookback=122)


def test_window_within_max():
    st = S.plan(S.WindowConfig(**CFG), 1000, 2200)
    assert st.kind == WIN_LATEST and not st.skip
    assert (st.fetch_start, st.fetch_n, st.buf_row, st.batch_row, st.batch_n) == (1000, 1200, 0, 0, 1300)
    assert (st.latest_ct, st.save_start, st.save_end, st.next_processed_ct) == (2200, 901, 1200, 2078)
    # processedCt ahead of the stream: the unsigned fetch length is capped at the maximum
    st = S.plan(S.WindowConfig(**CFG), 5000, 4000)
    assert st.kind == WIN_LATEST and (st.fetch_start, st.fetch_n, st.batch_n) == (1000, 3000, 100)


def test_window_behind_cases():
    st = S.plan(S.WindowConfig(**CFG, time_behind_ms=0), 1000, 10000)
    assert st.kind == WIN_SKIP_LATEST and st.skip
    assert (st.fetch_start, st.fetch_n, st.buf_row, st.batch_row, st.batch_n) == (7000, 3000, 100, 100, 3000)
    assert (st.latest_ct, st.save_start, st.save_end, st.next_processed_ct) == (10000, 7001, 2900, 9878)
    # 500 ms allowed: allowedCt = 10000 - 15000 < processedCt -> catch up from processedCt
    st = S.plan(S.WindowConfig(**CFG, time_behind_ms=500), 1000, 10000)
    assert st.kind == WIN_CATCH_UP and not st.skip and st.allowed_ct == -5000
    assert (st.fetch_start, st.fetch_n, st.buf_row, st.batch_row, st.batch_n) == (1000, 3000, 100, 0, 3100)
    assert (st.latest_ct, st.save_start, st.save_end, st.next_processed_ct) == (4000, 901, 3000, 3878)
    # 100 ms allowed: allowedCt = 7000 > processedCt -> small skip from allowedCt, or big skip
    st = S.plan(S.WindowConfig(**CFG, time_behind_ms=100, small_skip=True), 1000, 10000)
    assert st.kind == WIN_SMALL_SKIP and st.skip and st.allowed_ct == 7000
    assert (st.fetch_start, st.fetch_n, st.batch_row, st.batch_n, st.latest_ct) == (7000, 3000, 100, 3000, 10000)
    st = S.plan(S.WindowConfig(**CFG, time_behind_ms=100, small_skip=True), 1000, 8500)
    assert st.kind == WIN_SMALL_SKIP and (st.fetch_start, st.fetch_n, st.batch_n, st.latest_ct) == (5500, 3000, 3000, 8500)
    st = S.plan(S.WindowConfig(**CFG, time_behind_ms=100), 1000, 10000)
    assert st.kind == WIN_BIG_SKIP and (st.fetch_start, st.fetch_n, st.batch_row, st.batch_n) == (7000, 3000, 100, 3000)


def test_window_allowed_count_is_computed_in_fp32():
    # 2**24 + 1 is not a float: latestCt rounds to 2**24 before the 30 samples are taken off
    st = S.plan(S.WindowConfig(**CFG, time_behind_ms=1), 0, (1 << 24) + 1)
    assert st.allowed_ct == (1 << 24) - 30


def test_save_spikes_rules():
    ev = np.zeros(5, EVENT_DTYPE)
    # batch-relative samples s: index = s - M - lookback = s - 183, time = 1000 + index
    ev["sample"] = [210, 200, 205, 700, 190]
    ev["template"] = [3, 3, 3, 4, 6]
    ev["amplitude"] = [1, 2, 3, 4, 5]
    last = np.zeros(8, np.int64)
    last[6] = 1005
    out, src = S.save_spikes(ev, M=61, lookback=122, dt=7, save_start=1000, save_end=500, last_spike=last,
                             with_index=True)
    # template 3 in the order given: 1027 kept; 1017 and 1022 come after it and the signed
    # difference to the last kept time (-10, -5) is below dt, so both are dropped;
    # template 4 past the end: dropped, but its last time is updated first; template 6: 1007 - 1005 < 7
    assert list(out["sample"]) == [1027]
    assert list(out["amplitude"]) == [1] and list(src) == [0]
    assert last[3] == 1027 and last[4] == 1517 and last[6] == 1005
    # in time order the 1017 and 1027 spikes (10 apart) both survive, 1022 does not
    ev_t = ev[[1, 2, 0]]
    out_t = S.save_spikes(ev_t, M=61, lookback=122, dt=7, save_start=1000, save_end=500,
                          last_spike=np.zeros(8, np.int64))
    assert list(out_t["sample"]) == [1017, 1027]
    # the next batch: template 4 again 3 samples after the suppressed tail spike -> dropped
    ev2 = np.zeros(1, EVENT_DTYPE)
    ev2["sample"], ev2["template"] = 183 + 20, 4
    assert len(S.save_spikes(ev2, M=61, lookback=122, dt=7, save_start=1500, save_end=500, last_spike=last)) == 0
    # zero-initialised last times suppress spikes in the stream's first dt samples
    ev3 = np.zeros(1, EVENT_DTYPE)
    ev3["sample"], ev3["template"] = 183 + 3, 0
    assert len(S.save_spikes(ev3, M=61, lookback=122, dt=7, save_start=0, save_end=500,
                             last_spike=np.zeros(8, np.int64))) == 0
    # equal times keep insertion order (upper_bound)
    ev4 = np.zeros(3, EVENT_DTYPE)
    ev4["sample"] = [300, 300, 250]
    ev4["template"] = [1, 2, 5]
    out4 = S.save_spikes(ev4, M=61, lookback=122, dt=7, save_start=1000, save_end=500,
                         last_spike=np.zeros(8, np.int64))
    assert list(out4["template"]) == [5, 1, 2]
    with pytest.raises(LssError):
        S.save_spikes(ev4, M=61, lookback=122, dt=7, save_start=0, save_end=500, last_spike=np.zeros(2, np.int64))
