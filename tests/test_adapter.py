"""The C++ OnlineSpikesV2 adapter (include/lss_online_spikes.hpp) against the
Python live loop (livespikesorter_amd.stream.run_replay) on the same replayed
recording and stream clock: the same payloads (decoded, apart from the
measured processing time) and the same spikeOutput.txt.

The adapter mirrors OnlineSpikesV2's constructor and runSpikeSorting
(src/SpikeSorter/OnlineSpikesV2.h:29-34, OnlineSpikesV2.cu:662-850) over a
DataSocket (src/SpikeSorter/dataSocket.h:23-57).  tests/cpp/adapter_replay
drives it.  With its host "stub" matcher the loop's bookkeeping (window
policy, the fetch buffer's stale rows, saveSpikes, payload, spike lines) is
compared on the CPU through all five window kinds; the GPU test runs the
library's matcher in both loops.
"""
import os
import subprocess
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

from livespikesorter_amd import stream as S
from livespikesorter_amd._capi import (EVENT_DTYPE, SEM_KILOSORT, WIN_BIG_SKIP, WIN_CATCH_UP, WIN_LATEST,
                                       WIN_SKIP_LATEST, WIN_SMALL_SKIP, LssError)

BIN = Path(__file__).resolve().parent / "cpp" / "adapter_replay"


def _driver():
    if not BIN.exists():
        raise RuntimeError(f"{BIN} is not built (make -C tests/cpp, part of __graft_entry__.build())")
    return str(BIN)


class _StubSorter:
    """The host rule of adapter_replay's StubMatcher: row t is an event when its
    channel-0 value exceeds 300 and is the strict maximum of rows t-2 .. t+2;
    template = channel-1 value & 7, amplitude = value / 8, sample = t - M/2 + nt0min."""

    def __init__(self, C):
        self.info = SimpleNamespace(C=C, T=8, M=61, nt0min=20, dt=15)
        self._p2p = 0.0

    def process_batch(self, batch):
        v = batch[:, 0].astype(np.int32)
        n = len(v)
        peak = v > 300
        for d in (-2, -1, 1, 2):
            nb = np.full(n, np.iinfo(np.int32).min, np.int32)
            if d < 0:
                nb[-d:] = v[:n + d]
            else:
                nb[:n - d] = v[d:]
            peak &= nb < v
        t = np.flatnonzero(peak)
        ev = np.zeros(len(t), EVENT_DTYPE)
        ev["sample"] = t - self.info.M // 2 + self.info.nt0min
        ev["template"] = batch[t, 1].astype(np.int32) & 7
        ev["amplitude"] = v[t].astype(np.float32) / np.float32(8)
        self._p2p = float(v.max() - v.min()) if n else 0.0
        return ev, None

    def last_positions(self):
        raise LssError(-5, "no probe geometry")

    def query(self):
        return SimpleNamespace(p2p=self._p2p)


def _recording(n, C, seed):
    r = np.random.default_rng(seed)
    x = r.integers(-100, 100, size=(n, C)).astype(np.int16)
    for t in r.choice(n, size=n // 150, replace=False):
        x[t, 0] = r.integers(301, 900)
    return x


def _clock(r, n, start, steps, jumps):
    """start, then per batch 1500 + U(0, 300) samples, with the given large jumps
    at random batches; stops before the recording's end."""
    counts, c = [start], start
    for k in range(steps):
        c += 1500 + int(r.integers(0, 300)) + (jumps[k % len(jumps)] if r.random() < 0.25 else 0)
        if c > n:
            break
        counts.append(c)
    return counts


def _run_cpp(mode, oss, ap, n_chan_file, cfg, semantics, preprocess, counts, out):
    out.mkdir(parents=True, exist_ok=True)
    cf = out / "counts.txt"
    cf.write_text("\n".join(str(c) for c in counts) + "\n")
    # The driver is a second process on the GPU while this one holds its own HIP
    # context: its cooperative peel launch is refused there (hipErrorIllegalState),
    # so it launches the same co-resident grid plainly (this process is idle meanwhile).
    env = dict(os.environ, LSS_PEEL_LAUNCH="plain")
    res = subprocess.run([_driver(), mode, str(oss), str(ap), str(n_chan_file), str(cfg.min_window),
                          str(cfg.max_window), str(cfg.time_behind_ms), str(int(cfg.small_skip)), str(semantics),
                          str(preprocess), str(cf), str(out)], capture_output=True, text=True, timeout=300, env=env)
    assert res.returncode == 0, res.stderr
    data = (out / "payloads.bin").read_bytes()
    pays, i = [], 0
    while i < len(data):
        k = int(np.frombuffer(data[i:i + 8], np.uint64)[0])
        pays.append(data[i + 8:i + 8 + k])
        i += 8 + k
    text = (out / "spikeOutput.txt").read_text() if (out / "spikeOutput.txt").exists() else ""
    return pays, text


def _no_time(payload: bytes) -> bytes:
    p = S.decode_payload(payload)
    p.process_time = 0
    return S.encode_payload(p)


def _compare(py_out, py_text, cpp_pays, cpp_text):
    assert len(cpp_pays) == len(py_out)
    for b, q in zip(py_out, cpp_pays):
        assert _no_time(q) == b.payload
    assert cpp_text == py_text


@pytest.mark.parametrize("time_behind,small_skip,jumps,kinds", [
    (0, False, [4000], {WIN_LATEST, WIN_SKIP_LATEST}),
    (200, False, [3500, 9000], {WIN_LATEST, WIN_CATCH_UP, WIN_BIG_SKIP}),
    (200, True, [3500, 9000], {WIN_LATEST, WIN_CATCH_UP, WIN_SMALL_SKIP}),
], ids=["skip-latest", "catch-up-big-skip", "small-skip"])
def test_adapter_loop_matches_python_loop_stub(tmp_path, time_behind, small_skip, jumps, kinds):
    C, n_chan_file, n = 8, 9, 240000
    x = _recording(n, C, seed=11 + time_behind + int(small_skip))
    ap = S.write_ap_bin(tmp_path / "rec.ap.bin", x, n_chan_file)
    cfg = S.WindowConfig(min_window=1500, max_window=3000, time_behind_ms=time_behind, small_skip=small_skip,
                         lookback=122)
    counts = _clock(np.random.default_rng(5), n, 2000, 60, jumps)
    with S.ReplaySource(ap, n_chan_file) as src:
        py = S.run_replay(_StubSorter(C), src, cfg, counts[1:], spike_file=tmp_path / "py.txt", start_ct=counts[0])
    assert kinds <= {b.step.kind for b in py}, "the clock did not reach every window kind of this setting"
    assert sum(len(b.events) for b in py) > 100
    pays, text = _run_cpp("stub", "-", ap, n_chan_file, cfg, SEM_KILOSORT, 0, counts, tmp_path / "cpp")
    _compare(py, (tmp_path / "py.txt").read_text(), pays, text)


@pytest.mark.gpu
def test_adapter_loop_matches_python_loop_gpu(model_dirs, tmp_path):
    """Both loops with the library's matcher on the GPU (Kilosort semantics, no
    preprocessing) over the same replayed cfg1 recording."""
    from livespikesorter_amd import OnlineSpikesV2, synth

    m, d = model_dirs("cfg1")
    n = 60000
    x, _ = synth.make_recording(m, n, seed=synth.TEST_SEED + 9)
    ap = S.write_ap_bin(tmp_path / "rec.ap.bin", x, m.C + 1)
    cfg = S.WindowConfig(min_window=1500, max_window=3000, lookback=2 * m.M)
    counts = _clock(np.random.default_rng(8), n - 3000, 0, 40, [4000])
    print(f"[adapter gpu] {len(counts) - 1} batches", flush=True)
    with S.ReplaySource(ap, m.C + 1) as src, OnlineSpikesV2(d, max_batch=4500) as sp:
        py = S.run_replay(sp, src, cfg, counts[1:], spike_file=tmp_path / "py.txt", start_ct=counts[0])
    assert sum(len(b.events) for b in py) > 100
    pays, text = _run_cpp("gpu", d, ap, m.C + 1, cfg, SEM_KILOSORT, 0, counts, tmp_path / "cpp")
    _compare(py, (tmp_path / "py.txt").read_text(), pays, text)
