"""The live loop around the matcher, over the C ABI's host-side pieces.

Mirrors ``OnlineSpikesV2::runSpikeSorting`` (src/SpikeSorter/OnlineSpikesV2.cu:
662-850) for a recorded stream: the window and skip policy
(:698-745, :class:`WindowConfig` / :func:`plan`), the persistent fetch buffer
the batch is cut from (:750-757), ``saveSpikes`` (:1173-1206,
:func:`save_spikes`), the ``OnlineSpikesPayload`` sent per batch
(:825-838, :func:`encode_payload`) and ``spikeOutput.txt``
(:1148-1166, :func:`format_spike_lines`).  :class:`ReplaySource` plays an
``.ap.bin`` recording in place of the SpikeGLX socket; :func:`run_replay`
drives the loop with a stream clock given by the caller.
"""
from __future__ import annotations

import ctypes as ct
import dataclasses
import os
from pathlib import Path

import numpy as np

from ._capi import (EVENT_DTYPE, PayloadHeader, SorterParams, WindowCfg, WindowStep, check, lib)

LONG_LLP64 = 4      # `long` of the reference's Windows build
LONG_LP64 = 8


@dataclasses.dataclass
class WindowConfig:
    """runSpikeSorting's window parameters (OnlineSpikesV2.cu:254-267)."""
    min_window: int
    max_window: int
    time_behind_ms: int = 0
    samp_rate: float = 30000.0
    small_skip: bool = False
    lookback: int = 122          # 2M (OnlineSpikesV2.cu:461)

    def _c(self) -> WindowCfg:
        return WindowCfg(self.min_window, self.max_window, self.time_behind_ms, self.samp_rate,
                         int(self.small_skip), self.lookback)


def plan(cfg: WindowConfig, processed_ct: int, stream_ct: int) -> WindowStep:
    """The next batch (lss_window_plan)."""
    st = WindowStep()
    check(lib().lss_window_plan(ct.byref(cfg._c()), int(processed_ct), int(stream_ct), ct.byref(st)))
    return st


def save_spikes(events: np.ndarray, *, M: int, lookback: int, dt: int, save_start: int, save_end: int,
                last_spike: np.ndarray, with_index: bool = False):
    """saveSpikes over one batch's events (first_sample 0); ``last_spike`` (int64 per
    template, zero-initialised) is updated in place.  With ``with_index`` also
    the index in ``events`` of each kept spike."""
    ev = np.ascontiguousarray(events, dtype=EVENT_DTYPE)
    if last_spike.dtype != np.int64 or not last_spike.flags.c_contiguous:
        raise TypeError("last_spike must be a contiguous int64 array")
    out = np.zeros(max(len(ev), 1), EVENT_DTYPE)
    src = np.zeros(max(len(ev), 1), np.int64)
    n = ct.c_int64()
    check(lib().lss_save_spikes(ev.ctypes.data, len(ev), M, lookback, dt, int(save_start), int(save_end),
                                last_spike.ctypes.data, len(last_spike), out.ctypes.data, src.ctypes.data, len(out),
                                ct.byref(n)))
    if with_index:
        return out[:n.value].copy(), src[:n.value].copy()
    return out[:n.value].copy()


def format_spike_lines(events: np.ndarray, y=None, recording_offset: int = 0) -> str:
    """spikeOutput.txt lines for ``events`` (lss_format_spikes)."""
    ev = np.ascontiguousarray(events, dtype=EVENT_DTYPE)
    yy = None if y is None else np.ascontiguousarray(y, dtype=np.float32)
    size = ct.c_int64()
    yp = None if yy is None else yy.ctypes.data
    check(lib().lss_format_spikes(int(recording_offset), ev.ctypes.data, yp, len(ev), None, 0, ct.byref(size)))
    buf = ct.create_string_buffer(size.value + 1)
    check(lib().lss_format_spikes(int(recording_offset), ev.ctypes.data, yp, len(ev), buf, size.value,
                                  ct.byref(size)))
    return buf.raw[:size.value].decode()


@dataclasses.dataclass
class Payload:
    """OnlineSpikesPayload (src/Networking/onlineSpikesPayload.h:8-48)."""
    recording_offset: int = 0
    stream_sample_ct: int = 0
    times: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int64))
    templates: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int64))
    amplitudes: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.float32))
    vrms: float = 0.0
    p2p: float = 0.0
    process_time: int = 0
    event_stream_sample_ct: int = 0
    predict_label: int = 0
    label: int = 0
    n_trials: int = 0
    n_correct: int = 0
    confidence: float = 0.0

    def _header(self) -> PayloadHeader:
        return PayloadHeader(self.recording_offset, self.stream_sample_ct, self.vrms, self.p2p, self.process_time,
                             self.event_stream_sample_ct, self.predict_label, self.label, self.n_trials,
                             self.n_correct, self.confidence)


def encode_payload(p: Payload, long_bytes: int = LONG_LLP64) -> bytes:
    t = np.ascontiguousarray(p.times, dtype=np.int64)
    k = np.ascontiguousarray(p.templates, dtype=np.int64)
    a = np.ascontiguousarray(p.amplitudes, dtype=np.float32)
    if not (len(t) == len(k) == len(a)):
        raise ValueError("times, templates and amplitudes must have one length")
    h = p._header()
    size = ct.c_int64()
    check(lib().lss_payload_encode(ct.byref(h), t.ctypes.data, k.ctypes.data, a.ctypes.data, len(t), long_bytes,
                                   None, 0, ct.byref(size)))
    buf = (ct.c_uint8 * size.value)()
    check(lib().lss_payload_encode(ct.byref(h), t.ctypes.data, k.ctypes.data, a.ctypes.data, len(t), long_bytes,
                                   buf, size.value, ct.byref(size)))
    return bytes(buf)


def decode_payload(data: bytes, long_bytes: int = LONG_LLP64) -> Payload:
    src = (ct.c_uint8 * len(data)).from_buffer_copy(data)
    h = PayloadHeader()
    n = ct.c_int64()
    check(lib().lss_payload_decode(src, len(data), long_bytes, ct.byref(h), None, None, None, 0, ct.byref(n)))
    t = np.zeros(n.value, np.int64)
    k = np.zeros(n.value, np.int64)
    a = np.zeros(n.value, np.float32)
    check(lib().lss_payload_decode(src, len(data), long_bytes, ct.byref(h), t.ctypes.data, k.ctypes.data,
                                   a.ctypes.data, n.value, ct.byref(n)))
    return Payload(h.recording_offset, h.stream_sample_ct, t, k, a, h.vrms, h.p2p, h.process_time,
                   h.event_stream_sample_ct, h.predict_label, h.label, h.n_trials, h.n_correct, h.confidence)


@dataclasses.dataclass
class SorterParameters:
    """SorterParameters (src/Networking/sorterParameters.h:6-29)."""
    T: int
    C: int
    M: int
    N: int
    samp_rate: float
    chanpos: np.ndarray
    neuron_indices: np.ndarray


def encode_sorter_params(p: SorterParameters, long_bytes: int = LONG_LLP64) -> bytes:
    cp = np.ascontiguousarray(p.chanpos, dtype=np.float64)
    ni = np.ascontiguousarray(p.neuron_indices, dtype=np.int32)
    sp = SorterParams(p.T, p.C, p.M, p.N, p.samp_rate)
    size = ct.c_int64()
    check(lib().lss_sorter_params_encode(ct.byref(sp), cp.ctypes.data, len(cp), ni.ctypes.data, len(ni), long_bytes,
                                         None, 0, ct.byref(size)))
    buf = (ct.c_uint8 * size.value)()
    check(lib().lss_sorter_params_encode(ct.byref(sp), cp.ctypes.data, len(cp), ni.ctypes.data, len(ni), long_bytes,
                                         buf, size.value, ct.byref(size)))
    return bytes(buf)


def decode_sorter_params(data: bytes, long_bytes: int = LONG_LLP64) -> SorterParameters:
    src = (ct.c_uint8 * len(data)).from_buffer_copy(data)
    sp = SorterParams()
    nc, ni = ct.c_int64(), ct.c_int64()
    check(lib().lss_sorter_params_decode(src, len(data), long_bytes, ct.byref(sp), None, 0, ct.byref(nc), None, 0,
                                         ct.byref(ni)))
    cp = np.zeros(nc.value, np.float64)
    ix = np.zeros(ni.value, np.int32)
    check(lib().lss_sorter_params_decode(src, len(data), long_bytes, ct.byref(sp), cp.ctypes.data, len(cp),
                                         ct.byref(nc), ix.ctypes.data, len(ix), ct.byref(ni)))
    return SorterParameters(sp.T, sp.C, sp.M, sp.N, sp.samp_rate, cp, ix)


class ReplaySource:
    """A recorded ``.ap.bin`` (interleaved int16) played as the stream (lss_replay_*)."""

    def __init__(self, path: str | os.PathLike, n_chan_file: int = 385, channels=None):
        self._r = ct.c_void_p()
        ch = None if channels is None else np.ascontiguousarray(channels, dtype=np.int32)
        check(lib().lss_replay_open(os.fsencode(str(path)), n_chan_file, None if ch is None else ch.ctypes.data,
                                    0 if ch is None else len(ch), ct.byref(self._r)))
        n, c = ct.c_int64(), ct.c_int32()
        check(lib().lss_replay_samples(self._r, ct.byref(n), ct.byref(c)))
        self.n_samples, self.n_channels = n.value, c.value

    def fetch(self, start: int, n: int) -> np.ndarray:
        out = np.zeros((max(n, 0), self.n_channels), np.int16)
        got = ct.c_int64()
        check(lib().lss_replay_fetch(self._r, int(start), int(n), out.ctypes.data, ct.byref(got)))
        return out[:got.value]

    def close(self) -> None:
        if self._r:
            lib().lss_replay_close(self._r)
            self._r = ct.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


def write_ap_bin(path: str | os.PathLike, x: np.ndarray, n_chan_file: int = 385, sync=None) -> Path:
    """Write int16 samples [n][C] as an interleaved .ap.bin of n_chan_file channels
    (the extra channels, the sync channel last, hold ``sync`` or zeros)."""
    x = np.asarray(x, dtype=np.int16)
    n, C = x.shape
    if C > n_chan_file:
        raise ValueError("more channels than the file holds")
    full = np.zeros((n, n_chan_file), np.int16)
    full[:, :C] = x
    if sync is not None:
        full[:, -1] = np.asarray(sync, dtype=np.int16)
    p = Path(path)
    full.tofile(p)
    return p


@dataclasses.dataclass
class LoopBatch:
    step: WindowStep
    events: np.ndarray        # saved spikes (reference times), after saveSpikes
    y: np.ndarray             # their positions' y (0 without probe geometry)
    payload: bytes


def run_replay(sorter, source: ReplaySource, cfg: WindowConfig, stream_counts, *, recording_offset: int = 0,
               long_bytes: int = LONG_LLP64, spike_file=None, start_ct: int = 0,
               stale_rows: str = "keep") -> list:
    """Run the live loop over ``source`` with the stream's sample count taking the
    successive values of ``stream_counts`` at each batch (the SpikeGLX clock):
    plan the window, fetch into the persistent buffer, match the batch cut from
    it with ``sorter`` (an :class:`~livespikesorter_amd.OnlineSpikesV2`, built
    with max_batch >= max_window + min_window), save the spikes, encode the
    payload and append spikeOutput.txt lines.  ``start_ct`` seeds processedCt
    and latestCt (OnlineSpikesV2.cu:693-694).

    Within max_window the reference's batch is the fetch plus min_window more
    buffer rows (currBatchNumSamples, :715), which still hold an earlier
    fetch's samples.  ``stale_rows="keep"`` matches those rows as the reference
    does; ``"zero"`` clears them first (a departure from the reference)."""
    if stale_rows not in ("keep", "zero"):
        raise ValueError("stale_rows is 'keep' or 'zero'")
    info = sorter.info
    C = source.n_channels
    rows = cfg.min_window + cfg.max_window
    buf = np.zeros((rows, C), np.int16)          # fetchBuf (its stale rows are part of the behaviour)
    # lastSpikeTime, zero-initialised; indexed by template (or final cluster when relabelling)
    last_spike = np.zeros(max(info.T, 1 << 16), np.int64)
    processed = latest = int(start_ct)
    with_pos = _has_positions(sorter)
    out = []
    fh = open(spike_file, "a") if spike_file is not None else None
    try:
        for ct_now in stream_counts:
            ct_now = int(ct_now)
            if ct_now < latest + cfg.min_window:
                raise ValueError("the stream count must pass latestCt + min_window (waitUntil)")
            st = plan(cfg, processed, ct_now)
            data = source.fetch(st.fetch_start, st.fetch_n)
            buf[st.buf_row:st.buf_row + len(data)] = data
            if stale_rows == "zero":
                buf[st.buf_row + len(data):] = 0
            batch = buf[st.batch_row:st.batch_row + st.batch_n]
            ev, _ = sorter.process_batch(batch)
            y_all = sorter.last_positions()[:, 1] if with_pos else np.zeros(len(ev), np.float32)
            saved, src = save_spikes(ev, M=info.M, lookback=cfg.lookback, dt=info.dt, save_start=st.save_start,
                                     save_end=st.save_end, last_spike=last_spike, with_index=True)
            y = y_all[src] if len(src) else np.zeros(0, np.float32)
            payload = encode_payload(Payload(recording_offset, st.latest_ct, saved["sample"].astype(np.int64),
                                             saved["template"].astype(np.int64), saved["amplitude"],
                                             p2p=float(sorter.query().p2p)), long_bytes)
            if fh is not None:
                fh.write(format_spike_lines(saved, y, recording_offset))
            out.append(LoopBatch(st, saved, y, payload))
            latest = st.latest_ct
            processed = st.next_processed_ct
    finally:
        if fh is not None:
            fh.close()
    return out


def _has_positions(sorter) -> bool:
    from ._capi import LssError
    try:
        sorter.last_positions()
        return True
    except LssError:
        return False


class ZScoreBaseline:
    """Long-run z-score baseline of binned spike counts on the GPU, FP64
    (lss_zscore_*; the Decoder's ZScoreSdmProcessor, src/Decoder/
    ZScoreSdmProcessor.cpp).  Series 0 is the total count, 1 + t template t."""

    def __init__(self, n_templates: int, *, bin_ms: float = 50.0, samp_rate: float = 30000.0,
                 bin_samples: int | None = None, ring_bins: int = 4096, device: int = 0):
        if bin_samples is None:
            b = ct.c_int64()
            check(lib().lss_zscore_bin_samples(bin_ms, samp_rate, ct.byref(b)))
            bin_samples = b.value
        self.bin_samples = int(bin_samples)
        self.S = 1 + n_templates
        self._z = ct.c_void_p()
        check(lib().lss_zscore_create(device, n_templates, self.bin_samples, ring_bins, ct.byref(self._z)))

    def add(self, events: np.ndarray) -> None:
        ev = np.ascontiguousarray(events, dtype=EVENT_DTYPE)
        check(lib().lss_zscore_add(self._z, ev.ctypes.data, len(ev)))

    def complete(self, end_sample: int, cap_bins: int = 4096):
        """Close the bins ending at or before ``end_sample``: (z [bins][S] f32, direction int8)."""
        z = np.zeros((cap_bins, self.S), np.float32)
        d = np.zeros((cap_bins, self.S), np.int8)
        n = ct.c_int64()
        check(lib().lss_zscore_complete(self._z, int(end_sample), z.ctypes.data, d.ctypes.data, cap_bins,
                                        ct.byref(n)))
        return z[:n.value].copy(), d[:n.value].copy()

    def freeze(self, mean=None, sd=None, trigger_z: float = 1.0) -> None:
        if (mean is None) != (sd is None):
            raise ValueError("give both mean and sd, or neither")
        m = None if mean is None else np.ascontiguousarray(mean, dtype=np.float64)
        s = None if sd is None else np.ascontiguousarray(sd, dtype=np.float64)
        check(lib().lss_zscore_freeze(self._z, None if m is None else m.ctypes.data,
                                      None if s is None else s.ctypes.data, trigger_z))

    def stats(self) -> dict:
        seen = np.zeros(self.S, np.int64)
        mean = np.zeros(self.S, np.float64)
        m2 = np.zeros(self.S, np.float64)
        late, first = ct.c_int64(), ct.c_int64()
        check(lib().lss_zscore_stats(self._z, seen.ctypes.data, mean.ctypes.data, m2.ctypes.data, ct.byref(late),
                                     ct.byref(first)))
        return {"bins_seen": seen, "mean": mean, "m2": m2, "late": late.value, "first_open": first.value}

    def close(self) -> None:
        if self._z:
            lib().lss_zscore_destroy(self._z)
            self._z = ct.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


def sorter_parameters(oss_input_dir, batch_samples: int, samp_rate: float = 30000.0) -> SorterParameters:
    """getSorterParams (OnlineSpikesV2.cu:1076-1086) from an oss_input directory:
    T final clusters, C channels and M samples from templates.npy, N the batch
    size, the active channels 0..C-1 as positions and templateMap.npy
    (loadTemplateMap, :568-577)."""
    d = Path(oss_input_dir)
    tpl = np.load(d / "templates.npy", mmap_mode="r")
    T, M, C = tpl.shape
    tmap = np.load(d / "templateMap.npy").ravel().astype(np.int32)
    if len(tmap) != T:
        raise ValueError("templateMap.npy: incorrect template map size")
    return SorterParameters(int(T), int(C), int(M), int(batch_samples), float(samp_rate),
                            np.arange(C, dtype=np.float64), tmap)
