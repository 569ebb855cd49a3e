"""Python mirror of the reference's sorter interface over the C ABI.

``OnlineSpikesV2`` here keeps the reference's names and argument meaning for
the matcher path (src/SpikeSorter/OnlineSpikesV2.h:29-56): it is built from an
``oss_input`` directory and matches batches of samples, returning the
``(sample, template, amplitude)`` events and their PC features.  All work is
done by ``liblss.so``; errors come back as :class:`LssError` with the
library's message (the reference throws ``std::runtime_error`` via
``_RUN_ERROR``, src/Helpers/Utils.h:41-46).
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

from ._capi import (EVENT_DTYPE, Config, Info, LssError, PREC_BF16_RESCORE, PREC_BF16X3, PREC_FP32, SEM_KILOSORT, SEM_LSS, check,
                    default_config, lib)

__all__ = ["OnlineSpikesV2", "StreamGroup", "LssError", "SEM_KILOSORT", "SEM_LSS", "PREC_FP32",
           "PREC_BF16_RESCORE", "PREC_BF16X3"]

(STAGE_XW, STAGE_BPC, STAGE_B, STAGE_CFMAX, STAGE_IMAX, STAGE_NM, STAGE_PHASES, STAGE_CTC, STAGE_PRE,
 STAGE_PRE_MEAN, STAGE_PRE_MED) = range(11)

# the peel kernel's phase clock (STAGE_PHASES): per phase the sum over workgroups,
# then the largest workgroup's, then scan passes and workgroups
PHASE_NAMES = ("exchange", "scan", "publish", "accept", "wait", "epilogue", "subtract", "barrier", "reduce")
NPHASE = len(PHASE_NAMES)


class OnlineSpikesV2:
    """One matcher instance (one probe stream) on one HIP device."""

    def __init__(self, oss_input_dir: str | os.PathLike, *, device: int = 0, chunk: int = 3000,
                 max_batch: int = 0, max_iter: int = 0, semantics: int = SEM_KILOSORT,
                 precision: int = PREC_FP32, dedup: int = -1, max_events: int = 65536,
                 keep_stages: bool = False, sparse: int = 2, profile: int = 0,
                 peel_split: int = 0, relabel: bool = False, preprocess: int = 0, keep=None):
        """``keep``: ascending Wall3 template indices to load (a template subset,
        lss_create_subset); event template ids are then positions in ``keep``."""
        self._h = ct.c_void_p()
        cfg = default_config(device=device, chunk=chunk, max_batch=max_batch, max_iter=max_iter,
                             semantics=semantics, precision=precision, dedup=dedup,
                             max_events=max_events, keep_stages=int(keep_stages), sparse=int(sparse),
                             profile=int(profile), peel_split=peel_split, relabel=int(relabel),
                             preprocess=preprocess)
        if keep is None:
            check(lib().lss_create(os.fsencode(str(oss_input_dir)), ct.byref(cfg), ct.byref(self._h)))
        else:
            k = np.ascontiguousarray(keep, dtype=np.int32)
            check(lib().lss_create_subset(os.fsencode(str(oss_input_dir)), ct.byref(cfg), k.ctypes.data, len(k),
                                          ct.byref(self._h)))
        self.info = self.query()
        self.last_W = 0

    def last_positions(self) -> np.ndarray:
        """(x, y) per event of the last call, in event order (lss_last_positions)."""
        n = ct.c_int64()
        cap = 65536
        out = np.empty((cap, 2), np.float32)
        check(lib().lss_last_positions(self._h, out.ctypes.data, cap, ct.byref(n)))
        return out[:n.value].copy()

    def autotune(self, dev_ptr: int, n_samples: int, ld: int, reps: int = 5) -> list:
        """Time each peel split on a device batch and keep the fastest for this
        width (lss_autotune).  Returns [(split, rows, ms, chosen)]."""
        from ._capi import TuneResult
        r = TuneResult()
        check(lib().lss_autotune(self._h, ct.c_void_p(dev_ptr), n_samples, ld, reps, ct.byref(r)))
        return [(r.split[i], r.rows[i], r.ms[i], i == r.chosen) for i in range(r.n)]

    def set_proposal(self, net, lo: float = 0.25, hi: float = 0.75, reach: int = 4) -> None:
        """Gate score tiles with a :class:`~livespikesorter_amd.proposal.ProposalNet`
        (``None`` detaches it).  Keep the network alive while it is attached."""
        check(lib().lss_set_proposal(self._h, net._p if net is not None else None, lo, hi, reach))
        self._prop = net

    # -- lifecycle ---------------------------------------------------------
    def close(self) -> None:
        if self._h:
            check(lib().lss_destroy(self._h))
            self._h = ct.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- queries -----------------------------------------------------------
    def query(self) -> Info:
        info = Info()
        check(lib().lss_query(self._h, ct.byref(info)))
        return info

    def _out(self, cap: int):
        # output buffers are allocated once per capacity and reused (a fresh
        # 16 MB feature buffer per call costs more than a whole chunk); results
        # are returned as copies of the filled part
        if getattr(self, "_bufs", None) is None or self._bufs[0] != cap:
            nC, K = self.info.nC, self.info.K
            self._bufs = (cap, np.zeros(cap, EVENT_DTYPE), np.zeros((cap, nC, K), np.float32))
        return self._bufs[1], self._bufs[2]

    # -- matching ------------------------------------------------------------
    def kilosortMatchingPursuit(self, batch: np.ndarray, first_sample: int = 0, *, chan_map=None,
                                cap: int = 65536):
        """Match one framed batch ``[W][ld]`` int16 (columns [M, W-M) are eligible).

        Counterpart of ``long kilosortMatchingPursuit(float* d_batch, long W)``
        (OnlineSpikesV2.h:56); returns ``(events, features)``.
        """
        x = np.ascontiguousarray(batch, dtype=np.int16)
        W, ld = x.shape
        cm = None if chan_map is None else np.ascontiguousarray(chan_map, dtype=np.int32)
        ev, feat = self._out(cap)
        n = ct.c_int64(0)
        check(lib().lss_process_batch(self._h, x.ctypes.data, W, ld,
                                      None if cm is None else cm.ctypes.data, int(first_sample),
                                      ev.ctypes.data, cap, ct.byref(n), feat.ctypes.data))
        self.last_W = W
        return ev[:n.value].copy(), feat[:n.value].copy()

    process_batch = kilosortMatchingPursuit

    def process_batch_device(self, dev_ptr: int, n_samples: int, ld: int, first_sample: int = 0, *,
                             chan_map=None, cap: int = 65536, features: bool = True):
        """Match a batch already resident in device memory (``dev_ptr`` from the caller)."""
        cm = None if chan_map is None else np.ascontiguousarray(chan_map, dtype=np.int32)
        ev, feat = self._out(cap)
        n = ct.c_int64(0)
        check(lib().lss_process_batch_device(self._h, ct.c_void_p(dev_ptr), n_samples, ld,
                                             None if cm is None else cm.ctypes.data,
                                             int(first_sample), ev.ctypes.data, cap, ct.byref(n),
                                             feat.ctypes.data if features else None))
        self.last_W = n_samples
        return ev[:n.value].copy(), feat[:n.value].copy()

    def process_batch_device_static(self, dev_ptr: int, n_samples: int, ld: int, first_sample: int = 0, *,
                                    chan_map=None, cap: int = 65536, features: bool = True):
        """:meth:`process_batch_device` through the allocation-free entry point
        (lss_process_batch_device_static: the batch's device buffers come from the
        library's static scratch array)."""
        cm = None if chan_map is None else np.ascontiguousarray(chan_map, dtype=np.int32)
        ev, feat = self._out(cap)
        n = ct.c_int64(0)
        check(lib().lss_process_batch_device_static(self._h, ct.c_void_p(dev_ptr), n_samples, ld,
                                                    None if cm is None else cm.ctypes.data,
                                                    int(first_sample), ev.ctypes.data, cap, ct.byref(n),
                                                    feat.ctypes.data if features else None))
        self.last_W = n_samples
        return ev[:n.value].copy(), feat[:n.value].copy()

    def process_chunk(self, chunk: np.ndarray, first_sample: int, *, chan_map=None,
                      cap: int = 65536):
        """Streaming form: next contiguous chunk ``[n][ld]`` of the stream."""
        x = np.ascontiguousarray(chunk, dtype=np.int16)
        n_s, ld = x.shape
        cm = None if chan_map is None else np.ascontiguousarray(chan_map, dtype=np.int32)
        ev, feat = self._out(cap)
        n = ct.c_int64(0)
        check(lib().lss_process_chunk(self._h, x.ctypes.data, n_s, ld,
                                      None if cm is None else cm.ctypes.data, int(first_sample),
                                      ev.ctypes.data, cap, ct.byref(n), feat.ctypes.data))
        return ev[:n.value].copy(), feat[:n.value].copy()

    def flush(self, *, cap: int = 65536):
        ev, feat = self._out(cap)
        n = ct.c_int64(0)
        check(lib().lss_flush(self._h, ev.ctypes.data, cap, ct.byref(n), feat.ctypes.data))
        return ev[:n.value].copy(), feat[:n.value].copy()

    def stage_times(self, reset: bool = False):
        """Accumulated device ms per stage (needs ``profile=1``; ``profile=2`` times the peel only): (dict, batches)."""
        from ._capi import NSTAGES, STAGE_NAMES
        ms = (ct.c_double * NSTAGES)()
        nb = ct.c_int64(0)
        check(lib().lss_stage_times(self._h, ms, ct.byref(nb), int(reset)))
        return dict(zip(STAGE_NAMES, list(ms))), nb.value

    def debug_stage(self, stage: int) -> np.ndarray:
        """Per-stage device buffers of the last batch (stages B, CFMAX and IMAX
        need ``keep_stages=True``; the PRE stages need preprocessing on)."""
        info = self.info
        W = self.last_W
        shapes = {STAGE_XW: ((info.C, W), np.float32), STAGE_BPC: ((info.K, info.C, W), np.float32),
                  STAGE_B: ((info.T, W), np.float32), STAGE_CFMAX: ((W,), np.float32),
                  STAGE_IMAX: ((W,), np.int32), STAGE_NM: ((info.T,), np.float32),
                  STAGE_PHASES: ((2 * NPHASE + 2,), np.uint64), STAGE_PRE: ((info.C, W), np.float32),
                  STAGE_PRE_MEAN: ((info.C,), np.float32), STAGE_PRE_MED: ((W,), np.float32)}
        shape, dt = shapes[stage]
        out = np.zeros(shape, dt)
        nb = ct.c_int64(out.nbytes)
        check(lib().lss_debug_stage(self._h, stage, out.ctypes.data, ct.byref(nb)))
        return out

    def ctc(self) -> np.ndarray:
        """The ctc table the matcher uses, [T][T][2M+1] f32 (file layout of ctc.npy)."""
        info = self.info
        out = np.zeros((info.T, info.T, 2 * info.M + 1), np.float32)
        nb = ct.c_int64(out.nbytes)
        check(lib().lss_debug_stage(self._h, STAGE_CTC, out.ctypes.data, ct.byref(nb)))
        return out


class StreamGroup:
    """Several probe streams matched by one graph launch per tick (lss_group_*).

    No counterpart in the reference, which runs one sorter thread per probe
    (main.cpp:77-85).  Each call to :meth:`process` matches one framed batch
    of ``batch`` columns of every stream.  ``peel_split``: 0 lets the group
    split each stream's templates over several peel workgroups per column
    block when that grid is co-resident, 1 keeps one, n > 1 asks for n ways
    for the largest stream (``info[s].peel_split`` reports the choice).
    ``profile``: 1 (or True) times the stages in front of the peel and the peel
    with HIP events and runs the peel's phase clock (:meth:`phases`); 2 times
    them without the clock.
    """

    def __init__(self, oss_input_dirs, *, batch: int, device: int = 0, max_iter: int = 0,
                 semantics: int = SEM_KILOSORT, max_events: int = 65536, sparse: int = 2,
                 precision: int = PREC_FP32, preprocess: int = 0, peel_split: int = 0, profile: int = 0):
        self._g = ct.c_void_p()
        self.n = len(oss_input_dirs)
        cfg = default_config(device=device, chunk=batch, max_batch=batch, max_iter=max_iter,
                             semantics=semantics, max_events=max_events, sparse=int(sparse),
                             precision=precision, preprocess=preprocess, peel_split=peel_split,
                             profile=int(profile))
        dirs = (ct.c_char_p * self.n)(*[os.fsencode(str(d)) for d in oss_input_dirs])
        check(lib().lss_group_create(self.n, dirs, ct.byref(cfg), ct.byref(self._g)))
        self.batch = batch
        self.info = [self.query(s) for s in range(self.n)]

    def query(self, stream: int) -> Info:
        info = Info()
        check(lib().lss_group_query(self._g, stream, ct.byref(info)))
        return info

    def phases(self, stream: int) -> np.ndarray:
        """The peel kernel's phase clock of the last launch for one stream (STAGE_PHASES layout)."""
        out = np.zeros(2 * NPHASE + 2, np.uint64)
        nb = ct.c_int64(out.nbytes)
        check(lib().lss_group_debug_stage(self._g, stream, STAGE_PHASES, out.ctypes.data, ct.byref(nb)))
        return out

    def stage_times(self, lane: int = 0, reset: bool = False):
        """(front-stage ms, peel ms, launches) accumulated on one lane (needs ``profile=True``)."""
        ms = (ct.c_double * 2)()
        n = ct.c_int64(0)
        check(lib().lss_group_stage_times(self._g, lane, ms, ct.byref(n), int(reset)))
        return ms[0], ms[1], n.value

    def set_proposal(self, stream: int, net, lo: float = 0.25, hi: float = 0.75, reach: int = 4) -> None:
        """Gate one stream's score tiles with a proposal network (``None`` detaches)."""
        check(lib().lss_group_set_proposal(self._g, stream, net._p if net is not None else None, lo, hi, reach))
        self._props = getattr(self, "_props", {})
        self._props[stream] = net

    def set_pipelined(self, on: bool = True) -> None:
        """Pipelined lanes (lss_group_set_pipelined): consecutive batches of the same
        streams may be in flight on two lanes at once, from two host threads; the
        next batch's front stages run beside the current peel.  Device input only."""
        check(lib().lss_group_set_pipelined(self._g, int(on)))

    def process(self, batches, first_samples, *, device: bool = False, lds=None, cap: int = 65536, lane: int = 0,
                copy: bool = True):
        """``batches``: per stream an int16 array ``[batch][ld]`` (or, with
        ``device=True``, a device pointer and ``lds`` the row strides); ``None``
        for a stream with no batch this call (it returns no events).  With
        ``copy=False`` the returned arrays are views of the lane's buffers, valid
        until the lane's next call."""
        n = self.n
        if device:
            ptrs = [int(b) if b is not None else None for b in batches]
            ld = list(lds)
            keep = None
        else:
            keep = [np.ascontiguousarray(b, dtype=np.int16) if b is not None else None for b in batches]
            for b in keep:
                if b is not None and b.shape[0] != self.batch:
                    raise ValueError("every batch must have the group's width")
            ptrs = [b.ctypes.data if b is not None else None for b in keep]
            # an idle stream keeps the row stride it was last given (no re-capture)
            last = getattr(self, "_ld", None) or [next(b for b in keep if b is not None).shape[1]] * n
            ld = list(lds) if lds is not None else [b.shape[1] if b is not None else last[s]
                                                     for s, b in enumerate(keep)]
            self._ld = ld
        bufs = getattr(self, "_bufs", {})                 # reused across calls, one set per lane
        if lane not in bufs or bufs[lane][0] != cap:
            evs = [np.zeros(cap, EVENT_DTYPE) for _ in range(n)]
            feats = [np.zeros((cap, self.info[s].nC, self.info[s].K), np.float32) for s in range(n)]
            # the output arguments of the call, fixed per lane
            bufs[lane] = (cap, evs, feats, (ct.c_void_p * n)(*[e.ctypes.data for e in evs]),
                          (ct.c_int64 * n)(*([cap] * n)), (ct.c_void_p * n)(*[f.ctypes.data for f in feats]))
        self._bufs = bufs
        _, evs, feats, c_ev, c_cap, c_feat = bufs[lane]
        c_ptrs = (ct.c_void_p * n)(*ptrs)
        c_ld = (ct.c_int32 * n)(*ld)
        c_first = (ct.c_int64 * n)(*[int(f) for f in first_samples])
        c_n = (ct.c_int64 * n)()
        check(lib().lss_group_process_lane(self._g, lane, c_ptrs, c_ld, int(device), c_first, c_ev, c_cap, c_n,
                                           c_feat))
        del keep
        if not copy:
            return [(evs[s][:c_n[s]], feats[s][:c_n[s]]) for s in range(n)]
        return [(evs[s][:c_n[s]].copy(), feats[s][:c_n[s]].copy()) for s in range(n)]

    def close(self) -> None:
        if self._g:
            check(lib().lss_group_destroy(self._g))
            self._g = ct.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
