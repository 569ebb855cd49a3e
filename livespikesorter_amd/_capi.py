"""ctypes binding of ``liblss.so`` (the C ABI declared in include/lss.h).

The library is the product: it holds every HIP kernel.  This module only
loads it and converts numpy arrays to pointers.  If the library is missing
or cannot be loaded, every entry point raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as ct
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "liblss.so"

LSS_OK = 0
LSS_ERR_ARG = -1
LSS_ERR_IO = -2
LSS_ERR_HIP = -3
LSS_ERR_CAPACITY = -4
LSS_ERR_STATE = -5
LSS_ERR_TIMEOUT = -6
LSS_ERR_LIMIT = -7
STATIC_SCRATCH_BYTES = 128 << 20     # LSS_STATIC_SCRATCH_BYTES

SEM_KILOSORT = 0
SEM_LSS = 1
PREC_FP32 = 0
PREC_BF16_RESCORE = 1
PREC_BF16X3 = 2
PRE_MEAN, PRE_CAR, PRE_HIGHPASS = 1, 2, 4
PRE_ALL = PRE_MEAN | PRE_CAR | PRE_HIGHPASS

ABI_FUNCTIONS = ("lss_config_init", "lss_create", "lss_process_batch", "lss_process_batch_device",
                 "lss_process_batch_device_static",
                 "lss_process_chunk", "lss_flush", "lss_query", "lss_debug_stage", "lss_stage_times",
                 "lss_destroy", "lss_last_error", "lss_group_create", "lss_group_process",
                 "lss_group_query", "lss_group_destroy", "lss_group_stage_times", "lss_group_debug_stage", "lss_proposal_n_params", "lss_proposal_create",
                 "lss_proposal_set_params", "lss_proposal_get_params", "lss_proposal_forward",
                 "lss_proposal_grad", "lss_proposal_train_step", "lss_proposal_destroy", "lss_set_proposal",
                 "lss_autotune", "lss_group_set_proposal", "lss_group_set_pipelined", "lss_group_process_lane",
                 "lss_last_positions",
                 "lss_create_subset", "lss_window_plan", "lss_save_spikes", "lss_format_spikes",
                 "lss_payload_encode", "lss_payload_decode", "lss_sorter_params_encode",
                 "lss_sorter_params_decode", "lss_replay_open", "lss_replay_samples", "lss_replay_fetch",
                 "lss_replay_close", "lss_zscore_bin_samples", "lss_zscore_create", "lss_zscore_add",
                 "lss_zscore_complete", "lss_zscore_freeze", "lss_zscore_stats", "lss_zscore_destroy")
NSTAGES = 7
STAGE_NAMES = ("h2d_copy", "preprocess", "whiten", "proposal", "pcconv", "project", "peel")


class Config(ct.Structure):
    _fields_ = [("device", ct.c_int32), ("chunk", ct.c_int32), ("max_batch", ct.c_int32),
                ("max_iter", ct.c_int32), ("semantics", ct.c_int32), ("precision", ct.c_int32),
                ("dedup", ct.c_int32), ("max_events", ct.c_int32), ("keep_stages", ct.c_int32),
                ("sparse", ct.c_int32), ("profile", ct.c_int32), ("peel_split", ct.c_int32),
                ("relabel", ct.c_int32), ("preprocess", ct.c_int32), ("reserved", ct.c_int32 * 2)]


class Event(ct.Structure):
    _fields_ = [("sample", ct.c_int64), ("template_id", ct.c_int32), ("amplitude", ct.c_float)]


EVENT_DTYPE = np.dtype([("sample", np.int64), ("template", np.int32), ("amplitude", np.float32)])
assert EVENT_DTYPE.itemsize == ct.sizeof(Event) == 16


class Info(ct.Structure):
    _fields_ = [("C", ct.c_int32), ("T", ct.c_int32), ("M", ct.c_int32), ("K", ct.c_int32),
                ("nC", ct.c_int32), ("nt0min", ct.c_int32), ("dt", ct.c_int32),
                ("semantics", ct.c_int32), ("Th", ct.c_float), ("max_batch", ct.c_int32),
                ("chunk", ct.c_int32), ("max_iter", ct.c_int32), ("batches", ct.c_int64),
                ("events", ct.c_int64), ("last_iters", ct.c_int32), ("last_spikes", ct.c_int32),
                ("last_ms", ct.c_double), ("white_kblocks", ct.c_int64), ("proj_kblocks", ct.c_int64),
                ("proj_kblocks_dense", ct.c_int64), ("precision", ct.c_int32), ("p2p", ct.c_float),
                ("rescored", ct.c_int64), ("rescore_flips", ct.c_int64), ("peel_cw", ct.c_int32),
                ("peel_rows", ct.c_int32), ("peel_split", ct.c_int32), ("peel_groups", ct.c_int32),
                ("gate_tiles_scored", ct.c_int64), ("gate_tiles", ct.c_int64),
                ("gate_fallback_blocks", ct.c_int64), ("gate_blocks", ct.c_int64)]


class TuneResult(ct.Structure):
    _fields_ = [("n", ct.c_int32), ("chosen", ct.c_int32), ("split", ct.c_int32 * 8), ("rows", ct.c_int32 * 8),
                ("ms", ct.c_double * 8)]


class LssError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class WindowCfg(ct.Structure):
    _fields_ = [("min_window", ct.c_int64), ("max_window", ct.c_int64), ("time_behind_ms", ct.c_int32),
                ("samp_rate", ct.c_float), ("small_skip", ct.c_int32), ("lookback", ct.c_int32)]


class WindowStep(ct.Structure):
    _fields_ = [("kind", ct.c_int32), ("skip", ct.c_int32), ("fetch_start", ct.c_int64), ("fetch_n", ct.c_int64),
                ("buf_row", ct.c_int64), ("batch_row", ct.c_int64), ("batch_n", ct.c_int64),
                ("latest_ct", ct.c_int64), ("allowed_ct", ct.c_int64), ("save_start", ct.c_int64),
                ("save_end", ct.c_int64), ("next_processed_ct", ct.c_int64)]


class PayloadHeader(ct.Structure):
    _fields_ = [("recording_offset", ct.c_int64), ("stream_sample_ct", ct.c_int64), ("vrms", ct.c_double),
                ("p2p", ct.c_float), ("process_time", ct.c_int64), ("event_stream_sample_ct", ct.c_int64),
                ("predict_label", ct.c_int16), ("label", ct.c_int16), ("n_trials", ct.c_int16),
                ("n_correct", ct.c_int16), ("confidence", ct.c_double)]


class SorterParams(ct.Structure):
    _fields_ = [("T", ct.c_int64), ("C", ct.c_int64), ("M", ct.c_int64), ("N", ct.c_int64),
                ("samp_rate", ct.c_float)]


WIN_LATEST, WIN_SKIP_LATEST, WIN_CATCH_UP, WIN_SMALL_SKIP, WIN_BIG_SKIP = range(5)

_LIB = None


def lib() -> ct.CDLL:
    """Load liblss.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(the HIP library is required; there is no CPU fallback)")
        L = ct.CDLL(str(LIB_PATH))
        L.lss_last_error.restype = ct.c_char_p
        L.lss_config_init.argtypes = [ct.POINTER(Config)]
        L.lss_create.argtypes = [ct.c_char_p, ct.POINTER(Config), ct.POINTER(ct.c_void_p)]
        L.lss_create_subset.argtypes = [ct.c_char_p, ct.POINTER(Config), ct.c_void_p, ct.c_int32,
                                        ct.POINTER(ct.c_void_p)]
        batch_args = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_int32, ct.c_void_p, ct.c_int64,
                      ct.c_void_p, ct.c_int64, ct.POINTER(ct.c_int64), ct.c_void_p]
        L.lss_process_batch.argtypes = batch_args
        L.lss_process_batch_device.argtypes = batch_args
        L.lss_process_batch_device_static.argtypes = batch_args
        L.lss_process_chunk.argtypes = batch_args
        L.lss_flush.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.POINTER(ct.c_int64), ct.c_void_p]
        L.lss_query.argtypes = [ct.c_void_p, ct.POINTER(Info)]
        L.lss_debug_stage.argtypes = [ct.c_void_p, ct.c_int32, ct.c_void_p, ct.POINTER(ct.c_int64)]
        L.lss_stage_times.argtypes = [ct.c_void_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_int64), ct.c_int32]
        L.lss_destroy.argtypes = [ct.c_void_p]
        L.lss_group_create.argtypes = [ct.c_int32, ct.POINTER(ct.c_char_p), ct.POINTER(Config),
                                       ct.POINTER(ct.c_void_p)]
        L.lss_group_process.argtypes = [ct.c_void_p, ct.POINTER(ct.c_void_p), ct.POINTER(ct.c_int32), ct.c_int32,
                                        ct.POINTER(ct.c_int64), ct.POINTER(ct.c_void_p), ct.POINTER(ct.c_int64),
                                        ct.POINTER(ct.c_int64), ct.POINTER(ct.c_void_p)]
        L.lss_group_process_lane.argtypes = [ct.c_void_p, ct.c_int32] + L.lss_group_process.argtypes[1:]
        L.lss_group_query.argtypes = [ct.c_void_p, ct.c_int32, ct.POINTER(Info)]
        L.lss_last_positions.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.POINTER(ct.c_int64)]
        L.lss_group_destroy.argtypes = [ct.c_void_p]
        L.lss_group_debug_stage.argtypes = [ct.c_void_p, ct.c_int32, ct.c_int32, ct.c_void_p, ct.POINTER(ct.c_int64)]
        L.lss_group_stage_times.argtypes = [ct.c_void_p, ct.c_int32, ct.POINTER(ct.c_double),
                                            ct.POINTER(ct.c_int64), ct.c_int32]
        L.lss_proposal_n_params.restype = ct.c_int32
        L.lss_proposal_n_params.argtypes = []
        L.lss_proposal_create.argtypes = [ct.c_int32, ct.c_int32, ct.c_int32, ct.c_uint64, ct.POINTER(ct.c_void_p)]
        L.lss_proposal_set_params.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64]
        L.lss_proposal_get_params.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64]
        L.lss_proposal_forward.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int32, ct.c_void_p]
        L.lss_proposal_grad.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int32, ct.c_float,
                                        ct.POINTER(ct.c_float), ct.c_void_p]
        L.lss_proposal_train_step.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_int32,
                                              ct.c_void_p, ct.c_float, ct.c_float, ct.POINTER(ct.c_float)]
        L.lss_proposal_destroy.argtypes = [ct.c_void_p]
        L.lss_autotune.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_int32, ct.c_int32,
                                   ct.POINTER(TuneResult)]
        L.lss_group_set_proposal.argtypes = [ct.c_void_p, ct.c_int32, ct.c_void_p, ct.c_float, ct.c_float,
                                             ct.c_int32]
        L.lss_set_proposal.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_float, ct.c_float, ct.c_int32]
        L.lss_group_set_pipelined.argtypes = [ct.c_void_p, ct.c_int32]
        P64 = ct.POINTER(ct.c_int64)
        L.lss_window_plan.argtypes = [ct.POINTER(WindowCfg), ct.c_int64, ct.c_int64, ct.POINTER(WindowStep)]
        L.lss_save_spikes.argtypes = [ct.c_void_p, ct.c_int64, ct.c_int32, ct.c_int32, ct.c_int32, ct.c_int64,
                                      ct.c_int64, ct.c_void_p, ct.c_int32, ct.c_void_p, ct.c_void_p, ct.c_int64, P64]
        L.lss_format_spikes.argtypes = [ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_int64,
                                        P64]
        L.lss_payload_encode.argtypes = [ct.POINTER(PayloadHeader), ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                         ct.c_int64, ct.c_int32, ct.c_void_p, ct.c_int64, P64]
        L.lss_payload_decode.argtypes = [ct.c_void_p, ct.c_int64, ct.c_int32, ct.POINTER(PayloadHeader),
                                         ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int64, P64]
        L.lss_sorter_params_encode.argtypes = [ct.POINTER(SorterParams), ct.c_void_p, ct.c_int64, ct.c_void_p,
                                               ct.c_int64, ct.c_int32, ct.c_void_p, ct.c_int64, P64]
        L.lss_sorter_params_decode.argtypes = [ct.c_void_p, ct.c_int64, ct.c_int32, ct.POINTER(SorterParams),
                                               ct.c_void_p, ct.c_int64, P64, ct.c_void_p, ct.c_int64, P64]
        L.lss_replay_open.argtypes = [ct.c_char_p, ct.c_int32, ct.c_void_p, ct.c_int32, ct.POINTER(ct.c_void_p)]
        L.lss_replay_samples.argtypes = [ct.c_void_p, P64, ct.POINTER(ct.c_int32)]
        L.lss_replay_fetch.argtypes = [ct.c_void_p, ct.c_int64, ct.c_int64, ct.c_void_p, P64]
        L.lss_replay_close.argtypes = [ct.c_void_p]
        L.lss_zscore_bin_samples.argtypes = [ct.c_double, ct.c_float, P64]
        L.lss_zscore_create.argtypes = [ct.c_int32, ct.c_int32, ct.c_int64, ct.c_int32, ct.POINTER(ct.c_void_p)]
        L.lss_zscore_add.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64]
        L.lss_zscore_complete.argtypes = [ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_int64, P64]
        L.lss_zscore_freeze.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_double]
        L.lss_zscore_stats.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, P64, P64]
        L.lss_zscore_destroy.argtypes = [ct.c_void_p]
        _LIB = L
    return _LIB


def check(code: int) -> None:
    if code != LSS_OK:
        raise LssError(code, lib().lss_last_error().decode(errors="replace"))


def default_config(**kw) -> Config:
    cfg = Config()
    check(lib().lss_config_init(ct.byref(cfg)))
    for k, v in kw.items():
        if not hasattr(cfg, k):
            raise TypeError(f"unknown config field {k}")
        setattr(cfg, k, int(v))
    return cfg
