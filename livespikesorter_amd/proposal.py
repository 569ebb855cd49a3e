"""Spike-proposal network (the learned gate in front of the matcher).

A 4-layer 1-D convolution (32 channels, kernel 9, BF16 weights and
activations, FP32 accumulation) over the whitened batch, run per 16-channel
group of the probe with shared weights; its output is P(spike) per (group,
sample).  Forward, backward and Adam run as the library's own HIP kernels
(``lss_proposal_*`` in include/lss.h).  No counterpart in the reference.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

from ._capi import check, lib

GROUP = 16


def n_params() -> int:
    return int(lib().lss_proposal_n_params())


class ProposalNet:
    """One network on one device.  ``channels`` and ``max_batch`` size the
    host-input entry points (:meth:`forward`, :meth:`grad`)."""

    def __init__(self, channels: int, max_batch: int, *, device: int = 0, seed: int = 1):
        self._p = ct.c_void_p()
        check(lib().lss_proposal_create(device, channels, max_batch, seed, ct.byref(self._p)))
        self.channels = channels
        self.groups = (channels + GROUP - 1) // GROUP
        self.max_batch = max_batch

    # -- parameters ------------------------------------------------------
    @property
    def params(self) -> np.ndarray:
        out = np.empty(n_params(), np.float32)
        check(lib().lss_proposal_get_params(self._p, out.ctypes.data, out.size))
        return out

    @params.setter
    def params(self, values) -> None:
        v = np.ascontiguousarray(values, dtype=np.float32)
        check(lib().lss_proposal_set_params(self._p, v.ctypes.data, v.size))

    def save(self, path: str | os.PathLike) -> None:
        np.save(path, self.params)

    def load(self, path: str | os.PathLike) -> None:
        self.params = np.load(path, allow_pickle=False)

    # -- compute ---------------------------------------------------------
    def forward(self, xw: np.ndarray) -> np.ndarray:
        """``xw`` [channels][n] whitened samples -> P(spike) [groups][n]."""
        x = np.ascontiguousarray(xw, dtype=np.float32)
        assert x.shape[0] == self.channels
        out = np.empty((self.groups, x.shape[1]), np.float32)
        check(lib().lss_proposal_forward(self._p, x.ctypes.data, x.shape[1], out.ctypes.data))
        return out

    def grad(self, xw: np.ndarray, labels: np.ndarray, pos_weight: float):
        """(loss, gradient [n_params]) of the weighted BCE for 0/1 ``labels`` [groups][n]."""
        x = np.ascontiguousarray(xw, dtype=np.float32)
        y = np.ascontiguousarray(labels, dtype=np.uint8)
        assert y.shape == (self.groups, x.shape[1])
        g = np.empty(n_params(), np.float32)
        loss = ct.c_float()
        check(lib().lss_proposal_grad(self._p, x.ctypes.data, y.ctypes.data, x.shape[1], pos_weight,
                                      ct.byref(loss), g.ctypes.data))
        return float(loss.value), g

    def train_step(self, matcher, dev_ptr: int, n: int, ld: int, labels: np.ndarray, pos_weight: float,
                   lr: float) -> float:
        """One Adam step on a framed int16 batch in device memory, whitened by ``matcher``."""
        y = np.ascontiguousarray(labels, dtype=np.uint8)
        loss = ct.c_float()
        check(lib().lss_proposal_train_step(self._p, matcher._h, ct.c_void_p(dev_ptr), n, ld, y.ctypes.data,
                                            pos_weight, lr, ct.byref(loss)))
        return float(loss.value)

    # -- lifecycle -------------------------------------------------------
    def close(self) -> None:
        if self._p:
            check(lib().lss_proposal_destroy(self._p))
            self._p = ct.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def spike_labels(events, template_group, first_sample: int, n: int, groups: int, half: int, nt0min: int,
                 tol: int = 1) -> np.ndarray:
    """0/1 labels [groups][n] of a framed batch from a spike list: the batch
    column of each event (sample - first_sample + half - nt0min) and the group
    of its template's main channel, widened by +-tol columns."""
    y = np.zeros((groups, n), np.uint8)
    if len(events):
        cols = events["sample"] - first_sample + half - nt0min
        grp = template_group[events["template"]]
        for d in range(-tol, tol + 1):
            c = cols + d
            ok = (c >= 0) & (c < n)
            y[grp[ok], c[ok]] = 1
    return y
