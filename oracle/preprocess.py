"""CPU restatement of the preprocessing in front of whitening -- TEST
INFRASTRUCTURE ONLY (imported by ``tests/`` and ``bench.py``'s cpu_baseline
leg as the checker; the product never imports it).

The chain, per batch of W samples x C channels (already gathered through the
channel map), in the order both references run it:

1. channel means  -- OnlineSpikesV2.cu:774-779 / myGPUhelpers.cu:99-119,704-712
   (``compute_channel_means``: FP32 sum of the int16-valued samples, divided
   by W); Kilosort ``X - X.mean(1)`` (src/Python/OnlineKilosort.py:118).
   The integer sum is exact here and divided once, which equals the FP32
   result of both references whenever their running sums stay below 2**24
   (then every partial sum is exact in FP32).
2. CAR            -- OnlineSpikesV2.cu:782-786 / myGPUhelpers.cu:480-543
   (``median_remove_kernel``: per sample, subtract the lower median
   sorted[(C-1)//2] over channels); Kilosort ``X - torch.median(X, 0)[0]``
   (OnlineKilosort.py:119-121; torch.median returns the lower median).
3. high-pass      -- OnlineSpikesV2.cu:1006-1056 (cuFFT, ``multiplyByConjugate
   Kernel``, ``shift_kernel`` myGPUhelpers.cu:72-83) / OnlineKilosort.py:124-
   128 with ``fft_highpass`` (KS/preprocessing.py:139-154): the impulse
   response cropped (or zero-padded) to W, y = real(ifft(fft(X) conj(fft(h)))),
   then a circular shift by W//2: Kilosort's torch ``fftshift`` gives
   out[k] = y[(k - W//2) mod W], LSS's ``shift_kernel`` out[k] = y[(k + W//2)
   mod W]; they agree for even W.

Everything runs in FP64 from the FP32 stage inputs; stages 1-2 round to FP32
as the references do (they are exact selections/differences there).

Pinning (tests/test_preprocess_oracle.py): the FFT form against the direct
circular-correlation sum; both shift conventions against torch.fft /
torch.fft.fftshift executing Kilosort's formula and against the LSS index
rule; the filter's response to periodic sinusoids against |H(f)|^2 of the
Butterworth design (scipy.signal.freqz); the median against a sort and
torch.median; the mean against exact rational arithmetic.
"""
from __future__ import annotations

import numpy as np

SEM_KILOSORT = 0
SEM_LSS = 1

PRE_MEAN = 1
PRE_CAR = 2
PRE_HP = 4


def highpass_filter(fs: float = 30000.0, cutoff: float = 300.0, NT: int = 30122) -> np.ndarray:
    """Impulse response of scipy's filtfilt with butter(3, cutoff, 'high') at
    the middle of NT samples, as FP32 (KS/preprocessing.py:121-136)."""
    from scipy.signal import butter, filtfilt
    b, a = butter(3, cutoff, fs=fs, btype="high")
    x = np.zeros(NT)
    x[NT // 2] = 1.0
    return filtfilt(b, a, x).astype(np.float32)


def channel_means(x: np.ndarray) -> np.ndarray:
    """x [W][C] int16 -> FP32 means [C] (exact integer sum, one rounding)."""
    s = x.astype(np.int64).sum(0)
    return (s / x.shape[0]).astype(np.float32)


def lower_median(x1: np.ndarray) -> np.ndarray:
    """x1 [W][C] FP32 -> per-sample lower median [W] FP32 (sorted[(C-1)//2])."""
    C = x1.shape[1]
    return np.sort(x1, axis=1)[:, (C - 1) // 2]


def cropped_taps(hp: np.ndarray, W: int) -> np.ndarray:
    """h_sub [W] FP64: hp cropped by (F-W)//2 when F >= W, else zero-padded with
    (W-F)//2 zeros in front (fft_highpass, KS/preprocessing.py:143-153;
    OnlineSpikesV2.cu:1016-1017 for the crop)."""
    F = len(hp)
    h = hp.astype(np.float64)
    if F >= W:
        c = (F - W) // 2
        return h[c:c + W].copy()
    pad = (W - F) // 2
    out = np.zeros(W)
    out[pad:pad + F] = h
    return out


def highpass(X: np.ndarray, hp: np.ndarray, semantics: int = SEM_KILOSORT) -> np.ndarray:
    """X [C][W] -> high-passed [C][W] FP64 by the FFT form of the references."""
    X = np.asarray(X, np.float64)
    W = X.shape[1]
    fh = np.fft.fft(cropped_taps(hp, W))
    y = np.real(np.fft.ifft(np.fft.fft(X, axis=1) * np.conj(fh)[None, :], axis=1))
    k = np.arange(W)
    src = (k - W // 2) % W if semantics == SEM_KILOSORT else (k + W // 2) % W
    return y[:, src]


def highpass_direct(X: np.ndarray, hp: np.ndarray, semantics: int = SEM_KILOSORT) -> np.ndarray:
    """Literal circular correlation out[k] = sum_m X[(k + s + m) mod W] h_sub[m]
    with s = -(W//2) (Kilosort) or +(W//2) (LSS); O(W^2), tiny cases only."""
    X = np.asarray(X, np.float64)
    C, W = X.shape
    h = cropped_taps(hp, W)
    s = -(W // 2) if semantics == SEM_KILOSORT else W // 2
    out = np.zeros((C, W))
    for k in range(W):
        for m in range(W):
            out[:, k] += X[:, (k + s + m) % W] * h[m]
    return out


def preprocess(x: np.ndarray, flags: int, hp: np.ndarray | None = None,
               semantics: int = SEM_KILOSORT) -> dict:
    """x [W][C] int16 (channel-gathered) -> stages: mean [C], med [W],
    x1 [C][W] FP32 (after mean and CAR), out [C][W] (FP64 when high-passed,
    else x1), p2p (max - min of the raw batch)."""
    x = np.asarray(x)
    W, C = x.shape
    xf = x.astype(np.float32)
    mean = channel_means(x) if flags & PRE_MEAN else np.zeros(C, np.float32)
    x1 = (xf - mean[None, :]).astype(np.float32)
    med = lower_median(x1) if flags & PRE_CAR else np.zeros(W, np.float32)
    x1 = (x1 - med[:, None]).astype(np.float32)
    out = highpass(x1.T, hp, semantics) if flags & PRE_HP else x1.T.astype(np.float64)
    return {"mean": mean, "med": med, "x1": np.ascontiguousarray(x1.T), "out": out,
            "p2p": float(x.max()) - float(x.min())}
