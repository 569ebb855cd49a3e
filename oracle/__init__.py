"""CPU oracle for the matcher -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
(``livespikesorter_amd``) never imports it and has no CPU fallback.

``run_batch`` wraps the C restatement in ``oracle.c`` (FP64 build ``liborc64``
or FP32 build ``liborc32``) of Kilosort 4.0.13's ``run_matching`` + feature
extraction (KS/template_matching.py:50-51,128-202, KS = the vendored
src/Python/run_online_spikes/Lib/site-packages/kilosort).  The pure-numpy
functions below restate single stages in the most literal form (loops over the
definition) so tiny cases can be brute-forced against the C code.
"""
from __future__ import annotations

import ctypes as ct
import dataclasses
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent

SEM_KILOSORT = 0
SEM_LSS = 1


class _Params(ct.Structure):
    _fields_ = [("C", ct.c_int), ("T", ct.c_int), ("M", ct.c_int), ("K", ct.c_int),
                ("nC", ct.c_int), ("Th", ct.c_double), ("max_iter", ct.c_int),
                ("semantics", ct.c_int), ("threads", ct.c_int)]


def build(force: bool = False) -> None:
    """Compile liborc64.so / liborc32.so with the committed Makefile."""
    args = ["make", "-C", str(HERE)]
    if force:
        args.append("-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


_LIBS: dict = {}


def _lib(bits: int):
    if bits not in _LIBS:
        path = HERE / f"liborc{bits}.so"
        if not path.exists():
            build()
        lib = ct.CDLL(str(path))
        _LIBS[bits] = lib
    return _LIBS[bits]


def _p(a, ctype):
    return a.ctypes.data_as(ct.POINTER(ctype)) if a is not None else None


@dataclasses.dataclass
class OracleResult:
    w: np.ndarray          # batch column of each spike (int64)
    t: np.ndarray          # template index (int32)
    amp: np.ndarray        # amplitude
    it: np.ndarray         # iteration that found it
    feat: np.ndarray       # [S][nC][K]
    n_iter: int
    Xw: np.ndarray | None = None
    Bpc: np.ndarray | None = None
    B: np.ndarray | None = None
    Cfmax: np.ndarray | None = None
    imax: np.ndarray | None = None
    margins: np.ndarray | None = None   # [3][W] decision margins (threshold, template, pooling)


def run_batch(x: np.ndarray, Wr, wPCA, U, ctc, iCC, iU, *, Th: float = 8.0, nC: int = 10,
              max_iter: int = 100, semantics: int = SEM_KILOSORT, bits: int = 64,
              threads: int = 0, chan_map=None, stages: bool = False, cap: int | None = None,
              features: bool = True, xf: np.ndarray | None = None, margins: bool = False) -> OracleResult:
    """Run one batch (``x`` int16 [W][ld]) through the oracle.  With ``xf``
    (preprocessed input [C][W], FP64, channel-gathered) ``x`` is ignored and
    whitening starts from ``xf`` (see oracle.preprocess).  ``margins``: also
    return the peel's decision margins per column (oracle.c, ``peel_m``)."""
    lib = _lib(bits)
    real = np.float64 if bits == 64 else np.float32
    creal = ct.c_double if bits == 64 else ct.c_float
    if xf is not None:
        xf = np.ascontiguousarray(xf, dtype=np.float64)
        x = np.zeros((xf.shape[1], 1), np.int16)
    x = np.ascontiguousarray(x, dtype=np.int16)
    W, ld = x.shape
    T, K, C = U.shape
    M = wPCA.shape[1]
    Wr = np.ascontiguousarray(Wr, dtype=np.float32)
    wPCA = np.ascontiguousarray(wPCA, dtype=np.float32)
    U = np.ascontiguousarray(U, dtype=np.float32)
    ctc = np.ascontiguousarray(ctc, dtype=np.float32)
    iCC = np.ascontiguousarray(iCC, dtype=np.int64)
    iU = np.ascontiguousarray(iU, dtype=np.int64)
    cm = None if chan_map is None else np.ascontiguousarray(chan_map, dtype=np.int32)
    if cap is None:
        cap = max(1024, W * max_iter // (M + 1) + 16)
    prm = _Params(C, T, M, K, nC, float(Th), max_iter, semantics, threads)
    ow = np.zeros(cap, np.int64)   # C long is 64-bit on LP64 Linux
    ot = np.zeros(cap, np.int32)
    oa = np.zeros(cap, real)
    oi = np.zeros(cap, np.int32)
    of = np.zeros((cap, nC, K), real) if features else None
    nit = ct.c_int(0)
    sXw = np.zeros((C, W), real) if stages else None
    sBpc = np.zeros((C, K, W), real) if stages else None
    sB = np.zeros((T, W), real) if stages else None
    sCf = np.zeros(W, real) if stages else None
    sIm = np.zeros(W, np.int32) if stages else None
    mg = np.zeros((3, W), real) if margins else None
    outs = (ct.c_long(cap), _p(ow, ct.c_long), _p(ot, ct.c_int), _p(oa, creal), _p(oi, ct.c_int), _p(of, creal),
            ct.byref(nit), _p(sXw, creal), _p(sBpc, creal), _p(sB, creal), _p(sCf, creal), _p(sIm, ct.c_int),
            _p(mg, creal))
    tabs = (_p(Wr, ct.c_float), _p(wPCA, ct.c_float), _p(U, ct.c_float), _p(ctc, ct.c_float),
            _p(iCC, ct.c_int64), _p(iU, ct.c_int64))
    if xf is not None:
        if xf.shape != (C, W):
            raise ValueError(f"xf must be [C][W] = [{C}][{W}]")
        fn = getattr(lib, f"orc{bits}_run_batch_f")
        fn.restype = ct.c_long
        n = fn(ct.byref(prm), _p(xf, ct.c_double), ct.c_long(W), *tabs, *outs)
    else:
        fn = getattr(lib, f"orc{bits}_run_batch")
        fn.restype = ct.c_long
        n = fn(ct.byref(prm), _p(x, ct.c_int16), ct.c_long(W), ct.c_long(ld), _p(cm, ct.c_int), *tabs, *outs)
    if n < 0:
        raise RuntimeError("oracle spike capacity exceeded")
    return OracleResult(ow[:n].copy(), ot[:n].copy(), oa[:n].copy(), oi[:n].copy(),
                        of[:n].copy() if features else np.zeros((n, nC, K), real),
                        int(nit.value), sXw, sBpc, sB, sCf, sIm, mg)


def run_model_batch(model_dir_tables: dict, x: np.ndarray, **kw) -> OracleResult:
    """``run_batch`` with the arrays of ``load_tables``."""
    t = model_dir_tables
    return run_batch(x, t["whiteningMat"], t["wPCA"], t["Wall3"], t["ctc"], t["iCC"], t["iU"],
                     Th=t["Th"], nC=t["nC"], **kw)


def load_tables(path: str | os.PathLike, ctc: np.ndarray | None = None) -> dict:
    """Read the oss_input arrays the oracle needs (plain numpy, no pickles).
    ``ctc``: use this table instead of the directory's ctc.npy (for a
    directory written without it)."""
    p = Path(path)
    names = ("whiteningMat", "wPCA", "Wall3", "iCC", "iU", "Ucc") + (("ctc",) if ctc is None else ())
    out = {k: np.load(p / f"{k}.npy", allow_pickle=False) for k in names}
    if ctc is not None:
        out["ctc"] = np.asarray(ctc, np.float32)
    misc = {}
    for line in (p / "misc.txt").read_text().splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            misc[k.strip()] = v.strip()
    out["Th"] = float(misc["Th_learned"])
    out["nC"] = int(misc["numNearestChans"])
    out["nt0min"] = int(misc["nt0min"])
    out["dt"] = int(misc["duplicate_spike_bins"])
    return out


# ---------------------------------------------------------------------------
# Literal single-stage restatements (pure Python loops; tiny inputs only)
# ---------------------------------------------------------------------------

def lit_conv1d_same(X, wPCA):
    """Bpc[c][k][w] = sum_m X[c][w+m-M//2] wPCA[k][m], zero padded (torch conv1d, padding=M//2)."""
    C, W = len(X), len(X[0])
    K, M = len(wPCA), len(wPCA[0])
    out = [[[0.0] * W for _ in range(K)] for _ in range(C)]
    for c in range(C):
        for k in range(K):
            for w in range(W):
                s = 0.0
                for m in range(M):
                    wi = w + m - M // 2
                    if 0 <= wi < W:
                        s += float(X[c][wi]) * float(wPCA[k][m])
                out[c][k][w] = s
    return out


def lit_project(U, Bpc):
    """B[t][w] = sum_{k,c} U[t][k][c] Bpc[c][k][w] (einsum 'ijk, kjl -> il')."""
    T, K, C = len(U), len(U[0]), len(U[0][0])
    W = len(Bpc[0][0])
    return [[sum(float(U[t][k][c]) * Bpc[c][k][w] for k in range(K) for c in range(C))
             for w in range(W)] for t in range(T)]


def lit_scores(B, nm, M):
    """Cf = relu(B)^2/nm, zeroed in the first and last M columns; max/argmax over t; max-pool."""
    T, W = len(B), len(B[0])
    Cf = [[(max(B[t][w], 0.0) ** 2 / nm[t]) if M <= w < W - M else 0.0 for w in range(W)]
          for t in range(T)]
    Cfmax, imax = [], []
    for w in range(W):
        best, bi = Cf[0][w], 0
        for t in range(1, T):
            if Cf[t][w] > best:
                best, bi = Cf[t][w], t
        Cfmax.append(best)
        imax.append(bi)
    Cmax = [max(Cfmax[max(0, w - M):min(W, w + M + 1)]) for w in range(W)]
    return Cf, Cfmax, imax, Cmax


def peel(B: np.ndarray, nm: np.ndarray, ctc: np.ndarray, U: np.ndarray, wPCA: np.ndarray, *,
         Th: float = 8.0, max_iter: int = 100, semantics: int = SEM_KILOSORT, bits: int = 64,
         cap: int = 100000, margins: bool = False):
    """The peel loop alone on a given score matrix ``B [T][W]`` (modified copy returned).

    Returns (w, t, amp, iter, n_iter, first_Cfmax, first_imax, B_after), and the
    decision margins [3][W] after them when ``margins`` is set.
    """
    lib = _lib(bits)
    real = np.float64 if bits == 64 else np.float32
    creal = ct.c_double if bits == 64 else ct.c_float
    Bc = np.ascontiguousarray(B, dtype=real).copy()
    T, W = Bc.shape
    K, M = wPCA.shape
    C = U.shape[2]
    nm = np.ascontiguousarray(nm, dtype=real)
    ctc = np.ascontiguousarray(ctc, dtype=np.float32)
    U = np.ascontiguousarray(U, dtype=np.float32)
    wPCA = np.ascontiguousarray(wPCA, dtype=np.float32)
    prm = _Params(C, T, M, K, 1, float(Th), max_iter, semantics, 0)
    ow = np.zeros(cap, np.int64)
    ot = np.zeros(cap, np.int32)
    oa = np.zeros(cap, real)
    oi = np.zeros(cap, np.int32)
    nit = ct.c_int(0)
    cf = np.zeros(W, real)
    im = np.zeros(W, np.int32)
    mg = np.zeros((3, W), real) if margins else None
    fn = getattr(lib, f"orc{bits}_peel_m")
    fn.restype = ct.c_long
    n = fn(ct.byref(prm), _p(Bc, creal), None, ct.c_long(W), _p(nm, creal), _p(ctc, ct.c_float),
           _p(U, ct.c_float), _p(wPCA, ct.c_float), ct.c_long(cap), _p(ow, ct.c_long),
           _p(ot, ct.c_int), _p(oa, creal), _p(oi, ct.c_int), ct.byref(nit), _p(cf, creal),
           _p(im, ct.c_int), _p(mg, creal))
    if n < 0:
        raise RuntimeError("oracle spike capacity exceeded")
    out = (ow[:n], ot[:n], oa[:n], oi[:n], int(nit.value), cf, im, Bc)
    return out + (mg,) if margins else out


def stage_conv_project(x: np.ndarray, Wr, wPCA, U, *, bits: int = 64, threads: int = 0):
    """Whitening, PC conv and projection stages of the C oracle. Returns (Xw, Bpc, B)."""
    lib = _lib(bits)
    real = np.float64 if bits == 64 else np.float32
    creal = ct.c_double if bits == 64 else ct.c_float
    x = np.ascontiguousarray(x, dtype=np.int16)
    W, ld = x.shape
    T, K, C = U.shape
    M = wPCA.shape[1]
    Wr = np.ascontiguousarray(Wr, dtype=np.float32)
    wPCA = np.ascontiguousarray(wPCA, dtype=np.float32)
    U = np.ascontiguousarray(U, dtype=np.float32)
    prm = _Params(C, T, M, K, 1, 8.0, 100, 0, threads)
    Xw = np.zeros((C, W), real)
    Bpc = np.zeros((C, K, W), real)
    B = np.zeros((T, W), real)
    getattr(lib, f"orc{bits}_whiten")(ct.byref(prm), _p(x, ct.c_int16), ct.c_long(W), ct.c_long(ld),
                                      None, _p(Wr, ct.c_float), _p(Xw, creal))
    getattr(lib, f"orc{bits}_pcconv")(ct.byref(prm), _p(Xw, creal), ct.c_long(W), _p(wPCA, ct.c_float),
                                      _p(Bpc, creal))
    getattr(lib, f"orc{bits}_project")(ct.byref(prm), _p(Bpc, creal), ct.c_long(W), _p(U, ct.c_float),
                                       _p(B, creal))
    return Xw, Bpc, B


# ---------------------------------------------------------------------------
# Spike positions and final-cluster relabelling (pure numpy, float64)
# ---------------------------------------------------------------------------

def spike_positions(feat: np.ndarray, templates: np.ndarray, iCC: np.ndarray, iU: np.ndarray,
                    xc: np.ndarray, yc: np.ndarray):
    """Kilosort's compute_spike_positions (KS/postprocessing.py:33-46), as LSS's
    compute_spike_positions_kernel computes it (OnlineSpikesV2.cu:190-245):
    feat [n][nC][K] PC features, templates [n] file indices; the mass of
    nearest channel j is sum_k feat^2, the position the mass-weighted mean of
    the nearest channels' coordinates."""
    feat = np.asarray(feat, np.float64)
    n, nC, K = feat.shape
    mass = (feat ** 2).sum(-1)
    tot = mass.sum(1, keepdims=True)
    w = np.divide(mass, tot, out=np.zeros_like(mass), where=tot > 0)
    chs = np.asarray(iCC)[:nC, np.asarray(iU)[np.asarray(templates)]].T      # [n][nC]
    return (np.asarray(xc, np.float64)[chs] * w).sum(1), (np.asarray(yc, np.float64)[chs] * w).sum(1)


def closest_cluster(x, y, centroids: np.ndarray) -> np.ndarray:
    """OnlineSpikesV2::closestCluster (OnlineSpikesV2.cu:1058-1072): index of the
    nearest (x, y) centroid, the first among equals."""
    c = np.asarray(centroids, np.float64).reshape(-1, 2)
    d = np.hypot(np.asarray(x, np.float64)[:, None] - c[None, :, 0], np.asarray(y, np.float64)[:, None] - c[None, :, 1])
    return np.argmin(d, axis=1)
