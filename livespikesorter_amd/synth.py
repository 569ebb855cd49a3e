"""Seeded synthetic recordings and Kilosort-style template sets.

This module builds the inputs the matcher consumes: an ``oss_input``
directory (the file set read by ``OnlineSpikesV2::initializeStaticMemory``,
src/SpikeSorter/OnlineSpikesV2.cu:500-510 and :580-638, written by
src/Python/run_online_spikes.py:637-672) and int16 recordings with ground
truth.  It is input generation, not part of the matched path.

Generative model (all draws from numpy PCG64 streams split per component):

* temporal PCs ``wPCA[K][M]``: QR of Hann-tapered Gaussian rows, so rows are
  orthonormal;
* spatial templates ``U[t,k,c] = g_t(c) * beta[t,k]`` with a Gaussian
  footprint ``g_t(c) = exp(-(c-mu_t)^2 / (2 sigma^2))``, ``mu_t ~ U[0,C)`` and
  ``beta ~ N(0,1)``;
* whitened-domain signal ``x = sum_i a_i * wave[t_i] (centred on s_i) + N(0,1)``,
  ``wave[t][m][c] = sum_k U[t,k,c] wPCA[k][m]``;
* raw signal ``Sigma^{1/2} x`` quantised to int16 with a gain of ``gain``
  counts per whitened unit; the whitening matrix handed to the matcher is
  ``Sigma^{-1/2} / gain`` (``I / gain`` for identity whitening).

The template-derived tables (``ctc``, ``iCC``, ``iU``, ``Ucc``, waveforms)
follow KS/template_matching.py:20-26 (prepare_extract) and :110-121
(prepare_matching); ``KS/`` abbreviates
src/Python/run_online_spikes/Lib/site-packages/kilosort/.
"""
from __future__ import annotations

import dataclasses
import os
from pathlib import Path

import numpy as np

TEST_SEED = 20261203
TRAIN_SEED = 20261204

# Sub-stream identifiers for SeedSequence spawning (one stream per component).
_S_PCS, _S_TEMPL, _S_WHITEN, _S_RATES, _S_SPIKES, _S_NOISE = range(6)


@dataclasses.dataclass
class ProbeModel:
    C: int
    T: int
    M: int
    K: int
    fs: float
    gain: float
    wPCA: np.ndarray          # [K][M] float32, orthonormal rows
    U: np.ndarray             # [T][K][C] float32 (Kilosort Wall3)
    mu: np.ndarray            # [T] footprint centres (channels)
    whiten: np.ndarray        # [C][C] float32, matrix handed to the matcher
    mix: np.ndarray | None    # [C][C] float64 Sigma^{1/2}, None for identity
    rates: np.ndarray         # [T] Hz
    amp_mean: float
    amp_sd: float
    isi: str                  # "poisson" or "gamma"
    gamma_shape: float
    nC: int = 10
    nt0min: int = 20
    Th: float = 8.0
    dup_bins: int = 7
    seed: int = TEST_SEED

    @property
    def waves(self) -> np.ndarray:
        """wave[t][m][c] = sum_k U[t,k,c] wPCA[k][m] (run_online_spikes.py:635)."""
        return np.einsum("tkc,km->tmc", self.U.astype(np.float64),
                         self.wPCA.astype(np.float64)).astype(np.float32)


def _rngs(seed: int, n: int = 8):
    ss = np.random.SeedSequence(seed)
    return [np.random.Generator(np.random.PCG64(s)) for s in ss.spawn(n)]


def make_pcs(rng: np.random.Generator, K: int = 6, M: int = 61) -> np.ndarray:
    g = rng.standard_normal((K, M)) * np.hanning(M + 2)[1:-1]
    q, _ = np.linalg.qr(g.T)
    return np.ascontiguousarray(q.T).astype(np.float32)


def banded_cov(rng: np.random.Generator, C: int, decay: float = 0.9):
    s = rng.uniform(0.5, 2.0, size=C)
    idx = np.arange(C)
    corr = decay ** np.abs(idx[:, None] - idx[None, :])
    return corr * s[:, None] * s[None, :]


def make_model(C: int, T: int, seed: int = TEST_SEED, *, whitening: str = "identity",
               sigma: float = 2.0, rate: float | tuple = 5.0, amp=(10.0, 2.0),
               isi: str = "poisson", gamma_shape: float = 0.5, gain: float = 64.0,
               M: int = 61, K: int = 6, fs: float = 30000.0, Th: float = 8.0) -> ProbeModel:
    r = _rngs(seed)
    wPCA = make_pcs(r[_S_PCS], K, M)
    rt = r[_S_TEMPL]
    mu = rt.uniform(0.0, C, size=T)
    beta = rt.standard_normal((T, K))
    ch = np.arange(C)
    g = np.exp(-((ch[None, :] - mu[:, None]) ** 2) / (2.0 * sigma ** 2))
    U = (beta[:, :, None] * g[:, None, :]).astype(np.float32)
    if whitening == "identity":
        whiten = (np.eye(C) / gain).astype(np.float32)
        mix = None
    elif whitening == "banded":
        cov = banded_cov(r[_S_WHITEN], C)
        ev, V = np.linalg.eigh(cov)
        mix = (V * np.sqrt(ev)) @ V.T
        whiten = (((V / np.sqrt(ev)) @ V.T) / gain).astype(np.float32)
    else:
        raise ValueError(f"unknown whitening {whitening!r}")
    if isinstance(rate, tuple):
        lo, hi = rate
        rates = np.exp(r[_S_RATES].uniform(np.log(lo), np.log(hi), size=T))
    else:
        rates = np.full(T, float(rate))
    return ProbeModel(C=C, T=T, M=M, K=K, fs=fs, gain=gain, wPCA=wPCA, U=U, mu=mu,
                      whiten=whiten, mix=mix, rates=rates, amp_mean=amp[0], amp_sd=amp[1],
                      isi=isi, gamma_shape=gamma_shape, Th=Th, seed=seed)


# ---------------------------------------------------------------------------
# Template-derived tables (Kilosort definitions, computed in float64)
# ---------------------------------------------------------------------------

def wtw(wPCA: np.ndarray) -> np.ndarray:
    """WtW[k1][k2][l] = sum_m W[k1][m - (l - M)] W[k2][m] (KS/template_matching.py:113-114).

    conv1d(W[k1], W[k2], padding=M) followed by a flip along the lag axis.
    """
    K, M = wPCA.shape
    W = wPCA.astype(np.float64)
    out = np.zeros((K, K, 2 * M + 1))
    for l in range(2 * M + 1):
        d = l - M                      # W[k1][m - d] pairs with W[k2][m]
        lo, hi = max(0, d), min(M, M + d)
        if lo < hi:
            out[:, :, l] = W[:, lo - d:hi - d] @ W[:, lo:hi].T
    return out


def ctc_table(U: np.ndarray, wPCA: np.ndarray, block: int = 128) -> np.ndarray:
    """ctc[i][j][l] = sum_{k,m} (sum_c U[i,k,c] U[j,m,c]) WtW[k][m][l] (:119-120)."""
    T, K, C = U.shape
    M = wPCA.shape[1]
    Wt = wtw(wPCA).reshape(K * K, 2 * M + 1)
    U64 = U.astype(np.float64)
    Uf = U64.reshape(T * K, C)
    out = np.empty((T, T, 2 * M + 1), dtype=np.float32)
    for i0 in range(0, T, block):
        i1 = min(T, i0 + block)
        utu = (Uf[i0 * K:i1 * K] @ Uf.T).reshape(i1 - i0, K, T, K)      # [i,k,j,m]
        utu = utu.transpose(0, 2, 1, 3).reshape((i1 - i0) * T, K * K)
        out[i0:i1] = (utu @ Wt).reshape(i1 - i0, T, 2 * M + 1).astype(np.float32)
    return out


def extract_tables(U: np.ndarray, xc: np.ndarray, yc: np.ndarray, nC: int):
    """iCC, iU, Ucc as in prepare_extract (KS/template_matching.py:20-26).

    ``np.argsort`` there is not stable; equidistant channels are ordered here by
    channel index (stable sort), which is the rule this build uses throughout.
    """
    ds = (xc - xc[:, None]) ** 2 + (yc - yc[:, None]) ** 2
    iCC = np.argsort(ds, axis=0, kind="stable")[:nC].astype(np.int64)   # [nC][C]
    iU = np.argmax((U.astype(np.float64) ** 2).sum(1), axis=-1).astype(np.int64)  # [T]
    T = U.shape[0]
    Ucc = U[np.arange(T)[None, :], :, iCC[:, iU]]                      # [nC][T][K]
    return iCC, iU, np.ascontiguousarray(Ucc.astype(np.float32))


def highpass_impulse(fs: float = 30000.0, cutoff: float = 300.0, NT: int = 30122) -> np.ndarray:
    """hp_filter.npy: impulse response of filtfilt with a 3rd-order Butterworth
    high-pass, from a unit impulse at NT//2 of NT samples, as FP32
    (get_highpass_filter, KS/preprocessing.py:121-136)."""
    from scipy.signal import butter, filtfilt
    b, a = butter(3, cutoff, fs=fs, btype="high")
    x = np.zeros(NT)
    x[NT // 2] = 1.0
    return filtfilt(b, a, x).astype(np.float32)


def write_oss_input(model: ProbeModel, path: str | os.PathLike, *, with_ctc: bool = True) -> Path:
    """Write an oss_input directory (file names as read at OnlineSpikesV2.cu:500-510,580-598)."""
    p = Path(path)
    p.mkdir(parents=True, exist_ok=True)
    C, T, M, K = model.C, model.T, model.M, model.K
    xc = np.zeros(C, dtype=np.float32)
    yc = (20.0 * np.arange(C)).astype(np.float32)
    iCC, iU, Ucc = extract_tables(model.U, xc, yc, model.nC)
    waves = model.waves
    np.save(p / "templates.npy", waves)                       # [T][M][C]
    np.save(p / "whiteningMat.npy", model.whiten.astype(np.float32))
    np.save(p / "channelMap.npy", np.arange(C, dtype=np.int32))
    np.save(p / "templateMap.npy", np.arange(T, dtype=np.int32))
    np.save(p / "Wall3.npy", model.U.astype(np.float32))
    if with_ctc:
        ctc = ctc_table(model.U, model.wPCA)
        np.save(p / "ctc.npy", ctc)
    np.save(p / "wPCA.npy", model.wPCA.astype(np.float32))
    np.save(p / "wPCA_permuted.npy", np.ascontiguousarray(model.wPCA.T).astype(np.float32))
    np.save(p / "preclustered_template_waveforms.npy", waves)
    np.save(p / "drift_matrix.npy", np.eye(C, dtype=np.float32))
    np.save(p / "iCC.npy", iCC)
    np.save(p / "iU.npy", iU)
    np.save(p / "Ucc.npy", Ucc)
    np.save(p / "xc.npy", xc)
    np.save(p / "yc.npy", yc)
    # final clusters: one per template, centred on its footprint (x = 0, y = 20 um x centre)
    np.save(p / "cluster_centroids.npy",
            np.stack([np.zeros(T), 20.0 * model.mu], axis=1).astype(np.float32))
    np.save(p / "hp_filter.npy", highpass_impulse(model.fs))
    with open(p / "misc.txt", "w") as f:
        f.write(f"nt0min:{model.nt0min}\n")
        f.write(f"numNearestChans:{model.nC}\n")
        f.write(f"Th_learned:{model.Th:g}\n")
        f.write(f"duplicate_spike_bins:{model.dup_bins}\n")
    return p


# ---------------------------------------------------------------------------
# Recordings
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class Truth:
    centre: np.ndarray     # int64 sample at which the waveform is centred
    template: np.ndarray   # int32
    amplitude: np.ndarray  # float32 (whitened units; B/nm recovers it)


def draw_spikes(model: ProbeModel, n_samples: int, seed: int) -> Truth:
    """Spike trains: Poisson (exponential ISI) or gamma ISI of the given shape."""
    rng = _rngs(seed)[_S_SPIKES]
    M = model.M
    dur = n_samples / model.fs
    cs, ts, amps = [], [], []
    for t in range(model.T):
        rate = model.rates[t]
        mean_isi = 1.0 / rate
        n_draw = int(rate * dur * 3 + 20)
        if model.isi == "gamma":
            k = model.gamma_shape
            isis = rng.gamma(k, mean_isi / k, size=n_draw)
        else:
            isis = rng.exponential(mean_isi, size=n_draw)
        tt = np.cumsum(isis)
        tt = tt[tt < dur]
        s = np.round(tt * model.fs).astype(np.int64)
        s = s[(s >= M) & (s < n_samples - M)]
        a = rng.normal(model.amp_mean, model.amp_sd, size=len(s))
        a = np.maximum(a, 1.0)
        cs.append(s)
        ts.append(np.full(len(s), t, dtype=np.int32))
        amps.append(a.astype(np.float32))
    c = np.concatenate(cs)
    t = np.concatenate(ts)
    a = np.concatenate(amps)
    order = np.lexsort((t, c))
    return Truth(c[order], t[order], a[order])


def _render_block(model: ProbeModel, truth: Truth, b0: int, b1: int, seed_seq, noise: float) -> np.ndarray:
    """int16 samples [b0, b1) of a recording (one noise stream per block)."""
    C, M = model.C, model.M
    half = M // 2
    waves = model.waves                                   # [T][M][C]
    rng = np.random.Generator(np.random.PCG64(seed_seq))
    x = rng.standard_normal((b1 - b0, C), dtype=np.float32) * np.float32(noise)
    lo = np.searchsorted(truth.centre, b0 - half, side="left")
    hi = np.searchsorted(truth.centre, b1 + half, side="left")
    for i in range(lo, hi):
        s = int(truth.centre[i])
        w0, w1 = s - half, s + half + 1
        a0, a1 = max(w0, b0), min(w1, b1)
        if a0 >= a1:
            continue
        x[a0 - b0:a1 - b0] += truth.amplitude[i] * waves[truth.template[i], a0 - w0:a1 - w0]
    if model.mix is not None:
        x = x @ model.mix.T.astype(np.float32)
    q = np.rint(x * np.float32(model.gain))
    np.clip(q, -32768, 32767, out=q)
    return q.astype(np.int16)


_POOL_STATE = {}


def _render_job(args):
    b0, b1, seed_seq = args
    st = _POOL_STATE
    return b0, _render_block(st["model"], st["truth"], b0, b1, seed_seq, st["noise"])


def make_recording(model: ProbeModel, n_samples: int, seed: int | None = None,
                   block: int = 30000, noise: float = 1.0, workers: int = 1):
    """int16 recording ``[n_samples][C]`` (channel-fastest, as SpikeGLX streams it).

    ``workers`` > 1 renders the blocks in forked processes; the result is the
    same array (each block has its own noise stream)."""
    seed = model.seed if seed is None else seed
    truth = draw_spikes(model, n_samples, seed)
    out = np.empty((n_samples, model.C), dtype=np.int16)
    noise_ss = np.random.SeedSequence([seed, _S_NOISE])
    starts = list(range(0, n_samples, block))
    seeds = noise_ss.spawn(len(starts))
    jobs = [(b0, min(n_samples, b0 + block), seeds[bi]) for bi, b0 in enumerate(starts)]
    if workers > 1 and len(jobs) > 1:
        import multiprocessing as mp
        _POOL_STATE.update(model=model, truth=truth, noise=noise)
        # the workers are let finish and joined (close + join), never terminated:
        # terminate() signals them, and a signal handler they inherited (a
        # profiler's, say) would report that as an abort
        pool = mp.get_context("fork").Pool(workers)
        try:
            for b0, q in pool.imap_unordered(_render_job, jobs, chunksize=2):
                out[b0:b0 + len(q)] = q
            pool.close()
        except BaseException:
            pool.terminate()
            raise
        finally:
            pool.join()
            _POOL_STATE.clear()
    else:
        for b0, b1, sq in jobs:
            out[b0:b1] = _render_block(model, truth, b0, b1, sq, noise)
    return out, truth


def kilosort_time(centre: np.ndarray, M: int = 61, nt0min: int = 20) -> np.ndarray:
    """Kilosort's reported spike time for a waveform centred at ``centre``.

    st = (w - nt) + ibatch*NT - nt//2 + nt0min (KS/template_matching.py:68): the
    column w of a batch whose column 0 is absolute sample ibatch*NT - nt.
    """
    return centre - M // 2 + nt0min


CONFIGS = {
    # Plain matcher on CPU: 64 ch, 32 templates, 1 s as one 30000-sample batch.
    "cfg1": dict(C=64, T=32, whitening="identity", rate=5.0, amp=(10.0, 2.0)),
    # Plain FP32 matcher on one GPU: 384 ch, 512 templates, 3000-sample chunks.
    "cfg2": dict(C=384, T=512, whitening="banded", rate=(0.5, 20.0), amp=(12.0, 3.0)),
    # BF16 projection: 384 ch, 1024 templates, bursty gamma(0.5) ISIs.
    "cfg3": dict(C=384, T=1024, whitening="banded", rate=(0.5, 20.0), amp=(12.0, 3.0),
                 isi="gamma", gamma_shape=0.5),
    # Learned proposal + BF16 matcher: 384 ch, 1024 templates, 1500-sample chunks
    # (the probe model is drawn from the test seed; training recordings from TRAIN_SEED).
    "cfg5": dict(C=384, T=1024, whitening="banded", rate=(0.5, 20.0), amp=(12.0, 3.0)),
}
