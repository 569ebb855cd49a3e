"""Parity at the configurations' own shapes, and the ingest forms a live
SpikeGLX stream delivers, against the FP64 oracle (through the C ABI).

* cfg3: 384 channels x 1024 templates, one 60000-sample batch with its
  61-sample margins (W = 60122, Kilosort's batch, KS/parameters.py:42-44),
  bursty gamma(0.5) firing: too wide for register-resident score tiles, so the
  peel keeps its tile in global memory (``peel_rows == 0``).  FP32 and BF16
  projection with FP32 rescoring.
* cfg4: 8 probe streams x 384 channels, templates 256 ... 1024, 1500-sample
  chunks (W = 1622) in one group launch: every stream against the oracle.
* ingest: rows of 385 int16 (the sync channel last, src/SpikeSorter/
  dataSocket.cpp:103-126), a permuted channel map, and a drift matrix saved the
  way the launcher saves it, ``np.save(.., get_drift_matrix(..).T)``
  (src/Python/run_online_spikes.py:628), applied as X <- D (Wr X)
  (src/Python/OnlineKilosort.py:139-140; OnlineSpikesV2.cu:804-808).

The bar is the one of tests/test_gpu_matcher.py (tests/parity.py
check_events): (sample, template) identical for >= 99.5% of events, the rest
within +-1 sample and each difference a near tie of the oracle's own decision
margins (below parity.TIE_RTOL); amplitudes 1e-3 relative; features 1e-3 of
the spike's feature magnitude.
"""
import numpy as np
import pytest

import oracle
from livespikesorter_amd import StreamGroup, synth
from livespikesorter_amd.matcher import LssError, OnlineSpikesV2, PREC_BF16_RESCORE, PREC_FP32
from tests.parity import agreement, check_events

pytestmark = pytest.mark.gpu

CFG4_T = (256, 384, 512, 640, 768, 896, 1024, 1024)


@pytest.fixture(scope="module")
def cfg3_wide(model_dirs):
    m, d = model_dirs("cfg3")
    x, _ = synth.make_recording(m, 60122, seed=synth.TEST_SEED + 7)
    tab = oracle.load_tables(d)
    print("[cfg3 wide] FP64 oracle, W=60122", flush=True)
    ref = oracle.run_model_batch(tab, x, margins=True)
    return m, d, x, ref


@pytest.mark.parametrize("prec", [PREC_FP32, PREC_BF16_RESCORE], ids=["fp32", "bf16"])
def test_cfg3_wide_batch_global_tile(cfg3_wide, prec):
    m, d, x, ref = cfg3_wide
    print(f"[cfg3 wide] GPU precision {prec}", flush=True)
    with OnlineSpikesV2(d, max_batch=len(x), precision=prec, max_events=1 << 17) as sp:
        ev, feat = sp.process_batch(x, cap=1 << 17)
        info = sp.query()
    assert info.peel_rows == 0, "a 60122-column batch runs the global-tile peel"
    assert len(ref.w) > 5000
    check_events(ref, ev, feat, M=m.M, nt0min=m.nt0min, label=f"cfg3 W=60122 prec={prec}")
    if prec == PREC_BF16_RESCORE:
        assert info.rescored >= len(ev)


def _cfg4_models():
    ms = []
    for k, T in enumerate(CFG4_T):
        ms.append(synth.make_model(C=384, T=T, seed=synth.TEST_SEED + 10 + k, whitening="banded",
                                   rate=(0.5, 20.0), amp=(12.0, 3.0)))
    return ms


def test_cfg4_group_against_oracle(tmp_path):
    """8 x 384-channel streams with 256 ... 1024 templates in one graph launch
    (one sorter per probe, as main.cpp:77-85 runs them): each stream's events
    and features against the FP64 oracle, two ticks."""
    N, M = 1500, 61
    W = N + 2 * M
    models = _cfg4_models()
    dirs, recs = [], []
    for k, m in enumerate(models):
        d = tmp_path / f"s{k}"
        # ctc is computed at load time on the GPU; the oracle gets it from the generator
        synth.write_oss_input(m, d, with_ctc=False)
        dirs.append(d)
        recs.append(synth.make_recording(m, N + W, seed=synth.TEST_SEED + 20 + k)[0])
    print("[cfg4] group create", flush=True)
    with StreamGroup(dirs, batch=W) as grp:
        geo = [(int(i.peel_rows), int(i.peel_split), int(i.peel_groups)) for i in grp.info]
        print(f"[cfg4] peel geometry {geo}", flush=True)
        outs = [grp.process([x[j * N:j * N + W] for x in recs], [j * N] * len(recs)) for j in range(2)]
    for k, m in enumerate(models):
        tab = oracle.load_tables(dirs[k], ctc=synth.ctc_table(m.U, m.wPCA))
        for j in range(2):
            ref = oracle.run_model_batch(tab, recs[k][j * N:j * N + W], margins=True)
            ev, feat = outs[j][k]
            check_events(ref, ev, feat, M=M, nt0min=m.nt0min, first_sample=j * N,
                         label=f"cfg4 stream {k} T={m.T} tick {j}")


def _drift(C, rng):
    """An interpolating drift matrix: row i mixes channels i-1, i, i+1 with
    unequal weights (so D and its transpose differ)."""
    D = np.zeros((C, C))
    for i in range(C):
        w = {i - 1: 0.05, i: 0.75 + 0.1 * rng.random(), i + 1: 0.2}
        for j, v in w.items():
            if 0 <= j < C:
                D[i, j] = v
    return (D / D.sum(1, keepdims=True)).astype(np.float32)


@pytest.fixture(scope="module")
def ingest(tmp_path_factory):
    C, LD = 384, 385
    m = synth.make_model(C=C, T=96, seed=synth.TEST_SEED + 3, whitening="banded", rate=(2.0, 20.0),
                         amp=(12.0, 3.0))
    d = tmp_path_factory.mktemp("ingest")
    synth.write_oss_input(m, d)
    rng = np.random.default_rng(5)
    D = _drift(C, rng)
    np.save(d / "drift_matrix.npy", D.T)              # Fortran-ordered on disk, as the launcher saves it
    assert np.load(d / "drift_matrix.npy").flags.f_contiguous
    n = 4 * 1500
    x, _ = synth.make_recording(m, n, seed=synth.TEST_SEED + 4)
    perm = rng.permutation(C).astype(np.int32)        # matcher channel c sits in file column perm[c]
    rows = rng.integers(-2000, 2000, size=(n, LD)).astype(np.int16)   # junk in unused columns and sync
    rows[:, perm] = x
    return m, d, D, x, rows, perm


def _oracle_drift(tab, D, xg, **kw):
    """FP64 oracle of X <- D (Wr X): Wr X in FP64 on the channel-gathered batch
    xg [W][C], then the oracle whitens that with D."""
    xf = tab["whiteningMat"].astype(np.float64) @ xg.T.astype(np.float64)
    return oracle.run_batch(None, D, tab["wPCA"], tab["Wall3"], tab["ctc"], tab["iCC"], tab["iU"],
                            Th=tab["Th"], nC=tab["nC"], xf=xf, **kw)


def test_ingest_batch_row_stride_chan_map_drift(ingest):
    """lss_process_batch on 385-wide rows with a permuted channel map (the
    whitening's scalar gather) and a Fortran-ordered drift matrix."""
    m, d, D, x, rows, perm = ingest
    tab = oracle.load_tables(d)
    W = 3122
    with OnlineSpikesV2(d, max_batch=W) as sp:
        for j in range(2):
            b = rows[j * 1500:j * 1500 + W]
            ev, feat = sp.process_batch(b, first_sample=j * 1500, chan_map=perm)
            ref = _oracle_drift(tab, D, x[j * 1500:j * 1500 + W], margins=True)
            check_events(ref, ev, feat, M=m.M, nt0min=m.nt0min, first_sample=j * 1500, label=f"ingest batch {j}")
            assert len(ev) > 20
    # the drift is applied as D (Wr X), not as its transpose: the transposed matrix gives other spikes
    ref_t = _oracle_drift(tab, np.ascontiguousarray(D.T), x[:W])
    ref = _oracle_drift(tab, D, x[:W])
    assert not (len(ref_t.w) == len(ref.w) and np.array_equal(ref_t.w, ref.w) and np.allclose(ref_t.amp, ref.amp))


def test_ingest_stream_row_stride_chan_map_drift(ingest):
    """lss_process_chunk on 385-wide rows with a permuted channel map: the
    library gathers the channels into its frames; events against the oracle
    on the same frames."""
    m, d, D, x, rows, perm = ingest
    tab = oracle.load_tables(d)
    N, M = 1500, m.M
    got = []
    with OnlineSpikesV2(d, chunk=N) as sp:
        for j in range(4):
            ev, _ = sp.process_chunk(rows[j * N:(j + 1) * N], j * N, chan_map=perm)
            got.append(ev)
        got.append(sp.flush()[0])
    got = np.concatenate(got)
    n = 4 * N
    frames = [np.concatenate([np.repeat(x[:1], M, 0), x[:N]])]
    starts = [-M]
    for j in range(1, 4):
        frames.append(x[j * N - 2 * M:(j + 1) * N])
        starts.append(j * N - 2 * M)
    frames.append(np.concatenate([x[n - 2 * M:], np.repeat(x[-1:], M, 0)]))
    starts.append(n - 2 * M)
    ref_s, ref_t = [], []
    for fr, s0 in zip(frames, starts):
        r = _oracle_drift(tab, D, fr, features=False)
        keep = (r.w >= M) & (r.w < len(fr) - M)
        ref_s.append(r.w[keep] + s0 - M // 2 + m.nt0min)
        ref_t.append(r.t[keep])
    ref_s, ref_t = np.concatenate(ref_s), np.concatenate(ref_t)
    ident, within, _ = agreement(ref_s, ref_t, got["sample"], got["template"])
    print(f"[ingest stream] oracle {len(ref_s)} gpu {len(got)} identical {ident:.5f} within1 {within:.5f}")
    assert len(got) > 50 and ident >= 0.995 and within >= 0.999


def test_capacity_error_leaves_stream_state(ingest):
    """LSS_ERR_CAPACITY from lss_process_chunk with duplicate suppression on:
    the call can be repeated with a larger buffer and gives what a fresh
    handle gives (the dedup table and the stream position were not advanced)."""
    m, d, D, x, rows, perm = ingest
    N = 1500
    with OnlineSpikesV2(d, chunk=N, dedup=1) as ref_h:
        want = [ref_h.process_chunk(rows[j * N:(j + 1) * N], j * N, chan_map=perm)[0] for j in range(3)]
    assert len(want[1]) > 2
    with OnlineSpikesV2(d, chunk=N, dedup=1) as sp:
        got = [sp.process_chunk(rows[:N], 0, chan_map=perm)[0]]
        with pytest.raises(LssError) as ei:
            sp.process_chunk(rows[N:2 * N], N, chan_map=perm, cap=1)
        assert ei.value.code == -4
        got.append(sp.process_chunk(rows[N:2 * N], N, chan_map=perm)[0])
        got.append(sp.process_chunk(rows[2 * N:3 * N], 2 * N, chan_map=perm)[0])
        assert sp.query().events == sum(len(e) for e in want)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
