"""Parity of the HIP matcher (through the C ABI) against the FP64 oracle.

Bar (stated per north-star parity rules): spike (sample, template) identical
for >= 99.5% of events, the rest within +-1 sample; amplitudes within 1e-3
relative error; PC features within 1e-3 of the spike's feature-vector
magnitude (max |f| over its nC x K entries; single entries can cancel to ~0,
where a relative error means nothing).  Intermediate stages are FP32 against
FP64 and are checked to FP32 accumulation accuracy.
"""
import numpy as np
import pytest

import oracle
from livespikesorter_amd import synth
from livespikesorter_amd.matcher import (OnlineSpikesV2, STAGE_B, STAGE_BPC, STAGE_CFMAX, STAGE_IMAX,
                                         STAGE_NM, STAGE_XW)
from tests.parity import agreement, check_events

pytestmark = pytest.mark.gpu


def _oracle(tables, x, **kw):
    kw.setdefault("margins", True)
    return oracle.run_model_batch(tables, x, **kw)


@pytest.fixture(scope="module")
def small(model_dirs):
    m, d = model_dirs("cfg1")
    x, truth = synth.make_recording(m, 3122 * 2, seed=synth.TEST_SEED)
    return m, d, oracle.load_tables(d), x


@pytest.fixture(scope="module")
def cfg2(model_dirs):
    m, d = model_dirs("cfg2")
    x, truth = synth.make_recording(m, 3122 * 3, seed=synth.TEST_SEED)
    return m, d, oracle.load_tables(d), x


def test_stages_against_oracle(small):
    m, d, tab, x = small
    batch = x[:3122]
    ref = _oracle(tab, batch, stages=True)
    with OnlineSpikesV2(d, keep_stages=True) as sp:
        ev, feat = sp.process_batch(batch)
        Xw = sp.debug_stage(STAGE_XW)
        Bpc = sp.debug_stage(STAGE_BPC)
        B = sp.debug_stage(STAGE_B)
        cf = sp.debug_stage(STAGE_CFMAX)
        im = sp.debug_stage(STAGE_IMAX)
        nm = sp.debug_stage(STAGE_NM)
    tol = lambda a: 2e-6 * np.abs(a).max() + 1e-7
    assert np.abs(Xw - ref.Xw).max() <= tol(ref.Xw)
    rb = ref.Bpc.transpose(1, 0, 2)
    assert np.abs(Bpc - rb).max() <= tol(rb) * 4
    assert np.abs(B - ref.B).max() <= 1e-5 * np.abs(ref.B).max()
    nm64 = (tab["Wall3"].astype(np.float64) ** 2).sum((1, 2))
    assert np.array_equal(nm, nm64.astype(np.float32))
    assert np.abs(cf - ref.Cfmax).max() <= 1e-4 * np.abs(ref.Cfmax).max()
    assert np.mean(im == ref.imax) >= 0.999
    check_events(ref, ev, feat, label="cfg1-model 3122")


def test_cfg1_full_batch(model_dirs):
    """cfg1: 64 channels, 32 templates, the whole 1 s as one 30000-sample batch."""
    m, d = model_dirs("cfg1")
    x, _ = synth.make_recording(m, 30000 + 122, seed=synth.TEST_SEED)
    tab = oracle.load_tables(d)
    ref = _oracle(tab, x)
    with OnlineSpikesV2(d, max_batch=30122) as sp:
        ev, feat = sp.process_batch(x)
    check_events(ref, ev, feat, label="cfg1 30122")


def test_cfg2_chunk(cfg2):
    m, d, tab, x = cfg2
    with OnlineSpikesV2(d) as sp:
        for j in range(2):
            batch = x[j * 3000:j * 3000 + 3122]
            ref = _oracle(tab, batch)
            ev, feat = sp.process_batch(batch, first_sample=j * 3000)
            check_events(ref, ev, feat, first_sample=j * 3000, label=f"cfg2 chunk {j}")


def test_sparse_skipping_is_exact(cfg2):
    """Skipping exact-zero blocks and reordering templates changes no bit of the output."""
    m, d, tab, x = cfg2
    batch = x[:3122]
    with OnlineSpikesV2(d, sparse=1) as a, OnlineSpikesV2(d, sparse=0) as b:
        ea, fa = a.process_batch(batch)
        eb, fb = b.process_batch(batch)
    assert np.array_equal(ea, eb)
    assert np.array_equal(fa, fb)


def test_negligible_tails(cfg2):
    """Sparse level 2 (the default) also drops projection weights at most 2^-40 of their
    template's largest and ctc rows at most 2^-40 sqrt(nm_s nm_t): every spike keeps its
    sample and template, amplitudes move by less than 1e-6 relative (a score near
    threshold moves by far less than one FP32 rounding), and so do the features (they
    read the amplitudes, not the tables).  The projection's K-block lists shrink."""
    m, d, tab, x = cfg2
    with OnlineSpikesV2(d, sparse=2) as a, OnlineSpikesV2(d, sparse=1) as b:
        ia, ib = a.query(), b.query()
        for j in range(3):
            batch = x[j * 3000:j * 3000 + 3122]
            ea, fa = a.process_batch(batch, first_sample=j * 3000)
            eb, fb = b.process_batch(batch, first_sample=j * 3000)
            assert np.array_equal(ea["sample"], eb["sample"]) and np.array_equal(ea["template"], eb["template"])
            rel = np.abs(ea["amplitude"] - eb["amplitude"]) / np.abs(eb["amplitude"])
            assert rel.max() <= 1e-6, rel.max()
            mag = np.abs(fb).reshape(len(fb), -1).max(1)
            assert (np.abs(fa - fb).reshape(len(fb), -1).max(1) <= 1e-6 * np.maximum(mag, 1e-30)).all()
    print(f"projection K-blocks: {ib.proj_kblocks} exact zeros, {ia.proj_kblocks} with negligible tails")
    assert ia.proj_kblocks <= ib.proj_kblocks


def test_template_split_is_exact(cfg2, model_dirs):
    """Splitting a column block's templates over several workgroups changes no bit."""
    m, d, tab, x = cfg2
    batch = x[:3122]
    with OnlineSpikesV2(d, peel_split=1) as a, OnlineSpikesV2(d) as b, \
            OnlineSpikesV2(d, peel_split=2, keep_stages=True) as c:
        ea, fa = a.process_batch(batch)
        eb, fb = b.process_batch(batch)
        ec, fc = c.process_batch(batch)
        qa, qb, qc = a.query(), b.query(), c.query()
        cf_c = c.debug_stage(STAGE_CFMAX)
        im_c = c.debug_stage(STAGE_IMAX)
    print(f"split {qa.peel_split}/{qb.peel_split}/{qc.peel_split}, rows {qa.peel_rows}/{qb.peel_rows}/"
          f"{qc.peel_rows}, groups {qa.peel_groups}/{qb.peel_groups}/{qc.peel_groups}")
    assert qa.peel_split == 1 and qb.peel_split > 1 and qc.peel_split == 2
    assert np.array_equal(ea, eb) and np.array_equal(fa, fb)
    assert np.array_equal(ea, ec) and np.array_equal(fa, fc)
    # first-iteration column maxima assembled from the subsets' partials
    ref = _oracle(tab, batch, stages=True, features=False)
    assert np.abs(cf_c - ref.Cfmax).max() <= 1e-4 * np.abs(ref.Cfmax).max()
    assert np.mean(im_c == ref.imax) >= 0.999
    # 1024 templates, 1500-sample chunks: eight subsets of 128 rows, and the
    # default (four subsets of 256 rows)
    m5, d5 = model_dirs("cfg2", T=1024)
    x5, _ = synth.make_recording(m5, 1622 * 2, seed=synth.TEST_SEED)
    with OnlineSpikesV2(d5, max_batch=1622, peel_split=1) as a, OnlineSpikesV2(d5, max_batch=1622) as b, \
            OnlineSpikesV2(d5, max_batch=1622, peel_split=8) as c8:
        for j in range(2):
            ea, fa = a.process_batch(x5[j * 1500:j * 1500 + 1622])
            eb, fb = b.process_batch(x5[j * 1500:j * 1500 + 1622])
            ec, fc = c8.process_batch(x5[j * 1500:j * 1500 + 1622])
            assert np.array_equal(ea, eb) and np.array_equal(fa, fb)
            assert np.array_equal(ea, ec) and np.array_equal(fa, fc)
        assert b.query().peel_split == 4 and b.query().peel_rows == 16
        assert c8.query().peel_split == 8 and c8.query().peel_rows == 8


def test_autotune_keeps_results(cfg2):
    """lss_autotune times every split on a device batch, keeps the fastest for
    that width, and the output stays bit-identical."""
    from livespikesorter_amd._hip import DeviceBuffer, synchronize
    m, d, tab, x = cfg2
    batch = np.ascontiguousarray(x[:3122])
    dx = DeviceBuffer.from_array(batch)
    synchronize()
    with OnlineSpikesV2(d) as a, OnlineSpikesV2(d) as b:
        ea, fa = a.process_batch_device(dx.ptr, 3122, 384)
        res = b.autotune(dx.ptr, 3122, 384, reps=3)
        print("autotune:", res)
        eb, fb = b.process_batch_device(dx.ptr, 3122, 384)
        chosen = [r for r in res if r[3]][0]
        assert b.query().peel_split == chosen[0]
    dx.free()
    assert len(res) >= 2 and res[0][0] == 1
    assert np.array_equal(ea, eb) and np.array_equal(fa, fb)


def test_positions_and_relabel(cfg2):
    """Spike positions from the GPU features against the position oracle, and
    the nearest-cluster relabelling (saveSpikes' template index) against the
    oracle's nearest centroid."""
    m, d, tab, x = cfg2
    batch = x[:3122]
    xc = np.load(d / "xc.npy")
    yc = np.load(d / "yc.npy")
    cent = np.load(d / "cluster_centroids.npy")
    with OnlineSpikesV2(d) as a, OnlineSpikesV2(d, relabel=True) as b:
        ea, fa = a.process_batch(batch)
        pa = a.last_positions()
        eb, fb = b.process_batch(batch)
        pb = b.last_positions()
    assert len(pa) == len(ea) > 0
    px, py = oracle.spike_positions(fa, ea["template"], tab["iCC"], tab["iU"], xc, yc)
    err = np.abs(pa - np.stack([px, py], 1)).max()
    scale = max(np.abs(xc).max(), np.abs(yc).max())
    print(f"positions: {len(ea)} events, max |gpu - oracle(gpu features)| {err:.2e} um (scale {scale:.0f} um)")
    assert err <= 1e-6 * scale                                  # FP32 rounding of the weighted sums
    # against the FP64 oracle's own features
    ref = _oracle(tab, batch)
    rx, ry = oracle.spike_positions(ref.feat, ref_templates(ref), tab["iCC"], tab["iU"], xc, yc)
    ref_s = ref.w - 61 // 2 + 20
    _, _, pairs = agreement(ref_s, ref_templates(ref), ea["sample"], ea["template"], tol=0)
    ri = np.array([p[0] for p in pairs]); gi = np.array([p[1] for p in pairs])
    e64 = max(np.abs(pa[gi, 0] - rx[ri]).max(), np.abs(pa[gi, 1] - ry[ri]).max())
    print(f"positions vs FP64 oracle features: max error {e64:.2e} um over {len(pairs)} events")
    assert e64 <= 1e-5 * scale
    # relabelling: same spikes (features identical), template = nearest centroid
    assert np.array_equal(eb["sample"], ea["sample"]) and np.array_equal(fb, fa)
    want = oracle.closest_cluster(pa[:, 0], pa[:, 1], cent)
    same = np.mean(eb["template"] == want)
    print(f"relabel: {same:.5f} of events on the oracle's nearest centroid")
    assert same >= 0.999


def ref_templates(ref):
    return np.asarray(ref.t)


def test_epoch_wrap_keeps_results(cfg2, monkeypatch):
    """The peel's published column maxima carry a per-batch epoch in their tags;
    when it wraps the per-batch reset clears every word.  A handle started three
    batches before the wrap gives the same events over batches of changing
    width (words left from other layouts) as one started at zero."""
    m, d, tab, x = cfg2
    widths = [3122, 2000, 3122, 1500, 3122, 2500]
    monkeypatch.setenv("LSS_EPOCH_START", str((1 << 19) - 3))
    a = OnlineSpikesV2(d)
    monkeypatch.delenv("LSS_EPOCH_START")
    with a, OnlineSpikesV2(d) as b:
        for j, wd in enumerate(widths):
            batch = x[j * 1000:j * 1000 + wd]
            ea, fa = a.process_batch(batch)
            eb, fb = b.process_batch(batch)
            assert len(ea) > 0
            assert np.array_equal(ea, eb) and np.array_equal(fa, fb), f"batch {j} (width {wd})"


def test_profile_levels(cfg2):
    """Stage timing changes no result; level 2 times the peel stage only."""
    m, d, tab, x = cfg2
    batch = x[:3122]
    with OnlineSpikesV2(d) as a, OnlineSpikesV2(d, profile=1) as b, OnlineSpikesV2(d, profile=2) as c:
        ea, fa = a.process_batch(batch)
        eb, fb = b.process_batch(batch)
        ec, fc = c.process_batch(batch)
        sb, nb = b.stage_times()
        sc, nc = c.stage_times()
    assert np.array_equal(ea, eb) and np.array_equal(fa, fb)
    assert np.array_equal(ea, ec) and np.array_equal(fa, fc)
    assert nb == 1 and nc == 1
    assert sb["whiten"] > 0 and sb["peel"] > 0
    assert sc["peel"] > 0 and all(v == 0 for k, v in sc.items() if k != "peel")


def test_deterministic(cfg2):
    m, d, tab, x = cfg2
    batch = x[3000:6122]
    with OnlineSpikesV2(d) as sp:
        e1, f1 = sp.process_batch(batch)
        e2, f2 = sp.process_batch(batch)
    assert np.array_equal(e1, e2) and np.array_equal(f1, f2)


def test_streaming_equals_framed_batches(small):
    """lss_process_chunk frames [previous 2M | chunk] and emits each sample once."""
    m, d, tab, x = small
    N, M = 1000, m.M
    n_total = 5 * N
    rec = x[:n_total]
    got = []
    with OnlineSpikesV2(d, chunk=N) as sp:
        for j in range(5):
            ev, _ = sp.process_chunk(rec[j * N:(j + 1) * N], j * N)
            got.append(ev)
        ev, _ = sp.flush()
        got.append(ev)
    got = np.concatenate(got)
    # oracle on the same frames
    frames = [np.concatenate([np.repeat(rec[:1], M, 0), rec[:N]])]
    starts = [-M]
    for j in range(1, 5):
        frames.append(rec[j * N - 2 * M:(j + 1) * N])
        starts.append(j * N - 2 * M)
    frames.append(np.concatenate([rec[n_total - 2 * M:], np.repeat(rec[-1:], M, 0)]))
    starts.append(n_total - 2 * M)
    ref_s, ref_t = [], []
    for fr, s0 in zip(frames, starts):
        r = _oracle(tab, fr, features=False)
        ref_s.append(r.w + s0 - M // 2 + m.nt0min)
        ref_t.append(r.t)
    ref_s = np.concatenate(ref_s)
    ref_t = np.concatenate(ref_t)
    assert np.all(np.diff(got["sample"]) >= 0)
    ident, within, _ = agreement(ref_s, ref_t, got["sample"], got["template"])
    assert ident >= 0.995 and within >= 0.999
    # every emitted sample lies in the stream
    assert got["sample"].min() >= -M and got["sample"].max() < n_total


def test_lss_semantics(small):
    m, d, tab, x = small
    batch = x[:3122]
    ref = oracle.run_model_batch(tab, batch, semantics=oracle.SEM_LSS, max_iter=50, bits=32)
    with OnlineSpikesV2(d, semantics=1) as sp:
        ev, feat = sp.process_batch(batch)
        assert sp.info.max_iter == 50
    check_events(ref, ev, None, label="LSS semantics")


def test_device_resident_input(small):
    from livespikesorter_amd._hip import DeviceBuffer
    m, d, tab, x = small
    batch = np.ascontiguousarray(x[:3122])
    dx = DeviceBuffer.from_array(batch)
    assert np.array_equal(dx.download(batch.shape, batch.dtype), batch)
    with OnlineSpikesV2(d) as sp:
        e1, f1 = sp.process_batch(batch)
        e2, f2 = sp.process_batch_device(dx.ptr, batch.shape[0], batch.shape[1])
    dx.free()
    assert np.array_equal(e1, e2) and np.array_equal(f1, f2)


def test_errors(small):
    m, d, tab, x = small
    from livespikesorter_amd import LssError
    with OnlineSpikesV2(d, chunk=1000) as sp:
        with pytest.raises(LssError) as ei:
            sp.process_batch(np.zeros((5000, m.C), np.int16))
        assert ei.value.code == -1
        ev, _ = sp.process_batch(np.zeros((200, m.C), np.int16))
        assert len(ev) == 0
        with pytest.raises(LssError):
            sp.flush()
    with pytest.raises(LssError) as ei:
        OnlineSpikesV2(str(d) + "_missing")
    assert ei.value.code == -2


def test_stream_group_matches_single_handles(model_dirs):
    """Streams of different template counts in one graph launch give, bit for
    bit, the events and features of separate single-stream handles."""
    from livespikesorter_amd import StreamGroup
    specs = [dict(C=64, T=32, seed=31), dict(C=64, T=96, seed=32), dict(C=96, T=160, seed=33)]
    dirs, recs = [], []
    for k, sp in enumerate(specs):
        m, d = model_dirs("grp%d" % k, C=sp["C"], T=sp["T"], seed=sp["seed"], whitening="banded",
                          rate=(2.0, 20.0), amp=(11.0, 3.0))
        x, _ = synth.make_recording(m, 4 * 1500 + 1622, seed=sp["seed"] + 100)
        dirs.append(d)
        recs.append(x)
    W = 1622
    singles = [OnlineSpikesV2(d, max_batch=W) for d in dirs]
    with StreamGroup(dirs, batch=W) as grp:
        for tick in range(3):
            bs = [x[tick * 1500:tick * 1500 + W] for x in recs]
            out = grp.process(bs, [tick * 1500] * len(bs))
            for k, (ev, feat) in enumerate(out):
                e1, f1 = singles[k].process_batch(bs[k], first_sample=tick * 1500)
                assert len(e1) > 0
                assert np.array_equal(ev, e1), f"stream {k} tick {tick}"
                assert np.array_equal(feat, f1)
        # a tick with stream 1 idle: the others are unchanged, stream 1 returns nothing
        bs = [x[4 * 1500:4 * 1500 + W] for x in recs]
        idle = [bs[0], None, bs[2]]
        out = grp.process(idle, [4 * 1500] * 3)
        assert len(out[1][0]) == 0
        for k in (0, 2):
            e1, f1 = singles[k].process_batch(bs[k], first_sample=4 * 1500)
            assert np.array_equal(out[k][0], e1) and np.array_equal(out[k][1], f1)
        # and the next full tick still matches (stream 1's state was untouched)
        out = grp.process(bs, [4 * 1500] * 3)
        e1, f1 = singles[1].process_batch(bs[1], first_sample=4 * 1500)
        assert np.array_equal(out[1][0], e1)
        # two lanes driven from two threads at once, on disjoint streams
        from concurrent.futures import ThreadPoolExecutor
        bs = [x[2 * 1500:2 * 1500 + W] for x in recs]
        with ThreadPoolExecutor(2) as pool:
            fa = pool.submit(grp.process, [bs[0], None, bs[2]], [3000] * 3, lane=0)
            fb = pool.submit(grp.process, [None, bs[1], None], [3000] * 3, lane=1)
            oa, ob = fa.result(), fb.result()
        for k, o in ((0, oa[0]), (1, ob[1]), (2, oa[2])):
            e1, f1 = singles[k].process_batch(bs[k], first_sample=3000)
            assert np.array_equal(o[0], e1) and np.array_equal(o[1], f1), f"lane run, stream {k}"
    for sp in singles:
        sp.close()


def test_stream_group_template_split(model_dirs):
    """A group whose largest stream has its templates split over several peel
    workgroups per column block (a shorter register tile for every stream)
    gives, bit for bit, the events and features of the unsplit group and of
    single handles; peel_split 1 keeps one subset, 2 asks for two ways."""
    from livespikesorter_amd import StreamGroup
    specs = [dict(C=64, T=96, seed=41), dict(C=64, T=256, seed=42), dict(C=64, T=520, seed=43)]
    dirs, recs = [], []
    for k, sp in enumerate(specs):
        m, d = model_dirs("grps%d" % k, C=sp["C"], T=sp["T"], seed=sp["seed"], whitening="banded",
                          rate=(2.0, 20.0), amp=(11.0, 3.0))
        x, _ = synth.make_recording(m, 2 * 1500 + 1622, seed=sp["seed"] + 100)
        dirs.append(d)
        recs.append(x)
    W = 1622
    ticks = [[x[t * 1500:t * 1500 + W] for x in recs] for t in range(2)]
    singles = [OnlineSpikesV2(d, max_batch=W) for d in dirs]
    ref = [[singles[k].process_batch(ticks[t][k], first_sample=t * 1500) for k in range(3)] for t in range(2)]
    for sp in singles:
        sp.close()
    want = {0: [(16, 1), (16, 1), (16, 3)], 1: [(64, 1)] * 3, 2: [(32, 1), (32, 1), (32, 2)]}
    for split, geo in want.items():
        with StreamGroup(dirs, batch=W, peel_split=split) as grp:
            assert [(int(i.peel_rows), int(i.peel_split)) for i in grp.info] == geo, f"peel_split={split}"
            for t in range(2):
                out = grp.process(ticks[t], [t * 1500] * 3)
                for k, (ev, feat) in enumerate(out):
                    e1, f1 = ref[t][k]
                    assert len(e1) > 0
                    assert np.array_equal(ev, e1), f"peel_split={split} stream {k} tick {t}"
                    assert np.array_equal(feat, f1)


def test_two_handles_on_alternating_chunks(cfg2):
    """Two handles of one probe driven from two threads on alternating chunks
    (a caller with a backlog) give, bit for bit, what one handle gives."""
    from concurrent.futures import ThreadPoolExecutor
    m, d, tab, x = cfg2
    W = 3122
    starts = list(range(0, len(x) - W + 1, 750))
    with OnlineSpikesV2(d, max_batch=W) as one:
        ref = [one.process_batch(x[s:s + W], first_sample=s) for s in starts]
    hs = [OnlineSpikesV2(d, max_batch=W) for _ in range(2)]
    out = [None] * len(starts)

    def worker(k):
        for i in range(k, len(starts), 2):
            out[i] = hs[k].process_batch(x[starts[i]:starts[i] + W], first_sample=starts[i])

    try:
        for _ in range(2):
            with ThreadPoolExecutor(2) as pool:
                list(pool.map(worker, range(2)))
            for i, (e1, f1) in enumerate(ref):
                assert len(e1) > 0
                assert np.array_equal(out[i][0], e1), f"chunk {i}"
                assert np.array_equal(out[i][1], f1), f"chunk {i}"
    finally:
        for h in hs:
            h.close()


def test_stream_group_profile_timing(model_dirs):
    """A group created with profile times its front stages and its peel with
    HIP events around two graph launches per tick; the results are the ones of
    the unprofiled group, bit for bit."""
    from livespikesorter_amd import StreamGroup
    dirs, recs = [], []
    for k, T in enumerate((64, 160)):
        m, d = model_dirs("grpp%d" % k, C=64, T=T, seed=51 + k, whitening="banded", rate=(2.0, 20.0),
                          amp=(11.0, 3.0))
        dirs.append(d)
        recs.append(synth.make_recording(m, 1500 + 1622, seed=151 + k)[0])
    W = 1622
    with StreamGroup(dirs, batch=W) as g0, StreamGroup(dirs, batch=W, profile=True) as g1:
        for t in range(2):
            bs = [x[t * 1500:t * 1500 + W] for x in recs]
            a = g0.process(bs, [t * 1500] * 2)
            b = g1.process(bs, [t * 1500] * 2)
            for (e0, f0), (e1, f1) in zip(a, b):
                assert np.array_equal(e0, e1) and np.array_equal(f0, f1)
        front, peel, n = g1.stage_times(0, reset=True)
        assert n == 2 and front > 0 and peel > 0
        assert g1.stage_times(0) == (0.0, 0.0, 0)
        with pytest.raises(Exception):
            g0.stage_times(0)
