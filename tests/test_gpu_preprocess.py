"""Preprocessing in front of whitening (channel means, CAR, high-pass) on the
GPU against the oracle (oracle/preprocess.py, pinned in
test_preprocess_oracle.py).

Bar: means, medians and the mean/CAR output are bit-exact (integer sums with
one rounding, an exact selection, FP32 differences, as the references
compute them); the high-pass output is FP32 accumulation over the filter's
taps against the FP64 FFT form, within 2e-6 of the batch's largest
magnitude; events from the preprocessed batch meet the matcher's parity bar
against the FP64 oracle run on the oracle's preprocessed batch.
"""
import numpy as np
import pytest

import oracle
from livespikesorter_amd import LssError, PRE_ALL, PRE_CAR, PRE_HIGHPASS, PRE_MEAN, SEM_KILOSORT, SEM_LSS, synth
from livespikesorter_amd.matcher import STAGE_PRE, STAGE_PRE_MEAN, STAGE_PRE_MED, OnlineSpikesV2, StreamGroup
from oracle import preprocess as pp
from tests.parity import check_events

pytestmark = pytest.mark.gpu

HP_ATOL = 2e-6


def raw_recording(m, n, seed=synth.TEST_SEED):
    """Synthetic int16 recording with per-channel offsets, a common-mode 50 Hz
    hum and a slow drift: what the three steps remove."""
    x, _ = synth.make_recording(m, n, seed=seed)
    r = np.random.default_rng(seed + 17)
    t = np.arange(n) / m.fs
    off = r.integers(-600, 600, m.C)
    common = 25.0 * np.sin(2 * np.pi * 50.0 * t + 0.4) + 40.0 * np.sin(2 * np.pi * 0.7 * t)
    y = x.astype(np.float64) + off[None, :] + np.round(common)[:, None]
    return np.clip(np.round(y), -32768, 32767).astype(np.int16)


@pytest.fixture(scope="module")
def small(model_dirs):
    m, d = model_dirs("cfg1")
    return m, d, oracle.load_tables(d), np.load(d / "hp_filter.npy")


@pytest.mark.parametrize("flags", [PRE_MEAN, PRE_MEAN | PRE_CAR, PRE_ALL, PRE_CAR | PRE_HIGHPASS])
@pytest.mark.parametrize("W", [3122, 3121, 700])
@pytest.mark.parametrize("sem", [SEM_KILOSORT, SEM_LSS])
def test_stages_match_oracle(small, flags, W, sem):
    m, d, tab, hp = small
    x = raw_recording(m, W, seed=synth.TEST_SEED + W)
    print(f"[pre flags={flags} W={W} sem={sem}]", flush=True)
    with OnlineSpikesV2(d, max_batch=W, preprocess=flags, semantics=sem) as sp:
        sp.process_batch(x)
        got = sp.debug_stage(STAGE_PRE)
        mean = sp.debug_stage(STAGE_PRE_MEAN)
        med = sp.debug_stage(STAGE_PRE_MED)
        p2p = sp.query().p2p
    ref = pp.preprocess(x, flags, hp, semantics=sem)
    assert np.array_equal(mean, ref["mean"])
    assert np.array_equal(med, ref["med"])
    assert p2p == ref["p2p"]
    scale = np.abs(ref["x1"]).max()
    if flags & PRE_HIGHPASS:
        err = np.abs(got - ref["out"]).max() / scale
        print(f"[pre flags={flags} W={W} sem={sem}] high-pass max err {err:.2e} of {scale:.1f}")
        assert err <= HP_ATOL
    else:
        assert np.array_equal(got, ref["x1"])


@pytest.mark.parametrize("cfg,W", [("cfg1", 30122), ("cfg2", 3122), ("cfg2", 3121)])
def test_events_after_preprocessing_match_oracle(model_dirs, cfg, W):
    m, d = model_dirs(cfg)
    tab = oracle.load_tables(d)
    hp = np.load(d / "hp_filter.npy")
    x = raw_recording(m, W)
    print(f"[pre events {cfg} W={W}]", flush=True)
    with OnlineSpikesV2(d, max_batch=W, preprocess=PRE_ALL) as sp:
        ev, feat = sp.process_batch(x)
    xf = pp.preprocess(x, PRE_ALL, hp)["out"]
    ref = oracle.run_model_batch(tab, x, xf=xf)
    assert len(ref.w) > 20
    check_events(ref, ev, feat, M=m.M, nt0min=m.nt0min, label=f"pre {cfg} W={W}")


def test_group_with_preprocessing_matches_single_handles(model_dirs):
    m, d = model_dirs("cfg1")
    W = 3122
    xs = [raw_recording(m, W, seed=synth.TEST_SEED + 100 + i) for i in range(3)]
    single = []
    with OnlineSpikesV2(d, max_batch=W, preprocess=PRE_ALL) as sp:
        for x in xs:
            single.append(sp.process_batch(x))
    print("[pre group]", flush=True)
    with StreamGroup([d] * 3, batch=W, preprocess=PRE_ALL) as grp:
        out = grp.process(xs, [0, 0, 0])
    for (e0, f0), (e1, f1) in zip(single, out):
        assert np.array_equal(e0, e1) and np.array_equal(f0, f1)


def test_short_filter_lss_error_kilosort_padding(small, tmp_path):
    import shutil
    m, d, tab, _ = small
    d2 = tmp_path / "short"
    shutil.copytree(d, d2)
    hp = pp.highpass_filter(NT=1001)
    np.save(d2 / "hp_filter.npy", hp)
    W = 3122
    x = raw_recording(m, W)
    with OnlineSpikesV2(d2, max_batch=W, preprocess=PRE_ALL, semantics=SEM_LSS) as sp:
        with pytest.raises(LssError, match="longer than"):
            sp.process_batch(x)
    with OnlineSpikesV2(d2, max_batch=W, preprocess=PRE_ALL) as sp:
        sp.process_batch(x)
        got = sp.debug_stage(STAGE_PRE)
    ref = pp.preprocess(x, PRE_ALL, hp)
    assert np.abs(got - ref["out"]).max() / np.abs(ref["x1"]).max() <= HP_ATOL


def test_preprocessing_requires_filter_file(small, tmp_path):
    import shutil
    m, d, _, _ = small
    d2 = tmp_path / "nofilter"
    shutil.copytree(d, d2)
    (d2 / "hp_filter.npy").unlink()
    with pytest.raises(LssError, match="hp_filter"):
        OnlineSpikesV2(d2, preprocess=PRE_ALL)
    with OnlineSpikesV2(d2, max_batch=3122, preprocess=PRE_MEAN | PRE_CAR) as sp:
        sp.process_batch(raw_recording(m, 3122))
