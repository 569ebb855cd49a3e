"""Template-subset selection and the subset directory writer (host logic,
following src/Python/subset_templates.py), checked against hand-built
directories whose expected selections are worked out by hand."""
import json

import numpy as np
import pytest

from livespikesorter_amd import subset


def make_dir(tmp_path, *, unclu_T=6, T=None, K=2, C=8, M=5, peaks=None, spikes=None, cmap=None, waves=False):
    """Wall3 template t: PC energy on channel peaks[t] (weight 2) and its right neighbour (weight 1)."""
    d = tmp_path / "oss"
    d.mkdir()
    peaks = list(range(unclu_T)) if peaks is None else peaks
    W = np.zeros((unclu_T, K, C), np.float32)
    for t, c in enumerate(peaks):
        W[t, :, c] = 2.0 * (1 + t)
        if c + 1 < C:
            W[t, :, c + 1] = 1.0 * (1 + t)
    np.save(d / "Wall3.npy", W)
    np.save(d / "templates.npy", np.zeros((T or unclu_T, M, C), np.float32))
    np.save(d / "wPCA.npy", np.zeros((K, M), np.float32))
    np.save(d / "iU.npy", np.asarray(peaks, np.int64))
    np.save(d / "Ucc.npy", np.arange(3 * unclu_T * K, dtype=np.float32).reshape(3, unclu_T, K))
    np.save(d / "ctc.npy", np.arange(unclu_T * unclu_T * (2 * M + 1), dtype=np.float32)
            .reshape(unclu_T, unclu_T, 2 * M + 1))
    np.save(d / "channelMap.npy", np.arange(C, dtype=np.int32) if cmap is None else np.asarray(cmap, np.int32))
    if spikes is not None:
        np.save(d / "spike_templates.npy", np.asarray(spikes, np.int64))
    if waves:
        wv = np.zeros((unclu_T, M, C), np.float32)
        for t, c in enumerate(peaks):
            wv[t, 2, c] = 1.0
            wv[t, 2, (c + 3) % C] = 0.05      # leakage below a 10% overlap threshold
        np.save(d / "preclustered_template_waveforms.npy", wv)
    return d


def test_parse_channel_range():
    assert subset.parse_channel_range("100-150") == (100, 150)
    assert subset.parse_channel_range(" 150 - 100 ") == (100, 150)
    assert subset.parse_channel_range("") is None and subset.parse_channel_range(None) is None
    with pytest.raises(ValueError):
        subset.parse_channel_range("100")


def test_channels_in_range_follows_channel_map(tmp_path):
    d = make_dir(tmp_path, cmap=[7, 6, 5, 4, 3, 2, 1, 0])
    assert list(subset.channels_in_range(d, 1, 2)) == [5, 6]


def test_templates_on_channels_peak_and_overlap(tmp_path):
    d = make_dir(tmp_path, peaks=[0, 2, 4, 6, 1, 3])
    # Wall3 norm support: template t covers peaks[t] and peaks[t] + 1 (half the peak)
    assert subset.templates_on_channels(d, [2, 3], "peak") == [1, 5]
    assert subset.templates_on_channels(d, [2, 3], "overlap") == [1, 4, 5]
    with pytest.raises(ValueError):
        subset.templates_on_channels(d, [2], "nearest")


def test_templates_on_channels_prefers_waveforms(tmp_path):
    d = make_dir(tmp_path, peaks=[0, 2, 4, 6, 1, 3], waves=True)
    # waveform support is the peak channel only (leakage is 5% < 10%)
    assert subset.templates_on_channels(d, [3], "overlap") == [5]
    assert subset.templates_on_channels(d, [3], "peak") == [5]


def test_choose_keep_by_spike_count_then_norm(tmp_path):
    spikes = [3] * 5 + [1] * 4 + [5] * 2 + [0] + [-1, 99]      # out-of-range ids are dropped
    d = make_dir(tmp_path, T=3, spikes=spikes)
    dims = subset.shapes(d)
    assert subset.choose_keep(d, dims, 2, "spikes") == ([1, 3], "spike count")
    assert subset.choose_keep(d, dims, 3, "spikes", candidates=[0, 1, 2, 5]) == ([0, 1, 5], "spike count")
    # norm: Wall3 rows scale with t + 1, so the highest indices rank first
    assert subset.choose_keep(d, dims, 2, "norm") == ([4, 5], "template L2 norm")
    assert subset.choose_keep(d, dims, 0) == (list(range(6)), "all")
    assert subset.choose_keep(d, dims, 9, candidates=[4, 2]) == ([2, 4], "channel range")
    (d / "spike_templates.npy").unlink()
    assert subset.choose_keep(d, dims, 2, "spikes") == ([4, 5], "template L2 norm")


def test_choose_keep_rejects_final_cluster_ids(tmp_path):
    d = make_dir(tmp_path, unclu_T=6, T=4, spikes=[0, 1, 2, 3])
    with pytest.raises(ValueError, match="final-cluster"):
        subset.choose_keep(d, subset.shapes(d), 2, "spikes")


def test_compute_keep_filters_channels_first(tmp_path):
    d = make_dir(tmp_path, T=2, peaks=[0, 2, 4, 6, 1, 3], spikes=[1] * 3 + [5] * 9 + [4] * 7)
    keep, ranked, meta = subset.compute_keep(d, 2, "spikes", (1, 3), "overlap")
    # overlap candidates on channels 1..3: templates 0 (ch 0-1), 1 (2-3), 4 (1-2), 5 (3-4)
    assert keep == [4, 5] and ranked == "spike count"
    assert meta["templates_on_channels"] == 4 and meta["probe_channels_in_range"] == 3
    with pytest.raises(ValueError):
        subset.compute_keep(d, 2, "spikes", (50, 60))


def test_write_subset_takes_every_preclustered_axis(tmp_path):
    d = make_dir(tmp_path, unclu_T=6, T=2)
    out = tmp_path / "sub"
    keep = [1, 4, 5]
    man = subset.write_subset(d, out, keep)
    assert man["new_unclu_T"] == 3 and man["kept_indices"] == keep
    assert json.loads((out / "subset_manifest.json").read_text())["original_unclu_T"] == 6
    W, W2 = np.load(d / "Wall3.npy"), np.load(out / "Wall3.npy")
    assert np.array_equal(W2, W[keep])
    ctc, ctc2 = np.load(d / "ctc.npy"), np.load(out / "ctc.npy")
    assert np.array_equal(ctc2, ctc[np.ix_(keep, keep)])
    assert np.array_equal(np.load(out / "iU.npy"), np.load(d / "iU.npy")[keep])
    assert np.array_equal(np.load(out / "Ucc.npy"), np.load(d / "Ucc.npy")[:, keep])
    # final-cluster files are untouched
    assert np.load(out / "templates.npy").shape[0] == 2
    with pytest.raises(ValueError, match="final-cluster"):
        subset.write_subset(d, tmp_path / "sub2", [0])
