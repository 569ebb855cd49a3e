"""Template-subset selection and the reduced-directory writer (host logic).

Selection rule: src/Python/subset_templates.py:87-137,163-256 (channel filter,
then the activity cap ranked by spike_detection_templates.npy, norm fallback).
The fixture lays out the ranking files as the launcher writes them
(src/Python/run_online_spikes.py:646-647): ``spike_templates.npy`` holds
Kilosort's final-cluster ids (KS/io.py:248) and
``spike_detection_templates.npy`` the preclustered template of each spike
(KS/io.py:251).  Expected selections are worked out by hand."""
import json

import numpy as np
import pytest

from livespikesorter_amd import subset


def make_dir(tmp_path, *, unclu_T=6, T=None, K=2, C=8, M=5, peaks=None, detect=None, clusters=None,
             cmap=None, waves=False, name="oss"):
    """Wall3 template t: support on channel peaks[t] (weight 2 (t+1)) and its right
    neighbour (weight t+1)."""
    d = tmp_path / name
    d.mkdir()
    T = T or unclu_T
    peaks = list(range(unclu_T)) if peaks is None else peaks
    W = np.zeros((unclu_T, K, C), np.float32)
    for t, c in enumerate(peaks):
        W[t, :, c] = 2.0 * (1 + t)
        if c + 1 < C:
            W[t, :, c + 1] = 1.0 * (1 + t)
    np.save(d / "Wall3.npy", W)
    np.save(d / "templates.npy", np.zeros((T, M, C), np.float32))
    np.save(d / "wPCA.npy", np.zeros((K, M), np.float32))
    np.save(d / "iU.npy", np.asarray(peaks, np.int64))
    np.save(d / "Ucc.npy", np.arange(3 * unclu_T * K, dtype=np.float32).reshape(3, unclu_T, K))
    np.save(d / "ctc.npy", np.arange(unclu_T * unclu_T * (2 * M + 1), dtype=np.float32)
            .reshape(unclu_T, unclu_T, 2 * M + 1))
    np.save(d / "channelMap.npy", np.arange(C, dtype=np.int32) if cmap is None else np.asarray(cmap, np.int32))
    if detect is not None:
        detect = np.asarray(detect, np.int64)
        np.save(d / "spike_detection_templates.npy", detect)
        # final-cluster ids of the same spikes: always below T
        cl = np.asarray(clusters, np.int64) if clusters is not None else detect % T
        np.save(d / "spike_templates.npy", cl)
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


def test_probe_dims(tmp_path):
    d = make_dir(tmp_path, unclu_T=6, T=3, K=2, C=8, M=5)
    assert subset.probe_dims(d) == subset.ProbeDims(n_pre=6, n_final=3, K=2, C=8, M=5)
    (d / "templates.npy").unlink()
    assert subset.probe_dims(d) == subset.ProbeDims(n_pre=6, n_final=6, K=2, C=8, M=5)


def test_channel_positions_follow_channel_map(tmp_path):
    d = make_dir(tmp_path, cmap=[7, 6, 5, 4, 3, 2, 1, 0])
    assert list(subset.channel_positions(d, 1, 2)) == [5, 6]
    (d / "channelMap.npy").rename(d / "channel_map.npy")
    assert list(subset.channel_positions(d, 7, 9)) == [0]
    (d / "channel_map.npy").unlink()
    with pytest.raises(ValueError, match="channel"):
        subset.channel_positions(d, 0, 3)


def test_templates_on_channels_peak_and_overlap(tmp_path):
    d = make_dir(tmp_path, peaks=[0, 2, 4, 6, 1, 3])
    # Wall3 support: template t on peaks[t] and, at half of it, peaks[t] + 1
    assert subset.templates_on_channels(d, [2, 3], "peak") == [1, 5]
    assert subset.templates_on_channels(d, [2, 3], "overlap") == [1, 4, 5]
    with pytest.raises(ValueError):
        subset.templates_on_channels(d, [2], "nearest")


def test_templates_on_channels_prefers_waveforms(tmp_path):
    d = make_dir(tmp_path, peaks=[0, 2, 4, 6, 1, 3], waves=True)
    # waveform support is the peak channel only (leakage 5% < 10%)
    assert subset.templates_on_channels(d, [3], "overlap") == [5]
    assert subset.templates_on_channels(d, [3], "peak") == [5]


def test_ranking_uses_detection_ids_not_cluster_ids(tmp_path):
    """A realistic directory: 6 preclustered templates merged into 2 final
    clusters.  The cluster ids (spike_templates.npy, all < 2) must not be used;
    the detection ids rank templates 3 (5 spikes) and 1 (4 spikes) first."""
    detect = [3] * 5 + [1] * 4 + [5] * 2 + [0]
    clusters = [1] * 5 + [0] * 4 + [1] * 2 + [0]
    d = make_dir(tmp_path, unclu_T=6, T=2, detect=detect, clusters=clusters)
    sel = subset.select_templates(d, 2, "spikes")
    assert sel.keep == [1, 3] and sel.ranked_by == "spike count"
    assert subset.select_templates(d, 3, "spikes").keep == [1, 3, 5]


def test_ranking_drops_out_of_range_ids_and_breaks_ties_low(tmp_path):
    # ids -1 and 99 are ignored; templates 2 and 4 tie at 2 spikes: the lower index wins
    detect = [5] * 3 + [4] * 2 + [2] * 2 + [-1, 99, 99, 99]
    d = make_dir(tmp_path, unclu_T=6, T=3, detect=detect)
    assert subset.select_templates(d, 2, "spikes").keep == [2, 5]
    assert subset.select_templates(d, 3, "spikes").keep == [2, 4, 5]


def test_norm_ranking_and_fallback(tmp_path):
    d = make_dir(tmp_path, unclu_T=6, T=3, detect=[0] * 9 + [1, 3])
    # Wall3 rows scale with t + 1: the highest indices have the largest norm
    sel = subset.select_templates(d, 2, "norm")
    assert sel.keep == [4, 5] and sel.ranked_by == "template L2 norm"
    assert subset.select_templates(d, 2, "spikes").keep == [0, 1]
    (d / "spike_detection_templates.npy").unlink()
    sel = subset.select_templates(d, 2, "spikes")
    assert sel.keep == [4, 5] and sel.ranked_by == "template L2 norm"
    with pytest.raises(ValueError):
        subset.select_templates(d, 2, "loudness")


def test_no_cap_keeps_every_candidate(tmp_path):
    d = make_dir(tmp_path, unclu_T=6, T=3, peaks=[0, 2, 4, 6, 1, 3])
    assert subset.select_templates(d, 0) == subset.Selection([0, 1, 2, 3, 4, 5], "all", None)
    assert subset.select_templates(d, 6).keep == [0, 1, 2, 3, 4, 5]
    sel = subset.select_templates(d, 9, channel_range=(2, 3))
    assert sel.keep == [1, 5] and sel.ranked_by == "channel range"


def test_rejects_final_cluster_ids_in_detection_file(tmp_path):
    # a detection file whose largest id (3) is below the 4 final clusters
    d = make_dir(tmp_path, unclu_T=6, T=4, detect=[0, 1, 2, 3])
    with pytest.raises(ValueError, match="final-cluster"):
        subset.select_templates(d, 2, "spikes")


def test_channel_filter_then_cap(tmp_path):
    d = make_dir(tmp_path, T=2, peaks=[0, 2, 4, 6, 1, 3], detect=[1] * 3 + [5] * 9 + [4] * 7)
    sel = subset.select_templates(d, 2, "spikes", (1, 3), "overlap")
    # overlap candidates on channels 1..3: templates 0 (ch 0-1), 1 (2-3), 4 (1-2), 5 (3-4)
    assert sel.keep == [4, 5] and sel.ranked_by == "spike count"
    assert sel.channel_filter["templates_on_channels"] == 4
    assert sel.channel_filter["probe_channels_in_range"] == 3
    with pytest.raises(ValueError):
        subset.select_templates(d, 2, "spikes", (50, 60))


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
    with pytest.raises(ValueError, match="final clusters"):
        subset.write_subset(d, tmp_path / "sub2", [0])
    with pytest.raises(ValueError, match="indices"):
        subset.write_subset(d, tmp_path / "sub3", [0, 6])
