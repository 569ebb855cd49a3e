"""Choosing which preclustered templates a matcher instance loads.

The matcher's cost grows with the number of preclustered templates (the rows
of ``Wall3.npy``), so a probe can be run on fewer of them.  The selection
rule is the reference's template-subsetting tool
(src/Python/subset_templates.py:87-137, 163-256), and the files it reduces
are the ones its directory rewrite touches (:291-347):

1. *Channel filter* (optional).  A probe-channel interval ``[lo, hi]`` is
   turned into positions on the templates' channel axis through the
   directory's channel map; a template is a candidate when its dominant
   channel lies there (``"peak"``) or when any channel carrying at least a
   tenth of its largest support does (``"overlap"``).  A channel's support is
   the peak-to-peak amplitude of the preclustered waveform, or the L2 norm
   over PCs of ``Wall3`` when the waveforms are not in the directory.
2. *Activity cap* (optional).  Of the candidates, the ``n`` most active are
   kept: activity is the number of spikes Kilosort detected with each
   preclustered template (``spike_detection_templates.npy``, written by
   KS/io.py:251 and copied into ``oss_input`` by run_online_spikes.py:647),
   or, when that file is absent, the L2 norm of the template's ``Wall3`` row.
   Equal activity is broken toward the lower template index (the reference
   leaves ties to numpy's unstable ``argsort``).

The kept indices either go to ``OnlineSpikesV2(..., keep=...)``
(``lss_create_subset``: the tensors are reduced at load time) or to
:func:`write_subset`, which writes a reduced copy of the directory.  Every
problem is reported as :class:`ValueError`.
"""
from __future__ import annotations

import dataclasses
import json
import shutil
from pathlib import Path

import numpy as np

#: preclustered spike ids used for the activity ranking
DETECTION_IDS = "spike_detection_templates.npy"

#: (file, axes whose length is the preclustered template count); every other
#: file of the directory describes final clusters, channels or PCs and is
#: copied unchanged
PRECLUSTER_TENSORS = (
    ("Wall3.npy", (0,)),                                   # [unclu_T][K][C]
    ("ctc.npy", (0, 1)),                                   # [unclu_T][unclu_T][2M+1]
    ("ctc_permuted.npy", (0, 1)),
    ("iU.npy", (0,)),                                      # [unclu_T]
    ("Ucc.npy", (1,)),                                     # [nC][unclu_T][K]
    ("preclustered_template_waveforms.npy", (0,)),         # [unclu_T][M][C]
)


@dataclasses.dataclass(frozen=True)
class ProbeDims:
    """Sizes of an ``oss_input`` directory.  ``n_pre`` preclustered templates
    (Wall3 rows), ``n_final`` final clusters (templates.npy rows; equal to
    ``n_pre`` when templates.npy is absent)."""
    n_pre: int
    n_final: int
    K: int
    C: int
    M: int


@dataclasses.dataclass
class Selection:
    keep: list                 # ascending preclustered indices
    ranked_by: str             # "spike count", "template L2 norm", or why nothing was ranked
    channel_filter: dict | None = None


def probe_dims(oss_dir) -> ProbeDims:
    d = Path(oss_dir)
    wall3 = np.load(d / "Wall3.npy", mmap_mode="r")
    if wall3.ndim != 3:
        raise ValueError(f"Wall3.npy has shape {wall3.shape}; a [templates][PCs][channels] array is needed")
    n_pre, K, C = (int(v) for v in wall3.shape)
    tpath = d / "templates.npy"
    if tpath.exists():
        tmpl = np.load(tpath, mmap_mode="r")
        if tmpl.ndim != 3 or tmpl.shape[2] != C:
            raise ValueError(f"templates.npy {tuple(tmpl.shape)} and Wall3.npy {tuple(wall3.shape)} "
                             f"disagree on the channel count")
        n_final, M = int(tmpl.shape[0]), int(tmpl.shape[1])
    else:
        n_final, M = n_pre, int(np.load(d / "wPCA.npy", mmap_mode="r").shape[1])
    return ProbeDims(n_pre=n_pre, n_final=n_final, K=K, C=C, M=M)


def parse_channel_range(text):
    """``"a-b"`` (inclusive probe channel numbers, either order) -> ``(lo, hi)``;
    ``None`` or blank text -> ``None``."""
    if text is None:
        return None
    text = str(text).strip()
    if not text:
        return None
    lo_txt, sep, hi_txt = text.partition("-")
    if not sep:
        raise ValueError(f"cannot read {text!r} as a probe-channel interval; write it as two numbers "
                         f"joined by '-', such as 96-191")
    a, b = int(lo_txt), int(hi_txt)
    return min(a, b), max(a, b)


def _channel_map(d: Path) -> np.ndarray:
    for name in ("channelMap.npy", "channel_map.npy"):
        if (d / name).exists():
            return np.load(d / name, allow_pickle=False).reshape(-1).astype(np.int64)
    raise ValueError(f"{d} holds no channelMap.npy or channel_map.npy, so probe channel numbers cannot "
                     f"be placed on the templates' channel axis")


def channel_positions(oss_dir, lo: int, hi: int) -> np.ndarray:
    """Positions j on the templates' channel axis whose probe channel
    (channel map entry j) lies in ``[lo, hi]``."""
    cmap = _channel_map(Path(oss_dir))
    return np.flatnonzero((lo <= cmap) & (cmap <= hi))


def channel_support(oss_dir) -> np.ndarray:
    """Support of each preclustered template on each channel, ``[n_pre][C]``."""
    d = Path(oss_dir)
    wpath = d / "preclustered_template_waveforms.npy"
    if wpath.exists():
        wave = np.load(wpath, allow_pickle=False)                  # [n_pre][M][C]
        return np.ptp(wave, axis=1)
    U = np.load(d / "Wall3.npy", allow_pickle=False).astype(np.float64)   # [n_pre][K][C]
    return np.sqrt((U * U).sum(axis=1))


def templates_on_channels(oss_dir, positions, rule: str = "peak", frac: float = 0.1) -> list:
    """Preclustered templates whose support reaches the channel positions given."""
    sup = channel_support(oss_dir)
    inside = np.isin(np.arange(sup.shape[1]), np.asarray(positions, np.int64))
    if rule == "peak":
        hit = inside[np.argmax(sup, axis=1)]
    elif rule == "overlap":
        strong = sup >= frac * sup.max(axis=1)[:, None]
        hit = (strong & inside).any(axis=1)
    else:
        raise ValueError(f"channel rule {rule!r} is neither 'peak' nor 'overlap'")
    return np.flatnonzero(hit).tolist()


def activity(oss_dir, dims: ProbeDims, metric: str = "spikes"):
    """(score per preclustered template, label): detection counts, or the
    ``Wall3`` row norm when asked for or when the detection ids are absent."""
    d = Path(oss_dir)
    if metric not in ("spikes", "norm"):
        raise ValueError(f"activity metric {metric!r} is neither 'spikes' nor 'norm'")
    ids_path = d / DETECTION_IDS
    if metric == "spikes" and ids_path.exists():
        ids = np.load(ids_path, allow_pickle=False).reshape(-1).astype(np.int64)
        # preclustered ids run up to n_pre - 1 >= n_final; a largest id below the final
        # cluster count means the file holds final-cluster ids
        if ids.size and ids.max() < dims.n_final:
            raise ValueError(f"{DETECTION_IDS}: largest id {int(ids.max())} is below the {dims.n_final} final "
                             f"clusters, so these are final-cluster ids; preclustered ids are needed")
        ids = ids[(ids >= 0) & (ids < dims.n_pre)]
        return np.bincount(ids, minlength=dims.n_pre).astype(np.float64), "spike count"
    rows = np.load(d / "Wall3.npy", allow_pickle=False).astype(np.float64).reshape(dims.n_pre, -1)
    return np.sqrt((rows * rows).sum(axis=1)), "template L2 norm"


def select_templates(oss_dir, n: int = 0, metric: str = "spikes", channel_range=None,
                     channel_rule: str = "peak") -> Selection:
    """Channel filter first, then the activity cap of ``n`` (``n <= 0``: no cap)."""
    dims = probe_dims(oss_dir)
    pool = np.arange(dims.n_pre)
    info = None
    if channel_range:
        lo, hi = channel_range
        pos = channel_positions(oss_dir, lo, hi)
        if pos.size == 0:
            raise ValueError(f"the channel map has no probe channel in [{lo}, {hi}]")
        pool = np.asarray(templates_on_channels(oss_dir, pos, channel_rule), np.int64)
        if pool.size == 0:
            raise ValueError(f"no preclustered template has its support on probe channels [{lo}, {hi}]")
        info = {"channel_range": [int(lo), int(hi)], "channel_criterion": channel_rule,
                "probe_channels_in_range": int(pos.size), "templates_on_channels": int(pool.size)}
    if n is None or n <= 0 or n >= pool.size:
        why = "channel range" if pool.size < dims.n_pre else "all"
        return Selection(pool.tolist(), why, info)
    score, label = activity(oss_dir, dims, metric)
    # most active first; equal scores: lower index first
    order = np.lexsort((pool, -score[pool]))
    return Selection(np.sort(pool[order[:n]]).tolist(), label, info)


def write_subset(src, dst, keep, *, allow_below_t: bool = False) -> dict:
    """Copy directory ``src`` to ``dst`` and cut every preclustered tensor down
    to the templates ``keep``; the final-cluster files stay as they are.  A
    ``subset_manifest.json`` records the selection."""
    src, dst = Path(src), Path(dst)
    dims = probe_dims(src)
    idx = np.unique(np.asarray(list(keep), np.int64))
    if idx.size and (idx[0] < 0 or idx[-1] >= dims.n_pre):
        raise ValueError(f"template indices must lie in [0, {dims.n_pre})")
    if idx.size < dims.n_final and not allow_below_t:
        raise ValueError(f"keeping {idx.size} preclustered templates leaves fewer than the {dims.n_final} "
                         f"final clusters; pass allow_below_t=True to accept that (relabelling here scans "
                         f"every centroid, so it stays correct)")
    shutil.copytree(src, dst)
    for name, axes in PRECLUSTER_TENSORS:
        path = dst / name
        if not path.exists():
            continue
        arr = np.load(path, allow_pickle=False)
        for ax in axes:
            if arr.shape[ax] != dims.n_pre:
                raise ValueError(f"{name}: axis {ax} has length {arr.shape[ax]}, not the {dims.n_pre} "
                                 f"preclustered templates")
            arr = arr.take(idx, axis=ax)
        np.save(path, arr)
    manifest = {"original_unclu_T": dims.n_pre, "new_unclu_T": int(idx.size), "T_unchanged": dims.n_final,
                "kept_indices": idx.tolist()}
    (dst / "subset_manifest.json").write_text(json.dumps(manifest, indent=2))
    return manifest
