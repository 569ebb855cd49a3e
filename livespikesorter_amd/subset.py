"""Template subsetting: pick which preclustered templates a matcher loads.

Mirrors src/Python/subset_templates.py: a channel filter narrows the
candidates first (:224-256, :183-221), then an activity cap keeps the ``n``
most active survivors (:87-137), ranked by spike count from
``spike_templates.npy`` or, without that file, by the L2 norm of each
template's Wall3 rows.  The kept indices go to
``OnlineSpikesV2(..., keep=...)`` (``lss_create_subset``), which subsets
Wall3, ctc and iU at load time; :func:`write_subset` instead writes a reduced
copy of the directory, subsetting the same tensors on disk as
``subset_oss_input_inplace`` (:291-347) does.  Errors that the reference
reports by ``sys.exit`` or ``RuntimeError`` raise :class:`ValueError` here.
"""
from __future__ import annotations

import json
import shutil
from pathlib import Path

import numpy as np

RANK_SOURCE = "spike_templates.npy"

# preclustered-template axes of each file (subset_templates.py:47-53)
UNCLU_T_AXES = {
    "Wall3.npy": (0,),
    "ctc.npy": (0, 1),
    "ctc_permuted.npy": (0, 1),
    "iU.npy": (0,),
    "Ucc.npy": (1,),
    "preclustered_template_waveforms.npy": (0,),
}


def shapes(oss_dir) -> dict:
    """T (final clusters), M, C from templates.npy and unclu_T, K from Wall3.npy (:69-84)."""
    d = Path(oss_dir)
    w = np.load(d / "Wall3.npy", mmap_mode="r")
    if w.ndim != 3:
        raise ValueError(f"Wall3.npy expected [unclu_T, K, C], got {w.shape}")
    unclu_T, K, C = w.shape
    if (d / "templates.npy").exists():
        t = np.load(d / "templates.npy", mmap_mode="r")
        if t.ndim != 3 or t.shape[2] != C:
            raise ValueError(f"templates.npy {t.shape} does not match Wall3 {w.shape}")
        T, M = t.shape[0], t.shape[1]
    else:
        T, M = unclu_T, np.load(d / "wPCA.npy", mmap_mode="r").shape[1]
    return dict(T=int(T), M=int(M), C=int(C), unclu_T=int(unclu_T), K=int(K))


def parse_channel_range(s):
    """``'lo-hi'`` (inclusive probe channel numbers) -> (lo, hi), None when blank (:140-160)."""
    if s is None or not str(s).strip():
        return None
    s = str(s).strip()
    if "-" not in s:
        raise ValueError(f"channel range '{s}' must look like 'lo-hi', e.g. '100-150'")
    a, b = s.split("-", 1)
    lo, hi = int(a.strip()), int(b.strip())
    return (lo, hi) if lo <= hi else (hi, lo)


def channels_in_range(oss_dir, c_lo: int, c_hi: int) -> np.ndarray:
    """C-axis positions j whose probe channel channelMap[j] lies in [c_lo, c_hi] (:163-180)."""
    d = Path(oss_dir)
    p = d / "channelMap.npy"
    if not p.exists():
        p = d / "channel_map.npy"
    if not p.exists():
        raise ValueError("need channelMap.npy (or channel_map.npy) to map probe channels onto the template axis")
    cmap = np.load(p).ravel().astype(np.int64)
    return np.where((cmap >= c_lo) & (cmap <= c_hi))[0]


def templates_on_channels(oss_dir, allowed_chan_idx, criterion: str = "peak", overlap_frac: float = 0.1) -> list:
    """Preclustered templates with support on the allowed C-axis positions (:183-221).

    Support per channel: peak-to-peak of preclustered_template_waveforms.npy
    [unclu_T, M, C], else the L2 norm over PCs of Wall3 [unclu_T, K, C].
    ``peak``: the largest-support channel is allowed; ``overlap``: some
    channel with support >= overlap_frac x the template's peak is allowed.
    """
    d = Path(oss_dir)
    pre = d / "preclustered_template_waveforms.npy"
    if pre.exists():
        w = np.load(pre)
        amp = w.max(axis=1) - w.min(axis=1)
    else:
        amp = np.linalg.norm(np.load(d / "Wall3.npy"), axis=1)
    allowed = np.zeros(amp.shape[1], dtype=bool)
    allowed[np.asarray(allowed_chan_idx, dtype=np.int64)] = True
    if criterion == "peak":
        keep = np.where(allowed[amp.argmax(axis=1)])[0]
    elif criterion == "overlap":
        on = amp >= overlap_frac * amp.max(axis=1, keepdims=True)
        keep = np.where((on & allowed[None, :]).any(axis=1))[0]
    else:
        raise ValueError(f"unknown channel criterion '{criterion}' (peak|overlap)")
    return [int(t) for t in keep]


def choose_keep(oss_dir, dims: dict, n: int, metric: str = "spikes", candidates=None):
    """Sorted kept indices and the ranking used (:87-137).

    Without a cap (``n <= 0``) or with ``n`` at least the number of
    candidates, every candidate is kept.  Otherwise the candidates are walked
    in decreasing activity and the first ``n`` kept.
    """
    d = Path(oss_dir)
    unclu_T = dims["unclu_T"]
    cand = sorted({int(c) for c in (range(unclu_T) if candidates is None else candidates)})
    if not n or n <= 0 or n >= len(cand):
        return cand, ("all" if len(cand) == unclu_T else "channel range")
    order = None
    if metric == "spikes":
        p = d / RANK_SOURCE
        if p.exists():
            sp = np.load(p).astype(np.int64).ravel()
            if sp.size and int(sp.max()) < dims["T"]:
                raise ValueError(f"{RANK_SOURCE} max index {int(sp.max())} < T ({dims['T']}): it looks like "
                                 f"final-cluster space, not preclustered")
            sp = sp[(sp >= 0) & (sp < unclu_T)]
            order = np.argsort(np.bincount(sp, minlength=unclu_T))[::-1]
            ranked = "spike count"
        else:
            metric = "norm"
    if metric == "norm":
        energy = np.linalg.norm(np.load(d / "Wall3.npy").reshape(unclu_T, -1), axis=1)
        order = np.argsort(energy)[::-1]
        ranked = "template L2 norm"
    if order is None:
        raise ValueError(f"unknown metric '{metric}' (spikes|norm)")
    cs = set(cand)
    return sorted([int(i) for i in order if int(i) in cs][:n]), ranked


def compute_keep(oss_dir, n: int = 0, metric: str = "spikes", channel_range=None,
                 channel_criterion: str = "peak"):
    """(keep, ranked, channel metadata): channel filter first, then the cap (:224-256)."""
    dims = shapes(oss_dir)
    candidates, meta = None, None
    if channel_range:
        c_lo, c_hi = channel_range
        idx = channels_in_range(oss_dir, c_lo, c_hi)
        if idx.size == 0:
            raise ValueError(f"no probe channels fall in range [{c_lo}, {c_hi}]")
        candidates = templates_on_channels(oss_dir, idx, channel_criterion)
        if not candidates:
            raise ValueError(f"no preclustered templates have support on channels [{c_lo}, {c_hi}]")
        meta = {"channel_range": [int(c_lo), int(c_hi)], "channel_criterion": channel_criterion,
                "probe_channels_in_range": int(idx.size), "templates_on_channels": len(candidates)}
    keep, ranked = choose_keep(oss_dir, dims, n, metric, candidates)
    return keep, ranked, meta


def subset_axes(arr: np.ndarray, axes, keep, unclu_T: int, name: str) -> np.ndarray:
    """Take ``keep`` along each listed axis, checking its length is unclu_T (:278-288)."""
    for ax in axes:
        ra = ax if ax >= 0 else arr.ndim + ax
        if arr.shape[ra] != unclu_T:
            raise ValueError(f"{name}: axis {ax} length {arr.shape[ra]} != unclu_T {unclu_T}")
        arr = np.take(arr, np.asarray(keep, np.int64), axis=ra)
    return arr


def write_subset(src, dst, keep, *, allow_below_t: bool = False) -> dict:
    """Copy ``src`` to ``dst`` with the preclustered tensors subset to ``keep``
    (subset_oss_input_inplace, :291-347, on a copy); writes subset_manifest.json."""
    src, dst = Path(src), Path(dst)
    dims = shapes(src)
    keep = sorted(int(k) for k in keep)
    if len(keep) < dims["T"] and not allow_below_t:
        raise ValueError(f"subset of {len(keep)} templates is below the final-cluster count {dims['T']}; "
                         f"pass allow_below_t=True (this matcher's relabelling scans every centroid)")
    shutil.copytree(src, dst)
    for name, axes in UNCLU_T_AXES.items():
        p = dst / name
        if p.exists():
            np.save(p, subset_axes(np.load(p, allow_pickle=False), axes, keep, dims["unclu_T"], name))
    manifest = {"original_unclu_T": dims["unclu_T"], "new_unclu_T": len(keep), "T_unchanged": dims["T"],
                "kept_indices": keep}
    (dst / "subset_manifest.json").write_text(json.dumps(manifest, indent=2))
    return manifest
