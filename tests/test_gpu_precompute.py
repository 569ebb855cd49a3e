"""Load-time precompute from the raw Kilosort outputs (run_online_spikes.py:609-667;
KS/template_matching.py:20-26,110-121): an oss_input directory reduced to
{misc.txt, whiteningMat, wPCA, Wall3, xc, yc} loads, its ctc table is built on
the GPU, and the matcher gives the same spikes as with the launcher's files.

The ctc table is checked against synth.ctc_table (FP64 numpy, pinned in
test_oracle by its closed form, symmetry and the waveform-subtraction
identity): both accumulate in FP64 and round once, so entries agree to one
FP32 unit in the last place.
"""
import shutil

import numpy as np
import pytest

from livespikesorter_amd import synth
from livespikesorter_amd.matcher import OnlineSpikesV2
from tests.parity import agreement

pytestmark = pytest.mark.gpu

DERIVED = ("ctc.npy", "templates.npy", "iCC.npy", "iU.npy", "Ucc.npy", "preclustered_template_waveforms.npy")


def reduced_copy(src, dst, drop=DERIVED):
    shutil.copytree(src, dst)
    for name in drop:
        (dst / name).unlink(missing_ok=True)
    return dst


def ulp_err(got, ref):
    ref = ref.astype(np.float32)
    return np.abs(got.astype(np.float64) - ref) / np.spacing(np.abs(ref).astype(np.float32))


@pytest.mark.parametrize("K,M,T", [(6, 61, 64), (3, 41, 40), (8, 63, 24)])
def test_gpu_ctc_matches_fp64_table(tmp_path, K, M, T):
    m = synth.make_model(C=48, T=T, K=K, M=M, whitening="banded")
    m.nC = min(m.nC, 64 // K)          # the feature vector holds at most 64 values
    full = synth.write_oss_input(m, tmp_path / "full")
    red = reduced_copy(full, tmp_path / "red")
    want = synth.ctc_table(m.U, m.wPCA)
    print(f"[ctc K={K} M={M} T={T}] loading reduced directory", flush=True)
    with OnlineSpikesV2(red) as sp:
        got = sp.ctc()
    assert got.shape == want.shape
    err = ulp_err(got, want)
    print(f"[ctc K={K} M={M} T={T}] max ulp {err.max():.2f}, exact {np.mean(got == want):.6f}")
    assert err.max() <= 1.0
    assert np.mean(got == want) >= 0.999
    # exact zeros where the templates share no channel
    assert np.array_equal(got == 0, want == 0)
    with OnlineSpikesV2(full) as sp:
        assert np.array_equal(sp.ctc(), np.load(full / "ctc.npy"))


@pytest.mark.parametrize("sparse", [True, False])
def test_events_identical_without_launcher_tables(model_dirs, tmp_path, sparse):
    m, full = model_dirs("cfg2")
    red = reduced_copy(full, tmp_path / "red")
    x, _ = synth.make_recording(m, 3000 * 4 + 122, seed=synth.TEST_SEED)
    out, tabs = [], []
    for d in (full, red):
        print(f"[reduced sparse={sparse}] {d.name}", flush=True)
        with OnlineSpikesV2(d, sparse=sparse, max_batch=len(x)) as sp:
            out.append(sp.process_batch(x))
            tabs.append(sp.ctc())
            q = sp.info
    (e0, f0), (e1, f1) = out
    same = np.array_equal(tabs[0], tabs[1])
    print(f"[reduced sparse={sparse}] {len(e0)} vs {len(e1)} events, ctc tables identical: {same}")
    assert len(e0) > 100
    assert q.M == m.M and q.C == m.C and q.T == m.T and q.K == m.K
    if same:
        assert np.array_equal(e0, e1)
        assert np.array_equal(f0, f1)
    else:
        # a 1-ulp ctc entry moves later residues by rounding-level amounts only
        assert ulp_err(tabs[1], tabs[0]).max() <= 1.0
        ident, within, _ = agreement(e0["sample"], e0["template"], e1["sample"], e1["template"], tol=1)
        assert ident >= 0.999 and within >= 0.9999


def test_load_time_subset_matches_subset_directory(model_dirs, tmp_path):
    """A template subset applied at load (lss_create_subset) gives the events of
    a directory subset on disk the way subset_oss_input_inplace writes it."""
    from livespikesorter_amd import subset
    from livespikesorter_amd.matcher import LssError
    m, full = model_dirs("cfg2")
    sel = subset.select_templates(full, 200, "norm")
    keep = sel.keep
    assert len(keep) == 200 and sel.ranked_by == "template L2 norm"
    sub = tmp_path / "sub"
    subset.write_subset(full, sub, keep, allow_below_t=True)
    x, _ = synth.make_recording(m, 3000 * 4 + 122, seed=synth.TEST_SEED)
    print("[subset] load-time keep list", flush=True)
    with OnlineSpikesV2(full, keep=keep, max_batch=len(x)) as sp:
        e0, f0 = sp.process_batch(x)
        assert sp.info.T == 200
        assert np.array_equal(sp.ctc(), np.load(full / "ctc.npy")[np.ix_(keep, keep)])
    print("[subset] subset directory", flush=True)
    with OnlineSpikesV2(sub, max_batch=len(x)) as sp:
        e1, f1 = sp.process_batch(x)
    print(f"[subset] {len(e0)} vs {len(e1)} events")
    assert len(e0) > 100
    assert np.array_equal(e0, e1) and np.array_equal(f0, f1)
    with pytest.raises(LssError):
        OnlineSpikesV2(full, keep=[5, 3])
    with pytest.raises(LssError):
        OnlineSpikesV2(full, keep=[0, m.T])
