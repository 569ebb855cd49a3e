"""The allocation-free entry point (lss_process_batch_device_static) against
the entry point it varies, lss_process_batch_device: events and features
bit-identical on consecutive calls (the second with a smaller batch), with a
call of another handle (another probe, another layout of the static array)
in between, for the FP32 and the BF16 matcher; and the rejection of a handle
whose batches do not fit the static array."""
import numpy as np
import pytest

from livespikesorter_amd import synth
from livespikesorter_amd._capi import STATIC_SCRATCH_BYTES
from livespikesorter_amd._hip import DeviceBuffer
from livespikesorter_amd.matcher import LssError, OnlineSpikesV2, PREC_BF16_RESCORE, PREC_FP32

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [PREC_FP32, PREC_BF16_RESCORE], ids=["fp32", "bf16"])
def test_static_entry_matches_device_entry(model_dirs, prec):
    m, d = model_dirs("cfg2")                        # 384 channels, 512 templates
    m1, d1 = model_dirs("cfg1")                      # 64 channels, 32 templates
    x = synth.make_recording(m, 2 * 3000 + 122, seed=synth.TEST_SEED + 31)[0]
    x1 = synth.make_recording(m1, 3122, seed=synth.TEST_SEED + 32)[0]
    C = x.shape[1]
    dx, dx1 = DeviceBuffer.from_array(x), DeviceBuffer.from_array(x1)
    try:
        with OnlineSpikesV2(d, max_batch=3122, precision=prec) as ref, \
                OnlineSpikesV2(d, max_batch=3122, precision=prec) as st, \
                OnlineSpikesV2(d1, max_batch=3122) as other:
            calls = [(0, 3122), (3000, 2000)]          # the second call is smaller
            n_static = 0
            for k, (start, W) in enumerate(calls):
                print(f"[no-alloc prec={prec}] call {k}: columns {W}", flush=True)
                ptr = dx.ptr + start * C * 2
                e1, f1 = ref.process_batch_device(ptr, W, C, first_sample=start)
                e2, f2 = st.process_batch_device_static(ptr, W, C, first_sample=start)
                assert len(e1) > 20
                assert np.array_equal(e1, e2), f"call {k}: events differ"
                assert np.array_equal(f1, f2), f"call {k}: features differ"
                n_static += len(e2)
                # another handle through the same static array, then the first again
                o1, g1 = other.process_batch_device(dx1.ptr, 3122, x1.shape[1])
                o2, g2 = other.process_batch_device_static(dx1.ptr, 3122, x1.shape[1])
                assert len(o1) > 0 and np.array_equal(o1, o2) and np.array_equal(g1, g2)
            e3, f3 = st.process_batch_device_static(dx.ptr, 3122, C)
            e4, f4 = ref.process_batch_device(dx.ptr, 3122, C)
            assert np.array_equal(e3, e4) and np.array_equal(f3, f4)
            assert st.query().events == n_static + len(e3)
    finally:
        dx.free()
        dx1.free()


def test_static_entry_rejects_oversized_handles(model_dirs):
    m, d = model_dirs("cfg3")                        # 384 channels, 1024 templates
    need = 4 * 384 * 7 * 60122 + 4 * 1024 * 60122   # whitened batch + PC projection + scores alone
    assert need > STATIC_SCRATCH_BYTES
    x = np.zeros((64, 384), np.int16)
    dx = DeviceBuffer.from_array(x)
    try:
        with OnlineSpikesV2(d, max_batch=60122) as sp:
            with pytest.raises(LssError) as ei:
                sp.process_batch_device_static(dx.ptr, 64, 384)
            assert ei.value.code == -1 and "static array" in str(ei.value)
    finally:
        dx.free()
