"""MI355X-native online template matcher (Kilosort4-style), drop-in for the
matcher path of LiveSpikeSorter's OnlineSpikesV2.

The compute path is ``liblss.so`` (hand-written HIP kernels for gfx950 behind
the C ABI in include/lss.h); this package holds its ctypes binding, the
Python mirror of the sorter interface and the synthetic data generator.
"""
from ._capi import LssError, SEM_KILOSORT, SEM_LSS, PREC_FP32, PREC_BF16_RESCORE, PREC_BF16X3, EVENT_DTYPE  # noqa: F401
from ._capi import PRE_MEAN, PRE_CAR, PRE_HIGHPASS, PRE_ALL  # noqa: F401
from .matcher import OnlineSpikesV2, StreamGroup  # noqa: F401
