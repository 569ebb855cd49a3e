"""C-ABI library checks that need no GPU: it loads, exports every declared
symbol, and its host-only entry points behave."""
import ctypes as ct
import re
from pathlib import Path

import pytest

from livespikesorter_amd import _capi

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "lss.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|int32_t|const char\*)\s+(lss_\w+)\s*\(", text, flags=re.M)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(_capi.ABI_FUNCTIONS)


def test_library_exports_every_declared_symbol():
    lib = _capi.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_config_defaults():
    cfg = _capi.default_config()
    assert cfg.chunk == 3000 and cfg.max_events == 65536 and cfg.sparse == 2
    assert cfg.semantics == _capi.SEM_KILOSORT and cfg.dedup == -1 and cfg.profile == 0
    assert ct.sizeof(_capi.Config) == 16 * 4


def test_null_arguments_are_errors_not_crashes():
    lib = _capi.lib()
    assert lib.lss_config_init(None) == _capi.LSS_ERR_ARG
    h = ct.c_void_p()
    assert lib.lss_create(None, None, ct.byref(h)) == _capi.LSS_ERR_ARG
    assert b"null" in lib.lss_last_error()
    n = ct.c_int64(5)
    assert lib.lss_process_batch(None, None, 0, 0, None, 0, None, 0, ct.byref(n), None) == _capi.LSS_ERR_ARG
    assert lib.lss_destroy(None) == 0


def test_event_layout():
    assert ct.sizeof(_capi.Event) == 16
    assert _capi.Event.template_id.offset == 8 and _capi.Event.amplitude.offset == 12
