"""CPU-only checks of the C ABI: the library loads, exports every entry point
declared in include/ltpg.h, and reports errors without a GPU."""
import ctypes
import os
import re
import subprocess

import pytest

import ltpg_amd as L
from ltpg_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(ltpg_[a-z0-9_]+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported():
    declared = _declared()
    assert len(declared) > 20
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\b(ltpg_[a-z0-9_]+)\b", out))
    missing = declared - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"


def test_ctypes_signatures_cover_the_header():
    declared = _declared()
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert declared <= bound, f"not bound in _lib.SIGNATURES: {sorted(declared - bound)}"


def test_abi_version_and_config_defaults():
    assert L.lib.ltpg_abi_version() == 1
    cfg = L.default_config()
    assert cfg.size == ctypes.sizeof(L.Config)
    assert cfg.kv_hot_row == L.NO_HOT_ROW
    assert cfg.values == L.VALUES_RMW and cfg.rules == L.RULES_REFERENCE


def test_config_struct_layout_matches_c():
    # offsets of the C struct (include/ltpg.h) on x86-64
    assert ctypes.sizeof(L.Config) == 72
    assert L.Config.kv_rows.offset == 24
    assert L.Config.kv_hot_row.offset == 48
    assert L.Config.hbm_budget.offset == 64
    assert ctypes.sizeof(L.KvBatch) == 32


def test_python_constants_match_the_header():
    """Every integer #define LTPG_X of include/ltpg.h that the Python binding
    mirrors as X has the header's value."""
    text = open(os.path.join(ROOT, "include", "ltpg.h")).read()
    defines = dict(re.findall(r"#define LTPG_([A-Z0-9_]+) (0x[0-9a-fA-F]+|\d+)u?\b", text))
    checked = set()
    for name, value in defines.items():
        if hasattr(L, name):
            assert getattr(L, name) == int(value, 0), name
            checked.add(name)
    assert {"TUNE_ALL", "TUNE_SMALL_EPOCH", "SMALL_EPOCH_OPS", "BATCH_REQUEST_MAJOR"} <= checked
    assert len(checked) >= 20


def test_serve_config_defaults_match_c():
    cfg = L.ServeConfig()
    L.lib.ltpg_serve_config_init(ctypes.byref(cfg))
    assert cfg.size == ctypes.sizeof(L.ServeConfig) == 48
    assert (cfg.slots, cfg.kv_ops, cfg.epoch_growth, cfg.epoch_floor, cfg.plain) == (2, 16, 8, 32, 0)


@pytest.mark.skipif(L.device_count() > 0, reason="only meaningful without a GPU")
def test_create_without_gpu_reports_device_error():
    with pytest.raises(L.LtpgError) as ei:
        L.Context(kv_rows=10, kv_fields=2, max_txns=16, kv_max_slots=4)
    assert ei.value.code == L.ERR_DEVICE
    assert "device" in str(ei.value).lower()


def test_create_rejects_bad_config_before_touching_the_device():
    cfg = L.default_config()
    cfg.size = 3
    h = ctypes.c_void_p()
    assert L.lib.ltpg_create(ctypes.byref(cfg), ctypes.byref(h)) == L.ERR_INVALID
    cfg = L.default_config(values=9)
    assert L.lib.ltpg_create(ctypes.byref(cfg), ctypes.byref(h)) == L.ERR_INVALID
    assert L.lib.ltpg_create(None, ctypes.byref(h)) == L.ERR_INVALID


def test_null_context_is_rejected():
    assert L.lib.ltpg_epoch_run(None, L.PHASE_ALL) == L.ERR_INVALID
    assert L.lib.ltpg_sync(None) == L.ERR_INVALID
    assert L.lib.ltpg_destroy(None) == L.OK
