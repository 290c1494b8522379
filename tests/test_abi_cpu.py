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
    assert L.lib.ltpg_abi_version() == 2
    cfg = L.default_config()
    assert cfg.size == ctypes.sizeof(L.Config)
    assert cfg.kv_hot_row == L.NO_HOT_ROW
    assert cfg.values == L.VALUES_RMW and cfg.rules == L.RULES_REFERENCE


# ctypes mirror -> C typedef of include/ltpg.h
C_STRUCTS = {"Config": "ltpg_config", "KvBatch": "ltpg_kv_batch", "KvGen": "ltpg_kv_gen",
             "NeworderBatch": "ltpg_neworder_batch", "PaymentBatch": "ltpg_payment_batch",
             "OrderstatusBatch": "ltpg_orderstatus_batch", "StocklevelBatch": "ltpg_stocklevel_batch",
             "DeliveryBatch": "ltpg_delivery_batch",
             "TpccGenConfig": "ltpg_tpcc_gen_config", "BatchDesc": "ltpg_batch_desc",
             "TuningEntry": "ltpg_tuning_entry", "ServeConfig": "ltpg_serve_config", "ServeDone": "ltpg_serve_done",
             "ServeStats": "ltpg_serve_stats", "ServeDeviceStats": "ltpg_serve_device_stats"}


def test_struct_layouts_match_a_compiled_c_caller(tmp_path):
    """sizeof and offsetof of every field of every ABI struct, as gcc lays
    them out for a C caller that includes include/ltpg.h, equal those of the
    Python binding's ctypes mirrors (a field missing in C fails to compile)."""
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "ltpg.h"', 'int main(void) {']
    for py, c in C_STRUCTS.items():
        lines.append(f'printf("{py} sizeof() %zu\\n", sizeof({c}));')
        for name, *_ in getattr(_lib, py)._fields_:
            lines.append(f'printf("{py} {name} %zu\\n", offsetof({c}, {name}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {(a, b): int(v) for a, b, v in (ln.split() for ln in out if ln)}
    for py in C_STRUCTS:
        cls = getattr(_lib, py)
        assert got[(py, "sizeof()")] == ctypes.sizeof(cls), py
        for name, *_ in cls._fields_:
            assert got[(py, name)] == getattr(cls, name).offset, f"{py}.{name}"


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
    assert cfg.size == ctypes.sizeof(L.ServeConfig)
    assert (cfg.slots, cfg.kv_ops, cfg.epoch_growth, cfg.epoch_floor, cfg.plain, cfg.kv_loop, cfg.ring_requests,
            cfg.first_seq) == (2, 16, 8, 32, 0, 1, 0, 0)


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
