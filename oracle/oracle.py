"""ctypes wrapper of oracle/liboracle.so.

TEST INFRASTRUCTURE ONLY.  The oracle is the CPU checker the HIP path is
compared against; only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.  The product (ltpg_amd/) never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    u64p = ctypes.POINTER(ctypes.c_uint64)

    class KvResult(ctypes.Structure):
        _fields_ = [("commit_bitmap", u64p), ("conflicts", u32p), ("read_values", u32p),
                    ("commit_count", ctypes.c_uint32), ("range_errors", ctypes.c_uint32)]

    lib.oracle_kv_epoch.restype = ctypes.c_int
    lib.oracle_kv_epoch.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, u32p,
                                    ctypes.c_uint32, ctypes.c_uint32, u64p, u32p, u32p, ctypes.c_uint32,
                                    u64p, ctypes.POINTER(KvResult)]
    lib.KvResult = KvResult
    return lib


_lib = _load()


def _u32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def _u64p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


class KvEpoch:
    """Outputs of one key-value epoch of the oracle."""

    def __init__(self, bitmap, conflicts, read_values, commit_count, range_errors, table):
        self.bitmap = bitmap
        self.conflicts = conflicts
        self.read_values = read_values
        self.commit_count = commit_count
        self.range_errors = range_errors
        self.table = table

    def committed(self, t):
        return bool((int(self.bitmap[t // 64]) >> (t % 64)) & 1)


def kv_epoch(table, rows, fields, field_words, keys, meta, tids, n_slots, values=1, want_reads=True,
             inplace=False):
    """Run one epoch on a copy of `table` (row-major u32), or on `table`
    itself when inplace=True; return KvEpoch."""
    if inplace:
        assert table.dtype == np.uint32 and table.flags.c_contiguous
        table = table.reshape(-1)
    else:
        table = np.array(table, dtype=np.uint32, copy=True).reshape(-1)
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    meta = np.ascontiguousarray(meta, dtype=np.uint32)
    tids = np.ascontiguousarray(tids, dtype=np.uint32)
    B = int(tids.shape[0])
    S = int(n_slots)
    assert table.size == rows * fields * field_words
    assert keys.size == B * S and meta.size == B * S
    bitmap = np.zeros(max(1, (B + 63) // 64), dtype=np.uint64)
    conflicts = np.zeros(max(B, 1), dtype=np.uint32)
    reads = np.zeros(max(B * S * field_words, 1), dtype=np.uint32) if want_reads else None
    regs = np.empty(rows * 2, dtype=np.uint64)
    res = _lib.KvResult(_u64p(bitmap), _u32p(conflicts), _u32p(reads) if want_reads else None, 0, 0)
    rc = _lib.oracle_kv_epoch(rows, fields, field_words, _u32p(table), B, S, _u64p(keys), _u32p(meta),
                              _u32p(tids), values, _u64p(regs), ctypes.byref(res))
    assert rc == 0
    return KvEpoch(bitmap[: (B + 63) // 64], conflicts[:B],
                   reads[: B * S * field_words] if want_reads else None,
                   res.commit_count, res.range_errors, table)
