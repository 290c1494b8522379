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


def _load(path=None):
    if path is None:
        if not os.path.exists(LIB_PATH):
            build()
        path = LIB_PATH
    lib = ctypes.CDLL(path)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    u64p = ctypes.POINTER(ctypes.c_uint64)

    class KvResult(ctypes.Structure):
        _fields_ = [("commit_bitmap", u64p), ("conflicts", u32p), ("read_values", u32p),
                    ("commit_count", ctypes.c_uint32), ("range_errors", ctypes.c_uint32)]

    lib.oracle_kv_epoch.restype = ctypes.c_int
    lib.oracle_kv_epoch.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, u32p,
                                    ctypes.c_uint32, ctypes.c_uint32, u64p, u32p, u32p, ctypes.c_uint32,
                                    u64p, ctypes.POINTER(KvResult), ctypes.c_uint64, ctypes.c_uint32]
    lib.KvResult = KvResult
    lib.oracle_kv_epoch_deltas.restype = ctypes.c_int
    lib.oracle_kv_epoch_deltas.argtypes = lib.oracle_kv_epoch.argtypes + [u32p]

    class TpccResult(ctypes.Structure):
        _fields_ = [("no_bitmap", u64p), ("pay_bitmap", u64p), ("no_conflicts", u32p), ("pay_conflicts", u32p),
                    ("no_reads", u32p), ("pay_reads", u32p), ("pay_cid", u32p), ("no_commits", ctypes.c_uint32),
                    ("pay_commits", ctypes.c_uint32), ("range_errors", ctypes.c_uint32)]

    lib.TpccResult = TpccResult
    tables_t = u32p * 9
    lib.tables_t = tables_t
    lib.oracle_tpcc_shape.restype = None
    lib.oracle_tpcc_shape.argtypes = [ctypes.c_uint32, ctypes.c_uint32, u64p, u32p]
    lib.oracle_tpcc_pop_uniform.restype = ctypes.c_uint32
    lib.oracle_tpcc_pop_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32,
                                            ctypes.c_uint32]
    lib.oracle_tpcc_populate.restype = None
    lib.oracle_tpcc_populate.argtypes = [ctypes.c_uint32, ctypes.c_uint64, tables_t]
    lib.oracle_tpcc_create.restype = ctypes.c_void_p
    lib.oracle_tpcc_create.argtypes = [ctypes.c_uint32]
    lib.oracle_tpcc_destroy.restype = None
    lib.oracle_tpcc_destroy.argtypes = [ctypes.c_void_p]
    lib.oracle_tpcc_epoch.restype = ctypes.c_int
    lib.oracle_tpcc_epoch.argtypes = [ctypes.c_void_p, tables_t, ctypes.c_uint32, ctypes.c_uint32, u32p,
                                      ctypes.c_uint32, ctypes.c_uint32, u32p, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.POINTER(TpccResult)]

    class TpccRoResult(ctypes.Structure):
        _fields_ = [("os_bitmap", u64p), ("sl_bitmap", u64p), ("os_conflicts", u32p), ("sl_conflicts", u32p),
                    ("os_reads", u32p), ("sl_reads", u32p), ("os_commits", ctypes.c_uint32),
                    ("sl_commits", ctypes.c_uint32)]

    lib.TpccRoResult = TpccRoResult
    if hasattr(lib, "oracle_tpcc_epoch_ro"):
        lib.oracle_tpcc_epoch_ro.restype = ctypes.c_int
        lib.oracle_tpcc_epoch_ro.argtypes = [ctypes.c_void_p, tables_t, ctypes.c_uint32, ctypes.c_uint32, u32p,
                                             ctypes.c_uint32, ctypes.c_uint32, u32p, ctypes.c_uint32,
                                             ctypes.c_uint32, u32p, ctypes.c_uint32, ctypes.c_uint32, u32p,
                                             ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(TpccResult),
                                             ctypes.POINTER(TpccRoResult)]
    class TpccDlResult(ctypes.Structure):
        _fields_ = [("bitmap", u64p), ("conflicts", u32p), ("reads", u32p), ("commits", ctypes.c_uint32)]

    lib.TpccDlResult = TpccDlResult
    if hasattr(lib, "oracle_tpcc_epoch_all"):
        lib.oracle_tpcc_epoch_all.restype = ctypes.c_int
        lib.oracle_tpcc_epoch_all.argtypes = (lib.oracle_tpcc_epoch_ro.argtypes[:-2] +
                                              [u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(TpccResult),
                                               ctypes.POINTER(TpccRoResult), ctypes.POINTER(TpccDlResult)])
    return lib


_lib = _load()


def load_variant(path):
    """A separately built copy of the oracle library (tests that plant a
    mistake in a copy of the oracle load it with this)."""
    return _load(path)


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


NO_HOT_ROW = (1 << 64) - 1


def kv_epoch(table, rows, fields, field_words, keys, meta, tids, n_slots, values=1, want_reads=True,
             inplace=False, hot_row=NO_HOT_ROW, max_scan=0, deltas=None):
    """Run one epoch on a copy of `table` (row-major u32), or on `table`
    itself when inplace=True; return KvEpoch.  Read values hold
    max(1, max_scan) fields per op slot.  deltas: the RMW delta of each
    transaction (default tids | 1)."""
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
    L = max(1, int(max_scan))
    reads = np.zeros(max(B * S * L * field_words, 1), dtype=np.uint32) if want_reads else None
    regs = np.empty(rows * 2, dtype=np.uint64)
    res = _lib.KvResult(_u64p(bitmap), _u32p(conflicts), _u32p(reads) if want_reads else None, 0, 0)
    args = (rows, fields, field_words, _u32p(table), B, S, _u64p(keys), _u32p(meta), _u32p(tids), values,
            _u64p(regs), ctypes.byref(res), hot_row, int(max_scan))
    if deltas is None:
        rc = _lib.oracle_kv_epoch(*args)
    else:
        deltas = np.ascontiguousarray(deltas, dtype=np.uint32)
        assert deltas.shape == (B,)
        rc = _lib.oracle_kv_epoch_deltas(*args, _u32p(deltas))
    assert rc == 0
    return KvEpoch(bitmap[: (B + 63) // 64], conflicts[:B],
                   reads[: B * S * L * field_words] if want_reads else None,
                   res.commit_count, res.range_errors, table)


# ---- order-entry -----------------------------------------------------------

TPCC_COLS = [9, 11, 21, 5, 3, 8, 10, 17, 5]


def tpcc_shape(warehouses, table):
    rows, cols = ctypes.c_uint64(), ctypes.c_uint32()
    _lib.oracle_tpcc_shape(warehouses, table, ctypes.byref(rows), ctypes.byref(cols))
    return rows.value, cols.value


def tpcc_empty_tables(warehouses):
    return [np.zeros(tpcc_shape(warehouses, t)[0] * TPCC_COLS[t], dtype=np.uint32) for t in range(9)]


def tpcc_populate(warehouses, seed):
    """The loader's nine column-major tables (tbl[col * rows + row])."""
    tables = tpcc_empty_tables(warehouses)
    _lib.oracle_tpcc_populate(warehouses, seed, _lib.tables_t(*[_u32p(t) for t in tables]))
    return tables


def tpcc_pop_uniform(seed, stream, row, a, b):
    return int(_lib.oracle_tpcc_pop_uniform(seed, stream, row, a, b))


class TpccEpoch:
    """Outputs of one order-entry epoch of the oracle (per type)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    @staticmethod
    def bit(bitmap, i):
        return bool((int(bitmap[i // 64]) >> (i % 64)) & 1)


class TpccOracle:
    """oracle_tpcc: registers for W warehouses, reused across epochs."""

    def __init__(self, warehouses, lib=None):
        self.warehouses = warehouses
        self.lib = lib if lib is not None else _lib
        self.h = self.lib.oracle_tpcc_create(warehouses)
        assert self.h, "oracle_tpcc_create failed"

    def close(self):
        if self.h:
            self.lib.oracle_tpcc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def epoch(self, tables, batch, rules=0, values=1, ro=None, dl=None):
        """One epoch over `tables` (list of 9 column-major u32 arrays, updated
        in place) for `batch` (an object with no/no_cap/n_no/pay/pay_cap/n_pay:
        the packed NEWORDER_QUERY / PAYMENT_QUERY blocks) and, optionally,
        read-only transactions `ro` (an object with n_os / n_sl and field
        arrays os[name] / sl[name]) and Deliveries `dl` (an object with n and
        arrays O_ID, NO_W_ID, NO_O_ID, TID); their results come back in
        TpccEpoch.ro and TpccEpoch.dl."""
        for t in tables:
            assert t.dtype == np.uint32 and t.flags.c_contiguous
        n_no, n_pay = int(batch.n_no), int(batch.n_pay)
        nb = np.zeros(max(1, (n_no + 63) // 64), np.uint64)
        pb = np.zeros(max(1, (n_pay + 63) // 64), np.uint64)
        nc = np.zeros(max(1, n_no), np.uint32)
        pc = np.zeros(max(1, n_pay), np.uint32)
        nr = np.zeros(max(1, 51 * n_no), np.uint32)
        pr = np.zeros(max(1, 24 * n_pay), np.uint32)
        cid = np.zeros(max(1, n_pay), np.uint32)
        lib = self.lib
        res = lib.TpccResult(_u64p(nb), _u64p(pb), _u32p(nc), _u32p(pc), _u32p(nr), _u32p(pr), _u32p(cid), 0, 0, 0)
        no = np.ascontiguousarray(batch.no, dtype=np.uint32)
        pay = np.ascontiguousarray(batch.pay, dtype=np.uint32)
        dl_out = None
        if dl is not None:
            n_dl = int(dl.n)
            dlb = np.ascontiguousarray(np.stack([np.asarray(getattr(dl, f), np.uint32)[:max(n_dl, 1)]
                                                 if n_dl else np.zeros(1, np.uint32)
                                                 for f in ("O_ID", "NO_W_ID", "NO_O_ID", "TID")]), np.uint32)
            db = np.zeros(max(1, (n_dl + 63) // 64), np.uint64)
            dc, dr = np.zeros(max(1, n_dl), np.uint32), np.zeros(max(1, 2 * n_dl), np.uint32)
            dres = lib.TpccDlResult(_u64p(db), _u32p(dc), _u32p(dr), 0)
        if ro is None and dl is None:
            rc = lib.oracle_tpcc_epoch(self.h, lib.tables_t(*[_u32p(t) for t in tables]), rules, values,
                                        _u32p(no), batch.no_cap, n_no, _u32p(pay), batch.pay_cap, n_pay,
                                        ctypes.byref(res))
            ro_out = None
        elif ro is None:
            z = np.zeros(3, np.uint32)
            rres = lib.TpccRoResult()
            rc = lib.oracle_tpcc_epoch_all(self.h, lib.tables_t(*[_u32p(t) for t in tables]), rules, values,
                                           _u32p(no), batch.no_cap, n_no, _u32p(pay), batch.pay_cap, n_pay,
                                           _u32p(z), 1, 0, _u32p(z), 1, 0, _u32p(dlb), max(n_dl, 1), n_dl,
                                           ctypes.byref(res), ctypes.byref(rres), ctypes.byref(dres))
            ro_out = None
        else:
            n_os, n_sl = int(ro.n_os), int(ro.n_sl)
            osb = np.ascontiguousarray(np.stack([ro.os[f][:max(n_os, 1)] for f in
                                                 ("W_ID", "D_ID", "C_ID", "O_ID", "OL_ID", "TID")]), np.uint32)
            slb = np.ascontiguousarray(np.stack([ro.sl[f][:max(n_sl, 1)] for f in ("W_ID", "S_ID", "TID")]),
                                       np.uint32)
            ob = np.zeros(max(1, (n_os + 63) // 64), np.uint64)
            sb = np.zeros(max(1, (n_sl + 63) // 64), np.uint64)
            oc, sc = np.zeros(max(1, n_os), np.uint32), np.zeros(max(1, n_sl), np.uint32)
            orr, srr = np.zeros(max(1, 3 * n_os), np.uint32), np.zeros(max(1, n_sl), np.uint32)
            rres = lib.TpccRoResult(_u64p(ob), _u64p(sb), _u32p(oc), _u32p(sc), _u32p(orr), _u32p(srr), 0, 0)
            if dl is None:
                rc = lib.oracle_tpcc_epoch_ro(self.h, lib.tables_t(*[_u32p(t) for t in tables]), rules, values,
                                               _u32p(no), batch.no_cap, n_no, _u32p(pay), batch.pay_cap, n_pay,
                                               _u32p(osb), max(n_os, 1), n_os, _u32p(slb), max(n_sl, 1), n_sl,
                                               ctypes.byref(res), ctypes.byref(rres))
            else:
                rc = lib.oracle_tpcc_epoch_all(self.h, lib.tables_t(*[_u32p(t) for t in tables]), rules, values,
                                                _u32p(no), batch.no_cap, n_no, _u32p(pay), batch.pay_cap, n_pay,
                                                _u32p(osb), max(n_os, 1), n_os, _u32p(slb), max(n_sl, 1), n_sl,
                                                _u32p(dlb), max(n_dl, 1), n_dl, ctypes.byref(res),
                                                ctypes.byref(rres), ctypes.byref(dres))
            ro_out = dict(os_bitmap=ob[:(n_os + 63) // 64], sl_bitmap=sb[:(n_sl + 63) // 64],
                          os_conflicts=oc[:n_os], sl_conflicts=sc[:n_sl], os_reads=orr[:3 * n_os],
                          sl_reads=srr[:n_sl], os_commits=rres.os_commits, sl_commits=rres.sl_commits)
        assert rc == 0
        if dl is not None:
            dl_out = dict(bitmap=db[:(n_dl + 63) // 64], conflicts=dc[:n_dl], reads=dr[:2 * n_dl],
                          commits=dres.commits)
        return TpccEpoch(dl=dl_out, no_bitmap=nb[:(n_no + 63) // 64], pay_bitmap=pb[:(n_pay + 63) // 64],
                         no_conflicts=nc[:n_no], pay_conflicts=pc[:n_pay], no_reads=nr[:51 * n_no],
                         pay_reads=pr[:24 * n_pay], pay_cid=cid[:n_pay], no_commits=res.no_commits,
                         pay_commits=res.pay_commits, range_errors=res.range_errors, ro=ro_out)
