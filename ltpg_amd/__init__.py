"""ltpg_amd - MI355X-native deterministic concurrency control (LTPG's hot path).

Python mirror of the reference's host interface over the C ABI of
libltpg.so (include/ltpg.h).  The reference's host objects map as follows:

  YCSBDatabase / Database (snapshot, registers, clear_LOG)
      -> Context.kv_table_load / kv_table_read / kv_fill_reference /
         epoch_reset                 (LTPG/YCSB/0/Database.h:14-131)
  YCSBQuery / Query (batch upload, commit readback)
      -> Context.kv_submit / commit_bitmap / commit_count
                                     (LTPG/YCSB/0/Query.h:126-231)
  KERNEL_EXECUTE / KERNEL_CHECK / KERNEL_WRITEBACK
      -> Context.epoch_run(phases)   (LTPG/YCSB/0/ycsb.cu:9-58)

Every call goes to the HIP implementation; nothing here computes results.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import Config, KvBatch, KvGen, lib

OK = 0
ERR_INVALID, ERR_NOMEM, ERR_DEVICE, ERR_STATE, ERR_RANGE, ERR_BUDGET, ERR_UNSUPPORTED = (
    -1, -2, -3, -4, -5, -6, -7)

RULES_REFERENCE, RULES_INTENDED = 0, 1
VALUES_COPYBACK, VALUES_RMW = 0, 1
MEM_HOST, MEM_DEVICE = 0, 1
LAYOUT_ROW_MAJOR, LAYOUT_COLUMN_MAJOR = 0, 1
WORKLOAD_KV, WORKLOAD_TPCC = 1, 2
TYPE_KV, TYPE_NEWORDER, TYPE_PAYMENT = 0, 1, 2
PHASE_CID, PHASE_EXECUTE, PHASE_CHECK, PHASE_WRITEBACK, PHASE_ALL = 1, 2, 4, 8, 15
KV_NONE, KV_READ, KV_WRITE = 0, 1, 2
NO_HOT_ROW = (1 << 64) - 1
FLAG_CONFLICT_COUNTS = 1
TIMING_SLOTS = 5
TIMING_NAMES = ["cid", "execute", "check", "readout", "writeback"]


def kv_meta(kind, field):
    """LTPG_KV_META(kind, field)."""
    return (int(kind) | (int(field) << 8)) & 0xFFFFFFFF


class LtpgError(RuntimeError):
    def __init__(self, code, where):
        msg = lib.ltpg_last_error()
        self.code = code
        super().__init__(f"{where} failed ({code}): {msg.decode() if msg else ''}")


def _check(rc, where):
    if rc != OK:
        raise LtpgError(rc, where)


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def device_count():
    return lib.ltpg_device_count()


def default_config(**kw):
    cfg = Config()
    lib.ltpg_config_init(ctypes.byref(cfg))
    for k, v in kw.items():
        if not hasattr(cfg, k):
            raise AttributeError(f"ltpg_config has no field {k}")
        setattr(cfg, k, v)
    return cfg


class DeviceBuffer:
    """A library-owned device allocation (for LTPG_MEM_DEVICE batches)."""

    def __init__(self, ctx, nbytes):
        self._ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(lib.ltpg_device_alloc(ctx.handle, self.nbytes, ctypes.byref(p)), "ltpg_device_alloc")
        self.ptr = p.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes:
            raise ValueError("array larger than device buffer")
        _check(lib.ltpg_copy_to_device(self._ctx.handle, ctypes.c_void_p(self.ptr), _ptr(arr), arr.nbytes),
               "ltpg_copy_to_device")

    def download(self, arr):
        if arr.nbytes > self.nbytes or not arr.flags.c_contiguous:
            raise ValueError("bad destination array")
        _check(lib.ltpg_copy_to_host(self._ctx.handle, _ptr(arr), ctypes.c_void_p(self.ptr), arr.nbytes),
               "ltpg_copy_to_host")

    def copy_from(self, other, nbytes=None):
        """Device -> device copy (asynchronous on the context stream)."""
        n = min(self.nbytes, other.nbytes) if nbytes is None else int(nbytes)
        _check(lib.ltpg_device_copy(self._ctx.handle, ctypes.c_void_p(self.ptr), ctypes.c_void_p(other.ptr), n),
               "ltpg_device_copy")

    def free(self):
        if self.ptr:
            _check(lib.ltpg_device_free(self._ctx.handle, ctypes.c_void_p(self.ptr)), "ltpg_device_free")
            self.ptr = None


class DeviceKvBatch:
    """A key-value batch resident in device memory (inputs already in HBM)."""

    def __init__(self, ctx, keys, meta, tids, n_slots):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        meta = np.ascontiguousarray(meta, dtype=np.uint32)
        tids = np.ascontiguousarray(tids, dtype=np.uint32)
        self.n_txns = int(tids.shape[0])
        self.n_slots = int(n_slots)
        if keys.size != self.n_txns * self.n_slots or meta.size != keys.size:
            raise ValueError("keys/meta must hold n_slots * n_txns entries")
        self.bufs = [DeviceBuffer(ctx, max(a.nbytes, 8)) for a in (keys, meta, tids)]
        for b, a in zip(self.bufs, (keys, meta, tids)):
            if a.nbytes:
                b.upload(a)

    def struct(self):
        return KvBatch(self.n_txns, self.n_slots, self.bufs[0].ptr, self.bufs[1].ptr, self.bufs[2].ptr)

    def free(self):
        for b in self.bufs:
            b.free()


class Context:
    """One ltpg_ctx: device tables, registers, access sets and one stream."""

    def __init__(self, config=None, **kw):
        cfg = config if config is not None else default_config(**kw)
        self.config = cfg
        h = ctypes.c_void_p()
        _check(lib.ltpg_create(ctypes.byref(cfg), ctypes.byref(h)), "ltpg_create")
        self.handle = h
        self._cur_txns = 0
        self._cur_slots = 0

    def close(self):
        if self.handle:
            lib.ltpg_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- tables ---------------------------------------------------------
    @property
    def kv_shape(self):
        c = self.config
        return (int(c.kv_rows), int(c.kv_fields), int(c.kv_field_bytes) // 4)

    def kv_table_load(self, table, layout=LAYOUT_ROW_MAJOR):
        table = np.ascontiguousarray(table, dtype=np.uint32)
        rows, fields, fw = self.kv_shape
        if table.size != rows * fields * fw:
            raise ValueError("table size does not match the configured shape")
        _check(lib.ltpg_kv_table_load(self.handle, _ptr(table), rows, layout), "ltpg_kv_table_load")

    def kv_table_read(self, layout=LAYOUT_ROW_MAJOR):
        rows, fields, fw = self.kv_shape
        out = np.empty(rows * fields * fw, dtype=np.uint32)
        _check(lib.ltpg_kv_table_read(self.handle, _ptr(out), rows, layout), "ltpg_kv_table_read")
        return out

    def kv_fill_reference(self):
        _check(lib.ltpg_kv_table_fill_reference(self.handle), "ltpg_kv_table_fill_reference")

    # ---- batches and phases --------------------------------------------
    def kv_submit(self, keys, meta, tids, n_slots):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        meta = np.ascontiguousarray(meta, dtype=np.uint32)
        tids = np.ascontiguousarray(tids, dtype=np.uint32)
        n = int(tids.shape[0])
        if keys.size != n * n_slots or meta.size != keys.size:
            raise ValueError("keys/meta must hold n_slots * n_txns entries")
        b = KvBatch(n, int(n_slots), keys.ctypes.data if keys.size else None,
                    meta.ctypes.data if meta.size else None, tids.ctypes.data if n else None)
        _check(lib.ltpg_kv_submit(self.handle, ctypes.byref(b), MEM_HOST), "ltpg_kv_submit")
        self._cur_txns, self._cur_slots = n, int(n_slots)

    def kv_submit_device(self, batch):
        b = batch.struct()
        _check(lib.ltpg_kv_submit(self.handle, ctypes.byref(b), MEM_DEVICE), "ltpg_kv_submit")
        self._cur_txns, self._cur_slots = batch.n_txns, batch.n_slots

    def epoch_run(self, phases=PHASE_ALL):
        _check(lib.ltpg_epoch_run(self.handle, phases), "ltpg_epoch_run")

    def epoch_reset(self):
        _check(lib.ltpg_epoch_reset(self.handle), "ltpg_epoch_reset")

    def sync(self):
        _check(lib.ltpg_sync(self.handle), "ltpg_sync")

    def set_timing(self, on=True):
        _check(lib.ltpg_set_timing(self.handle, 1 if on else 0), "ltpg_set_timing")

    def phase_times(self):
        """(summed ms per timing slot [cid, execute, check, readout, writeback],
        number of timed epochs); starts a new measurement."""
        ms = (ctypes.c_float * TIMING_SLOTS)()
        n = ctypes.c_uint32()
        _check(lib.ltpg_phase_times(self.handle, ms, ctypes.byref(n)), "ltpg_phase_times")
        return list(ms), n.value

    def commit_totals(self, type_=TYPE_KV, reset=False):
        """(committed, attempted) transactions since the last reset."""
        c, a = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib.ltpg_commit_totals(self.handle, type_, ctypes.byref(c), ctypes.byref(a), 1 if reset else 0),
               "ltpg_commit_totals")
        return c.value, a.value

    # ---- results --------------------------------------------------------
    def commit_bitmap(self, type_=TYPE_KV, n_txns=None):
        n = self._cur_txns if n_txns is None else n_txns
        out = np.zeros(max(1, (n + 63) // 64), dtype=np.uint64)
        _check(lib.ltpg_commit_bitmap(self.handle, type_, out.ctypes.data_as(_lib.c_u64p), out.size),
               "ltpg_commit_bitmap")
        return out[: (n + 63) // 64]

    def commit_count(self, type_=TYPE_KV):
        v = ctypes.c_uint32()
        _check(lib.ltpg_commit_count(self.handle, type_, ctypes.byref(v)), "ltpg_commit_count")
        return v.value

    def conflict_counters(self, type_=TYPE_KV, n_txns=None):
        n = self._cur_txns if n_txns is None else n_txns
        out = np.zeros(n, dtype=np.uint32)
        _check(lib.ltpg_conflict_counters(self.handle, type_, out.ctypes.data_as(_lib.c_u32p), n),
               "ltpg_conflict_counters")
        return out

    def kv_read_values(self):
        """Read values of committed transactions' read slots (slot-major,
        field_words per op); other slots are left as they were."""
        _, _, fw = self.kv_shape
        n = self._cur_txns * self._cur_slots * fw
        out = np.zeros(max(n, 1), dtype=np.uint32)
        _check(lib.ltpg_kv_read_values(self.handle, out.ctypes.data_as(_lib.c_u32p), n), "ltpg_kv_read_values")
        return out[:n]


def kv_generate(seed=0x4C545047, rows=1_000_000, fields=10, n_txns=65536, n_slots=10, min_ops=None,
                max_ops=None, read_pct=50, theta=0.8, tid_base=0, distinct_keys=True, txn_base=0,
                threads=0):
    """Seeded key-value batch (slot-major keys/meta, tids) from libltpg's generator."""
    g = KvGen()
    lib.ltpg_kv_gen_init(ctypes.byref(g))
    g.seed, g.rows, g.fields, g.n_txns, g.n_slots = seed, rows, fields, n_txns, n_slots
    g.min_ops = n_slots if min_ops is None else min_ops
    g.max_ops = n_slots if max_ops is None else max_ops
    g.read_pct, g.theta, g.tid_base = read_pct, float(theta), tid_base
    g.distinct_keys, g.txn_base = 1 if distinct_keys else 0, txn_base
    keys = np.empty(n_txns * n_slots, dtype=np.uint64)
    meta = np.empty(n_txns * n_slots, dtype=np.uint32)
    tids = np.empty(n_txns, dtype=np.uint32)
    _check(lib.ltpg_kv_generate(ctypes.byref(g), keys.ctypes.data_as(_lib.c_u64p),
                                meta.ctypes.data_as(_lib.c_u32p), tids.ctypes.data_as(_lib.c_u32p), threads),
           "ltpg_kv_generate")
    return keys, meta, tids
