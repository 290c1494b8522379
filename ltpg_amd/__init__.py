"""ltpg_amd - MI355X-native deterministic concurrency control (LTPG's hot path).

Python mirror of the reference's host interface over the C ABI of
libltpg.so (include/ltpg.h).  The reference's host objects map as follows:

  YCSBDatabase / Database (snapshot, registers, clear_LOG)
      -> Context.kv_table_load / kv_table_read / kv_fill_reference /
         epoch_reset                 (LTPG/YCSB/0/Database.h:14-131)
  YCSBQuery / Query (batch upload, commit readback)
      -> Context.kv_submit / commit_bitmap / commit_count
                                     (LTPG/YCSB/0/Query.h:126-231)
  KERNEL_CID / KERNEL_EXECUTE / KERNEL_CHECK / KERNEL_WRITEBACK
      -> Context.epoch_run(phases)   (LTPG/YCSB/0/ycsb.cu:9-58,
                                      LTPG/TPCC-allopt-final/tp.cu:11-117)
  Database (order-entry snapshot, create_*, customer name index)
      -> Context.tpcc_table_load / tpcc_table_read / tpcc_populate
                                     (TPCC-allopt-final/Database.h:13-451)
  Query (NEWORDER_QUERY / PAYMENT_QUERY upload) and MakeNeworder /
  MakePayment
      -> Context.tpcc_submit, TpccGenerator
                                     (TPCC-allopt-final/Query.h:16-212,532-631)

Every call goes to the HIP implementation; nothing here computes results.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import BatchDesc, Config, ServeConfig, ServeDone, ServeStats, KvBatch, KvGen, NeworderBatch, PaymentBatch, TpccGenConfig, lib

OK = 0
ERR_INVALID, ERR_NOMEM, ERR_DEVICE, ERR_STATE, ERR_RANGE, ERR_BUDGET, ERR_UNSUPPORTED = (
    -1, -2, -3, -4, -5, -6, -7)

RULES_REFERENCE, RULES_INTENDED = 0, 1
VALUES_COPYBACK, VALUES_RMW = 0, 1
MEM_HOST, MEM_DEVICE = 0, 1
BATCH_REQUEST_MAJOR = 0x10
LAYOUT_ROW_MAJOR, LAYOUT_COLUMN_MAJOR = 0, 1
WORKLOAD_KV, WORKLOAD_TPCC = 1, 2
TYPE_KV, TYPE_NEWORDER, TYPE_PAYMENT, TYPE_ORDERSTATUS, TYPE_STOCKLEVEL, TYPE_DELIVERY = 0, 1, 2, 3, 4, 5
PHASE_CID, PHASE_EXECUTE, PHASE_CHECK, PHASE_WRITEBACK, PHASE_ALL = 1, 2, 4, 8, 15
KV_NONE, KV_READ, KV_WRITE, KV_SCAN = 0, 1, 2, 3
NO_HOT_ROW = (1 << 64) - 1
FLAG_CONFLICT_COUNTS = 1
TIMING_SLOTS = 5
TIMING_NAMES = ["cid", "execute", "check", "readout", "writeback"]
TPCC_WAREHOUSE, TPCC_DISTRICT, TPCC_CUSTOMER, TPCC_HISTORY, TPCC_NEWORDER, TPCC_ORDER, TPCC_ORDERLINE, \
    TPCC_STOCK, TPCC_ITEM = range(9)
TPCC_TABLES = ["warehouse", "district", "customer", "history", "neworder", "order", "orderline", "stock", "item"]
TPCC_MAX_LINES = 15
NEWORDER_READS, PAYMENT_READS, ORDERSTATUS_READS, STOCKLEVEL_READS, DELIVERY_READS = 51, 24, 3, 1, 2
ORDERSTATUS_FIELDS = ("W_ID", "D_ID", "C_ID", "O_ID", "OL_ID", "TID")
STOCKLEVEL_FIELDS = ("W_ID", "S_ID", "TID")
DELIVERY_FIELDS = ("O_ID", "NO_W_ID", "NO_O_ID", "TID")
# u32 words per transaction of the packed NEWORDER_QUERY / PAYMENT_QUERY blocks
NEWORDER_WORDS, PAYMENT_WORDS = 53, 10
RUN_PLAIN, RUN_GRAPH = 0, 1
TUNE_EXEC_PER_OP, TUNE_SPLIT_REGISTERS, TUNE_SMALL_CHUNKS, TUNE_LANE_COPY, TUNE_ALL = 1, 2, 4, 8, 15
TUNE_SMALL_EPOCH, SMALL_EPOCH_OPS = 16, 2048
TUNE_PARTITIONED = 32
TUNE_MORE_ITEMS = 64
TUNE_TPCC_SMALL_BLOCKS = 128


def tune_index(flags):
    """Index of a key-value launch shape in a tuning entry's ms (LTPG_TUNE_INDEX)."""
    return (flags & TUNE_ALL) | (16 if flags & TUNE_PARTITIONED else 0) | (32 if flags & TUNE_MORE_ITEMS else 0)


def kv_shapes(partitioned=True):
    """The distinct key-value launch shapes (the autotuner's set): every
    TUNE_ALL combination with and without MORE_ITEMS, and PARTITIONED with
    the SPLIT_REGISTERS / SMALL_CHUNKS / LANE_COPY combinations."""
    out = [f | m for m in (0, TUNE_MORE_ITEMS) for f in range(16)]
    if partitioned:
        out += [TUNE_PARTITIONED | f for f in range(16) if not f & TUNE_EXEC_PER_OP]
    return out
NOALLOC_MAX_ROWS = NOALLOC_MAX_TXNS = 1 << 20


def kv_meta(kind, field):
    """LTPG_KV_META(kind, field)."""
    return (int(kind) | (int(field) << 8)) & 0xFFFFFFFF


def kv_scan_meta(field, length):
    """LTPG_KV_SCAN_META(field, length): a scan of `length` rows (1..256)."""
    return (KV_SCAN | (int(field) << 8) | ((int(length) - 1) << 24)) & 0xFFFFFFFF


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


def random_access_rates(ctx, nbytes, n_ops=1 << 24):
    """(load, probe+atomicMin) random 8-byte access rates of the device in
    G ops/s over an array of nbytes (ltpg_random_access_rates)."""
    lo, at = ctypes.c_double(), ctypes.c_double()
    _check(lib.ltpg_random_access_rates(ctx.handle, int(nbytes), int(n_ops), ctypes.byref(lo), ctypes.byref(at)),
           "ltpg_random_access_rates")
    return lo.value, at.value


def random_gather_rate(ctx, nbytes, field_bytes, stride_bytes, n_ops):
    """GB/s (read + written) of n_ops field copies from random slots of an
    nbytes array into a dense one (ltpg_random_gather_rate)."""
    g = ctypes.c_double()
    _check(lib.ltpg_random_gather_rate(ctx.handle, int(nbytes), int(field_bytes), int(stride_bytes), int(n_ops),
                                       ctypes.byref(g)), "ltpg_random_gather_rate")
    return g.value


def sync_costs(ctx, blocks, reps=1000):
    """{grid_barrier_us, graph_boundary_us, stream_boundary_us} for
    `blocks` 256-thread workgroups (ltpg_sync_costs)."""
    out = (ctypes.c_double * 3)()
    _check(lib.ltpg_sync_costs(ctx.handle, int(blocks), int(reps), out), "ltpg_sync_costs")
    return {"grid_barrier_us": out[0], "graph_boundary_us": out[1], "stream_boundary_us": out[2]}


LATENCY_NAMES = ("barrier_ns", "lds_load_ns", "lds_atomic_ns", "pcie_load_ns", "publish_ns")


def latency_costs(ctx):
    """Step latencies of the served device loop's epoch chain, in ns
    (ltpg_latency_costs): LDS-only workgroup barrier, dependent LDS load,
    dependent LDS atomic, PCIe round trip, release publish to host memory."""
    out = (ctypes.c_double * 5)()
    _check(lib.ltpg_latency_costs(ctx.handle, out), "ltpg_latency_costs")
    return dict(zip(LATENCY_NAMES, out))


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

    def __init__(self, ctx, keys, meta, tids, n_slots, request_major=False):
        self.request_major = bool(request_major)
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
        self._n_no = 0
        self._n_pay = 0
        self._n_os = 0
        self._n_sl = 0
        self._n_dl = 0

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

    def kv_table_read_rows(self, rows):
        """Rows `rows` of the table, shape (n, fields, field_words)."""
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        _, fields, fw = self.kv_shape
        out = np.empty((rows.size, fields, fw), dtype=np.uint32)
        _check(lib.ltpg_kv_table_read_rows(self.handle, rows.ctypes.data_as(_lib.c_u64p), rows.size,
                                           out.ctypes.data_as(_lib.c_u32p)), "ltpg_kv_table_read_rows")
        return out

    def kv_set_tuning(self, flags=0):
        """LTPG_TUNE_* launch shape of the key-value kernels (results never change)."""
        _check(lib.ltpg_kv_set_tuning(self.handle, int(flags)), "ltpg_kv_set_tuning")

    def tpcc_set_tuning(self, flags=0):
        """Launch shape of the order-entry kernels: 0 or TUNE_TPCC_SMALL_BLOCKS."""
        _check(lib.ltpg_tpcc_set_tuning(self.handle, int(flags)), "ltpg_tpcc_set_tuning")

    def kv_fill_reference(self):
        _check(lib.ltpg_kv_table_fill_reference(self.handle), "ltpg_kv_table_fill_reference")

    # ---- batches and phases --------------------------------------------
    def kv_submit(self, keys, meta, tids, n_slots, request_major=False):
        """Submit a host batch: slot-major keys/meta ([s * B + t]), or with
        request_major ([t * n_slots + s])."""
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        meta = np.ascontiguousarray(meta, dtype=np.uint32)
        tids = np.ascontiguousarray(tids, dtype=np.uint32)
        n = int(tids.shape[0])
        if keys.size != n * n_slots or meta.size != keys.size:
            raise ValueError("keys/meta must hold n_slots * n_txns entries")
        b = KvBatch(n, int(n_slots), keys.ctypes.data if keys.size else None,
                    meta.ctypes.data if meta.size else None, tids.ctypes.data if n else None)
        _check(lib.ltpg_kv_submit(self.handle, ctypes.byref(b), MEM_HOST | (BATCH_REQUEST_MAJOR if request_major else 0)),
               "ltpg_kv_submit")
        self._cur_txns, self._cur_slots = n, int(n_slots)

    def kv_submit_device(self, batch):
        b = batch.struct()
        mem = MEM_DEVICE | (BATCH_REQUEST_MAJOR if batch.request_major else 0)
        _check(lib.ltpg_kv_submit(self.handle, ctypes.byref(b), mem), "ltpg_kv_submit")
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

    # ---- order-entry ------------------------------------------------------
    def tpcc_shape(self, table):
        rows, cols = ctypes.c_uint64(), ctypes.c_uint32()
        _check(lib.ltpg_tpcc_table_shape(self.handle, table, ctypes.byref(rows), ctypes.byref(cols)),
               "ltpg_tpcc_table_shape")
        return rows.value, cols.value

    def tpcc_table_load(self, table, data, layout=LAYOUT_COLUMN_MAJOR):
        rows, cols = self.tpcc_shape(table)
        data = np.ascontiguousarray(data, dtype=np.uint32)
        if data.size != rows * cols:
            raise ValueError("table size does not match its shape")
        _check(lib.ltpg_tpcc_table_load(self.handle, table, data.ctypes.data_as(_lib.c_u32p), rows, layout),
               "ltpg_tpcc_table_load")

    def tpcc_table_read(self, table, layout=LAYOUT_COLUMN_MAJOR):
        rows, cols = self.tpcc_shape(table)
        out = np.empty(rows * cols, dtype=np.uint32)
        _check(lib.ltpg_tpcc_table_read(self.handle, table, out.ctypes.data_as(_lib.c_u32p), rows, layout),
               "ltpg_tpcc_table_read")
        return out

    def tpcc_populate(self, seed=0x4C545047):
        _check(lib.ltpg_tpcc_populate(self.handle, seed), "ltpg_tpcc_populate")

    def tpcc_submit(self, batch, mem=MEM_HOST):
        """Submit a TpccBatch (host blocks), or with mem=MEM_DEVICE a pair of
        (NeworderBatch, PaymentBatch) structs of device pointers."""
        if mem == MEM_HOST:
            nb = neworder_struct(batch.no, batch.no_cap, batch.n_no)
            pb = payment_struct(batch.pay, batch.pay_cap, batch.n_pay)
            n_no, n_pay = batch.n_no, batch.n_pay
        else:
            nb, pb = batch
            n_no, n_pay = nb.n, pb.n
        _check(lib.ltpg_tpcc_submit(self.handle, ctypes.byref(nb), ctypes.byref(pb), mem), "ltpg_tpcc_submit")
        self._n_no, self._n_pay = n_no, n_pay
        self._n_os = self._n_sl = self._n_dl = 0

    def tpcc_submit_readonly(self, ro, mem=MEM_HOST):
        """Add a ReadOnlyBatch (host arrays), or with mem=MEM_DEVICE a pair
        of (OrderstatusBatch, StocklevelBatch) structs of device pointers, to
        the batch of the last tpcc_submit."""
        if mem == MEM_HOST:
            ob, sb = ro.structs()
        else:
            ob, sb = ro
        _check(lib.ltpg_tpcc_submit_readonly(self.handle, ctypes.byref(ob), ctypes.byref(sb), mem),
               "ltpg_tpcc_submit_readonly")
        self._n_os, self._n_sl = ob.n, sb.n

    def tpcc_submit_delivery(self, dl, mem=MEM_HOST):
        """Add a DeliveryBatch (host arrays), or with mem=MEM_DEVICE a
        DeliveryBatch struct of device pointers (upload_delivery_batch), to
        the batch of the last tpcc_submit."""
        st = dl.struct() if mem == MEM_HOST else dl
        _check(lib.ltpg_tpcc_submit_delivery(self.handle, ctypes.byref(st), mem), "ltpg_tpcc_submit_delivery")
        self._n_dl = st.n

    def _tpcc_n(self, type_):
        return {TYPE_NEWORDER: self._n_no, TYPE_PAYMENT: self._n_pay, TYPE_ORDERSTATUS: self._n_os,
                TYPE_STOCKLEVEL: self._n_sl, TYPE_DELIVERY: self._n_dl}[type_]

    def tpcc_payment_cid(self):
        out = np.zeros(max(self._n_pay, 1), dtype=np.uint32)
        _check(lib.ltpg_tpcc_payment_cid(self.handle, out.ctypes.data_as(_lib.c_u32p), self._n_pay),
               "ltpg_tpcc_payment_cid")
        return out[:self._n_pay]

    def tpcc_read_values(self, type_):
        per = {TYPE_NEWORDER: NEWORDER_READS, TYPE_PAYMENT: PAYMENT_READS, TYPE_ORDERSTATUS: ORDERSTATUS_READS,
               TYPE_STOCKLEVEL: STOCKLEVEL_READS, TYPE_DELIVERY: DELIVERY_READS}[type_]
        n = self._tpcc_n(type_) * per
        out = np.zeros(max(n, 1), dtype=np.uint32)
        _check(lib.ltpg_tpcc_read_values(self.handle, type_, out.ctypes.data_as(_lib.c_u32p), n),
               "ltpg_tpcc_read_values")
        return out[:n]

    def tpcc_bitmap(self, type_):
        return self.commit_bitmap(type_, self._tpcc_n(type_))

    def tpcc_conflicts(self, type_):
        return self.conflict_counters(type_, self._n_no if type_ == TYPE_NEWORDER else self._n_pay)

    def kv_read_values(self):
        """Read values of committed transactions' read and scan slots
        (slot-major, max(1, kv_max_scan) fields of field_words per op);
        other entries are left as they were."""
        _, _, fw = self.kv_shape
        n = self._cur_txns * self._cur_slots * fw * max(1, int(self.config.kv_max_scan))
        out = np.zeros(max(n, 1), dtype=np.uint32)
        _check(lib.ltpg_kv_read_values(self.handle, out.ctypes.data_as(_lib.c_u32p), n), "ltpg_kv_read_values")
        return out[:n]


class ReadOnlyBatch:
    """Host arrays of read-only order-entry transactions: Order-Status
    fields (W_ID, D_ID, C_ID, O_ID, OL_ID, TID) and Stock-Level fields
    (W_ID, S_ID, TID), one u32 array each."""

    def __init__(self, n_os, n_sl):
        self.n_os, self.n_sl = int(n_os), int(n_sl)
        self.os = {f: np.zeros(max(self.n_os, 1), dtype=np.uint32) for f in ORDERSTATUS_FIELDS}
        self.sl = {f: np.zeros(max(self.n_sl, 1), dtype=np.uint32) for f in STOCKLEVEL_FIELDS}

    def structs(self):
        ob, sb = _lib.OrderstatusBatch(), _lib.StocklevelBatch()
        ob.n, sb.n = self.n_os, self.n_sl
        for f in ORDERSTATUS_FIELDS:
            setattr(ob, f, self.os[f].ctypes.data)
        for f in STOCKLEVEL_FIELDS:
            setattr(sb, f, self.sl[f].ctypes.data)
        return ob, sb


class DeliveryBatch:
    """Host arrays of Deliveries: DeliveryQuery (O_ID, NO_W_ID, NO_O_ID)
    and TID, one u32 array each (attributes of those names)."""

    def __init__(self, n):
        self.n = int(n)
        for f in DELIVERY_FIELDS:
            setattr(self, f, np.zeros(max(self.n, 1), dtype=np.uint32))

    def struct(self):
        st = _lib.DeliveryBatch()
        st.n = self.n
        for f in DELIVERY_FIELDS:
            setattr(st, f, getattr(self, f).ctypes.data)
        return st


def neworder_struct(block, cap, n):
    """ltpg_neworder_batch viewing a packed NEWORDER_QUERY block (53 * cap u32)."""
    b = NeworderBatch()
    lib.ltpg_neworder_view(block.ctypes.data_as(_lib.c_u32p), cap, n, ctypes.byref(b))
    return b


def payment_struct(block, cap, n):
    """ltpg_payment_batch viewing a packed PAYMENT_QUERY block (10 * cap u32)."""
    b = PaymentBatch()
    lib.ltpg_payment_view(block.ctypes.data_as(_lib.c_u32p), cap, n, ctypes.byref(b))
    return b


class TpccBatch:
    """One order-entry batch in the reference's packed layout: `no` is a
    NEWORDER_QUERY block [53 * no_cap] and `pay` a PAYMENT_QUERY block
    [10 * pay_cap] (TPCC-allopt-final/Datastructure.h:211-249)."""

    NO_FIELDS = {"W_ID": 0, "D_ID": 1, "C_ID": 2, "O_ID": 3, "N_O_ID": 4, "O_OL_CNT": 5, "O_OL_ID": 6,
                 "OL_I_ID": 7, "OL_SUPPLY_W_ID": 22, "OL_QUANTITY": 37, "TID": 52}
    PAY_FIELDS = {"W_ID": 0, "D_ID": 1, "C_ID": 2, "C_LAST": 3, "isName": 4, "C_D_ID": 5, "C_W_ID": 6,
                  "H_AMOUNT": 7, "H_ID": 8, "TID": 9}

    def __init__(self, no, no_cap, n_no, pay, pay_cap, n_pay):
        self.no, self.no_cap, self.n_no = no, int(no_cap), int(n_no)
        self.pay, self.pay_cap, self.n_pay = pay, int(pay_cap), int(n_pay)

    @classmethod
    def empty(cls, n_no, n_pay, no_cap=None, pay_cap=None):
        no_cap = max(1, n_no if no_cap is None else no_cap)
        pay_cap = max(1, n_pay if pay_cap is None else pay_cap)
        return cls(np.zeros(NEWORDER_WORDS * no_cap, np.uint32), no_cap, n_no,
                   np.zeros(PAYMENT_WORDS * pay_cap, np.uint32), pay_cap, n_pay)

    def nf(self, name, line=None):
        """View of NewOrder field `name` (a line of it for OL_* fields)."""
        k = self.NO_FIELDS[name]
        if name.startswith("OL_"):
            if line is None:
                return self.no[k * self.no_cap:(k + 15) * self.no_cap].reshape(15, self.no_cap)[:, :self.n_no]
            k += line
        return self.no[k * self.no_cap:k * self.no_cap + self.n_no]

    def pf(self, name):
        """View of Payment field `name`."""
        k = self.PAY_FIELDS[name]
        return self.pay[k * self.pay_cap:k * self.pay_cap + self.n_pay]


class TpccGenerator:
    """Seeded NewOrder / Payment batches (ltpg_tpcc_gen_*)."""

    def __init__(self, **kw):
        cfg = TpccGenConfig()
        lib.ltpg_tpcc_gen_config_init(ctypes.byref(cfg))
        for k, v in kw.items():
            if not hasattr(cfg, k):
                raise AttributeError(f"ltpg_tpcc_gen_config has no field {k}")
            setattr(cfg, k, v)
        self.config = cfg
        h = ctypes.c_void_p()
        _check(lib.ltpg_tpcc_gen_create(ctypes.byref(cfg), ctypes.byref(h)), "ltpg_tpcc_gen_create")
        self.handle = h

    def next(self, n_txns, tid_base=0):
        n_no = n_txns * self.config.neworder_pct // 100
        b = TpccBatch.empty(n_no, n_txns - n_no)
        a, c = ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib.ltpg_tpcc_gen_next(self.handle, n_txns, tid_base, b.no.ctypes.data_as(_lib.c_u32p), b.no_cap,
                                      b.pay.ctypes.data_as(_lib.c_u32p), b.pay_cap, ctypes.byref(a),
                                      ctypes.byref(c)), "ltpg_tpcc_gen_next")
        assert a.value == b.n_no and c.value == b.n_pay
        return b

    def close(self):
        if self.handle:
            lib.ltpg_tpcc_gen_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def readonly_batch(b, warehouses, n_os, n_sl, seed, tid_lo=0, tid_hi=None, hit_pct=50):
    """Seeded Order-Status / Stock-Level transactions beside a TpccBatch b:
    hit_pct % of them read rows b's NewOrders write (the order, one of its
    order lines, the stock row of one of its lines), the rest uniform rows
    (Stock-Level items skewed as the NewOrder items: OR of two uniform
    draws).  TIDs are drawn in [tid_lo, tid_hi), so their priorities
    interleave with the batch's."""
    rng = np.random.default_rng(seed)
    W = warehouses
    tid_hi = tid_lo + b.n_no + b.n_pay + n_os + n_sl if tid_hi is None else tid_hi
    ro = ReadOnlyBatch(n_os, n_sl)
    nos = b.n_no
    hit_os = (rng.random(n_os) < hit_pct / 100.0) & (nos > 0)
    q = rng.integers(0, max(nos, 1), n_os)
    nl = np.maximum(b.nf("O_OL_CNT").astype(np.int64), 1) if nos else np.ones(1, np.int64)
    line = (rng.random(n_os) * nl[q]).astype(np.int64)
    uni = lambda lo, hi, n: rng.integers(lo, hi, n).astype(np.uint32)
    if n_os:
        ro.os["W_ID"][:n_os] = np.where(hit_os, b.nf("W_ID")[q] if nos else 0, uni(0, W, n_os))
        ro.os["D_ID"][:n_os] = np.where(hit_os, b.nf("D_ID")[q] if nos else 0, uni(0, 10, n_os))
        ro.os["C_ID"][:n_os] = np.where(hit_os, b.nf("C_ID")[q] if nos else 0, uni(0, 3000, n_os))
        ro.os["O_ID"][:n_os] = np.where(hit_os, b.nf("O_ID")[q] if nos else 0, uni(0, 30000, n_os))
        ro.os["OL_ID"][:n_os] = np.where(hit_os, (b.nf("O_OL_ID")[q] + line) if nos else 0, uni(0, 450000, n_os))
        ro.os["TID"][:n_os] = uni(tid_lo, tid_hi, n_os)
    if n_sl:
        hit_sl = (rng.random(n_sl) < hit_pct / 100.0) & (nos > 0)
        q2 = rng.integers(0, max(nos, 1), n_sl)
        l2 = (rng.random(n_sl) * nl[q2]).astype(np.int64)
        if nos:
            sw = np.array([b.nf("OL_SUPPLY_W_ID", int(k))[int(t)] for t, k in zip(q2, l2)], dtype=np.uint32)
            it = np.array([b.nf("OL_I_ID", int(k))[int(t)] for t, k in zip(q2, l2)], dtype=np.uint32)
        else:
            sw = it = np.zeros(n_sl, np.uint32)
        skew = ((rng.integers(0, 8192, n_sl) | rng.integers(1, 100001, n_sl)) % 100000).astype(np.uint32)
        ro.sl["W_ID"][:n_sl] = np.where(hit_sl, sw, uni(0, W, n_sl))
        ro.sl["S_ID"][:n_sl] = np.where(hit_sl, it, skew)
        ro.sl["TID"][:n_sl] = uni(tid_lo, tid_hi, n_sl)
    return ro


def upload_readonly_batch(ctx, ro):
    """Copy a ReadOnlyBatch into device buffers of ctx; returns
    ((OrderstatusBatch, StocklevelBatch) of device pointers, [buffers])."""
    bufs = []
    ob, sb = _lib.OrderstatusBatch(), _lib.StocklevelBatch()
    ob.n, sb.n = ro.n_os, ro.n_sl
    for st, arrays, fields in ((ob, ro.os, ORDERSTATUS_FIELDS), (sb, ro.sl, STOCKLEVEL_FIELDS)):
        for f in fields:
            buf = DeviceBuffer(ctx, max(arrays[f].nbytes, 16))
            buf.upload(arrays[f])
            bufs.append(buf)
            setattr(st, f, buf.ptr)
    return (ob, sb), bufs


def delivery_batch(b, warehouses, n, seed, tid_lo=0, tid_hi=None, hit_pct=50):
    """Seeded Deliveries beside a TpccBatch b.  hit_pct % of them take an
    O_ID equal to a row b's NewOrders write (an ORDER row, or an ORDERLINE
    row of one of its lines) or b's Payments write (a CUSTOMER row below
    the ORDER table's size), and half of those also delete the NEWORDER row
    a NewOrder inserts; the rest draw O_ID uniformly below W * 30000,
    NO_W_ID below W and NO_O_ID below 30000.  TIDs are drawn in
    [tid_lo, tid_hi), so priorities interleave with the batch's."""
    rng = np.random.default_rng(seed)
    W = warehouses
    tid_hi = tid_lo + b.n_no + b.n_pay + n if tid_hi is None else tid_hi
    dl = DeliveryBatch(n)
    if not n:
        return dl
    rows = W * 30000
    cand = []
    if b.n_no:
        w = b.nf("W_ID").astype(np.int64)
        cand.append(w * 30000 + b.nf("O_ID"))
        nl = np.maximum(np.minimum(b.nf("O_OL_CNT").astype(np.int64), 15), 1)
        cand.append(w * 450000 + b.nf("O_OL_ID") + (rng.random(b.n_no) * nl).astype(np.int64))
    if b.n_pay:
        cand.append(b.pf("C_W_ID").astype(np.int64) * 30000 + b.pf("C_D_ID").astype(np.int64) * 3000 +
                    b.pf("C_ID"))
    pool = np.concatenate(cand) if cand else np.zeros(0, np.int64)
    pool = pool[pool < rows]
    hit = (rng.random(n) < hit_pct / 100.0) & (pool.size > 0)
    pick = pool[rng.integers(0, max(pool.size, 1), n)] if pool.size else np.zeros(n, np.int64)
    dl.O_ID[:n] = np.where(hit, pick, rng.integers(0, rows, n)).astype(np.uint32)
    del_hit = hit & (rng.random(n) < 0.5) & (b.n_no > 0)
    q = rng.integers(0, max(b.n_no, 1), n)
    dl.NO_W_ID[:n] = np.where(del_hit, b.nf("W_ID")[q] if b.n_no else 0, rng.integers(0, W, n)).astype(np.uint32)
    dl.NO_O_ID[:n] = np.where(del_hit, b.nf("N_O_ID")[q] if b.n_no else 0,
                              rng.integers(0, 30000, n)).astype(np.uint32)
    dl.TID[:n] = rng.integers(tid_lo, tid_hi, n).astype(np.uint32)
    return dl


def upload_delivery_batch(ctx, dl):
    """Copy a DeliveryBatch into device buffers of ctx; returns
    (DeliveryBatch struct of device pointers, [buffers])."""
    st, bufs = _lib.DeliveryBatch(), []
    st.n = dl.n
    for f in DELIVERY_FIELDS:
        a = getattr(dl, f)
        buf = DeviceBuffer(ctx, max(a.nbytes, 16))
        buf.upload(a)
        bufs.append(buf)
        setattr(st, f, buf.ptr)
    return st, bufs


def upload_tpcc_batch(ctx, b):
    """Copy a TpccBatch into device buffers of ctx; returns
    ((NeworderBatch, PaymentBatch) of device pointers, [buffers])."""
    nob, pab = DeviceBuffer(ctx, max(b.no.nbytes, 16)), DeviceBuffer(ctx, max(b.pay.nbytes, 16))
    nob.upload(b.no)
    pab.upload(b.pay)
    hn, hp = neworder_struct(b.no, b.no_cap, b.n_no), payment_struct(b.pay, b.pay_cap, b.n_pay)
    nb, pb = NeworderBatch(), PaymentBatch()
    nb.n, nb.line_stride, pb.n = hn.n, hn.line_stride, hp.n
    for name, _ in NeworderBatch._fields_[2:]:
        setattr(nb, name, nob.ptr + (getattr(hn, name) - b.no.ctypes.data))
    for name, _ in PaymentBatch._fields_[2:]:
        setattr(pb, name, pab.ptr + (getattr(hp, name) - b.pay.ctypes.data))
    return (nb, pb), [nob, pab]


class Batched:
    """ltpg_batched: a list of device-resident batches of both workloads run
    as one submission (RUN_GRAPH) or one epoch at a time (RUN_PLAIN)."""

    def __init__(self, kv_ctx, tpcc_ctx, descs):
        arr = (BatchDesc * max(len(descs), 1))(*descs)
        h = ctypes.c_void_p()
        _check(lib.ltpg_batched_create(kv_ctx.handle if kv_ctx else None, tpcc_ctx.handle if tpcc_ctx else None,
                                       arr, len(descs), ctypes.byref(h)), "ltpg_batched_create")
        self.handle = h
        self.descs = list(descs)

    def run(self, mode=RUN_GRAPH):
        _check(lib.ltpg_batched_run(self.handle, mode), "ltpg_batched_run")

    def times(self):
        """Completion time (ms from the run's start) of every batch of the last run."""
        out = np.zeros(max(len(self.descs), 1), dtype=np.float64)
        _check(lib.ltpg_batched_times(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      len(self.descs)), "ltpg_batched_times")
        return out[:len(self.descs)]

    def autotune(self, reps=3):
        """Pick the launch shape per workload and size class; returns one
        dict per class (workload, log2 size, chosen TUNE_* flags, ms of every
        shape tried: key-value at tune_index(flags), order-entry [0] and [1])."""
        out = (_lib.TuningEntry * 66)()
        n = ctypes.c_uint32()
        _check(lib.ltpg_batched_autotune(self.handle, reps, out, 66, ctypes.byref(n)), "ltpg_batched_autotune")
        return [{"workload": e.workload, "log2_txns": e.log2_txns, "flags": e.flags, "ms": list(e.ms)}
                for e in out[:n.value]]

    def set_tuning(self, flags):
        """One launch shape for every batch (recaptures): the key-value bits
        for key-value batches, TUNE_TPCC_SMALL_BLOCKS for order-entry ones."""
        _check(lib.ltpg_batched_set_tuning(self.handle, flags), "ltpg_batched_set_tuning")

    def bitmap(self, i, type_):
        d = self.descs[i]
        n = {TYPE_KV: d.kv.n_txns, TYPE_NEWORDER: d.neworder.n, TYPE_PAYMENT: d.payment.n,
             TYPE_ORDERSTATUS: d.orderstatus.n, TYPE_STOCKLEVEL: d.stocklevel.n, TYPE_DELIVERY: d.delivery.n}[type_]
        out = np.zeros(max(1, (n + 63) // 64), dtype=np.uint64)
        _check(lib.ltpg_batched_bitmap(self.handle, i, type_, out.ctypes.data_as(_lib.c_u64p), out.size),
               "ltpg_batched_bitmap")
        return out[:(n + 63) // 64]

    def close(self):
        if self.handle:
            lib.ltpg_batched_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def kv_desc(batch):
    """BatchDesc of a DeviceKvBatch."""
    d = BatchDesc()
    d.workload = WORKLOAD_KV
    d.kv = batch.struct()
    return d


def tpcc_desc(dev, ro_dev=None, dl_dev=None):
    """BatchDesc of an uploaded order-entry batch ((nb, pb) of
    upload_tpcc_batch), with read-only transactions ((ob, sb) of
    upload_readonly_batch) and Deliveries (a struct of
    upload_delivery_batch) if given."""
    d = BatchDesc()
    d.workload = WORKLOAD_TPCC
    d.neworder, d.payment = dev
    if ro_dev is not None:
        d.orderstatus, d.stocklevel = ro_dev
    if dl_dev is not None:
        d.delivery = dl_dev
    return d


class Server:
    """ltpg_serve_*: continuous admission of arriving requests into epochs
    over resident tables; aborted requests re-queue with their TID."""

    def __init__(self, kv_ctx=None, tpcc_ctx=None, **kw):
        cfg = ServeConfig()
        lib.ltpg_serve_config_init(ctypes.byref(cfg))
        for k, v in kw.items():
            if not hasattr(cfg, k):
                raise AttributeError(f"ltpg_serve_config has no field {k}")
            setattr(cfg, k, v)
        self.config = cfg
        h = ctypes.c_void_p()
        _check(lib.ltpg_serve_start(kv_ctx.handle if kv_ctx else None, tpcc_ctx.handle if tpcc_ctx else None,
                                    ctypes.byref(cfg), ctypes.byref(h)), "ltpg_serve_start")
        self.handle = h

    def submit_kv(self, keys, meta):
        """Requests request-major: keys/meta of shape (n, kv_ops); returns the first id."""
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        meta = np.ascontiguousarray(meta, dtype=np.uint32)
        n = keys.shape[0]
        first = ctypes.c_uint64()
        _check(lib.ltpg_serve_submit_kv(self.handle, n, keys.ctypes.data_as(_lib.c_u64p),
                                        meta.ctypes.data_as(_lib.c_u32p), ctypes.byref(first)), "ltpg_serve_submit_kv")
        return first.value

    def submit_tpcc(self, batch):
        """Requests from a TpccBatch (NewOrders get the first ids, then Payments)."""
        nb = neworder_struct(batch.no, batch.no_cap, batch.n_no)
        pb = payment_struct(batch.pay, batch.pay_cap, batch.n_pay)
        first = ctypes.c_uint64()
        _check(lib.ltpg_serve_submit_tpcc(self.handle, ctypes.byref(nb), ctypes.byref(pb), ctypes.byref(first)),
               "ltpg_serve_submit_tpcc")
        return first.value

    def submit_readonly(self, ro):
        """Requests from a ReadOnlyBatch (Order-Status get the first ids, then Stock-Level)."""
        ob, sb = ro.structs()
        first = ctypes.c_uint64()
        _check(lib.ltpg_serve_submit_readonly(self.handle, ctypes.byref(ob), ctypes.byref(sb), ctypes.byref(first)),
               "ltpg_serve_submit_readonly")
        return first.value

    def submit_delivery(self, dl):
        """Requests from a DeliveryBatch (their TIDs are not taken)."""
        st = dl.struct()
        first = ctypes.c_uint64()
        _check(lib.ltpg_serve_submit_delivery(self.handle, ctypes.byref(st), ctypes.byref(first)),
               "ltpg_serve_submit_delivery")
        return first.value

    def poll(self, max_n=1 << 16):
        out = (ServeDone * max_n)()
        n = ctypes.c_uint32()
        _check(lib.ltpg_serve_poll(self.handle, out, max_n, ctypes.byref(n)), "ltpg_serve_poll")
        return [{"id": d.id, "type": d.type, "tid": d.tid, "seq": (d.tid_hi << 32) | d.tid, "attempts": d.attempts,
                 "epoch": d.epoch, "arrival_ns": d.arrival_ns, "commit_ns": d.commit_ns} for d in out[:n.value]]

    def read_values(self, rid, max_words=1 << 16):
        out = np.zeros(max_words, dtype=np.uint32)
        n = ctypes.c_uint32()
        _check(lib.ltpg_serve_read_values(self.handle, rid, out.ctypes.data_as(_lib.c_u32p), max_words,
                                          ctypes.byref(n)), "ltpg_serve_read_values")
        return out[:n.value]

    def epoch(self, workload, e, max_n=1 << 24):
        n = ctypes.c_uint32()
        _check(lib.ltpg_serve_epoch(self.handle, workload, e, None, 0, ctypes.byref(n)), "ltpg_serve_epoch")
        ids = np.zeros(max(n.value, 1), dtype=np.uint64)
        _check(lib.ltpg_serve_epoch(self.handle, workload, e, ids.ctypes.data_as(_lib.c_u64p), n.value,
                                    ctypes.byref(n)), "ltpg_serve_epoch")
        return ids[:n.value]

    def drain(self):
        _check(lib.ltpg_serve_drain(self.handle), "ltpg_serve_drain")

    def stats(self):
        st = ServeStats()
        _check(lib.ltpg_serve_stats_get(self.handle, ctypes.byref(st)), "ltpg_serve_stats_get")
        return {"committed": list(st.committed), "attempts": list(st.attempts), "epochs": list(st.epochs),
                "max_epoch_txns": list(st.max_epoch_txns)}

    def device_stats(self):
        """The key-value device loop as the device measured it (empty
        without one)."""
        st = _lib.ServeDeviceStats()
        st.size = ctypes.sizeof(st)
        _check(lib.ltpg_serve_device_stats_get(self.handle, ctypes.byref(st)), "ltpg_serve_device_stats_get")
        parts = ["control", "admit", "execute", "mark", "check", "unused", "write", "publish", "requeue", "idle"]
        return {"kv_loop": bool(st.kv_loop), "launches": st.launches, "epochs": st.epochs, "busy_ns": st.busy_ns,
                "exits": dict(zip(["other", "stop", "idle", "stall", "slice"], list(st.exits))),
                "profile_ns": dict(zip(parts, list(st.prof_ns))), "profile_epochs": st.prof_epochs}

    def stop(self):
        if self.handle:
            h, self.handle = self.handle, None
            _check(lib.ltpg_serve_stop(h), "ltpg_serve_stop")

    def cancel(self):
        """Drop what has not been admitted yet, finish the epochs in flight,
        stop; returns the number of requests that never committed."""
        dropped = ctypes.c_uint64()
        if self.handle:
            h, self.handle = self.handle, None
            _check(lib.ltpg_serve_cancel(h, ctypes.byref(dropped)), "ltpg_serve_cancel")
        return dropped.value

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass


def kv_generate(seed=0x4C545047, rows=1_000_000, fields=10, n_txns=65536, n_slots=10, min_ops=None,
                max_ops=None, read_pct=50, theta=0.8, tid_base=0, distinct_keys=True, txn_base=0,
                threads=0, scan_len=0):
    """Seeded key-value batch (slot-major keys/meta, tids) from libltpg's
    generator; with scan_len > 0 the reads are scans of that many rows."""
    g = KvGen()
    lib.ltpg_kv_gen_init(ctypes.byref(g))
    g.scan_len = scan_len
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
