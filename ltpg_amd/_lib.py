"""ctypes declarations for libltpg.so (include/ltpg.h).

The shared library is the product: this module only loads it and declares
the C signatures.  Importing fails loudly when the library has not been built
(there is no Python or CPU fallback for the hot path).
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libltpg.so")

c_u8p = ctypes.POINTER(ctypes.c_uint8)
c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_u64p = ctypes.POINTER(ctypes.c_uint64)


class Config(ctypes.Structure):
    """struct ltpg_config."""

    _fields_ = [
        ("size", ctypes.c_uint32),
        ("device", ctypes.c_int32),
        ("workload", ctypes.c_uint32),
        ("rules", ctypes.c_uint32),
        ("values", ctypes.c_uint32),
        ("max_txns", ctypes.c_uint32),
        ("kv_rows", ctypes.c_uint64),
        ("kv_fields", ctypes.c_uint32),
        ("kv_field_bytes", ctypes.c_uint32),
        ("kv_max_slots", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("kv_hot_row", ctypes.c_uint64),
        ("tpcc_warehouses", ctypes.c_uint32),
        ("kv_max_scan", ctypes.c_uint32),
        ("hbm_budget", ctypes.c_uint64),
    ]


class KvBatch(ctypes.Structure):
    """struct ltpg_kv_batch (slot-major, slot s of txn t at [s*B + t])."""

    _fields_ = [
        ("n_txns", ctypes.c_uint32),
        ("n_slots", ctypes.c_uint32),
        ("keys", ctypes.c_void_p),
        ("meta", ctypes.c_void_p),
        ("tids", ctypes.c_void_p),
    ]


class KvGen(ctypes.Structure):
    """struct ltpg_kv_gen."""

    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("rows", ctypes.c_uint64),
        ("fields", ctypes.c_uint32),
        ("n_txns", ctypes.c_uint32),
        ("n_slots", ctypes.c_uint32),
        ("min_ops", ctypes.c_uint32),
        ("max_ops", ctypes.c_uint32),
        ("read_pct", ctypes.c_uint32),
        ("theta", ctypes.c_double),
        ("tid_base", ctypes.c_uint32),
        ("distinct_keys", ctypes.c_uint32),
        ("txn_base", ctypes.c_uint64),
        ("scan_len", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class NeworderBatch(ctypes.Structure):
    """struct ltpg_neworder_batch (line l of txn q at [l * line_stride + q])."""

    _fields_ = [("n", ctypes.c_uint32), ("line_stride", ctypes.c_uint32)] + [
        (name, ctypes.c_void_p) for name in ("W_ID", "D_ID", "C_ID", "O_ID", "N_O_ID", "O_OL_CNT", "O_OL_ID",
                                             "OL_I_ID", "OL_SUPPLY_W_ID", "OL_QUANTITY", "TID")]


class PaymentBatch(ctypes.Structure):
    """struct ltpg_payment_batch."""

    _fields_ = [("n", ctypes.c_uint32), ("reserved", ctypes.c_uint32)] + [
        (name, ctypes.c_void_p) for name in ("W_ID", "D_ID", "C_ID", "C_LAST", "isName", "C_D_ID", "C_W_ID",
                                             "H_AMOUNT", "H_ID", "TID")]


class OrderstatusBatch(ctypes.Structure):
    """struct ltpg_orderstatus_batch."""

    _fields_ = [("n", ctypes.c_uint32), ("reserved", ctypes.c_uint32)] + [
        (name, ctypes.c_void_p) for name in ("W_ID", "D_ID", "C_ID", "O_ID", "OL_ID", "TID")]


class StocklevelBatch(ctypes.Structure):
    """struct ltpg_stocklevel_batch."""

    _fields_ = [("n", ctypes.c_uint32), ("reserved", ctypes.c_uint32)] + [
        (name, ctypes.c_void_p) for name in ("W_ID", "S_ID", "TID")]


class DeliveryBatch(ctypes.Structure):
    """struct ltpg_delivery_batch."""

    _fields_ = [("n", ctypes.c_uint32), ("reserved", ctypes.c_uint32)] + [
        (name, ctypes.c_void_p) for name in ("O_ID", "NO_W_ID", "NO_O_ID", "TID")]


class TpccGenConfig(ctypes.Structure):
    """struct ltpg_tpcc_gen_config."""

    _fields_ = [("seed", ctypes.c_uint64)] + [
        (name, ctypes.c_uint32) for name in ("warehouses", "neworder_pct", "remote_supply_pct",
                                             "remote_customer_pct", "byname_pct", "nurand_items", "min_lines",
                                             "max_lines", "max_amount", "max_quantity")]


class BatchDesc(ctypes.Structure):
    """struct ltpg_batch_desc."""

    _fields_ = [("workload", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("kv", KvBatch),
                ("neworder", NeworderBatch), ("payment", PaymentBatch), ("orderstatus", OrderstatusBatch),
                ("stocklevel", StocklevelBatch), ("delivery", DeliveryBatch)]


class TuningEntry(ctypes.Structure):
    """struct ltpg_tuning_entry."""

    _fields_ = [("workload", ctypes.c_uint32), ("log2_txns", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("ms", ctypes.c_float * 64)]


class ServeConfig(ctypes.Structure):
    """struct ltpg_serve_config."""

    _fields_ = [("size", ctypes.c_uint32), ("slots", ctypes.c_uint32), ("kv_ops", ctypes.c_uint32),
                ("max_epoch_txns", ctypes.c_uint32), ("hbm_budget", ctypes.c_uint64), ("keep_reads", ctypes.c_uint32),
                ("record_epochs", ctypes.c_uint32), ("epoch_growth", ctypes.c_uint32), ("plain", ctypes.c_uint32),
                ("epoch_floor", ctypes.c_uint32), ("kv_loop", ctypes.c_uint32), ("ring_requests", ctypes.c_uint32),
                ("first_seq", ctypes.c_uint64)]


class ServeDone(ctypes.Structure):
    """struct ltpg_serve_done."""

    _fields_ = [("id", ctypes.c_uint64), ("type", ctypes.c_uint32), ("tid", ctypes.c_uint32),
                ("attempts", ctypes.c_uint32), ("tid_hi", ctypes.c_uint32), ("epoch", ctypes.c_uint64),
                ("arrival_ns", ctypes.c_uint64), ("commit_ns", ctypes.c_uint64)]


class ServeStats(ctypes.Structure):
    """struct ltpg_serve_stats."""

    _fields_ = [("committed", ctypes.c_uint64 * 6), ("attempts", ctypes.c_uint64 * 6), ("epochs", ctypes.c_uint64 * 2),
                ("max_epoch_txns", ctypes.c_uint64 * 2)]


class ServeDeviceStats(ctypes.Structure):
    """struct ltpg_serve_device_stats."""

    _fields_ = [("size", ctypes.c_uint32), ("kv_loop", ctypes.c_uint32), ("launches", ctypes.c_uint64),
                ("epochs", ctypes.c_uint64), ("busy_ns", ctypes.c_uint64), ("exits", ctypes.c_uint64 * 5),
                ("prof_ns", ctypes.c_uint64 * 10), ("prof_epochs", ctypes.c_uint64)]


# (name, restype, argtypes) of every exported entry point.
SIGNATURES = [
    ("ltpg_config_init", None, [ctypes.POINTER(Config)]),
    ("ltpg_create", ctypes.c_int, [ctypes.POINTER(Config), ctypes.POINTER(ctypes.c_void_p)]),
    ("ltpg_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("ltpg_last_error", ctypes.c_char_p, []),
    ("ltpg_abi_version", ctypes.c_int, []),
    ("ltpg_device_count", ctypes.c_int, []),
    ("ltpg_set_stream", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("ltpg_sync", ctypes.c_int, [ctypes.c_void_p]),
    ("ltpg_device_alloc", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
    ("ltpg_device_free", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("ltpg_copy_to_device", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]),
    ("ltpg_copy_to_host", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]),
    ("ltpg_device_copy", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]),
    ("ltpg_random_access_rates", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                                 ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    ("ltpg_kv_table_load", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32]),
    ("ltpg_kv_table_read", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32]),
    ("ltpg_kv_table_fill_reference", ctypes.c_int, [ctypes.c_void_p]),
    ("ltpg_kv_table_read_rows", ctypes.c_int, [ctypes.c_void_p, c_u64p, ctypes.c_uint64, c_u32p]),
    ("ltpg_kv_epoch_noalloc", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(KvBatch), ctypes.c_uint32,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("ltpg_kv_set_tuning", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    ("ltpg_tpcc_set_tuning", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    ("ltpg_kv_submit", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(KvBatch), ctypes.c_uint32]),
    ("ltpg_epoch_run", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    ("ltpg_epoch_reset", ctypes.c_int, [ctypes.c_void_p]),
    ("ltpg_commit_bitmap", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, c_u64p, ctypes.c_uint64]),
    ("ltpg_commit_count", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, c_u32p]),
    ("ltpg_conflict_counters", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, c_u32p, ctypes.c_uint32]),
    ("ltpg_kv_read_values", ctypes.c_int, [ctypes.c_void_p, c_u32p, ctypes.c_uint64]),
    ("ltpg_phase_times", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), c_u32p]),
    ("ltpg_commit_totals", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, c_u64p, c_u64p, ctypes.c_int]),
    ("ltpg_set_timing", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("ltpg_tpcc_table_shape", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, c_u64p, c_u32p]),
    ("ltpg_tpcc_table_load", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_uint32, c_u32p, ctypes.c_uint64, ctypes.c_uint32]),
    ("ltpg_tpcc_table_read", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_uint32, c_u32p, ctypes.c_uint64, ctypes.c_uint32]),
    ("ltpg_tpcc_populate", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64]),
    ("ltpg_neworder_view", None,
     [c_u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(NeworderBatch)]),
    ("ltpg_payment_view", None,
     [c_u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(PaymentBatch)]),
    ("ltpg_tpcc_submit", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(NeworderBatch), ctypes.POINTER(PaymentBatch), ctypes.c_uint32]),
    ("ltpg_tpcc_submit_readonly", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(OrderstatusBatch), ctypes.POINTER(StocklevelBatch), ctypes.c_uint32]),
    ("ltpg_latency_costs", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    ("ltpg_sync_costs", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.POINTER(ctypes.c_double)]),
    ("ltpg_random_gather_rate", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
      ctypes.POINTER(ctypes.c_double)]),
    ("ltpg_tpcc_submit_delivery", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DeliveryBatch), ctypes.c_uint32]),
    ("ltpg_tpcc_payment_cid", ctypes.c_int, [ctypes.c_void_p, c_u32p, ctypes.c_uint32]),
    ("ltpg_tpcc_read_values", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, c_u32p, ctypes.c_uint64]),
    ("ltpg_tpcc_gen_config_init", None, [ctypes.POINTER(TpccGenConfig)]),
    ("ltpg_tpcc_gen_create", ctypes.c_int, [ctypes.POINTER(TpccGenConfig), ctypes.POINTER(ctypes.c_void_p)]),
    ("ltpg_tpcc_gen_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("ltpg_tpcc_gen_next", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, c_u32p, ctypes.c_uint32, c_u32p, ctypes.c_uint32,
      c_u32p, c_u32p]),
    ("ltpg_batched_create", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(BatchDesc), ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)]),
    ("ltpg_batched_run", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    ("ltpg_batched_bitmap", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, c_u64p, ctypes.c_uint64]),
    ("ltpg_batched_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("ltpg_batched_times", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_uint32]),
    ("ltpg_batched_set_tuning", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    ("ltpg_batched_autotune", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(TuningEntry), ctypes.c_uint32, c_u32p]),
    ("ltpg_serve_config_init", None, [ctypes.POINTER(ServeConfig)]),
    ("ltpg_serve_start", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ServeConfig), ctypes.POINTER(ctypes.c_void_p)]),
    ("ltpg_serve_submit_kv", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, c_u64p, c_u32p, c_u64p]),
    ("ltpg_serve_submit_tpcc", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(NeworderBatch), ctypes.POINTER(PaymentBatch), c_u64p]),
    ("ltpg_serve_submit_readonly", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(OrderstatusBatch), ctypes.POINTER(StocklevelBatch), c_u64p]),
    ("ltpg_serve_submit_delivery", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DeliveryBatch), c_u64p]),
    ("ltpg_serve_poll", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ServeDone), ctypes.c_uint32, c_u32p]),
    ("ltpg_serve_read_values", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, c_u32p, ctypes.c_uint32, c_u32p]),
    ("ltpg_serve_epoch", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, c_u64p, ctypes.c_uint32, c_u32p]),
    ("ltpg_serve_drain", ctypes.c_int, [ctypes.c_void_p]),
    ("ltpg_serve_stats_get", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ServeStats)]),
    ("ltpg_serve_device_stats_get", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ServeDeviceStats)]),
    ("ltpg_serve_stop", ctypes.c_int, [ctypes.c_void_p]),
    ("ltpg_serve_cancel", ctypes.c_int, [ctypes.c_void_p, c_u64p]),
    ("ltpg_kv_gen_init", None, [ctypes.POINTER(KvGen)]),
    ("ltpg_kv_generate", ctypes.c_int, [ctypes.POINTER(KvGen), c_u64p, c_u32p, c_u32p, ctypes.c_uint32]),
]


def load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C ltpg_amd/csrc` "
            "(or __graft_entry__.build()); there is no fallback for the HIP path")
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = load()
