"""Key-value parity tests: the HIP path (through the C ABI) against the oracle.

Bar: bit-exact commit bitmap, conflict counters, read values of committed
transactions and final table on the same seeded batch, and identical results
across repeated runs.
"""
import numpy as np
import pytest

import ltpg_amd as L
from oracle import oracle
from tests import kvcases as kc

pytestmark = pytest.mark.gpu

KAT = kc.load_kat()


def _ctx(rows, fields, fw, max_txns, max_slots, values=L.VALUES_RMW, counts=True, hot_row=L.NO_HOT_ROW,
         max_scan=0):
    return L.Context(workload=L.WORKLOAD_KV, kv_rows=rows, kv_fields=fields, kv_field_bytes=4 * fw,
                     max_txns=max_txns, kv_max_slots=max_slots, values=values,
                     flags=L.FLAG_CONFLICT_COUNTS if counts else 0, kv_hot_row=hot_row, kv_max_scan=max_scan)


def _request_major(a, S):
    """A slot-major [S * B] array as request-major [B * S]."""
    return np.ascontiguousarray(a.reshape(S, -1).T).reshape(-1)


def _gpu_epoch(ctx, keys, meta, tids, S, device=False, rm=False):
    """One epoch of a slot-major batch; with rm it goes over request-major."""
    if rm:
        keys, meta = _request_major(keys, S), _request_major(meta, S)
    if device:
        b = L.DeviceKvBatch(ctx, keys, meta, tids, S, request_major=rm)
        ctx.kv_submit_device(b)
    else:
        ctx.kv_submit(keys, meta, tids, S, request_major=rm)
    ctx.epoch_run(L.PHASE_ALL)
    ctx.sync()
    counts = bool(ctx.config.flags & L.FLAG_CONFLICT_COUNTS)
    out = dict(bitmap=ctx.commit_bitmap(), conflicts=ctx.conflict_counters() if counts else None,
               reads=ctx.kv_read_values(), count=ctx.commit_count(), table=ctx.kv_table_read())
    if device:
        b.free()
    return out


def _assert_reads(gpu_reads, ora_reads, meta, tids, fw, bitmap):
    """Read values must match on the read slots of committed transactions."""
    B = len(tids)
    S = meta.size // max(B, 1)
    committed = np.unpackbits(bitmap.view(np.uint8), bitorder="little")[:B].astype(bool)
    mask = ((meta.reshape(S, B) & 3) == 1) & committed[None, :]
    mask = np.repeat(mask.reshape(-1), fw)
    np.testing.assert_array_equal(gpu_reads[mask], ora_reads[mask])


def _assert_epoch(g, ora, meta, tids, fw):
    np.testing.assert_array_equal(g["bitmap"], ora.bitmap)
    if g["conflicts"] is not None:
        np.testing.assert_array_equal(g["conflicts"], ora.conflicts)
    assert g["count"] == ora.commit_count
    np.testing.assert_array_equal(g["table"], ora.table)
    _assert_reads(g["reads"], ora.read_values, meta, tids, fw, ora.bitmap)


@pytest.mark.parametrize("case", KAT["cases"], ids=[c["name"] for c in KAT["cases"]])
@pytest.mark.parametrize("form", ["counting", "fast", "small-epoch"])
def test_gpu_known_answers(gpu, case, form):
    counts = form == "counting"
    rows, fields, fw = case["rows"], case["fields"], case["field_words"]
    keys, meta, tids, S = kc.case_arrays(case)
    values = L.VALUES_RMW if case["values"] == "rmw" else L.VALUES_COPYBACK
    B = len(tids)
    hot = case.get("hot_row", L.NO_HOT_ROW)
    max_scan = case.get("max_scan", 0)
    Ls = max(1, max_scan)
    with _ctx(rows, fields, fw, max(B, 1), S, values, counts=counts, hot_row=hot, max_scan=max_scan) as ctx:
        ctx.kv_table_load(kc.initial_table(rows, fields, fw))
        if form == "small-epoch":
            ctx.kv_set_tuning(L.TUNE_SMALL_EPOCH)
        ctx.kv_submit(keys, meta, tids, S)
        ctx.epoch_run(L.PHASE_ALL)
        exp = case["expect"]
        if exp["range_errors"]:
            with pytest.raises(L.LtpgError) as ei:
                ctx.sync()
            assert ei.value.code == L.ERR_RANGE
        else:
            ctx.sync()
        assert kc.bitmap_to_list(ctx.commit_bitmap(), B) == exp["commit"]
        if counts:
            assert kc.unpack_conflicts(ctx.conflict_counters()).tolist() == exp["conflicts"]
        else:
            with pytest.raises(L.LtpgError) as ei:
                ctx.conflict_counters()
            assert ei.value.code == L.ERR_UNSUPPORTED
        assert ctx.commit_count() == sum(exp["commit"])
        np.testing.assert_array_equal(ctx.kv_table_read(), kc.expected_table(case))
        reads = ctx.kv_read_values()
        for t, s, words in exp["reads"]:
            i = (s * B + t) * Ls
            assert reads[i * fw:(i + 1) * fw].tolist() == words
        for t, s, j, words in exp.get("scan_reads", []):
            i = (s * B + t) * Ls + j
            assert reads[i * fw:(i + 1) * fw].tolist() == words


CASES = [
    # rows, fields, fw, B, S, theta, read_pct, min_ops, seed
    (1000, 4, 1, 5000, 8, 0.9, 50, 8, 1),
    (50_000, 10, 25, 20_000, 16, 0.99, 50, 16, 2),   # 100-byte fields, stride 64 words
    (50_000, 10, 25, 20_000, 16, 0.6, 50, 16, 3),
    (4096, 3, 7, 777, 64, 0.8, 30, 1, 4),            # ragged 1..64 ops, odd field width
    (100_000, 2, 80, 3000, 5, 0.5, 70, 2, 5),        # fields wider than 64 words
    (100, 2, 1, 65, 3, 0.0, 50, 3, 6),               # uniform keys, partial last wave
    (20_000, 3, 16, 4000, 10, 0.7, 50, 10, 7),       # exactly 64-byte fields
]


@pytest.mark.parametrize("rows,fields,fw,B,S,theta,read_pct,min_ops,seed", CASES)
@pytest.mark.parametrize("device,rm", [(False, False), (True, False), (True, True)],
                         ids=["host-batch", "device-batch", "device-request-major"])
@pytest.mark.parametrize("counts", [True, False], ids=["counting", "fast"])
def test_gpu_matches_oracle_seeded(gpu, rows, fields, fw, B, S, theta, read_pct, min_ops, seed, device, rm,
                                   counts):
    keys, meta, tids = L.kv_generate(seed=seed, rows=rows, fields=fields, n_txns=B, n_slots=S,
                                     min_ops=min_ops, max_ops=S, read_pct=read_pct, theta=theta)
    table0 = np.random.default_rng(seed).integers(0, 2**32, size=rows * fields * fw, dtype=np.uint64)
    table0 = table0.astype(np.uint32)
    ora = oracle.kv_epoch(table0, rows, fields, fw, keys, meta, tids, S, values=1)
    with _ctx(rows, fields, fw, B, S, counts=counts) as ctx:
        ctx.kv_table_load(table0)
        g = _gpu_epoch(ctx, keys, meta, tids, S, device=device, rm=rm)
    _assert_epoch(g, ora, meta, tids, fw)


@pytest.mark.parametrize("rm", [False, True], ids=["slot-major", "request-major"])
@pytest.mark.parametrize("counts", [True, False], ids=["counting", "fast"])
def test_gpu_empty_epochs_between_batches(gpu, counts, rm):
    """An epoch of no transactions commits nothing, changes no row and starts
    its commit count at zero; the epochs around it match the oracle."""
    rows, fields, fw, S = 5000, 4, 2, 8
    table = np.random.default_rng(11).integers(0, 2**32, size=rows * fields * fw, dtype=np.uint64).astype(np.uint32)
    empty = (np.zeros(0, np.uint64), np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    with _ctx(rows, fields, fw, 4000, S, counts=counts) as ctx:
        ctx.kv_table_load(table)
        for ep, B in enumerate([4000, 0, 0, 3000, 0]):
            if B:
                keys, meta, tids = L.kv_generate(seed=20 + ep, rows=rows, fields=fields, n_txns=B, n_slots=S,
                                                 min_ops=1, max_ops=S, read_pct=50, theta=0.8, tid_base=ep * 4000)
            else:
                keys, meta, tids = empty
            ora = oracle.kv_epoch(table, rows, fields, fw, keys, meta, tids, S, values=1)
            g = _gpu_epoch(ctx, keys, meta, tids, S, rm=rm)
            _assert_epoch(g, ora, meta, tids, fw)
            if B == 0:
                assert g["count"] == 0 and g["bitmap"].size == 0
            else:
                assert g["count"] > 0
            table = ora.table


@pytest.mark.parametrize("rm", [False, True], ids=["slot-major", "request-major"])
@pytest.mark.parametrize("hot_row", [0, 17])
@pytest.mark.parametrize("counts", [True, False], ids=["counting", "fast"])
def test_gpu_hot_row_reduce_matches_oracle(gpu, hot_row, counts, rm):
    """Ops on the hot row never register or conflict; committed writes to it
    add (TID|1) per word through the striped reduce.  Several epochs in place
    at Zipf 0.99 (row 0 is the most drawn key), 100-byte fields and a field
    count above the reduce's LDS bins per block pass."""
    rows, fields, fw, B, S = 20_000, 10, 25, 8000, 16
    rng = np.random.default_rng(hot_row + 3)
    table = rng.integers(0, 2**32, size=rows * fields * fw, dtype=np.uint64).astype(np.uint32)
    with _ctx(rows, fields, fw, B, S, counts=counts, hot_row=hot_row) as ctx:
        ctx.kv_table_load(table)
        for ep in range(4):
            n = B - ep * 1500
            keys, meta, tids = L.kv_generate(seed=70 + ep, rows=rows, fields=fields, n_txns=n, n_slots=S,
                                             min_ops=1, max_ops=S, theta=0.99, tid_base=ep * B)
            g = _gpu_epoch(ctx, keys, meta, tids, S, rm=rm)
            ora = oracle.kv_epoch(table, rows, fields, fw, keys, meta, tids, S, values=1, hot_row=hot_row)
            _assert_epoch(g, ora, meta, tids, fw)
            hot_ops = (keys.reshape(S, n) == hot_row) & ((meta.reshape(S, n) & 3) != 0)
            assert hot_ops.any()
            table = ora.table


@pytest.mark.parametrize("rm", [False, True], ids=["slot-major", "request-major"])
def test_gpu_small_epochs_across_tag_wraps(gpu, rm):
    """810 epochs through three epoch-tag wraps: single-workgroup epochs
    throughout, and separate-kernel epochs at the same tag in three cycles
    (submits 45, 300 and 810) with one cycle that has none.  The wrap
    refill, skipped when only single-workgroup epochs ran since the last
    one, never leaves an older epoch's register looking current.  Every
    epoch against the oracle."""
    rows, fields, fw, S = 300, 2, 1, 8
    table = kc.initial_table(rows, fields, fw)
    big_at = {45, 300, 810}
    with _ctx(rows, fields, fw, 400, S, counts=False) as ctx:
        ctx.kv_table_load(table)
        ctx.kv_set_tuning(L.TUNE_SMALL_EPOCH)
        for k in range(1, 811):
            n = 400 if k in big_at else 24
            keys, meta, tids = L.kv_generate(seed=1000 + k, rows=rows, fields=fields, n_txns=n, n_slots=S,
                                             min_ops=1, max_ops=S, theta=0.8, tid_base=k * 1000)
            g = _gpu_epoch(ctx, keys, meta, tids, S, rm=rm)
            ora = oracle.kv_epoch(table, rows, fields, fw, keys, meta, tids, S, values=1)
            np.testing.assert_array_equal(g["bitmap"], ora.bitmap, err_msg=f"submit {k}")
            np.testing.assert_array_equal(g["table"], ora.table, err_msg=f"submit {k}")
            table = ora.table


@pytest.mark.parametrize("values", [L.VALUES_RMW, L.VALUES_COPYBACK], ids=["rmw", "copyback"])
@pytest.mark.parametrize("rm", [False, True], ids=["slot-major", "request-major"])
@pytest.mark.parametrize("hot_row", [0, 17])
def test_gpu_small_epochs_with_hot_row_match_oracle(gpu, hot_row, rm, values):
    """The single-workgroup epoch with a hot row: its ops never register or
    conflict, committed writes to it add (TID|1) per word; 100-byte fields,
    Zipf 0.99 (row 0 in most transactions), scans over the hot row, epochs
    of 32 x 64 slots in place with a larger one in between."""
    rows, fields, fw, S, max_scan = 20_000, 10, 25, 64, 4
    rng = np.random.default_rng(hot_row + 11)
    table = rng.integers(0, 2**32, size=rows * fields * fw, dtype=np.uint64).astype(np.uint32)
    with _ctx(rows, fields, fw, 64, S, values, counts=False, hot_row=hot_row, max_scan=max_scan) as ctx:
        ctx.kv_table_load(table)
        ctx.kv_set_tuning(L.TUNE_SMALL_EPOCH)
        for ep, n in enumerate([32, 32, 64, 20]):
            keys, meta, tids = L.kv_generate(seed=90 + ep, rows=rows, fields=fields, n_txns=n, n_slots=S,
                                             min_ops=1, max_ops=S, theta=0.99, tid_base=ep * 100,
                                             scan_len=3 if ep == 3 else 0)
            if hot_row:  # put the hot row into some slots
                keys = np.where((np.arange(keys.size) % 7) == 0, hot_row, keys).astype(np.uint64)
            g = _gpu_epoch(ctx, keys, meta, tids, S, rm=rm)
            ora = oracle.kv_epoch(table, rows, fields, fw, keys, meta, tids, S, values=values, hot_row=hot_row,
                                  max_scan=max_scan)
            np.testing.assert_array_equal(g["bitmap"], ora.bitmap)
            assert g["count"] == ora.commit_count
            np.testing.assert_array_equal(g["table"], ora.table)
            _assert_scan_reads(g["reads"], ora, keys, meta, n, S, rows, fw, max_scan)
            hot_ops = (keys.reshape(S, n) == hot_row) & ((meta.reshape(S, n) & 3) != 0)
            assert hot_ops.any()
            table = ora.table


def _assert_scan_reads(gpu_reads, ora, keys, meta, B, S, rows, fw, max_scan):
    """Delivered rows of committed reads and scans match the oracle's."""
    L_ = max(1, max_scan)
    committed = np.unpackbits(ora.bitmap.view(np.uint8), bitorder="little")[:B].astype(bool)
    kind = (meta.reshape(S, B) & 3).astype(np.int64)
    ln = np.where(kind == 1, 1, np.where(kind == 3, (meta.reshape(S, B) >> 24).astype(np.int64) + 1, 0))
    ln = np.where(ln > max(max_scan, 1), 0, ln)
    ln = np.minimum(ln, rows - keys.reshape(S, B).astype(np.int64))
    ln = np.where(committed[None, :], ln, 0)
    mask = (np.arange(L_)[None, None, :] < ln[:, :, None]).reshape(-1)
    mask = np.repeat(mask, fw)
    np.testing.assert_array_equal(gpu_reads[mask], ora.read_values[mask])


@pytest.mark.parametrize("rm", [False, True], ids=["slot-major", "request-major"])
@pytest.mark.parametrize("fw", [2, 4], ids=["8B", "16B"])
@pytest.mark.parametrize("flags", [0, 1, 4, 7, 8, 15])
@pytest.mark.parametrize("counts", [True, False], ids=["counting", "fast"])
def test_gpu_scans_match_oracle(gpu, counts, flags, fw, rm):
    """Reads as scans (YCSB/2: each read a range of rows): ragged 1..12 ops,
    90 % scans of 10 rows, kv_max_scan 16 (so read-value slots are wider than
    the scans), several epochs in place, every launch shape family, both
    batch layouts; 16-byte fields take the per-lane copy under
    TUNE_LANE_COPY."""
    rows, fields, B, S, max_scan = 5000, 3, 3000, 12, 16
    table = kc.initial_table(rows, fields, fw)
    with _ctx(rows, fields, fw, B, S, counts=counts, max_scan=max_scan) as ctx:
        ctx.kv_table_load(table)
        ctx.kv_set_tuning(flags)
        for ep_i in range(3):
            keys, meta, tids = L.kv_generate(seed=300 + ep_i, rows=rows, fields=fields, n_txns=B, n_slots=S,
                                             min_ops=1, max_ops=S, read_pct=90, theta=0.7, scan_len=10,
                                             tid_base=ep_i * B)
            g = _gpu_epoch(ctx, keys, meta, tids, S, rm=rm)
            ora = oracle.kv_epoch(table, rows, fields, fw, keys, meta, tids, S, values=1, max_scan=max_scan)
            np.testing.assert_array_equal(g["bitmap"], ora.bitmap)
            if g["conflicts"] is not None:
                np.testing.assert_array_equal(g["conflicts"], ora.conflicts)
            np.testing.assert_array_equal(g["table"], ora.table)
            _assert_scan_reads(g["reads"], ora, keys, meta, B, S, rows, fw, max_scan)
            assert 0 < ora.commit_count < B
            table = ora.table


SMALL_CASES = [
    # rows, fields, fw, B, S, theta, read_pct, min_ops, seed, scan_len, max_scan
    (1000, 4, 1, 256, 8, 0.9, 50, 8, 1, 0, 0),          # exactly LTPG_SMALL_EPOCH_OPS slots
    (50_000, 10, 25, 32, 64, 0.99, 50, 1, 2, 0, 0),     # 1-64 ops, 100-byte fields
    (4096, 3, 7, 100, 20, 0.8, 30, 1, 4, 0, 0),         # odd field width
    (100_000, 2, 80, 20, 5, 0.5, 70, 2, 5, 0, 0),       # fields wider than 64 words
    (100, 2, 1, 65, 3, 0.0, 50, 3, 6, 0, 0),            # uniform keys
    (5000, 3, 2, 150, 12, 0.7, 90, 1, 7, 10, 16),       # scans of 10 rows, read slots of 16
    (200, 2, 1, 2048, 1, 0.9, 50, 1, 8, 0, 0),          # one op per transaction
]


@pytest.mark.parametrize("rows,fields,fw,B,S,theta,read_pct,min_ops,seed,scan_len,max_scan", SMALL_CASES)
@pytest.mark.parametrize("values", [L.VALUES_RMW, L.VALUES_COPYBACK], ids=["rmw", "copyback"])
@pytest.mark.parametrize("rm", [False, True], ids=["slot-major", "request-major"])
def test_gpu_small_epochs_match_oracle(gpu, rows, fields, fw, B, S, theta, read_pct, min_ops, seed, scan_len,
                                       max_scan, values, rm):
    """TUNE_SMALL_EPOCH: epochs of at most LTPG_SMALL_EPOCH_OPS slots run as
    one single-workgroup kernel; an epoch one transaction over the limit
    between them runs the multi-kernel path on the same context.  Every
    epoch bit-exact against the oracle (bitmap, count, committed reads,
    table)."""
    big = L.SMALL_EPOCH_OPS // S + 1
    table = np.random.default_rng(seed).integers(0, 2**32, size=rows * fields * fw, dtype=np.uint64)
    table = table.astype(np.uint32)
    with _ctx(rows, fields, fw, max(B, big), S, values, counts=False, max_scan=max_scan) as ctx:
        ctx.kv_table_load(table)
        ctx.kv_set_tuning(L.TUNE_SMALL_EPOCH)
        for ep, n in enumerate([B, big, B, B]):
            keys, meta, tids = L.kv_generate(seed=seed * 10 + ep, rows=rows, fields=fields, n_txns=n, n_slots=S,
                                             min_ops=min_ops, max_ops=S, read_pct=read_pct, theta=theta,
                                             scan_len=scan_len, tid_base=ep * 5000)
            ora = oracle.kv_epoch(table, rows, fields, fw, keys, meta, tids, S, values=values, max_scan=max_scan)
            g = _gpu_epoch(ctx, keys, meta, tids, S, rm=rm)
            np.testing.assert_array_equal(g["bitmap"], ora.bitmap)
            assert g["count"] == ora.commit_count
            np.testing.assert_array_equal(g["table"], ora.table)
            _assert_scan_reads(g["reads"], ora, keys, meta, n, S, rows, fw, max_scan)
            table = ora.table


def test_gpu_scan_reference_config(gpu):
    """The scan variant's configuration (YCSB/2/Predefine.h:6-13): 1M rows x
    10 u32 fields, batch 10,000 txns of 9 scans of 10 rows + 1 write, uniform
    keys; the reference's initial table."""
    rows, fields, fw, B, S = 1_000_000, 10, 1, 10000, 10
    keys, meta, tids = L.kv_generate(seed=0x4C545047, rows=rows, fields=fields, n_txns=B, n_slots=S,
                                     read_pct=90, theta=0.0, scan_len=10)
    for counts in (True, False):
        with _ctx(rows, fields, fw, B, S, counts=counts, max_scan=10) as ctx:
            ctx.kv_fill_reference()
            table0 = ctx.kv_table_read()
            g = _gpu_epoch(ctx, keys, meta, tids, S)
        ora = oracle.kv_epoch(table0, rows, fields, fw, keys, meta, tids, S, values=1, max_scan=10)
        np.testing.assert_array_equal(g["bitmap"], ora.bitmap)
        np.testing.assert_array_equal(g["table"], ora.table)
        _assert_scan_reads(g["reads"], ora, keys, meta, B, S, rows, fw, 10)


def test_gpu_scans_without_max_scan_are_range_errors(gpu):
    rows, fields, fw, B, S = 100, 2, 1, 50, 3
    keys, meta, tids = L.kv_generate(seed=5, rows=rows, fields=fields, n_txns=B, n_slots=S, read_pct=50,
                                     theta=0.0, scan_len=4)
    n_scans = int(((meta & 3) == 3).sum())
    assert n_scans > 0
    with _ctx(rows, fields, fw, B, S, max_scan=0) as ctx:
        ctx.kv_table_load(kc.initial_table(rows, fields, fw))
        ctx.kv_submit(keys, meta, tids, S)
        ctx.epoch_run()
        with pytest.raises(L.LtpgError) as ei:
            ctx.sync()
        assert ei.value.code == L.ERR_RANGE
    ora = oracle.kv_epoch(kc.initial_table(rows, fields, fw), rows, fields, fw, keys, meta, tids, S, max_scan=0)
    assert ora.range_errors == n_scans


def test_gpu_hot_row_beyond_table_rejected(gpu):
    with pytest.raises(L.LtpgError) as ei:
        _ctx(100, 2, 1, 10, 2, hot_row=100)
    assert ei.value.code == L.ERR_INVALID


def test_gpu_config1_reference_table(gpu):
    """1M rows x 10 u32 fields, 65,536 txns x 10 ops, 50 % read / 50 % RMW,
    Zipf 0.8, seed 0x4C545047; table holds the reference's initial values."""
    rows, fields, fw, B, S = 1_000_000, 10, 1, 65536, 10
    keys, meta, tids = L.kv_generate(seed=0x4C545047, rows=rows, fields=fields, n_txns=B, n_slots=S,
                                     read_pct=50, theta=0.8)
    with _ctx(rows, fields, fw, B, S) as ctx:
        ctx.kv_fill_reference()
        table0 = ctx.kv_table_read()
        ref = np.tile(np.arange(fields, dtype=np.uint32), rows)
        np.testing.assert_array_equal(table0, ref)
        g = _gpu_epoch(ctx, keys, meta, tids, S)
    ora = oracle.kv_epoch(table0, rows, fields, fw, keys, meta, tids, S, values=1)
    _assert_epoch(g, ora, meta, tids, fw)
    assert 0 < ora.commit_count < B


def test_gpu_equal_tids_ordered_by_position(gpu):
    """Equal TIDs (outside the reference's contract) follow the build's
    (TID, batch position) priority, on both check forms."""
    rows, fields, fw, B, S = 500, 2, 2, 6000, 6
    keys, meta, _ = L.kv_generate(seed=31, rows=rows, fields=fields, n_txns=B, n_slots=S, theta=0.8)
    tids = (np.random.default_rng(31).integers(0, 50, size=B)).astype(np.uint32)
    table0 = np.arange(rows * fields * fw, dtype=np.uint32)
    ora = oracle.kv_epoch(table0, rows, fields, fw, keys, meta, tids, S)
    for counts in (True, False):
        with _ctx(rows, fields, fw, B, S, counts=counts) as ctx:
            ctx.kv_table_load(table0)
            g = _gpu_epoch(ctx, keys, meta, tids, S)
        _assert_epoch(g, ora, meta, tids, fw)


def test_gpu_epochs_do_not_leak_registers(gpu):
    """Registers are epoch-tagged instead of cleared: consecutive epochs,
    including the refill when the 8-bit epoch counter wraps, must behave
    exactly like the oracle's fresh registers."""
    rows, fields, fw, B, S = 300, 3, 2, 400, 6
    rng = np.random.default_rng(11)
    table = rng.integers(0, 2**32, size=rows * fields * fw, dtype=np.uint64).astype(np.uint32)
    with _ctx(rows, fields, fw, B, S, counts=False) as ctx:
        ctx.kv_table_load(table)
        for e in range(300):
            n = B - (e % 7) * 40
            keys, meta, tids = L.kv_generate(seed=100 + e, rows=rows, fields=fields, n_txns=n, n_slots=S,
                                             theta=0.9, tid_base=(e % 5) * 1000)
            ctx.kv_submit(keys, meta, tids, S)
            ctx.epoch_run()
            ora = oracle.kv_epoch(table, rows, fields, fw, keys, meta, tids, S, want_reads=False)
            if e % 37 == 0 or 250 <= e <= 260:
                np.testing.assert_array_equal(ctx.commit_bitmap(), ora.bitmap, err_msg=f"epoch {e}")
                np.testing.assert_array_equal(ctx.kv_table_read(), ora.table, err_msg=f"epoch {e}")
            table = ora.table
            if e == 100:
                ctx.epoch_reset()  # an extra reset between epochs changes nothing
        np.testing.assert_array_equal(ctx.kv_table_read(), table)


def test_gpu_duplicate_keys_in_txn(gpu):
    """Distinct keys are not required: duplicate writes of one transaction
    apply in slot order, exactly as the oracle does."""
    rows, fields, fw, B, S = 64, 2, 3, 3000, 12
    keys, meta, tids = L.kv_generate(seed=21, rows=rows, fields=fields, n_txns=B, n_slots=S, theta=0.5,
                                     distinct_keys=False)
    table0 = np.arange(rows * fields * fw, dtype=np.uint32)
    ora = oracle.kv_epoch(table0, rows, fields, fw, keys, meta, tids, S)
    for counts in (True, False):
        with _ctx(rows, fields, fw, B, S, counts=counts) as ctx:
            ctx.kv_table_load(table0)
            g = _gpu_epoch(ctx, keys, meta, tids, S)
        _assert_epoch(g, ora, meta, tids, fw)


def test_gpu_copyback_keeps_table(gpu):
    rows, fields, fw, B, S = 10_000, 4, 4, 8000, 10
    keys, meta, tids = L.kv_generate(seed=3, rows=rows, fields=fields, n_txns=B, n_slots=S, theta=0.95)
    table0 = np.arange(rows * fields * fw, dtype=np.uint32) * 7
    ora = oracle.kv_epoch(table0, rows, fields, fw, keys, meta, tids, S, values=0)
    with _ctx(rows, fields, fw, B, S, values=L.VALUES_COPYBACK) as ctx:
        ctx.kv_table_load(table0)
        g = _gpu_epoch(ctx, keys, meta, tids, S)
    np.testing.assert_array_equal(g["table"], table0)
    np.testing.assert_array_equal(g["bitmap"], ora.bitmap)


def test_gpu_table_layout_roundtrips(gpu):
    rows, fields = 1234, 10
    colmajor = np.arange(rows * fields, dtype=np.uint32)  # column c of row r at c*rows + r
    with _ctx(rows, fields, 1, 64, 4) as ctx:
        ctx.kv_table_load(colmajor, layout=L.LAYOUT_COLUMN_MAJOR)
        rowmajor = ctx.kv_table_read(layout=L.LAYOUT_ROW_MAJOR)
        np.testing.assert_array_equal(rowmajor, colmajor.reshape(fields, rows).T.reshape(-1))
        np.testing.assert_array_equal(ctx.kv_table_read(layout=L.LAYOUT_COLUMN_MAJOR), colmajor)
    # strided device layout (100-byte fields start every 128 bytes) in chunks
    rows, fields, fw = 300_000, 3, 25
    dense = np.random.default_rng(5).integers(0, 2**32, size=rows * fields * fw, dtype=np.uint64)
    dense = dense.astype(np.uint32)
    with _ctx(rows, fields, fw, 64, 4) as ctx:
        ctx.kv_table_load(dense)
        np.testing.assert_array_equal(ctx.kv_table_read(), dense)
        ctx.kv_fill_reference()
        ref = np.repeat(np.tile(np.arange(fields, dtype=np.uint32), rows), fw)
        np.testing.assert_array_equal(ctx.kv_table_read(), ref)


def test_gpu_out_of_range_keys_reported(gpu):
    rows, fields, fw, B, S = 100, 2, 1, 256, 4
    keys, meta, tids = L.kv_generate(seed=9, rows=rows, fields=fields, n_txns=B, n_slots=S, theta=0.0)
    keys = keys.copy()
    keys[5] = rows + 7
    keys[300] = 2**63
    ora = oracle.kv_epoch(np.zeros(rows * fields * fw, np.uint32), rows, fields, fw, keys, meta, tids, S)
    assert ora.range_errors == 2
    with _ctx(rows, fields, fw, B, S) as ctx:
        ctx.kv_submit(keys, meta, tids, S)
        ctx.epoch_run()
        with pytest.raises(L.LtpgError) as ei:
            ctx.sync()
        assert ei.value.code == L.ERR_RANGE
        ctx.sync()  # the error is reported once
        np.testing.assert_array_equal(ctx.commit_bitmap(), ora.bitmap)
        np.testing.assert_array_equal(ctx.kv_table_read(), ora.table)


def test_gpu_rejects_bad_shapes(gpu):
    with pytest.raises(L.LtpgError) as ei:
        _ctx(10, 2, 1, 64, 65)
    assert ei.value.code == L.ERR_INVALID
    with pytest.raises(L.LtpgError) as ei:
        _ctx(10, 2, 1, (1 << 24) + 1, 4)
    assert ei.value.code == L.ERR_INVALID
    with _ctx(10, 2, 1, 64, 4) as ctx:
        keys, meta, tids = L.kv_generate(seed=1, rows=10, fields=2, n_txns=65, n_slots=4, theta=0.0)
        with pytest.raises(L.LtpgError):
            ctx.kv_submit(keys, meta, tids, 4)  # 65 > max_txns
        with pytest.raises(L.LtpgError) as ei:   # transactions without slots
            ctx.kv_submit(np.zeros(0, np.uint64), np.zeros(0, np.uint32), np.arange(3, dtype=np.uint32), 0)
        assert ei.value.code == L.ERR_INVALID
        with pytest.raises(L.LtpgError) as ei:
            ctx.epoch_run()
        assert ei.value.code == L.ERR_STATE


@pytest.mark.parametrize("theta", [0.6, 0.99])
def test_gpu_config2_full_size(gpu, theta):
    """10M rows x 10 fields x 100 B, 2^20 txns x 16 ops, 50/50: bit-exact
    against the oracle at full size, identical across two GPU runs, and the
    fast check agrees with the counting check."""
    rows, fields, fw, B, S = 10_000_000, 10, 25, 1 << 20, 16
    keys, meta, tids = L.kv_generate(seed=0x4C545047, rows=rows, fields=fields, n_txns=B, n_slots=S,
                                     theta=theta)
    with _ctx(rows, fields, fw, B, S, counts=False) as ctx:
        ctx.kv_fill_reference()
        table0 = ctx.kv_table_read()
        dev = L.DeviceKvBatch(ctx, keys, meta, tids, S)
        ctx.kv_submit_device(dev)
        ctx.epoch_run()
        ctx.sync()
        bm1, tab1, reads1 = ctx.commit_bitmap(), ctx.kv_table_read(), ctx.kv_read_values()
        # second run of the same batch on the same starting table
        ctx.kv_fill_reference()
        ctx.kv_submit_device(dev)
        ctx.epoch_run()
        ctx.sync()
        np.testing.assert_array_equal(ctx.commit_bitmap(), bm1)
        np.testing.assert_array_equal(ctx.kv_table_read(), tab1)
        dev.free()
    ora = oracle.kv_epoch(table0, rows, fields, fw, keys, meta, tids, S, values=1)
    np.testing.assert_array_equal(bm1, ora.bitmap)
    np.testing.assert_array_equal(tab1, ora.table)
    _assert_reads(reads1, ora.read_values, meta, tids, fw, ora.bitmap)
    del tab1, table0
    with _ctx(rows, 1, 1, B, S, counts=True) as ctx:  # counters only need the registers
        ctx.kv_submit(keys, (meta & 0xFF).astype(np.uint32), tids, S)
        ctx.epoch_run(L.PHASE_EXECUTE | L.PHASE_CHECK)
        ctx.sync()
        np.testing.assert_array_equal(ctx.commit_bitmap(), ora.bitmap)
        np.testing.assert_array_equal(ctx.conflict_counters(), ora.conflicts)
