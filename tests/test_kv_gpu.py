"""Key-value parity tests: the HIP path (through the C ABI) against the oracle.

Bar: bit-exact commit bitmap, conflict counters, staged read values and final
table on the same seeded batch, and identical results across repeated runs.
"""
import numpy as np
import pytest

import ltpg_amd as L
from oracle import oracle
from tests import kvcases as kc

pytestmark = pytest.mark.gpu

KAT = kc.load_kat()


def _ctx(rows, fields, fw, max_txns, max_slots, values=L.VALUES_RMW):
    return L.Context(workload=L.WORKLOAD_KV, kv_rows=rows, kv_fields=fields, kv_field_bytes=4 * fw,
                     max_txns=max_txns, kv_max_slots=max_slots, values=values)


def _gpu_epoch(ctx, keys, meta, tids, S, device=False):
    if device:
        b = L.DeviceKvBatch(ctx, keys, meta, tids, S)
        ctx.kv_submit_device(b)
    else:
        ctx.kv_submit(keys, meta, tids, S)
    ctx.epoch_run(L.PHASE_ALL)
    ctx.sync()
    out = dict(bitmap=ctx.commit_bitmap(), conflicts=ctx.conflict_counters(), reads=ctx.kv_read_values(),
               count=ctx.commit_count(), table=ctx.kv_table_read())
    if device:
        b.free()
    return out


def _assert_reads(gpu_reads, ora_reads, meta, fw):
    """Compare staged read values on read slots only."""
    is_read = np.repeat((meta & 3) == 1, fw)
    np.testing.assert_array_equal(gpu_reads[is_read], ora_reads[is_read])


@pytest.mark.parametrize("case", KAT["cases"], ids=[c["name"] for c in KAT["cases"]])
def test_gpu_known_answers(gpu, case):
    rows, fields, fw = case["rows"], case["fields"], case["field_words"]
    keys, meta, tids, S = kc.case_arrays(case)
    values = L.VALUES_RMW if case["values"] == "rmw" else L.VALUES_COPYBACK
    B = len(tids)
    with _ctx(rows, fields, fw, max(B, 1), S, values) as ctx:
        ctx.kv_table_load(kc.initial_table(rows, fields, fw))
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
        assert kc.unpack_conflicts(ctx.conflict_counters()).tolist() == exp["conflicts"]
        assert ctx.commit_count() == sum(exp["commit"])
        np.testing.assert_array_equal(ctx.kv_table_read(), kc.expected_table(case))
        reads = ctx.kv_read_values()
        for t, s, words in exp["reads"]:
            i = s * B + t
            assert reads[i * fw:(i + 1) * fw].tolist() == words


CASES = [
    # rows, fields, fw, B, S, theta, read_pct, min_ops, seed
    (1000, 4, 1, 5000, 8, 0.9, 50, 8, 1),
    (50_000, 10, 25, 20_000, 16, 0.99, 50, 16, 2),
    (50_000, 10, 25, 20_000, 16, 0.6, 50, 16, 3),
    (4096, 3, 7, 777, 64, 0.8, 30, 1, 4),        # ragged 1..64 ops, odd field width
    (100_000, 2, 80, 3000, 5, 0.5, 70, 2, 5),     # fields wider than 64 words
    (100, 2, 1, 65, 3, 0.0, 50, 3, 6),            # uniform keys, partial last wave
]


@pytest.mark.parametrize("rows,fields,fw,B,S,theta,read_pct,min_ops,seed", CASES)
@pytest.mark.parametrize("device", [False, True])
def test_gpu_matches_oracle_seeded(gpu, rows, fields, fw, B, S, theta, read_pct, min_ops, seed, device):
    keys, meta, tids = L.kv_generate(seed=seed, rows=rows, fields=fields, n_txns=B, n_slots=S,
                                     min_ops=min_ops, max_ops=S, read_pct=read_pct, theta=theta)
    table0 = np.random.default_rng(seed).integers(0, 2**32, size=rows * fields * fw, dtype=np.uint64)
    table0 = table0.astype(np.uint32)
    ora = oracle.kv_epoch(table0, rows, fields, fw, keys, meta, tids, S, values=1)
    with _ctx(rows, fields, fw, B, S) as ctx:
        ctx.kv_table_load(table0)
        g = _gpu_epoch(ctx, keys, meta, tids, S, device=device)
    np.testing.assert_array_equal(g["bitmap"], ora.bitmap)
    np.testing.assert_array_equal(g["conflicts"], ora.conflicts)
    assert g["count"] == ora.commit_count
    np.testing.assert_array_equal(g["table"], ora.table)
    _assert_reads(g["reads"], ora.read_values, meta, fw)


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
    np.testing.assert_array_equal(g["bitmap"], ora.bitmap)
    np.testing.assert_array_equal(g["conflicts"], ora.conflicts)
    np.testing.assert_array_equal(g["table"], ora.table)
    _assert_reads(g["reads"], ora.read_values, meta, fw)
    assert 0 < ora.commit_count < B


def test_gpu_epochs_do_not_leak_registers(gpu):
    """Registers are epoch-tagged instead of cleared: consecutive epochs must
    behave exactly like the oracle's fresh 0xFFFFFFFF registers."""
    rows, fields, fw, B, S = 2000, 3, 2, 4000, 6
    rng = np.random.default_rng(11)
    table = rng.integers(0, 2**32, size=rows * fields * fw, dtype=np.uint64).astype(np.uint32)
    with _ctx(rows, fields, fw, B, S) as ctx:
        ctx.kv_table_load(table)
        for e in range(5):
            keys, meta, tids = L.kv_generate(seed=100 + e, rows=rows, fields=fields, n_txns=B - 500 * e,
                                             n_slots=S, theta=0.9, tid_base=(7 - e) * 10000)
            ora = oracle.kv_epoch(table, rows, fields, fw, keys, meta, tids, S)
            g = _gpu_epoch(ctx, keys, meta, tids, S)
            np.testing.assert_array_equal(g["bitmap"], ora.bitmap)
            np.testing.assert_array_equal(g["conflicts"], ora.conflicts)
            np.testing.assert_array_equal(g["table"], ora.table)
            table = ora.table
        # an explicit reset between two runs of one batch
        ctx.epoch_reset()


def test_gpu_duplicate_keys_in_txn(gpu):
    """Distinct keys are not required: duplicate writes of one transaction
    apply in slot order, exactly as the oracle does."""
    rows, fields, fw, B, S = 64, 2, 3, 3000, 12
    keys, meta, tids = L.kv_generate(seed=21, rows=rows, fields=fields, n_txns=B, n_slots=S, theta=0.5,
                                     distinct_keys=False)
    table0 = np.arange(rows * fields * fw, dtype=np.uint32)
    ora = oracle.kv_epoch(table0, rows, fields, fw, keys, meta, tids, S)
    with _ctx(rows, fields, fw, B, S) as ctx:
        ctx.kv_table_load(table0)
        g = _gpu_epoch(ctx, keys, meta, tids, S)
    np.testing.assert_array_equal(g["bitmap"], ora.bitmap)
    np.testing.assert_array_equal(g["table"], ora.table)


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


def test_gpu_column_major_roundtrip(gpu):
    rows, fields = 1234, 10
    colmajor = np.arange(rows * fields, dtype=np.uint32)  # column c of row r at c*rows + r
    with _ctx(rows, fields, 1, 64, 4) as ctx:
        ctx.kv_table_load(colmajor, layout=L.LAYOUT_COLUMN_MAJOR)
        rowmajor = ctx.kv_table_read(layout=L.LAYOUT_ROW_MAJOR)
        np.testing.assert_array_equal(rowmajor, colmajor.reshape(fields, rows).T.reshape(-1))
        np.testing.assert_array_equal(ctx.kv_table_read(layout=L.LAYOUT_COLUMN_MAJOR), colmajor)


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
    with _ctx(10, 2, 1, 64, 4) as ctx:
        keys, meta, tids = L.kv_generate(seed=1, rows=10, fields=2, n_txns=65, n_slots=4, theta=0.0)
        with pytest.raises(L.LtpgError):
            ctx.kv_submit(keys, meta, tids, 4)  # 65 > max_txns
        with pytest.raises(L.LtpgError) as ei:
            ctx.epoch_run()
        assert ei.value.code == L.ERR_STATE


@pytest.mark.parametrize("theta", [0.6, 0.99])
def test_gpu_config2_full_size(gpu, theta):
    """10M rows x 10 fields x 100 B, 2^20 txns x 16 ops, 50/50: bit-exact
    against the oracle at full size, and identical across two GPU runs."""
    rows, fields, fw, B, S = 10_000_000, 10, 25, 1 << 20, 16
    keys, meta, tids = L.kv_generate(seed=0x4C545047, rows=rows, fields=fields, n_txns=B, n_slots=S,
                                     theta=theta)
    with _ctx(rows, fields, fw, B, S) as ctx:
        ctx.kv_fill_reference()
        table0 = ctx.kv_table_read()
        dev = L.DeviceKvBatch(ctx, keys, meta, tids, S)
        ctx.kv_submit_device(dev)
        ctx.epoch_run()
        ctx.sync()
        bm1, cf1, tab1 = ctx.commit_bitmap(), ctx.conflict_counters(), ctx.kv_table_read()
        reads1 = ctx.kv_read_values()
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
    np.testing.assert_array_equal(cf1, ora.conflicts)
    np.testing.assert_array_equal(tab1, ora.table)
    _assert_reads(reads1, ora.read_values, meta, fw)
