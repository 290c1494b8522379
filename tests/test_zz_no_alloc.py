"""ltpg_kv_epoch_noalloc, the key-value epoch that allocates nothing,
against the existing entry point (ltpg_kv_submit + ltpg_epoch_run on a
context) on the same table and batch: the same commit bitmap, committed
read values and final table, on two calls in a row, the second smaller.
Called through the C ABI; the test itself allocates the caller's buffers."""
import ctypes

import numpy as np
import pytest

import ltpg_amd as L
from tests import kvcases as kc


def _device_sync():
    hip = ctypes.CDLL("libamdhip64.so")
    assert hip.hipDeviceSynchronize() == 0


def _noalloc(table, rows, fields, fw, batch, values, bitmap, reads, result):
    return L.lib.ltpg_kv_epoch_noalloc(ctypes.c_void_p(table), rows, fields, 4 * fw, ctypes.byref(batch), values,
                                       ctypes.c_void_p(bitmap), ctypes.c_void_p(reads), ctypes.c_void_p(result), None)


@pytest.mark.gpu
def test_noalloc_epoch_matches_the_context_epoch(gpu):
    rows, fields, fw, S = 50_000, 4, 3, 12
    ctx = L.Context(workload=L.WORKLOAD_KV, kv_rows=rows, kv_fields=fields, kv_field_bytes=4 * fw,
                    max_txns=1 << 14, kv_max_slots=S, values=L.VALUES_RMW)
    t0 = kc.initial_table(rows, fields, fw)
    ctx.kv_table_load(t0)
    dtab = L.DeviceBuffer(ctx, t0.nbytes)
    dtab.upload(t0)
    tid = 0
    for B, seed in ((12000, 1), (3000, 2)):  # the second call smaller than the first
        keys, meta, tids = L.kv_generate(seed=seed, rows=rows, fields=fields, n_txns=B, n_slots=S, min_ops=1,
                                         max_ops=S, theta=0.9, tid_base=tid)
        tid += B
        ctx.kv_submit(keys, meta, tids, S)
        ctx.epoch_run()
        ctx.sync()
        want_bm = ctx.commit_bitmap()
        want_rv = ctx.kv_read_values()
        want_tab = ctx.kv_table_read()
        db = L.DeviceKvBatch(ctx, keys, meta, tids, S)
        nw = (B + 63) // 64
        dbm, drv, dres = L.DeviceBuffer(ctx, nw * 8), L.DeviceBuffer(ctx, S * B * fw * 4), L.DeviceBuffer(ctx, 8)
        rc = _noalloc(dtab.ptr, rows, fields, fw, db.struct(), L.VALUES_RMW, dbm.ptr, drv.ptr, dres.ptr)
        assert rc == L.OK, L.lib.ltpg_last_error()
        _device_sync()
        bm = np.empty(nw, np.uint64)
        dbm.download(bm)
        rv = np.empty(S * B * fw, np.uint32)
        drv.download(rv)
        res = np.empty(2, np.uint32)
        dres.download(res)
        tab = np.empty_like(t0)
        dtab.download(tab)
        np.testing.assert_array_equal(bm, want_bm[:nw])
        committed = np.unpackbits(bm.view(np.uint8), bitorder="little")[:B].astype(bool)
        reads = ((meta & 3) == L.KV_READ).reshape(S, B) & committed[None, :]
        np.testing.assert_array_equal(rv.reshape(S, B, fw)[reads], want_rv.reshape(S, B, fw)[reads])
        np.testing.assert_array_equal(tab, want_tab)
        assert res[0] == committed.sum() and res[1] == 0
        for x in (db, dbm, drv, dres):
            x.free()
    dtab.free()
    ctx.close()


def test_noalloc_refuses_sizes_beyond_its_array():
    """Checked before any GPU work: rows or transactions beyond the fixed
    array are LTPG_ERR_INVALID."""
    b = L.KvBatch()
    b.n_txns, b.n_slots = 16, 4
    assert _noalloc(1, L.NOALLOC_MAX_ROWS + 1, 2, 1, b, L.VALUES_RMW, 1, 1, 0) == L.ERR_INVALID
    b.n_txns = L.NOALLOC_MAX_TXNS + 1
    assert _noalloc(1, 1000, 2, 1, b, L.VALUES_RMW, 1, 1, 0) == L.ERR_INVALID
    b.n_txns, b.n_slots = 16, 65
    assert _noalloc(1, 1000, 2, 1, b, L.VALUES_RMW, 1, 1, 0) == L.ERR_INVALID
