"""Host-resident batches through the two device slots: batch e+1 is copied
while epoch e still runs, and the caller may reuse its buffers as soon as
the submit returns.  Epochs are submitted back to back with no
synchronisation between them, every host buffer is overwritten right after
its submit, and the running commit totals and final tables must equal the
oracle's replay of the same batches in order."""
import numpy as np
import pytest

import ltpg_amd as L
from oracle import oracle
from tests import kvcases as kc

pytestmark = pytest.mark.gpu

SEED = 0x4C545047


@pytest.mark.parametrize("mix_device", [False, True], ids=["all-host", "one-device-resident"])
def test_kv_host_batches_pipelined_match_oracle(gpu, mix_device):
    rows, fields, fw, S = 20000, 4, 3, 12
    sizes = [3000, 500, 4000, 1200, 2500, 800, 3600, 2000]
    ctx = L.Context(workload=L.WORKLOAD_KV, kv_rows=rows, kv_fields=fields, kv_field_bytes=4 * fw,
                    max_txns=max(sizes), kv_max_slots=S, values=L.VALUES_RMW)
    t0 = kc.initial_table(rows, fields, fw)
    ctx.kv_table_load(t0)
    ctx.sync()
    ctx.commit_totals(reset=True)
    batches, keep, tid = [], [], 0
    for i, B in enumerate(sizes):
        k, m, t = L.kv_generate(seed=SEED + i, rows=rows, fields=fields, n_txns=B, n_slots=S, min_ops=1, max_ops=S,
                                theta=0.9, tid_base=tid)
        tid += B
        batches.append((k.copy(), m.copy(), t.copy()))
        if mix_device and i == 4:
            db = L.DeviceKvBatch(ctx, k, m, t, S)
            keep.append(db)
            ctx.kv_submit_device(db)
        else:
            ctx.kv_submit(k, m, t, S)
            # the caller's buffers are free as soon as the submit returns
            k[:] = rows + 5
            m[:] = 0xFFFFFFFF
            t[:] = 0
        ctx.epoch_run()
    ctx.sync()
    commits, attempts = ctx.commit_totals()
    table = ctx.kv_table_read()
    want_commits, want = 0, t0.copy()
    for k, m, t in batches:
        ep = oracle.kv_epoch(want, rows, fields, fw, k, m, t, S, want_reads=False, inplace=True)
        want_commits += ep.commit_count
    assert (commits, attempts) == (want_commits, sum(sizes))
    np.testing.assert_array_equal(table, want)
    for b in keep:
        b.free()
    ctx.close()


def test_tpcc_host_batches_pipelined_match_oracle(gpu):
    W = 2
    ctx = L.Context(workload=L.WORKLOAD_TPCC, tpcc_warehouses=W, max_txns=4096, values=L.VALUES_RMW)
    ctx.tpcc_populate(SEED)
    ctx.sync()
    types = (L.TYPE_NEWORDER, L.TYPE_PAYMENT, L.TYPE_ORDERSTATUS, L.TYPE_STOCKLEVEL)
    for ty in types:
        ctx.commit_totals(ty, reset=True)
    gen = L.TpccGenerator(warehouses=W, seed=SEED)
    kept = []
    for i in range(6):
        b = gen.next(1500, tid_base=i * 2000)
        ro = L.readonly_batch(b, W, 150, 150, seed=SEED + i, tid_lo=i * 2000, tid_hi=i * 2000 + 1500)
        kept.append((L.TpccBatch(b.no.copy(), b.no_cap, b.n_no, b.pay.copy(), b.pay_cap, b.n_pay), ro))
        ctx.tpcc_submit(b)
        b.no[:] = 0xFFFFFFFF
        b.pay[:] = 0xFFFFFFFF
        ctx.tpcc_submit_readonly(ro)
        ctx.epoch_run()
    ctx.sync()
    got = {ty: ctx.commit_totals(ty) for ty in types}
    got_tables = [ctx.tpcc_table_read(t) for t in range(9)]
    tables = oracle.tpcc_populate(W, SEED)
    ora = oracle.TpccOracle(W)
    want = {ty: [0, 0] for ty in types}
    for b, ro in kept:
        ep = ora.epoch(tables, b, 0, 1, ro=ro)
        for ty, c, n in ((L.TYPE_NEWORDER, ep.no_commits, b.n_no), (L.TYPE_PAYMENT, ep.pay_commits, b.n_pay),
                         (L.TYPE_ORDERSTATUS, ep.ro["os_commits"], ro.n_os),
                         (L.TYPE_STOCKLEVEL, ep.ro["sl_commits"], ro.n_sl)):
            want[ty][0] += c
            want[ty][1] += n
    assert got == {ty: tuple(v) for ty, v in want.items()}
    for t in range(9):
        np.testing.assert_array_equal(got_tables[t], tables[t])
    ctx.close()
