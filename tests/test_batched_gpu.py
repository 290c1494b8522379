"""Batched form: the graph-captured submission of many uneven, mixed
batches gives the same commit sets and tables as the plain per-batch loop
(the reference's epoch loop), bit for bit, replay after replay."""
import numpy as np
import pytest

import ltpg_amd as L

pytestmark = pytest.mark.gpu

SEED = 0x4C545047
KV = dict(kv_rows=200_000, kv_fields=4, kv_field_bytes=16, kv_max_slots=24, max_txns=6000)
W = 2


def _contexts(hot_row=L.NO_HOT_ROW):
    kv = L.Context(workload=L.WORKLOAD_KV, kv_hot_row=hot_row, **KV)
    kv.kv_fill_reference()
    tp = L.Context(workload=L.WORKLOAD_TPCC, tpcc_warehouses=W, max_txns=6000)
    tp.tpcc_populate(SEED)
    return kv, tp


def _batches(rng, n, kv, tp, kinds=None):
    """n batches, sizes log-uniform in [64, 6000], type from the seed."""
    descs, keep, meta = [], [], []
    gen = L.TpccGenerator(warehouses=W, seed=5)
    for i in range(n):
        size = int(np.exp(rng.uniform(np.log(64), np.log(6000))))
        kind = kinds[i] if kinds else ("kv" if rng.random() < 0.5 else "tpcc")
        if kind == "kv":
            keys, m, tids = L.kv_generate(seed=SEED + i, rows=KV["kv_rows"], fields=KV["kv_fields"], n_txns=size,
                                          n_slots=KV["kv_max_slots"], min_ops=1, max_ops=KV["kv_max_slots"],
                                          theta=0.9, tid_base=i * 10000)
            db = L.DeviceKvBatch(kv, keys, m, tids, KV["kv_max_slots"])
            keep.append(db)
            descs.append(L.kv_desc(db))
        else:
            b = gen.next(size, tid_base=i * 10000)
            dev, bufs = L.upload_tpcc_batch(tp, b)
            keep.append(bufs)
            descs.append(L.tpcc_desc(dev))
        meta.append(kind)
    return descs, keep, meta


def _bitmaps(runner, meta):
    out = []
    for i, kind in enumerate(meta):
        if kind == "kv":
            out.append([runner.bitmap(i, L.TYPE_KV)])
        else:
            out.append([runner.bitmap(i, L.TYPE_NEWORDER), runner.bitmap(i, L.TYPE_PAYMENT)])
    return out


def _tables(kv, tp):
    return [kv.kv_table_read()] + [tp.tpcc_table_read(t) for t in range(9)]


def _run(mode, n, runs, kinds=None, hot_row=L.NO_HOT_ROW):
    kv, tp = _contexts(hot_row)
    descs, keep, meta = _batches(np.random.default_rng(42), n, kv, tp, kinds)
    r = L.Batched(kv, tp, descs)
    maps = []
    for _ in range(runs):
        r.run(mode)
        maps.append(_bitmaps(r, meta))
    tabs = _tables(kv, tp)
    r.close()
    return maps, tabs


def test_graph_form_matches_plain_loop(gpu):
    plain_maps, plain_tabs = _run(L.RUN_PLAIN, 40, 2)
    graph_maps, graph_tabs = _run(L.RUN_GRAPH, 40, 2)
    for rp, rg in zip(plain_maps, graph_maps):
        for bp, bg in zip(rp, rg):
            for x, y in zip(bp, bg):
                np.testing.assert_array_equal(x, y)
    for x, y in zip(plain_tabs, graph_tabs):
        np.testing.assert_array_equal(x, y)


def test_graph_replays_match_across_an_epoch_tag_wrap(gpu):
    """300 key-value epochs in one graph: the captured refill at the tag wrap
    keeps the results equal to the plain loop."""
    kinds = ["kv"] * 300
    plain_maps, plain_tabs = _run(L.RUN_PLAIN, 300, 1, kinds)
    graph_maps, graph_tabs = _run(L.RUN_GRAPH, 300, 1, kinds)
    for rp, rg in zip(plain_maps[0], graph_maps[0]):
        np.testing.assert_array_equal(rp[0], rg[0])
    np.testing.assert_array_equal(plain_tabs[0], graph_tabs[0])


def test_graph_form_with_hot_row_matches_plain_loop(gpu):
    """The hot-row reduce's stripes are cleared inside each epoch, so graph
    replays give the plain loop's results."""
    plain_maps, plain_tabs = _run(L.RUN_PLAIN, 24, 2, hot_row=0)
    graph_maps, graph_tabs = _run(L.RUN_GRAPH, 24, 2, hot_row=0)
    for rp, rg in zip(plain_maps, graph_maps):
        for bp, bg in zip(rp, rg):
            for x, y in zip(bp, bg):
                np.testing.assert_array_equal(x, y)
    for x, y in zip(plain_tabs, graph_tabs):
        np.testing.assert_array_equal(x, y)


def test_graph_form_with_scans_matches_plain_loop_and_oracle(gpu):
    """Key-value batches whose reads are 6-row scans: the pipelined graph
    gives the plain loop's commit sets and table, and both the oracle's."""
    from oracle import oracle
    from tests import kvcases as kc

    rows, fields, fw, S, max_scan = 20000, 3, 2, 16, 6
    rng = np.random.default_rng(3)
    sizes = [int(np.exp(rng.uniform(np.log(64), np.log(4000)))) for _ in range(24)]
    gens = [L.kv_generate(seed=50 + i, rows=rows, fields=fields, n_txns=n, n_slots=S, min_ops=1, max_ops=S,
                          read_pct=70, theta=0.9, scan_len=6, tid_base=i * 10000) for i, n in enumerate(sizes)]
    out = {}
    for mode in (L.RUN_PLAIN, L.RUN_GRAPH):
        kv = L.Context(workload=L.WORKLOAD_KV, kv_rows=rows, kv_fields=fields, kv_field_bytes=4 * fw,
                       kv_max_slots=S, max_txns=4000, kv_max_scan=max_scan)
        kv.kv_table_load(kc.initial_table(rows, fields, fw))
        keep = [L.DeviceKvBatch(kv, k, m, t, S) for (k, m, t) in gens]
        r = L.Batched(kv, None, [L.kv_desc(d) for d in keep])
        r.run(mode)
        out[mode] = ([r.bitmap(i, L.TYPE_KV) for i in range(len(gens))], kv.kv_table_read())
        r.close()
    for a, b in zip(out[L.RUN_PLAIN][0], out[L.RUN_GRAPH][0]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(out[L.RUN_PLAIN][1], out[L.RUN_GRAPH][1])
    table = kc.initial_table(rows, fields, fw)
    for i, (k, m, t) in enumerate(gens):
        ora = oracle.kv_epoch(table, rows, fields, fw, k, m, t, S, inplace=True, want_reads=False,
                              max_scan=max_scan)
        np.testing.assert_array_equal(out[L.RUN_GRAPH][0][i], ora.bitmap)
    np.testing.assert_array_equal(out[L.RUN_GRAPH][1], table)


def test_graph_form_with_readonly_matches_plain_loop_and_oracle(gpu):
    """Order-entry batches carrying Order-Status / Stock-Level transactions
    beside their NewOrders and Payments (priorities interleaved, half of the
    readers on rows the NewOrders write): the graph gives the plain loop's
    commit sets of all four types and its tables, and both the oracle's."""
    from oracle import oracle

    W = 2
    rng = np.random.default_rng(17)
    gen = L.TpccGenerator(warehouses=W, seed=23)
    batches = []
    for i in range(12):
        n = int(np.exp(rng.uniform(np.log(64), np.log(1500))))
        b = gen.next(n, tid_base=i * 10000)
        ro = L.readonly_batch(b, W, n // 6 + 1, n // 6 + 1, seed=80 + i, tid_lo=i * 10000)
        batches.append((b, ro))
    types = (L.TYPE_NEWORDER, L.TYPE_PAYMENT, L.TYPE_ORDERSTATUS, L.TYPE_STOCKLEVEL)
    out = {}
    for mode in (L.RUN_PLAIN, L.RUN_GRAPH):
        tp = L.Context(workload=L.WORKLOAD_TPCC, tpcc_warehouses=W, max_txns=4096)
        tp.tpcc_populate(SEED)
        keep, descs = [], []
        for b, ro in batches:
            dev, bufs = L.upload_tpcc_batch(tp, b)
            rdev, rbufs = L.upload_readonly_batch(tp, ro)
            keep += bufs + rbufs
            descs.append(L.tpcc_desc(dev, rdev))
        r = L.Batched(None, tp, descs)
        r.run(mode)
        out[mode] = ([[r.bitmap(i, t) for t in types] for i in range(len(batches))],
                     [tp.tpcc_table_read(t) for t in range(9)])
        r.close()
    for a, b in zip(out[L.RUN_PLAIN][0], out[L.RUN_GRAPH][0]):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
    for x, y in zip(out[L.RUN_PLAIN][1], out[L.RUN_GRAPH][1]):
        np.testing.assert_array_equal(x, y)
    tables = oracle.tpcc_populate(W, SEED)
    ora = oracle.TpccOracle(W)
    n_ro_commits = 0
    for i, (b, ro) in enumerate(batches):
        ep = ora.epoch(tables, b, ro=ro)
        got = out[L.RUN_GRAPH][0][i]
        np.testing.assert_array_equal(got[0], ep.no_bitmap)
        np.testing.assert_array_equal(got[1], ep.pay_bitmap)
        np.testing.assert_array_equal(got[2], ep.ro["os_bitmap"])
        np.testing.assert_array_equal(got[3], ep.ro["sl_bitmap"])
        n_ro_commits += int(sum(bin(int(w)).count("1") for w in got[2]))
    assert n_ro_commits > 0
    for t in range(9):
        np.testing.assert_array_equal(out[L.RUN_GRAPH][1][t], tables[t])


def test_plain_epochs_after_a_graph_run(gpu):
    """A context used by a graph run keeps working epoch by epoch."""
    kv, tp = _contexts()
    descs, keep, meta = _batches(np.random.default_rng(7), 10, kv, tp)
    r = L.Batched(kv, tp, descs)
    r.run(L.RUN_GRAPH)
    r.run(L.RUN_PLAIN)
    first = _bitmaps(r, meta)
    r.run(L.RUN_GRAPH)
    second = _bitmaps(r, meta)
    # the same batches on the tables as each run leaves them: commit sets
    # depend only on the batches (and CUSTOMER.C_LAST, never written)
    for a, b in zip(first, second):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
    r.close()


KINDS = ["kv", "tpcc", "kv", "kv", "tpcc", "tpcc", "kv", "tpcc"]
EMPTY_AT = {0: "kv", 3: "tpcc", 4: "kv", 11: "tpcc"}  # positions among the 12 batches


def _empty_desc(kind):
    d = L.BatchDesc()
    if kind == "kv":
        d.workload = L.WORKLOAD_KV
        d.kv = L.KvBatch(0, KV["kv_max_slots"], None, None, None)
    else:
        d.workload = L.WORKLOAD_TPCC  # every transaction count zero
    return d


@pytest.mark.parametrize("mode", [L.RUN_PLAIN, L.RUN_GRAPH], ids=["plain", "graph"])
def test_empty_batches_change_nothing(gpu, mode):
    """Batches with no transactions commit nothing and change no row: a run
    with them interleaved gives the same commit sets and tables as the run
    without them."""
    def run(empties):
        kv, tp = _contexts()
        descs, keep, meta = _batches(np.random.default_rng(7), len(KINDS), kv, tp, KINDS)
        real = list(range(len(descs)))
        if empties:
            src = iter(zip(descs, meta))
            descs, meta, real = [], [], []
            for pos in range(len(KINDS) + len(EMPTY_AT)):
                if pos in EMPTY_AT:
                    descs.append(_empty_desc(EMPTY_AT[pos]))
                    meta.append(EMPTY_AT[pos])
                else:
                    d, m = next(src)
                    descs.append(d)
                    meta.append(m)
                    real.append(pos)
        r = L.Batched(kv, tp, descs)
        r.run(mode)
        r.run(mode)
        maps = _bitmaps(r, meta)
        tabs = _tables(kv, tp)
        r.close()
        return maps, real, tabs

    base_maps, _, base_tabs = run(False)
    maps, real, tabs = run(True)
    for i, pos in enumerate(real):
        for x, y in zip(base_maps[i], maps[pos]):
            np.testing.assert_array_equal(x, y)
    for pos in EMPTY_AT:
        assert all(m.size == 0 for m in maps[pos])
    for x, y in zip(base_tabs, tabs):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("rm", [False, True], ids=["slot-major", "request-major"])
@pytest.mark.parametrize("fw", [2, 25], ids=["8B", "100B"])
@pytest.mark.parametrize("hot_row", [L.NO_HOT_ROW, 0], ids=["no-hot", "hot0"])
@pytest.mark.parametrize("flags", list(range(16)) + [L.TUNE_MORE_ITEMS | f for f in (0, 1, 3, 5, 15)] +
                         [L.TUNE_PARTITIONED | f for f in (0, 2, 4, 8, 14)])
def test_kv_launch_shapes_match_oracle(gpu, flags, hot_row, fw, rm):
    """Every launch shape against the oracle, batches in either layout;
    100-byte fields (stride 32 words) take the per-lane copy under
    TUNE_LANE_COPY, 8-byte ones keep the group copy."""
    from oracle import oracle
    from tests import kvcases as kc

    rows, fields, B, S = 5000, 3, 3000, 12
    ctx = L.Context(workload=L.WORKLOAD_KV, kv_rows=rows, kv_fields=fields, kv_field_bytes=4 * fw, max_txns=B,
                    kv_max_slots=S, kv_hot_row=hot_row)
    table = kc.initial_table(rows, fields, fw)
    ctx.kv_table_load(table)
    ctx.kv_set_tuning(flags)
    for ep in range(3):
        keys, meta, tids = L.kv_generate(seed=900 + ep, rows=rows, fields=fields, n_txns=B, n_slots=S, min_ops=1,
                                         max_ops=S, theta=0.9, tid_base=ep * B)
        if rm:
            ctx.kv_submit(keys.reshape(S, B).T.copy().reshape(-1), meta.reshape(S, B).T.copy().reshape(-1), tids, S,
                          request_major=True)
        else:
            ctx.kv_submit(keys, meta, tids, S)
        ctx.epoch_run()
        ctx.sync()
        ora = oracle.kv_epoch(table, rows, fields, fw, keys, meta, tids, S, inplace=True, hot_row=hot_row)
        bm = ctx.commit_bitmap()
        np.testing.assert_array_equal(bm, ora.bitmap)
        com = np.unpackbits(bm.view(np.uint8), bitorder="little")[:B].astype(bool)
        mask = np.repeat((((meta.reshape(S, B) & 3) == 1) & com[None, :]).reshape(-1), fw)
        np.testing.assert_array_equal(ctx.kv_read_values()[mask], ora.read_values[mask])
    np.testing.assert_array_equal(ctx.kv_table_read(), table)


def test_autotuned_graph_matches_untuned_plain(gpu):
    kv0, tp0 = _contexts()
    descs0, keep0, meta = _batches(np.random.default_rng(11), 30, kv0, tp0)
    plain = L.Batched(kv0, tp0, descs0)
    plain.run(L.RUN_PLAIN)
    want = _bitmaps(plain, meta)
    kv1, tp1 = _contexts()
    descs1, keep1, _ = _batches(np.random.default_rng(11), 30, kv1, tp1)
    tuned = L.Batched(kv1, tp1, descs1)
    table = tuned.autotune(reps=2)
    assert {e["workload"] for e in table} == {L.WORKLOAD_KV, L.WORKLOAD_TPCC}
    for e in table:
        assert len(e["ms"]) == 64
        if e["workload"] == L.WORKLOAD_KV:
            tried = [L.tune_index(f) for f in L.kv_shapes()]
            assert e["flags"] in L.kv_shapes()
        else:
            tried = [0, 1]
            assert e["flags"] in (0, L.TUNE_TPCC_SMALL_BLOCKS)
        assert all(e["ms"][k] > 0 for k in tried), e
        assert all(e["ms"][k] == 0 for k in range(64) if k not in tried), e
    tuned.run(L.RUN_GRAPH)
    got = _bitmaps(tuned, meta)
    for a, b in zip(want, got):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
    for x, y in zip(_tables(kv0, tp0), _tables(kv1, tp1)):
        np.testing.assert_array_equal(x, y)


def test_sync_costs_are_measured(gpu):
    """The grid-barrier and kernel-boundary costs behind the choice of
    graphs over a persistent kernel: a cooperative launch of co-resident
    workgroups passes its barriers (the poll limit is never reached), and
    every cost is a positive number of microseconds."""
    kv = L.Context(workload=L.WORKLOAD_KV, kv_rows=1024, kv_fields=1, kv_field_bytes=4, max_txns=64, kv_max_slots=1)
    for blocks in (64, 256):
        c = L.sync_costs(kv, blocks, 200)
        assert all(0.0 < x < 1000.0 for x in c.values()), c
    kv.close()


def test_latency_costs_are_measured(gpu):
    """The step latencies behind the served device loop's chain roofline:
    every chain ends (fixed counts, in-bounds indices) and gives a positive
    time per step; a PCIe round trip costs more than an LDS load."""
    kv = L.Context(workload=L.WORKLOAD_KV, kv_rows=1024, kv_fields=1, kv_field_bytes=4, max_txns=64, kv_max_slots=1)
    c = L.latency_costs(kv)
    assert set(c) == set(L.LATENCY_NAMES)
    assert all(0.0 < x < 100000.0 for x in c.values()), c
    assert c["pcie_load_ns"] > c["lds_load_ns"], c
    kv.close()


@pytest.mark.parametrize("rows", [5000, 200000], ids=["one-bucket", "25-buckets"])
@pytest.mark.parametrize("theta", [0.6, 0.99])
@pytest.mark.parametrize("hot_row", [L.NO_HOT_ROW, 0], ids=["no-hot", "hot0"])
@pytest.mark.parametrize("scan", [0, 4], ids=["point", "scans"])
def test_kv_partitioned_execute_matches_oracle(gpu, rows, theta, hot_row, scan):
    """The row-partitioned EXECUTE against the oracle: one bucket split over
    several slices (global W folding) and many single-slice buckets (LDS
    only), skew that overloads one bucket, the hot row, scans across bucket
    edges, and epochs one after the other on the same registers."""
    _partitioned_epochs(rows, theta, hot_row, scan)


def test_kv_partitioned_execute_many_buckets_matches_oracle(gpu):
    """More than 2048 buckets of 8192 rows (16.8M rows): the binning kernel
    whose LDS holds every bucket's counter up to the 16384-bucket limit."""
    _partitioned_epochs(2048 * 8192 + 5000, 0.6, L.NO_HOT_ROW, 0)


def _partitioned_epochs(rows, theta, hot_row, scan):
    from oracle import oracle
    from tests import kvcases as kc

    fields, fw, B, S = 3, 2, 6000, 10
    ctx = L.Context(workload=L.WORKLOAD_KV, kv_rows=rows, kv_fields=fields, kv_field_bytes=4 * fw, max_txns=B,
                    kv_max_slots=S, kv_hot_row=hot_row, kv_max_scan=scan)
    table = kc.initial_table(rows, fields, fw)
    ctx.kv_table_load(table)
    for ep in range(4):
        ctx.kv_set_tuning(L.TUNE_PARTITIONED | (L.TUNE_SPLIT_REGISTERS if ep % 2 else 0))
        keys, meta, tids = L.kv_generate(seed=700 + ep, rows=rows, fields=fields, n_txns=B, n_slots=S, min_ops=1,
                                         max_ops=S, theta=theta, tid_base=ep * B, scan_len=scan)
        if scan:
            keys[::97] = rows - 2  # scans clipped at the end of the table
        ctx.kv_submit(keys, meta, tids, S)
        ctx.epoch_run()
        ctx.sync()
        ora = oracle.kv_epoch(table, rows, fields, fw, keys, meta, tids, S, inplace=True, hot_row=hot_row,
                              max_scan=scan)
        bm = ctx.commit_bitmap()
        np.testing.assert_array_equal(bm, ora.bitmap)
        assert ctx.commit_count() == ora.commit_count
        com = np.unpackbits(bm.view(np.uint8), bitorder="little")[:B].astype(bool)
        L_ = max(1, scan)
        kinds = (meta & 3).astype(np.int64)
        lens = np.where(kinds == 3, (meta >> 24).astype(np.int64) + 1, (kinds == 1).astype(np.int64))
        j = np.arange(L_)[None, :]
        # row j of a read / scan slot of a committed transaction, inside the table
        valid = (j < lens[:, None]) & ((keys.astype(np.int64)[:, None] + j) < rows)
        valid &= np.broadcast_to(com[None, :], (S, B)).reshape(-1)[:, None]
        mask = np.repeat(valid.reshape(-1), fw)
        np.testing.assert_array_equal(ctx.kv_read_values()[mask], ora.read_values[mask])
    np.testing.assert_array_equal(ctx.kv_table_read(), table)


@pytest.mark.parametrize("flags", [L.TUNE_PARTITIONED, L.TUNE_PARTITIONED | L.TUNE_SPLIT_REGISTERS | 4,
                                   L.TUNE_MORE_ITEMS | L.TUNE_EXEC_PER_OP | L.TUNE_TPCC_SMALL_BLOCKS,
                                   L.TUNE_PARTITIONED | L.TUNE_TPCC_SMALL_BLOCKS],
                         ids=["partitioned", "partitioned-split-small", "more-items-small-blocks",
                              "partitioned-small-blocks"])
def test_graph_form_with_partitioned_execute_matches_plain_loop(gpu, flags):
    """One launch shape for every batch in the graph form -- the
    row-partitioned EXECUTE (EXECUTE of batch j + 1 overlapping CHECK of
    batch j, two register sets, one partition scratch), more items per lane,
    64-thread order-entry workgroups: the default-shape plain loop's commit
    sets and tables."""
    kv0, tp0 = _contexts()
    descs0, keep0, meta = _batches(np.random.default_rng(13), 30, kv0, tp0)
    plain = L.Batched(kv0, tp0, descs0)
    plain.run(L.RUN_PLAIN)
    want = _bitmaps(plain, meta)
    kv1, tp1 = _contexts()
    descs1, keep1, _ = _batches(np.random.default_rng(13), 30, kv1, tp1)
    part = L.Batched(kv1, tp1, descs1)
    part.set_tuning(flags)
    part.run(L.RUN_GRAPH)
    for a, b in zip(want, _bitmaps(part, meta)):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
    for x, y in zip(_tables(kv0, tp0), _tables(kv1, tp1)):
        np.testing.assert_array_equal(x, y)
    part.run(L.RUN_GRAPH)  # a replay of the captured graph decides the same
    for a, b in zip(want, _bitmaps(part, meta)):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
