"""Delivery (TPCC-alltxn/Execute.h:1772-2153) on the GPU against the
oracle, in the plain, batched and served forms.

Bar: bit-exact bitmaps of every type, Delivery read values, commit counts
and all nine tables, on the hand-derived known answers and on seeded
batches whose Deliveries hit the rows the NewOrders and Payments write
(ORDER / ORDERLINE / CUSTOMER rows through O_ID, NEWORDER rows a NewOrder
inserts).
"""
import time

import numpy as np
import pytest

import ltpg_amd as L
from oracle import oracle
from tests import tpcccases as tc

pytestmark = pytest.mark.gpu

SEED = 0x4C545047
KAT = tc.load_kat()
DL_CASES = [c for c in KAT["cases"] if "delivery" in c]


def _ctx(W, max_txns, rules=L.RULES_REFERENCE, values=L.VALUES_RMW):
    return L.Context(workload=L.WORKLOAD_TPCC, tpcc_warehouses=W, max_txns=max_txns, rules=rules, values=values)


def _kat_tables(case):
    W = case["W"]
    tables = oracle.tpcc_populate(W, KAT["seed"])
    for t, row, col, v in case["set"]:
        tables[t][col * tc.rows_of(W, t) + row] = np.uint32(v)
    return tables


def _kat_readonly(case):
    if "orderstatus" not in case and "stocklevel" not in case:
        return None
    os_, sl = case.get("orderstatus", []), case.get("stocklevel", [])
    ro = L.ReadOnlyBatch(len(os_), len(sl))
    for i, t in enumerate(os_):
        for f in L.ORDERSTATUS_FIELDS:
            ro.os[f][i] = t.get(f, 0)
    for i, t in enumerate(sl):
        for f in L.STOCKLEVEL_FIELDS:
            ro.sl[f][i] = t.get(f, 0)
    return ro


def _assert_tables(ctx, tables):
    for t in range(9):
        np.testing.assert_array_equal(ctx.tpcc_table_read(t), tables[t], err_msg=f"table {L.TPCC_TABLES[t]}")


@pytest.mark.parametrize("case", DL_CASES, ids=[c["name"] for c in DL_CASES])
def test_delivery_known_answers(gpu, case):
    """Every hand-derived epoch with Deliveries (fast check)."""
    W = case["W"]
    rules = L.RULES_REFERENCE if case["rules"] == "reference" else L.RULES_INTENDED
    values = L.VALUES_RMW if case["values"] == "rmw" else L.VALUES_COPYBACK
    tables = _kat_tables(case)
    ctx = _ctx(W, 64, rules, values)
    for t in range(9):
        ctx.tpcc_table_load(t, tables[t])
    nos = [dict(n, lines=[tuple(x) for x in n.get("lines", [])]) for n in case["neworders"]]
    b = tc.make_batch(nos, case["payments"])
    ro, dl = _kat_readonly(case), tc.kat_delivery(case)
    exp = case["expect"]
    ctx.tpcc_submit(b)
    ctx.tpcc_submit_delivery(dl)  # before the read-only ones: the order of the two submits is free
    if ro is not None:
        ctx.tpcc_submit_readonly(ro)
    ctx.epoch_run()
    if exp["range_errors"]:
        with pytest.raises(L.LtpgError) as ei:
            ctx.sync()
        assert ei.value.code == L.ERR_RANGE
    else:
        ctx.sync()
    assert tc.bits(ctx.tpcc_bitmap(L.TYPE_NEWORDER), b.n_no) == exp["no_commit"]
    assert tc.bits(ctx.tpcc_bitmap(L.TYPE_PAYMENT), b.n_pay) == exp["pay_commit"]
    assert tc.bits(ctx.tpcc_bitmap(L.TYPE_DELIVERY), dl.n) == exp["dl_commit"]
    assert ctx.commit_count(L.TYPE_DELIVERY) == sum(exp["dl_commit"])
    assert ctx.tpcc_payment_cid().tolist() == exp["cid"]
    dr = ctx.tpcc_read_values(L.TYPE_DELIVERY)
    for i, v, val in exp["dl_reads"]:
        assert int(dr[v * dl.n + i]) == val
    no_r = ctx.tpcc_read_values(L.TYPE_NEWORDER)
    for q, v, val in exp["no_reads"]:
        assert int(no_r[v * b.n_no + q]) == val
    if ro is not None:
        assert tc.bits(ctx.tpcc_bitmap(L.TYPE_ORDERSTATUS), ro.n_os) == exp["os_commit"]
        os_r = ctx.tpcc_read_values(L.TYPE_ORDERSTATUS)
        for i, v, val in exp["os_reads"]:
            assert int(os_r[v * ro.n_os + i]) == val
    for t, row, col, v in exp["cells"]:
        tables[t][col * tc.rows_of(W, t) + row] = np.uint32(v)
    _assert_tables(ctx, tables)


def _assert_dl(ctx, o, dl):
    np.testing.assert_array_equal(ctx.tpcc_bitmap(L.TYPE_DELIVERY), o.dl["bitmap"])
    assert ctx.commit_count(L.TYPE_DELIVERY) == o.dl["commits"]
    com = np.array(tc.bits(o.dl["bitmap"], dl.n), dtype=bool)
    got = ctx.tpcc_read_values(L.TYPE_DELIVERY).reshape(2, dl.n)[:, com]
    np.testing.assert_array_equal(got, o.dl["reads"].reshape(2, dl.n)[:, com])


@pytest.mark.parametrize("rules", [L.RULES_REFERENCE, L.RULES_INTENDED], ids=["reference", "intended"])
@pytest.mark.parametrize("values", [L.VALUES_RMW, L.VALUES_COPYBACK], ids=["rmw", "copyback"])
@pytest.mark.parametrize("device", [False, True], ids=["host-batch", "device-batch"])
@pytest.mark.parametrize("readonly", [False, True], ids=["no-readonly", "with-readonly"])
def test_delivery_seeded_epochs_match_oracle(gpu, rules, values, device, readonly):
    """Deliveries beside seeded batches (and read-only transactions), three
    epochs: every type's bitmap, the Delivery read values and counts, and the
    tables match the oracle; some Deliveries and some NewOrders lose to each
    other."""
    W = 2
    ctx = _ctx(W, 8192, rules, values)
    ctx.tpcc_populate(SEED)
    tables = oracle.tpcc_populate(W, SEED)
    ora = oracle.TpccOracle(W)
    gen = L.TpccGenerator(warehouses=W, seed=500)
    dl_aborts = no_lost_to_dl = 0
    for ep in range(3):
        b = gen.next(3000, tid_base=ep * 5000)
        ro = L.readonly_batch(b, W, 300, 300, seed=70 + ep, tid_lo=ep * 5000, tid_hi=ep * 5000 + 3600) \
            if readonly else None
        dl = L.delivery_batch(b, W, 600, seed=90 + ep, tid_lo=ep * 5000, tid_hi=ep * 5000 + 3600)
        keep = []
        if device:
            dev, bufs = L.upload_tpcc_batch(ctx, b)
            ddev, dbufs = L.upload_delivery_batch(ctx, dl)
            keep = bufs + dbufs
            ctx.tpcc_submit(dev, mem=L.MEM_DEVICE)
            ctx.tpcc_submit_delivery(ddev, mem=L.MEM_DEVICE)
            if ro is not None:
                rdev, rbufs = L.upload_readonly_batch(ctx, ro)
                keep += rbufs
                ctx.tpcc_submit_readonly(rdev, mem=L.MEM_DEVICE)
        else:
            ctx.tpcc_submit(b)
            if ro is not None:
                ctx.tpcc_submit_readonly(ro)
            ctx.tpcc_submit_delivery(dl)
        ctx.epoch_run()
        ctx.sync()
        o = ora.epoch(tables, b, rules, values, ro=ro, dl=dl)
        np.testing.assert_array_equal(ctx.tpcc_bitmap(L.TYPE_NEWORDER), o.no_bitmap)
        np.testing.assert_array_equal(ctx.tpcc_bitmap(L.TYPE_PAYMENT), o.pay_bitmap)
        assert ctx.commit_count(L.TYPE_NEWORDER) == o.no_commits
        if ro is not None:
            np.testing.assert_array_equal(ctx.tpcc_bitmap(L.TYPE_ORDERSTATUS), o.ro["os_bitmap"])
            np.testing.assert_array_equal(ctx.tpcc_bitmap(L.TYPE_STOCKLEVEL), o.ro["sl_bitmap"])
        _assert_dl(ctx, o, dl)
        dl_aborts += dl.n - o.dl["commits"]
        no_lost_to_dl += int((tc.unpack(o.no_conflicts)[:, 2] > 0).sum())
        for x in keep:
            x.free()
    _assert_tables(ctx, tables)
    assert dl_aborts > 0 and no_lost_to_dl > 0


def test_delivery_needs_the_fast_check(gpu):
    ctx = L.Context(workload=L.WORKLOAD_TPCC, tpcc_warehouses=1, max_txns=64, flags=L.FLAG_CONFLICT_COUNTS)
    ctx.tpcc_submit(tc.make_batch([], []))
    with pytest.raises(L.LtpgError) as ei:
        ctx.tpcc_submit_delivery(L.DeliveryBatch(1))
    assert ei.value.code == L.ERR_UNSUPPORTED


def test_delivery_c3_scale_matches_oracle(gpu):
    """W=64, 2^18 transactions with the reference's set sizes (TPCC-alltxn/
    Predefine.h:19-23): 1/128 of the batch each Order-Status, Delivery and
    Stock-Level, the rest NewOrder / Payment half and half."""
    W, n = 64, 1 << 18
    k = n // 128
    gen = L.TpccGenerator(warehouses=W, seed=SEED)
    b = gen.next(n - 3 * k)
    ro = L.readonly_batch(b, W, k, k, seed=7, tid_lo=0, tid_hi=n)
    dl = L.delivery_batch(b, W, k, seed=8, tid_lo=0, tid_hi=n)
    ctx = _ctx(W, n)
    ctx.tpcc_populate(SEED)
    tables = oracle.tpcc_populate(W, SEED)
    ctx.tpcc_submit(b)
    ctx.tpcc_submit_readonly(ro)
    ctx.tpcc_submit_delivery(dl)
    ctx.epoch_run()
    ctx.sync()
    o = oracle.TpccOracle(W).epoch(tables, b, 0, 1, ro=ro, dl=dl)
    np.testing.assert_array_equal(ctx.tpcc_bitmap(L.TYPE_NEWORDER), o.no_bitmap)
    np.testing.assert_array_equal(ctx.tpcc_bitmap(L.TYPE_PAYMENT), o.pay_bitmap)
    np.testing.assert_array_equal(ctx.tpcc_bitmap(L.TYPE_ORDERSTATUS), o.ro["os_bitmap"])
    _assert_dl(ctx, o, dl)
    _assert_tables(ctx, tables)


def test_graph_form_with_delivery_matches_plain_loop_and_oracle(gpu):
    """Order-entry batches of uneven size carrying Deliveries (and
    read-only transactions): the graph gives the plain loop's commit sets of
    every type and its tables, and both the oracle's."""
    W = 2
    rng = np.random.default_rng(27)
    gen = L.TpccGenerator(warehouses=W, seed=29)
    batches = []
    for i in range(10):
        n = int(np.exp(rng.uniform(np.log(64), np.log(1500))))
        b = gen.next(n, tid_base=i * 10000)
        ro = L.readonly_batch(b, W, n // 8 + 1, n // 8 + 1, seed=180 + i, tid_lo=i * 10000)
        dl = L.delivery_batch(b, W, n // 6 + 1, seed=280 + i, tid_lo=i * 10000, tid_hi=i * 10000 + n)
        batches.append((b, ro, dl))
    types = (L.TYPE_NEWORDER, L.TYPE_PAYMENT, L.TYPE_ORDERSTATUS, L.TYPE_STOCKLEVEL, L.TYPE_DELIVERY)
    out = {}
    for mode in (L.RUN_PLAIN, L.RUN_GRAPH):
        tp = _ctx(W, 4096)
        tp.tpcc_populate(SEED)
        keep, descs = [], []
        for b, ro, dl in batches:
            dev, bufs = L.upload_tpcc_batch(tp, b)
            rdev, rbufs = L.upload_readonly_batch(tp, ro)
            ddev, dbufs = L.upload_delivery_batch(tp, dl)
            keep += bufs + rbufs + dbufs
            descs.append(L.tpcc_desc(dev, rdev, ddev))
        r = L.Batched(None, tp, descs)
        r.run(mode)
        out[mode] = ([[r.bitmap(i, t) for t in types] for i in range(len(batches))],
                     [tp.tpcc_table_read(t) for t in range(9)])
        r.close()
    for a, b_ in zip(out[L.RUN_PLAIN][0], out[L.RUN_GRAPH][0]):
        for x, y in zip(a, b_):
            np.testing.assert_array_equal(x, y)
    for x, y in zip(out[L.RUN_PLAIN][1], out[L.RUN_GRAPH][1]):
        np.testing.assert_array_equal(x, y)
    tables = oracle.tpcc_populate(W, SEED)
    ora = oracle.TpccOracle(W)
    dl_commits = 0
    for i, (b, ro, dl) in enumerate(batches):
        ep = ora.epoch(tables, b, ro=ro, dl=dl)
        got = out[L.RUN_GRAPH][0][i]
        np.testing.assert_array_equal(got[0], ep.no_bitmap)
        np.testing.assert_array_equal(got[1], ep.pay_bitmap)
        np.testing.assert_array_equal(got[2], ep.ro["os_bitmap"])
        np.testing.assert_array_equal(got[3], ep.ro["sl_bitmap"])
        np.testing.assert_array_equal(got[4], ep.dl["bitmap"])
        dl_commits += ep.dl["commits"]
    assert dl_commits > 0
    for t in range(9):
        np.testing.assert_array_equal(out[L.RUN_GRAPH][1][t], tables[t])


@pytest.mark.parametrize("plain", [0, 1], ids=["served", "plain-form"])
def test_served_delivery_requests_replay_on_the_oracle(gpu, plain):
    """Delivery requests beside NewOrders and Payments (half of them on rows
    those write): the server's epochs -- NewOrders, Payments, Deliveries in
    member order, priorities the members' arrival ranks -- replayed through
    the oracle reproduce every decision, read value and table, and every
    request commits."""
    W = 2
    tp = _ctx(W, 2048)
    tp.tpcc_populate(SEED)
    gen = L.TpccGenerator(warehouses=W, seed=41)
    sv = L.Server(None, tp, slots=2, max_epoch_txns=700, keep_reads=1, record_epochs=1, plain=plain)
    index = {}  # request id -> (kind, batch, position)
    srcs = []
    for bi in range(3):
        b = gen.next(800)
        dl = L.delivery_batch(b, W, 200, seed=160 + bi)
        srcs.append((b, dl))
        first = sv.submit_tpcc(b)
        for q in range(b.n_no):
            index[first + q] = ("no", bi, q)
        for p in range(b.n_pay):
            index[first + b.n_no + p] = ("pay", bi, p)
        first = sv.submit_delivery(dl)
        for d in range(dl.n):
            index[first + d] = ("dl", bi, d)
        time.sleep(0.002)
    sv.drain()
    done = {d["id"]: d for d in sv.poll()}
    st = sv.stats()
    epochs = [sv.epoch(L.WORKLOAD_TPCC, e) for e in range(st["epochs"][1])]
    reads = {rid: sv.read_values(rid) for rid in done}
    sv.stop()
    assert len(done) == len(index)
    assert st["committed"][L.TYPE_DELIVERY] == 600
    assert all(done[rid]["type"] == L.TYPE_DELIVERY for rid in index if index[rid][0] == "dl")
    tables = oracle.tpcc_populate(W, SEED)
    ora = oracle.TpccOracle(W)
    requeued = 0
    for e, ids in enumerate(epochs):
        groups = {k: [int(i) for i in ids if index[int(i)][0] == k] for k in ("no", "pay", "dl")}
        assert [int(i) for i in ids] == groups["no"] + groups["pay"] + groups["dl"]
        eb = L.TpccBatch.empty(len(groups["no"]), len(groups["pay"]))
        for q, rid in enumerate(groups["no"]):
            _, bi, pos = index[rid]
            src = srcs[bi][0]
            for f in ("W_ID", "D_ID", "C_ID", "O_ID", "N_O_ID", "O_OL_CNT", "O_OL_ID"):
                eb.nf(f)[q] = src.nf(f)[pos]
            for f in ("OL_I_ID", "OL_SUPPLY_W_ID", "OL_QUANTITY"):
                for l in range(15):
                    eb.nf(f, l)[q] = src.nf(f, l)[pos]
            eb.nf("TID")[q] = done[rid]["tid"]
        for p, rid in enumerate(groups["pay"]):
            _, bi, pos = index[rid]
            src = srcs[bi][0]
            for f in ("W_ID", "D_ID", "C_ID", "C_LAST", "isName", "C_D_ID", "C_W_ID", "H_AMOUNT", "H_ID"):
                eb.pf(f)[p] = src.pf(f)[pos]
            eb.pf("TID")[p] = done[rid]["tid"]
        ed = L.DeliveryBatch(len(groups["dl"]))
        for d, rid in enumerate(groups["dl"]):
            _, bi, pos = index[rid]
            for f in L.DELIVERY_FIELDS:
                getattr(ed, f)[d] = done[rid]["tid"] if f == "TID" else getattr(srcs[bi][1], f)[pos]
        ep = ora.epoch(tables, eb, dl=ed)
        decided = [("no", ep.no_bitmap, ep.no_reads, 51), ("pay", ep.pay_bitmap, ep.pay_reads, 24),
                   ("dl", ep.dl["bitmap"], ep.dl["reads"], L.DELIVERY_READS)]
        for kind, bitmap, ora_reads, per in decided:
            n = len(groups[kind])
            for j, rid in enumerate(groups[kind]):
                c = bool((int(bitmap[j // 64]) >> (j % 64)) & 1)
                assert c == (done[rid]["epoch"] == e), f"{kind} request {rid} epoch {e}"
                requeued += not c
                if c:
                    np.testing.assert_array_equal(reads[rid], np.asarray(ora_reads).reshape(per, n)[:, j])
    assert requeued > 0
    for t in range(9):
        np.testing.assert_array_equal(tp.tpcc_table_read(t), tables[t])
