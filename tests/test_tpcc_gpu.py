"""Order-entry parity tests: the HIP path (through the C ABI) against the
oracle.

Bar: bit-exact NewOrder and Payment commit bitmaps, conflict counters,
resolved C_IDs, read values of committed transactions, range-error counts
and all nine tables after the epoch, on the same seeded batches; identical
results across repeated runs.
"""
import numpy as np
import pytest

import ltpg_amd as L
from oracle import oracle
from tests import tpcccases as tc

pytestmark = pytest.mark.gpu

SEED = 0x4C545047


def _ctx(W, max_txns, rules=L.RULES_REFERENCE, values=L.VALUES_RMW, counts=True):
    return L.Context(workload=L.WORKLOAD_TPCC, tpcc_warehouses=W, max_txns=max_txns, rules=rules, values=values,
                     flags=L.FLAG_CONFLICT_COUNTS if counts else 0)


def _gpu_epoch(ctx, batch, phases=(L.PHASE_ALL,), expect_range=False):
    ctx.tpcc_submit(batch)
    for ph in phases:
        ctx.epoch_run(ph)
    if expect_range:
        with pytest.raises(L.LtpgError) as ei:
            ctx.sync()
        assert ei.value.code == L.ERR_RANGE
    else:
        ctx.sync()
    counts = bool(ctx.config.flags & L.FLAG_CONFLICT_COUNTS)
    return dict(
        no_bitmap=ctx.tpcc_bitmap(L.TYPE_NEWORDER), pay_bitmap=ctx.tpcc_bitmap(L.TYPE_PAYMENT),
        no_conf=ctx.tpcc_conflicts(L.TYPE_NEWORDER) if counts else None,
        pay_conf=ctx.tpcc_conflicts(L.TYPE_PAYMENT) if counts else None,
        cid=ctx.tpcc_payment_cid(), no_reads=ctx.tpcc_read_values(L.TYPE_NEWORDER),
        pay_reads=ctx.tpcc_read_values(L.TYPE_PAYMENT),
        no_count=ctx.commit_count(L.TYPE_NEWORDER), pay_count=ctx.commit_count(L.TYPE_PAYMENT))


def _tables(ctx):
    return [ctx.tpcc_table_read(t) for t in range(9)]


def _assert_epoch(g, ora, batch):
    np.testing.assert_array_equal(g["no_bitmap"], ora.no_bitmap)
    np.testing.assert_array_equal(g["pay_bitmap"], ora.pay_bitmap)
    np.testing.assert_array_equal(g["cid"], ora.pay_cid)
    if g["no_conf"] is not None:
        np.testing.assert_array_equal(g["no_conf"], ora.no_conflicts)
        np.testing.assert_array_equal(g["pay_conf"], ora.pay_conflicts)
    assert g["no_count"] == ora.no_commits and g["pay_count"] == ora.pay_commits
    for key, nv, n, bm in (("no_reads", 51, batch.n_no, ora.no_bitmap), ("pay_reads", 24, batch.n_pay, ora.pay_bitmap)):
        if n == 0:
            continue
        com = np.array(tc.bits(bm, n), dtype=bool)
        gr = g[key].reshape(nv, n)[:, com]
        orr = getattr(ora, key).reshape(nv, n)[:, com]
        np.testing.assert_array_equal(gr, orr)


def _assert_tables(ctx, tables):
    for t in range(9):
        np.testing.assert_array_equal(ctx.tpcc_table_read(t), tables[t], err_msg=f"table {L.TPCC_TABLES[t]}")


def test_populate_matches_loader(gpu):
    ctx = _ctx(2, 1024)
    ctx.tpcc_populate(SEED)
    _assert_tables(ctx, oracle.tpcc_populate(2, SEED))


@pytest.mark.parametrize("layout", [L.LAYOUT_COLUMN_MAJOR, L.LAYOUT_ROW_MAJOR])
def test_table_load_read_round_trip(gpu, layout):
    ctx = _ctx(1, 64)
    rng = np.random.default_rng(3)
    for t in range(9):
        rows, cols = ctx.tpcc_shape(t)
        data = rng.integers(0, 2**32, size=rows * cols, dtype=np.uint64).astype(np.uint32)
        ctx.tpcc_table_load(t, data, layout)
        np.testing.assert_array_equal(ctx.tpcc_table_read(t, layout), data)
    # the two layouts are transposes of each other
    rows, cols = ctx.tpcc_shape(L.TPCC_STOCK)
    cm = ctx.tpcc_table_read(L.TPCC_STOCK, L.LAYOUT_COLUMN_MAJOR).reshape(cols, rows)
    rm = ctx.tpcc_table_read(L.TPCC_STOCK, L.LAYOUT_ROW_MAJOR).reshape(rows, cols)
    np.testing.assert_array_equal(cm.T, rm)


CASES = [(seed, rules, values, counts) for seed in range(3) for rules in (L.RULES_REFERENCE, L.RULES_INTENDED)
         for values in (L.VALUES_RMW, L.VALUES_COPYBACK) for counts in (True, False)]


@pytest.mark.parametrize("seed,rules,values,counts", CASES)
def test_seeded_epochs_match_oracle(gpu, seed, rules, values, counts):
    W = 2
    ctx = _ctx(W, 8192, rules, values, counts)
    ctx.tpcc_populate(SEED)
    tables = oracle.tpcc_populate(W, SEED)
    ora = oracle.TpccOracle(W)
    gen = L.TpccGenerator(warehouses=W, seed=100 + seed)
    for ep in range(3):
        b = gen.next(3000, tid_base=ep * 3000)
        g = _gpu_epoch(ctx, b)
        o = ora.epoch(tables, b, rules, values)
        assert o.range_errors == 0
        _assert_epoch(g, o, b)
    _assert_tables(ctx, tables)


@pytest.mark.parametrize("rules", [L.RULES_REFERENCE, L.RULES_INTENDED], ids=["reference", "intended"])
@pytest.mark.parametrize("counts", [True, False], ids=["counting", "fast"])
def test_small_blocks_shape_matches_oracle(gpu, rules, counts):
    """TUNE_TPCC_SMALL_BLOCKS (64-thread workgroups: 4 NewOrders or 64
    Payments each) against the oracle, switched on and off between epochs;
    odd batch sizes leave partial workgroups of both roles."""
    W = 2
    ctx = _ctx(W, 8192, rules, L.VALUES_RMW, counts)
    ctx.tpcc_populate(SEED)
    tables = oracle.tpcc_populate(W, SEED)
    ora = oracle.TpccOracle(W)
    gen = L.TpccGenerator(warehouses=W, seed=321)
    for ep, n in enumerate((3001, 777, 64, 2050)):
        ctx.tpcc_set_tuning(L.TUNE_TPCC_SMALL_BLOCKS if ep != 2 else 0)
        b = gen.next(n, tid_base=ep * 4000)
        g = _gpu_epoch(ctx, b)
        o = ora.epoch(tables, b, rules, L.VALUES_RMW)
        _assert_epoch(g, o, b)
    _assert_tables(ctx, tables)
    with pytest.raises(L.LtpgError):
        ctx.tpcc_set_tuning(L.TUNE_PARTITIONED)
    ctx.close()


@pytest.mark.parametrize("counts", [True, False], ids=["counting", "fast"])
def test_empty_and_one_sided_epochs_match_oracle(gpu, counts):
    """Epochs with no transactions, only NewOrders or only Payments between
    mixed ones: each matches the oracle, an empty one commits nothing and
    starts its counts at zero, and the tables carry over."""
    W = 2
    ctx = _ctx(W, 4096, counts=counts)
    ctx.tpcc_populate(SEED)
    tables = oracle.tpcc_populate(W, SEED)
    ora = oracle.TpccOracle(W)
    gens = {pct: L.TpccGenerator(warehouses=W, seed=60 + pct, neworder_pct=pct) for pct in (50, 100, 0)}
    tid = 0
    for pct, n in [(50, 2000), (None, 0), (100, 1500), (0, 1500), (None, 0), (50, 2000), (None, 0)]:
        b = L.TpccBatch.empty(0, 0) if pct is None else gens[pct].next(n, tid_base=tid)
        tid += n
        g = _gpu_epoch(ctx, b)
        o = ora.epoch(tables, b, L.RULES_REFERENCE, L.VALUES_RMW)
        assert o.range_errors == 0
        _assert_epoch(g, o, b)
        if pct is None:
            assert g["no_count"] == 0 and g["pay_count"] == 0
            assert g["no_bitmap"].size == 0 and g["pay_bitmap"].size == 0
    _assert_tables(ctx, tables)


def test_fast_and_counting_checks_agree_at_scale(gpu):
    W = 8
    gen_a = L.TpccGenerator(warehouses=W, seed=9)
    gen_b = L.TpccGenerator(warehouses=W, seed=9)
    a, b = _ctx(W, 1 << 16, counts=True), _ctx(W, 1 << 16, counts=False)
    a.tpcc_populate(SEED)
    b.tpcc_populate(SEED)
    for ep in range(2):
        ba, bb = gen_a.next(1 << 16, ep << 16), gen_b.next(1 << 16, ep << 16)
        ga, gb = _gpu_epoch(a, ba), _gpu_epoch(b, bb)
        for k in ("no_bitmap", "pay_bitmap", "cid", "no_count", "pay_count"):
            np.testing.assert_array_equal(ga[k], gb[k])
    for t in range(9):
        np.testing.assert_array_equal(a.tpcc_table_read(t), b.tpcc_table_read(t))


def test_phases_run_separately_match_all(gpu):
    W = 2
    gen = L.TpccGenerator(warehouses=W, seed=21)
    batch = gen.next(4000)
    res = []
    for phases in ((L.PHASE_ALL,), (L.PHASE_CID, L.PHASE_EXECUTE, L.PHASE_CHECK, L.PHASE_WRITEBACK)):
        ctx = _ctx(W, 4096)
        ctx.tpcc_populate(SEED)
        res.append((_gpu_epoch(ctx, batch, phases), _tables(ctx)))
    for k in ("no_bitmap", "pay_bitmap", "cid", "no_conf", "pay_conf"):
        np.testing.assert_array_equal(res[0][0][k], res[1][0][k])
    for t in range(9):
        np.testing.assert_array_equal(res[0][1][t], res[1][1][t])


EDGE = tc.make_batch(
    [
        # two NewOrders on one stock row (item 7 of warehouse 0): waw
        dict(W_ID=0, D_ID=1, C_ID=5, O_ID=10, N_O_ID=10, O_OL_ID=100, TID=3, lines=[(7, 0, 4), (8, 0, 1)]),
        dict(W_ID=0, D_ID=2, C_ID=6, O_ID=11, N_O_ID=11, O_OL_ID=110, TID=4, lines=[(7, 0, 2)]),
        # the same stock row twice in one order, small quantity (+91 rule)
        dict(W_ID=0, D_ID=3, C_ID=7, O_ID=12, N_O_ID=12, O_OL_ID=120, TID=5, lines=[(9, 0, 6), (9, 0, 7)]),
        # two inserts of one NEWORDER row, different ORDER rows
        dict(W_ID=0, D_ID=4, C_ID=8, O_ID=13, N_O_ID=20, O_OL_ID=130, TID=6, lines=[(11, 0, 1)]),
        dict(W_ID=0, D_ID=5, C_ID=9, O_ID=14, N_O_ID=20, O_OL_ID=140, TID=7, lines=[(12, 0, 1)]),
        # reads a customer a lower-TID Payment writes: raw only, commits
        dict(W_ID=0, D_ID=6, C_ID=42, O_ID=15, N_O_ID=21, O_OL_ID=150, TID=9, lines=[(13, 0, 1)]),
        # remote supply warehouse and 15 lines
        dict(W_ID=1, D_ID=0, C_ID=1, O_ID=1, N_O_ID=1, O_OL_ID=0, TID=10,
             lines=[(100 + i, i % 2, 1 + i % 3) for i in range(15)]),
        # out-of-range warehouse and item (range errors)
        dict(W_ID=5, D_ID=0, C_ID=0, O_ID=0, N_O_ID=0, O_OL_ID=0, TID=11, lines=[(200000, 0, 1)]),
    ],
    [
        dict(W_ID=0, D_ID=6, C_ID=42, C_D_ID=6, C_W_ID=0, H_AMOUNT=3, H_ID=1, TID=1),
        # same customer, higher TID: raw, war and waw
        dict(W_ID=0, D_ID=6, C_ID=42, C_D_ID=6, C_W_ID=0, H_AMOUNT=2, H_ID=2, TID=8),
        # by name: name 3, ordinal 4 -> C_ID 83; ordinal 200 -> none (C_ID 0)
        dict(W_ID=1, D_ID=2, isName=1, C_LAST=(3 << 8) + 4, C_D_ID=9, C_W_ID=1, H_AMOUNT=5, H_ID=3, TID=12),
        dict(W_ID=1, D_ID=2, isName=1, C_LAST=(3 << 8) + 200, C_D_ID=9, C_W_ID=0, H_AMOUNT=1, H_ID=4, TID=13),
        # history collision with the first Payment
        dict(W_ID=0, D_ID=0, C_ID=100, C_D_ID=0, C_W_ID=0, H_AMOUNT=1, H_ID=1, TID=14),
        # equal TIDs are ordered by batch position
        dict(W_ID=1, D_ID=1, C_ID=500, C_D_ID=1, C_W_ID=1, H_AMOUNT=4, H_ID=9, TID=12),
        dict(W_ID=1, D_ID=1, C_ID=500, C_D_ID=1, C_W_ID=1, H_AMOUNT=4, H_ID=10, TID=12),
    ])


@pytest.mark.parametrize("rules", [L.RULES_REFERENCE, L.RULES_INTENDED])
@pytest.mark.parametrize("counts", [True, False])
def test_edge_batch_matches_oracle(gpu, rules, counts):
    W = 2
    tables = oracle.tpcc_populate(W, SEED)
    # a stock row with a low quantity exercises the restock branch
    rows = tc.rows_of(W, tc.ST)
    tables[tc.ST][2 * rows + 9] = 12
    ctx = _ctx(W, 64, rules, L.VALUES_RMW, counts)
    for t in range(9):
        ctx.tpcc_table_load(t, tables[t])
    g = _gpu_epoch(ctx, EDGE, expect_range=True)
    o = oracle.TpccOracle(W).epoch(tables, EDGE, rules, L.VALUES_RMW)
    assert o.range_errors > 0
    _assert_epoch(g, o, EDGE)
    _assert_tables(ctx, tables)


def test_device_resident_batch(gpu):
    W = 2
    gen = L.TpccGenerator(warehouses=W, seed=33)
    batch = gen.next(2000)
    host = _ctx(W, 2048)
    host.tpcc_populate(SEED)
    ref = _gpu_epoch(host, batch)
    dev = _ctx(W, 2048)
    dev.tpcc_populate(SEED)
    nob = L.DeviceBuffer(dev, batch.no.nbytes)
    pab = L.DeviceBuffer(dev, batch.pay.nbytes)
    nob.upload(batch.no)
    pab.upload(batch.pay)
    nb = L.NeworderBatch()
    pb = L.PaymentBatch()
    # view the device blocks with the host view helper's offsets
    hv_n = L.neworder_struct(batch.no, batch.no_cap, batch.n_no)
    hv_p = L.payment_struct(batch.pay, batch.pay_cap, batch.n_pay)
    nb.n, nb.line_stride = hv_n.n, hv_n.line_stride
    for name, _ in L.NeworderBatch._fields_[2:]:
        setattr(nb, name, nob.ptr + (getattr(hv_n, name) - batch.no.ctypes.data))
    pb.n = hv_p.n
    for name, _ in L.PaymentBatch._fields_[2:]:
        setattr(pb, name, pab.ptr + (getattr(hv_p, name) - batch.pay.ctypes.data))
    dev.tpcc_submit((nb, pb), mem=L.MEM_DEVICE)
    dev.epoch_run(L.PHASE_ALL)
    dev.sync()
    np.testing.assert_array_equal(dev.tpcc_bitmap(L.TYPE_NEWORDER), ref["no_bitmap"])
    np.testing.assert_array_equal(dev.tpcc_bitmap(L.TYPE_PAYMENT), ref["pay_bitmap"])
    np.testing.assert_array_equal(dev.tpcc_payment_cid(), ref["cid"])
    for t in range(9):
        np.testing.assert_array_equal(dev.tpcc_table_read(t), host.tpcc_table_read(t))
    nob.free()
    pab.free()


def test_epoch_tag_wrap_keeps_results(gpu):
    """Register tags wrap after 255 epochs; results stay equal to the oracle."""
    W = 1
    ctx = _ctx(W, 512)
    ctx.tpcc_populate(SEED)
    tables = oracle.tpcc_populate(W, SEED)
    ora = oracle.TpccOracle(W)
    gen = L.TpccGenerator(warehouses=W, seed=77)
    for ep in range(300):
        b = gen.next(256, tid_base=ep * 256)
        g = _gpu_epoch(ctx, b)
        o = ora.epoch(tables, b)
        np.testing.assert_array_equal(g["no_bitmap"], o.no_bitmap)
        np.testing.assert_array_equal(g["pay_bitmap"], o.pay_bitmap)
    _assert_tables(ctx, tables)


def test_c3_full_size_matches_oracle(gpu):
    """W = 64, 2^18 transactions (131,072 NewOrder + 131,072 Payment), NURand
    items, 1 % remote supply, 15 % remote customers, 60 % by name."""
    W, n = 64, 1 << 18
    ctx = _ctx(W, n, counts=False)
    ctx.tpcc_populate(SEED)
    tables = oracle.tpcc_populate(W, SEED)
    ora = oracle.TpccOracle(W)
    gen = L.TpccGenerator(warehouses=W)
    for ep in range(2):
        b = gen.next(n, tid_base=ep * n)
        g = _gpu_epoch(ctx, b)
        o = ora.epoch(tables, b)
        assert o.range_errors == 0
        _assert_epoch(g, o, b)
    _assert_tables(ctx, tables)


def test_c3_determinism(gpu):
    W, n = 64, 1 << 18
    b = L.TpccGenerator(warehouses=W).next(n)
    out = []
    for _ in range(2):
        ctx = _ctx(W, n, counts=False)
        ctx.tpcc_populate(SEED)
        out.append((_gpu_epoch(ctx, b), [ctx.tpcc_table_read(t) for t in (tc.ST, tc.CU, tc.OL)]))
        ctx.close()
    for k in ("no_bitmap", "pay_bitmap", "cid", "no_reads", "pay_reads"):
        np.testing.assert_array_equal(out[0][0][k], out[1][0][k])
    for x, y in zip(out[0][1], out[1][1]):
        np.testing.assert_array_equal(x, y)


# ---- known answers and read-only transactions --------------------------------

def _kat_tables(case):
    W = case["W"]
    tables = oracle.tpcc_populate(W, tc.load_kat()["seed"])
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


KAT = tc.load_kat()


# cases with Deliveries run in tests/test_delivery_gpu.py
_KAT_PARAMS = [pytest.param(c, counts, id=f"{c['name']}-{'counting' if counts else 'fast'}")
               for c in KAT["cases"] if "delivery" not in c
               for counts in ((False,) if "orderstatus" in c or "stocklevel" in c else (True, False))]


@pytest.mark.parametrize("case,counts", _KAT_PARAMS)
def test_gpu_known_answers(gpu, case, counts):
    """Every hand-derived order-entry epoch, on both check forms (read-only
    transactions exist on the fast one only)."""
    ro = _kat_readonly(case)
    W = case["W"]
    rules = L.RULES_REFERENCE if case["rules"] == "reference" else L.RULES_INTENDED
    values = L.VALUES_RMW if case["values"] == "rmw" else L.VALUES_COPYBACK
    tables = _kat_tables(case)
    ctx = _ctx(W, 64, rules, values, counts)
    for t in range(9):
        ctx.tpcc_table_load(t, tables[t])
    nos = [dict(n, lines=[tuple(x) for x in n.get("lines", [])]) for n in case["neworders"]]
    b = tc.make_batch(nos, case["payments"])
    exp = case["expect"]
    ctx.tpcc_submit(b)
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
    if counts:
        assert tc.unpack(ctx.tpcc_conflicts(L.TYPE_NEWORDER)).tolist() == exp["no_conf"]
        assert tc.unpack(ctx.tpcc_conflicts(L.TYPE_PAYMENT)).tolist() == exp["pay_conf"]
    assert ctx.tpcc_payment_cid().tolist() == exp["cid"]
    no_r, pay_r = ctx.tpcc_read_values(L.TYPE_NEWORDER), ctx.tpcc_read_values(L.TYPE_PAYMENT)
    for q, v, val in exp["no_reads"]:
        assert int(no_r[v * b.n_no + q]) == val
    for p, v, val in exp["pay_reads"]:
        assert int(pay_r[v * b.n_pay + p]) == val
    if ro is not None:
        assert tc.bits(ctx.tpcc_bitmap(L.TYPE_ORDERSTATUS), ro.n_os) == exp["os_commit"]
        assert tc.bits(ctx.tpcc_bitmap(L.TYPE_STOCKLEVEL), ro.n_sl) == exp["sl_commit"]
        os_r, sl_r = ctx.tpcc_read_values(L.TYPE_ORDERSTATUS), ctx.tpcc_read_values(L.TYPE_STOCKLEVEL)
        for i, v, val in exp["os_reads"]:
            assert int(os_r[v * ro.n_os + i]) == val
        for i, v, val in exp["sl_reads"]:
            assert int(sl_r[v * ro.n_sl + i]) == val
        assert ctx.commit_count(L.TYPE_ORDERSTATUS) == ro.n_os
        assert ctx.commit_count(L.TYPE_STOCKLEVEL) == ro.n_sl
    for t, row, col, v in exp["cells"]:
        tables[t][col * tc.rows_of(W, t) + row] = np.uint32(v)
    _assert_tables(ctx, tables)


def _assert_readonly(ctx, ora, ro):
    r = ora.ro
    np.testing.assert_array_equal(ctx.tpcc_bitmap(L.TYPE_ORDERSTATUS), r["os_bitmap"])
    np.testing.assert_array_equal(ctx.tpcc_bitmap(L.TYPE_STOCKLEVEL), r["sl_bitmap"])
    np.testing.assert_array_equal(ctx.tpcc_read_values(L.TYPE_ORDERSTATUS), r["os_reads"])
    np.testing.assert_array_equal(ctx.tpcc_read_values(L.TYPE_STOCKLEVEL), r["sl_reads"])


@pytest.mark.parametrize("rules", [L.RULES_REFERENCE, L.RULES_INTENDED], ids=["reference", "intended"])
@pytest.mark.parametrize("values", [L.VALUES_RMW, L.VALUES_COPYBACK], ids=["rmw", "copyback"])
@pytest.mark.parametrize("device", [False, True], ids=["host-batch", "device-batch"])
def test_readonly_seeded_epochs_match_oracle(gpu, rules, values, device):
    """Order-Status / Stock-Level mixed into seeded batches, half of them on
    rows the batch's NewOrders write: every type's bitmap and read values,
    the counts and the tables match the oracle over three epochs."""
    W = 2
    ctx = _ctx(W, 8192, rules, values, counts=False)
    ctx.tpcc_populate(SEED)
    tables = oracle.tpcc_populate(W, SEED)
    ora = oracle.TpccOracle(W)
    gen = L.TpccGenerator(warehouses=W, seed=400)
    aborted_by_readers = 0  # NewOrders with war from the readers
    for ep in range(3):
        b = gen.next(3000, tid_base=ep * 4000)
        ro = L.readonly_batch(b, W, 500, 500, seed=40 + ep, tid_lo=ep * 4000, tid_hi=ep * 4000 + 3000)
        keep = []
        if device:
            dev, bufs = L.upload_tpcc_batch(ctx, b)
            rdev, rbufs = L.upload_readonly_batch(ctx, ro)
            keep = bufs + rbufs
            ctx.tpcc_submit(dev, mem=L.MEM_DEVICE)
            ctx.tpcc_submit_readonly(rdev, mem=L.MEM_DEVICE)
        else:
            ctx.tpcc_submit(b)
            ctx.tpcc_submit_readonly(ro)
        ctx.epoch_run()
        ctx.sync()
        g = dict(no_bitmap=ctx.tpcc_bitmap(L.TYPE_NEWORDER), pay_bitmap=ctx.tpcc_bitmap(L.TYPE_PAYMENT),
                 no_conf=None, pay_conf=None, cid=ctx.tpcc_payment_cid(),
                 no_reads=ctx.tpcc_read_values(L.TYPE_NEWORDER), pay_reads=ctx.tpcc_read_values(L.TYPE_PAYMENT),
                 no_count=ctx.commit_count(L.TYPE_NEWORDER), pay_count=ctx.commit_count(L.TYPE_PAYMENT))
        o = ora.epoch(tables, b, rules, values, ro=ro)
        _assert_epoch(g, o, b)
        _assert_readonly(ctx, o, ro)
        # NewOrders the readers gave war (an abort needs raw too: the KATs
        # pin that case; here raw is rare, from Payments of the same customer)
        c = tc.unpack(o.no_conflicts)
        aborted_by_readers += int((c[:, 1] > 0).sum())
        for x in keep:
            x.free()
    _assert_tables(ctx, tables)
    assert aborted_by_readers > 0


def test_readonly_needs_the_fast_check(gpu):
    ctx = _ctx(1, 64, counts=True)
    b = tc.make_batch([], [])
    ctx.tpcc_submit(b)
    with pytest.raises(L.LtpgError) as ei:
        ctx.tpcc_submit_readonly(L.ReadOnlyBatch(1, 0))
    assert ei.value.code == L.ERR_UNSUPPORTED


def test_readonly_c3_scale_matches_oracle(gpu):
    """W=64, 2^18 transactions: 45 % NewOrder, 43 % Payment, 6 % Order-Status,
    6 % Stock-Level."""
    W, n = 64, 1 << 18
    n_ro = n * 6 // 100
    gen = L.TpccGenerator(warehouses=W, seed=SEED, neworder_pct=51)
    b = gen.next(n - 2 * n_ro)
    ro = L.readonly_batch(b, W, n_ro, n_ro, seed=7, tid_lo=0, tid_hi=n - 2 * n_ro)
    ctx = _ctx(W, n, counts=False)
    ctx.tpcc_populate(SEED)
    tables = oracle.tpcc_populate(W, SEED)
    ctx.tpcc_submit(b)
    ctx.tpcc_submit_readonly(ro)
    ctx.epoch_run()
    ctx.sync()
    o = oracle.TpccOracle(W).epoch(tables, b, 0, 1, ro=ro)
    np.testing.assert_array_equal(ctx.tpcc_bitmap(L.TYPE_NEWORDER), o.no_bitmap)
    np.testing.assert_array_equal(ctx.tpcc_bitmap(L.TYPE_PAYMENT), o.pay_bitmap)
    _assert_readonly(ctx, o, ro)
    _assert_tables(ctx, tables)
