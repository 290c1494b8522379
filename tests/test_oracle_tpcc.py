"""Pins the order-entry oracle (CPU only).

The reference ships no tests, fixtures or golden vectors for this path, so
the oracle is pinned by (1) hand-derived known-answer epochs
(tests/golden/tpcc_kat.json), (2) conflict counters recomputed pairwise
from the reference's dispatch tables, (3) Aria's serializability guarantee:
the committed set is executed one transaction at a time in a precedence
order, which must reproduce the oracle's tables and read values, (4) brute
force over every TID order of tiny batches, (5) the closed form of the
by-name lookup on the loader's data, and (6) copies of the oracle with one
planted mistake each, which these pins must reject.
"""
import itertools
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle
from tests import tpcccases as tc

import ltpg_amd as L

KAT = tc.load_kat()
SEED = KAT["seed"]
_TABLE_CACHE = {}


def _tables(W):
    if W not in _TABLE_CACHE:
        _TABLE_CACHE[W] = oracle.tpcc_populate(W, SEED)
    return [t.copy() for t in _TABLE_CACHE[W]]


def _apply(W, tables, cells):
    for t, row, col, v in cells:
        tables[t][col * tc.rows_of(W, t) + row] = np.uint32(v)


def _kat_batch(case):
    nos = [dict(n, lines=[tuple(x) for x in n.get("lines", [])]) for n in case["neworders"]]
    return tc.make_batch(nos, case["payments"])


def _kat_readonly(case):
    """ReadOnlyBatch of a case's Order-Status / Stock-Level lists, or None."""
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


def _check_kat(case, run):
    """run(tables, batch, rules, values, ro, dl) -> TpccEpoch; compares every
    expectation of the case, return False on the first mismatch."""
    W = case["W"]
    tables = _tables(W)
    _apply(W, tables, case["set"])
    initial = [t.copy() for t in tables]
    b = _kat_batch(case)
    ro = _kat_readonly(case)
    dl = tc.kat_delivery(case)
    rules = 0 if case["rules"] == "reference" else 1
    values = 1 if case["values"] == "rmw" else 0
    ep = run(tables, b, rules, values, ro, dl)
    exp = case["expect"]
    ok = (tc.bits(ep.no_bitmap, b.n_no) == exp["no_commit"] and tc.bits(ep.pay_bitmap, b.n_pay) == exp["pay_commit"]
          and tc.unpack(ep.no_conflicts).tolist() == exp["no_conf"]
          and tc.unpack(ep.pay_conflicts).tolist() == exp["pay_conf"]
          and ep.pay_cid.tolist() == exp["cid"] and ep.range_errors == exp["range_errors"]
          and ep.no_commits == sum(exp["no_commit"]) and ep.pay_commits == sum(exp["pay_commit"]))
    expected = [t.copy() for t in initial]
    _apply(W, expected, exp["cells"])
    ok = ok and all(np.array_equal(tables[t], expected[t]) for t in range(9))
    for q, v, val in exp["no_reads"]:
        ok = ok and int(ep.no_reads[v * b.n_no + q]) == val
    for p, v, val in exp["pay_reads"]:
        ok = ok and int(ep.pay_reads[v * b.n_pay + p]) == val
    if ro is not None:
        r = ep.ro
        ok = ok and tc.bits(r["os_bitmap"], ro.n_os) == exp["os_commit"]
        ok = ok and tc.bits(r["sl_bitmap"], ro.n_sl) == exp["sl_commit"]
        if r.get("os_conflicts") is not None:
            ok = ok and tc.unpack(r["os_conflicts"]).tolist() == exp["os_conf"]
            ok = ok and tc.unpack(r["sl_conflicts"]).tolist() == exp["sl_conf"]
        for i, v, val in exp["os_reads"]:
            ok = ok and int(r["os_reads"][v * ro.n_os + i]) == val
        for i, v, val in exp["sl_reads"]:
            ok = ok and int(r["sl_reads"][v * ro.n_sl + i]) == val
    if dl is not None:
        x = ep.dl
        ok = ok and tc.bits(x["bitmap"], dl.n) == exp["dl_commit"] and x["commits"] == sum(exp["dl_commit"])
        ok = ok and tc.unpack(x["conflicts"]).tolist() == exp["dl_conf"]
        for i, v, val in exp["dl_reads"]:
            ok = ok and int(x["reads"][v * dl.n + i]) == val
    return ok


@pytest.mark.parametrize("case", KAT["cases"], ids=[c["name"] for c in KAT["cases"]])
def test_oracle_known_answers(case):
    ora = oracle.TpccOracle(case["W"])
    assert _check_kat(case, lambda tables, b, r, v, ro, dl: ora.epoch(tables, b, r, v, ro=ro, dl=dl))


def test_loader_contents():
    """Spot values of create_* (Database.h:104-327) and the draw ranges."""
    W = 2
    t = _tables(W)

    def get(tab, row, col):
        return int(t[tab][col * tc.rows_of(W, tab) + row])

    assert [get(tc.WH, 1, c) for c in range(9)] == [1, 1, 1, 0, 0, 0, 10, 0, 0]
    assert [get(tc.DI, 13, c) for c in range(5)] == [3, 1, 1, 1, 0]
    r = 30000 + 2 * 3000 + 17
    assert [get(tc.CU, r, c) for c in (0, 1, 2, 3, 9, 13)] == [17, 1, 2, 100000, 100000, r % 20]
    assert get(tc.OR, 33001, 1) == 11 and get(tc.OR, 33001, 2) == 1
    assert [get(tc.OL, 450007, c) for c in (0, 1, 2, 5)] == [7, 10, 1, 1]
    assert [get(tc.ST, 100003, c) for c in (0, 1, 2, 3, 4, 6, 15)] == [3, 1, 100000, 100000, 0, 1, 10]
    assert [get(tc.IT, 99999, c) for c in (0, 1, 4)] == [99999, 1000, 99999]
    assert not t[tc.HI].any() and not t[tc.NO].any()
    price = t[tc.IT][2 * 100000:3 * 100000]
    data = t[tc.IT][3 * 100000:4 * 100000]
    credit = t[tc.CU][10 * tc.rows_of(W, tc.CU):11 * tc.rows_of(W, tc.CU)]
    o_c = t[tc.OR][3 * tc.rows_of(W, tc.OR):4 * tc.rows_of(W, tc.OR)]
    assert price.min() == 1 and price.max() == 40000 and abs(price.mean() - 20000.5) < 300
    assert data.min() == 1 and data.max() == 10
    assert set(np.unique(credit).tolist()) == {0, 1}
    assert o_c.min() == 0 and o_c.max() == 2999


def _random_batch(rng, W, n_no, n_pay, dense=True):
    """Small batches whose ids collide often."""
    nos, pays = [], []
    tids = rng.permutation(n_no + n_pay) * 2 + 1
    for q in range(n_no):
        nl = int(rng.integers(0, 4))
        nos.append(dict(W_ID=int(rng.integers(0, W)), D_ID=int(rng.integers(0, 2)), C_ID=int(rng.integers(0, 4)),
                        O_ID=int(rng.integers(0, 5)), N_O_ID=int(rng.integers(0, 5)),
                        O_OL_ID=int(rng.integers(0, 12)), TID=int(tids[q]),
                        lines=[(int(rng.integers(0, 6)), int(rng.integers(0, W)), int(rng.integers(1, 11)))
                               for _ in range(nl)]))
    for p in range(n_pay):
        by_name = rng.random() < 0.4
        pays.append(dict(W_ID=int(rng.integers(0, W)), D_ID=int(rng.integers(0, 2)),
                         C_ID=0 if by_name else int(rng.integers(0, 4)),
                         C_LAST=(int(rng.integers(0, 4)) << 8) + int(rng.integers(0, 2)) if by_name else 20,
                         isName=int(by_name), C_D_ID=int(rng.integers(0, 2)), C_W_ID=int(rng.integers(0, W)),
                         H_AMOUNT=int(rng.integers(1, 6)), H_ID=int(rng.integers(0, 5)),
                         TID=int(tids[n_no + p])))
    return tc.make_batch(nos, pays)


def _pins_hold(W, tables0, b, ep, rules, values=1):
    """(ok, why): pairwise conflicts, Aria commit, serial execution."""
    no_c, pay_c, bad = tc.pairwise_conflicts(b, W, ep.pay_cid, rules)
    if not np.array_equal(tc.unpack(ep.no_conflicts), no_c) or not np.array_equal(tc.unpack(ep.pay_conflicts), pay_c):
        return False, "conflict counters differ from the pairwise count"
    if bad != ep.range_errors:
        return False, "range errors"
    no_commit, pay_commit = tc.bits(ep.no_bitmap, b.n_no), tc.bits(ep.pay_bitmap, b.n_pay)
    if no_commit != tc.aria(no_c).tolist() or pay_commit != tc.aria(pay_c).tolist():
        return False, "commit is not waw == 0 && (raw == 0 || war == 0)"
    ok, why, final, reads = tc.serial_check(W, tables0, b, ep.pay_cid, no_commit, pay_commit, rmw=values == 1)
    if not ok:
        return False, why
    for (typ, i), vals in reads.items():
        flat, n, nv = (ep.no_reads, b.n_no, 51) if typ == "no" else (ep.pay_reads, b.n_pay, 24)
        if [int(flat[v * n + i]) for v in range(nv)] != vals:
            return False, f"read values of {typ} {i}"
    return True, final


@pytest.mark.parametrize("seed", range(30))
def test_oracle_pairwise_and_serializable(seed):
    rng = np.random.default_rng(seed)
    W = 1 + seed % 2
    rules = seed % 3 == 2
    b = _random_batch(rng, W, int(rng.integers(0, 10)), int(rng.integers(0, 10)))
    tables0 = _tables(W)
    tables = [t.copy() for t in tables0]
    ep = oracle.TpccOracle(W).epoch(tables, b, int(rules), 1)
    ok, res = _pins_hold(W, tables0, b, ep, int(rules))
    assert ok, res
    for t in range(9):
        np.testing.assert_array_equal(tables[t], res[t])
    # the highest-priority transaction always commits
    prios = [((int(b.nf("TID")[q]), q), "no", q) for q in range(b.n_no)] + \
            [((int(b.pf("TID")[p]), b.n_no + p), "pay", p) for p in range(b.n_pay)]
    if prios:
        _, typ, i = min(prios)
        assert tc.bits(ep.no_bitmap if typ == "no" else ep.pay_bitmap, i + 1)[i] == 1


def _random_readonly(rng, W, n_os, n_sl, tid_pool):
    """Order-Status / Stock-Level transactions whose rows collide with
    _random_batch's (customers 0..3, orders 0..4, order lines 0..14, items
    0..5)."""
    ro = L.ReadOnlyBatch(n_os, n_sl)
    tids = rng.permutation(tid_pool)[:n_os + n_sl]
    for i in range(n_os):
        ro.os["W_ID"][i], ro.os["D_ID"][i] = rng.integers(0, W), rng.integers(0, 2)
        ro.os["C_ID"][i], ro.os["O_ID"][i] = rng.integers(0, 4), rng.integers(0, 5)
        ro.os["OL_ID"][i], ro.os["TID"][i] = rng.integers(0, 15), tids[i]
    for i in range(n_sl):
        ro.sl["W_ID"][i], ro.sl["S_ID"][i], ro.sl["TID"][i] = rng.integers(0, W), rng.integers(0, 6), tids[n_os + i]
    return ro


def _pins_hold_ro(W, tables0, b, ro, ep, rules, values=1):
    """_pins_hold with read-only transactions: their raw counters, Aria's
    rule for every type (read-only ones always commit), and the serial
    execution of all committed transactions with every read value."""
    no_c, pay_c, bad, os_c, sl_c = tc.pairwise_conflicts(b, W, ep.pay_cid, rules, ro)
    r = ep.ro
    got = [tc.unpack(ep.no_conflicts), tc.unpack(ep.pay_conflicts), tc.unpack(r["os_conflicts"]),
           tc.unpack(r["sl_conflicts"])]
    for g, e, name in zip(got, (no_c, pay_c, os_c, sl_c), ("no", "pay", "os", "sl")):
        if not np.array_equal(g, e):
            return False, f"{name} conflict counters differ from the pairwise count"
    if bad != ep.range_errors:
        return False, "range errors"
    commits = {"no": tc.bits(ep.no_bitmap, b.n_no), "pay": tc.bits(ep.pay_bitmap, b.n_pay),
               "os": tc.bits(r["os_bitmap"], ro.n_os), "sl": tc.bits(r["sl_bitmap"], ro.n_sl)}
    for name, conf in (("no", no_c), ("pay", pay_c), ("os", os_c), ("sl", sl_c)):
        if commits[name] != tc.aria(conf).tolist():
            return False, f"{name} commit is not waw == 0 && (raw == 0 || war == 0)"
    ok, why, final, reads = tc.serial_check(W, tables0, b, ep.pay_cid, commits["no"], commits["pay"],
                                            rmw=values == 1, ro=ro, ro_commit={"os": commits["os"],
                                                                               "sl": commits["sl"]})
    if not ok:
        return False, why
    flats = {"no": (ep.no_reads, b.n_no, 51), "pay": (ep.pay_reads, b.n_pay, 24),
             "os": (r["os_reads"], ro.n_os, 3), "sl": (r["sl_reads"], ro.n_sl, 1)}
    for (typ, i), vals in reads.items():
        flat, n, nv = flats[typ]
        if [int(flat[v * n + i]) for v in range(nv)] != vals:
            return False, f"read values of {typ} {i}"
    return True, final


@pytest.mark.parametrize("seed", range(30))
def test_oracle_readonly_pairwise_and_serializable(seed):
    """Order-Status / Stock-Level mixed into colliding batches: counters
    match the pairwise count from the dispatch tables (R registrations give
    NewOrders war), every type follows Aria's rule, the committed set
    executes serially to the oracle's tables and read values."""
    rng = np.random.default_rng(500 + seed)
    W = 1 + seed % 2
    n_no, n_pay = int(rng.integers(0, 8)), int(rng.integers(0, 8))
    b = _random_batch(rng, W, n_no, n_pay)
    ro = _random_readonly(rng, W, int(rng.integers(0, 6)), int(rng.integers(0, 6)),
                          np.arange(0, 2 * (n_no + n_pay) + 24, 2))
    tables0 = _tables(W)
    tables = [t.copy() for t in tables0]
    ep = oracle.TpccOracle(W).epoch(tables, b, seed % 2, 1, ro=ro)
    ok, res = _pins_hold_ro(W, tables0, b, ro, ep, seed % 2)
    assert ok, res
    for t in range(9):
        np.testing.assert_array_equal(tables[t], res[t])
    assert ep.ro["os_commits"] == ro.n_os and ep.ro["sl_commits"] == ro.n_sl


def test_oracle_readonly_bruteforce_tid_orders():
    """Every TID order of small batches with one Order-Status and one
    Stock-Level on colliding rows."""
    rng = np.random.default_rng(91)
    W = 1
    tables0 = _tables(W)
    ora = oracle.TpccOracle(W)
    checked = aborts = 0
    for pattern in range(4):
        base = _random_batch(rng, W, 2, 1)
        ro0 = _random_readonly(rng, W, 1, 1, np.arange(100, 110))
        # make the readers hit the NewOrders' rows
        ro0.os["O_ID"][0] = base.nf("O_ID")[0]
        ro0.os["W_ID"][0] = base.nf("W_ID")[0]
        if base.nf("O_OL_CNT")[1]:
            ro0.sl["W_ID"][0] = base.nf("OL_SUPPLY_W_ID", 0)[1]
            ro0.sl["S_ID"][0] = base.nf("OL_I_ID", 0)[1]
        for perm in itertools.permutations(range(5)):
            b = L.TpccBatch(base.no.copy(), base.no_cap, base.n_no, base.pay.copy(), base.pay_cap, base.n_pay)
            b.nf("TID")[:] = np.array(perm[:2], dtype=np.uint32) * 5 + 3
            b.pf("TID")[:] = np.array(perm[2:3], dtype=np.uint32) * 5 + 3
            ro0.os["TID"][0], ro0.sl["TID"][0] = perm[3] * 5 + 3, perm[4] * 5 + 3
            tables = [t.copy() for t in tables0]
            ep = ora.epoch(tables, b, pattern % 2, 1, ro=ro0)
            ok, res = _pins_hold_ro(W, tables0, b, ro0, ep, pattern % 2)
            assert ok, res
            aborts += b.n_no - sum(tc.bits(ep.no_bitmap, b.n_no))
            checked += 1
    assert checked == 480 and aborts > 0


def _random_delivery(rng, W, n, tid_pool):
    """Deliveries whose rows collide with _random_batch's: O_ID below 15
    reaches its orders (0..4), order lines (0..14) and customers (0..3) of
    warehouse 0, NO_O_ID below 5 its NEWORDER rows."""
    dl = L.DeliveryBatch(n)
    tids = rng.permutation(tid_pool)[:n]
    for i in range(n):
        dl.O_ID[i] = rng.integers(0, 15) if rng.random() < 0.9 else rng.integers(0, 40000)
        dl.NO_W_ID[i], dl.NO_O_ID[i] = rng.integers(0, W + (rng.random() < 0.1)), rng.integers(0, 5)
        dl.TID[i] = tids[i]
    return dl


def _pins_hold_dl(W, tables0, b, ro, dl, ep, rules, values=1):
    """_pins_hold_ro with Deliveries: every type's counters against the
    pairwise count from the dispatch tables, Aria's rule for every type, and
    the serial execution of the committed set with every read value."""
    no_c, pay_c, bad, os_c, sl_c, dl_c = tc.pairwise_conflicts(b, W, ep.pay_cid, rules, ro, dl)
    got = {"no": tc.unpack(ep.no_conflicts), "pay": tc.unpack(ep.pay_conflicts), "dl": tc.unpack(ep.dl["conflicts"])}
    want = {"no": no_c, "pay": pay_c, "dl": dl_c}
    if ro is not None:
        got.update(os=tc.unpack(ep.ro["os_conflicts"]), sl=tc.unpack(ep.ro["sl_conflicts"]))
        want.update(os=os_c, sl=sl_c)
    for name in got:
        if not np.array_equal(got[name], want[name]):
            return False, f"{name} conflict counters differ from the pairwise count"
    if bad != ep.range_errors:
        return False, "range errors"
    commits = {"no": tc.bits(ep.no_bitmap, b.n_no), "pay": tc.bits(ep.pay_bitmap, b.n_pay),
               "dl": tc.bits(ep.dl["bitmap"], dl.n)}
    if ro is not None:
        commits.update(os=tc.bits(ep.ro["os_bitmap"], ro.n_os), sl=tc.bits(ep.ro["sl_bitmap"], ro.n_sl))
    for name in commits:
        if commits[name] != tc.aria(want[name]).tolist():
            return False, f"{name} commit is not waw == 0 && (raw == 0 || war == 0)"
    ok, why, final, reads = tc.serial_check(
        W, tables0, b, ep.pay_cid, commits["no"], commits["pay"], rmw=values == 1, ro=ro,
        ro_commit={"os": commits["os"], "sl": commits["sl"]} if ro is not None else None, dl=dl,
        dl_commit=commits["dl"])
    if not ok:
        return False, why
    flats = {"no": (ep.no_reads, b.n_no, 51), "pay": (ep.pay_reads, b.n_pay, 24), "dl": (ep.dl["reads"], dl.n, 2)}
    if ro is not None:
        flats.update(os=(ep.ro["os_reads"], ro.n_os, 3), sl=(ep.ro["sl_reads"], ro.n_sl, 1))
    for (typ, i), vals in reads.items():
        flat, n, nv = flats[typ]
        if [int(flat[v * n + i]) for v in range(nv)] != vals:
            return False, f"read values of {typ} {i}"
    return True, final


@pytest.mark.parametrize("seed", range(40))
def test_oracle_delivery_pairwise_and_serializable(seed):
    """Deliveries mixed into colliding batches (with read-only transactions
    in half of them): every type's counters match the pairwise count from
    the dispatch tables (check_4_0 and check_5_0 never counted), every type
    follows Aria's rule, the committed set executes serially to the oracle's
    tables and read values."""
    rng = np.random.default_rng(900 + seed)
    W = 1 + seed % 2
    n_no, n_pay, n_dl = int(rng.integers(0, 7)), int(rng.integers(0, 7)), int(rng.integers(1, 7))
    b = _random_batch(rng, W, n_no, n_pay)
    pool = np.arange(0, 2 * (n_no + n_pay) + 40, 2)
    ro = _random_readonly(rng, W, int(rng.integers(0, 4)), int(rng.integers(0, 4)), pool) if seed % 2 else None
    dl = _random_delivery(rng, W, n_dl, pool + 1)
    tables0 = _tables(W)
    tables = [t.copy() for t in tables0]
    values = 0 if seed % 7 == 3 else 1
    ep = oracle.TpccOracle(W).epoch(tables, b, (seed // 2) % 2, values, ro=ro, dl=dl)
    ok, res = _pins_hold_dl(W, tables0, b, ro, dl, ep, (seed // 2) % 2, values)
    assert ok, res
    for t in range(9):
        np.testing.assert_array_equal(tables[t], res[t])


def test_oracle_delivery_bruteforce_tid_orders():
    """Every TID order of small batches (2 NewOrders, 1 Payment, 2
    Deliveries on their rows), both rule sets."""
    rng = np.random.default_rng(93)
    W = 1
    tables0 = _tables(W)
    ora = oracle.TpccOracle(W)
    checked = 0
    outcomes = set()
    for pattern in range(4):
        base = _random_batch(rng, W, 2, 1)
        dl0 = _random_delivery(rng, W, 2, np.arange(100, 110))
        # one Delivery on the first NewOrder's order and NEWORDER row, the
        # other on the Payment's customer
        dl0.O_ID[0] = base.nf("O_ID")[0] + 30000 * int(base.nf("W_ID")[0])
        dl0.NO_W_ID[0], dl0.NO_O_ID[0] = base.nf("W_ID")[0], base.nf("N_O_ID")[0]
        dl0.O_ID[1] = int(base.pf("C_W_ID")[0]) * 30000 + int(base.pf("C_D_ID")[0]) * 3000 + int(base.pf("C_ID")[0])
        for perm in itertools.permutations(range(5)):
            b = L.TpccBatch(base.no.copy(), base.no_cap, base.n_no, base.pay.copy(), base.pay_cap, base.n_pay)
            b.nf("TID")[:] = np.array(perm[:2], dtype=np.uint32) * 5 + 3
            b.pf("TID")[:] = np.array(perm[2:3], dtype=np.uint32) * 5 + 3
            dl0.TID[:2] = np.array(perm[3:5], dtype=np.uint32) * 5 + 3
            tables = [t.copy() for t in tables0]
            ep = ora.epoch(tables, b, pattern % 2, 1, dl=dl0)
            ok, res = _pins_hold_dl(W, tables0, b, None, dl0, ep, pattern % 2)
            assert ok, res
            outcomes.add(tuple(tc.bits(ep.dl["bitmap"], 2)) + tuple(tc.bits(ep.no_bitmap, 2)))
            checked += 1
    assert checked == 480 and len(outcomes) > 2


def test_oracle_bruteforce_all_tid_orders():
    """Every TID order of small colliding batches (<= 6 transactions)."""
    rng = np.random.default_rng(77)
    W = 1
    tables0 = _tables(W)
    ora = oracle.TpccOracle(W)
    checked = 0
    for pattern in range(8):
        n_no, n_pay = int(rng.integers(1, 4)), int(rng.integers(1, 4))
        base = _random_batch(rng, W, n_no, n_pay)
        for perm in itertools.permutations(range(n_no + n_pay)):
            b = L.TpccBatch(base.no.copy(), base.no_cap, base.n_no, base.pay.copy(), base.pay_cap, base.n_pay)
            b.nf("TID")[:] = np.array(perm[:n_no], dtype=np.uint32) * 5 + 3
            b.pf("TID")[:] = np.array(perm[n_no:], dtype=np.uint32) * 5 + 3
            tables = [t.copy() for t in tables0]
            ep = ora.epoch(tables, b, pattern % 2, 1)
            ok, res = _pins_hold(W, tables0, b, ep, pattern % 2)
            assert ok, res
            for t in range(9):
                np.testing.assert_array_equal(tables[t], res[t])
            checked += 1
    assert checked >= 100


def test_cid_closed_form_on_generated_batches():
    W = 4
    gen = L.TpccGenerator(warehouses=W, seed=11)
    b = gen.next(4000)
    ep = oracle.TpccOracle(W).epoch(_tables(W), b)
    np.testing.assert_array_equal(ep.pay_cid, tc.cid_closed_form(b, W))


def test_copyback_leaves_tables_and_matches_commits():
    W = 2
    gen = L.TpccGenerator(warehouses=W, seed=4)
    b = gen.next(3000)
    t0, t1 = _tables(W), _tables(W)
    e0 = oracle.TpccOracle(W).epoch(t0, b, 0, 0)
    e1 = oracle.TpccOracle(W).epoch(t1, b, 0, 1)
    for t in range(9):
        np.testing.assert_array_equal(t0[t], _TABLE_CACHE[W][t])
    np.testing.assert_array_equal(e0.no_bitmap, e1.no_bitmap)
    np.testing.assert_array_equal(e0.pay_bitmap, e1.pay_bitmap)
    np.testing.assert_array_equal(e0.no_reads, e1.no_reads)


def test_registers_reset_between_epochs():
    """A second epoch on a fresh oracle and on a reused one agree."""
    W = 1
    gen = L.TpccGenerator(warehouses=W, seed=8)
    b1, b2 = gen.next(500), gen.next(500, tid_base=500)
    ta, tb_ = _tables(W), _tables(W)
    reused = oracle.TpccOracle(W)
    reused.epoch(ta, b1)
    ea = reused.epoch(ta, b2)
    oracle.TpccOracle(W).epoch(tb_, b1)
    eb = oracle.TpccOracle(W).epoch(tb_, b2)
    np.testing.assert_array_equal(ea.no_bitmap, eb.no_bitmap)
    np.testing.assert_array_equal(ea.pay_conflicts, eb.pay_conflicts)


# ---- planted mistakes -------------------------------------------------------

ORACLE_SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")

MISTAKES = {
    # check_3_0 made reachable under the reference rules (App. quirk B1)
    "check_3_0_always": (r"if \(rules != 0\) war \+= lower\(o, NO, R, h.no, me\);",
                         "war += lower(o, NO, R, h.no, me);"),
    # the Payment warehouse update registered instead of isFrequent
    "ytd_registered": (r"if \(!ok\(o, WH, wh\)\) errors\+\+;( +)/\* op_1_0: isFrequent \*/",
                       "if (ok(o, WH, wh)) reg_min(o, WH, W, wh, me); else errors++;"),
    # by-name ordinal counted from 1
    "cid_one_based": (r"if \(seen\+\+ == k\)", "if (++seen == k)"),
    # restock branch with the wrong sign
    "restock_sign": (r"\*sq - qty \+ 91", "*sq + qty + 91"),
    # commit rule with OR
    "or_rule": (r"if \(waw == 0 && \(raw == 0 \|\| war == 0\)\) \{\n            res->pay_bitmap",
                "if (waw == 0 || raw == 0 || war == 0) {\n            res->pay_bitmap"),
    # war read from the W register of the customer
    "war_from_w": (r"war \+= lower\(o, CU, R, cu, me\);", "war += lower(o, CU, W, cu, me);"),
    # priority ties broken the wrong way (position descending)
    "tie_reversed": (r"const uint64_t me = \(\(uint64_t\)pv.TID\[p\] << 32\) \| \(n_no \+ p\);",
                     "const uint64_t me = ((uint64_t)pv.TID[p] << 32) | (0xffffffu - p);"),
    # an order line beyond the first not applied
    "first_line_only": (r"for \(uint32_t i = 0; i < nlines\(&nv, q\); i\+\+\) \{\n                const uint64_t ir = item_row\(&nv, q, i\), sr = stock_row\(&nv, q, i\), lr = line_row\(&nv, q, i\);\n                const uint32_t qty",
                        "for (uint32_t i = 0; i < 1 && i < nlines(&nv, q); i++) {\n                const uint64_t ir = item_row(&nv, q, i), sr = stock_row(&nv, q, i), lr = line_row(&nv, q, i);\n                const uint32_t qty"),
    # Delivery: check_4_0 made reachable (war from ORDERLINE W)
    "delivery_check_4_0": (r"/\* check_4_0 \(war, ORDERLINE W\) and check_5_0 \(raw, CUSTOMER R\)",
                           "if (ok(o, OL, od)) war += lower(o, OL, W, od, me);\n            /* check_4_0 (war, ORDERLINE W) and check_5_0 (raw, CUSTOMER R)"),
    # Delivery: the CUSTOMER waw (check_5_1) dropped
    "delivery_no_customer_waw": (r"if \(ok\(o, CU, od\)\) waw \+= lower\(o, CU, W, od, me\);",
                                 "if (ok(o, CU, od)) raw += 0;"),
    # Delivery: NEWORDER row with the warehouse offset omitted
    "delivery_neworder_row": (r"return \(uint64_t\)b\[\(uint64_t\)cap \+ d\] \* 30000 \+ b\[2 \* \(uint64_t\)cap \+ d\];",
                              "return (uint64_t)b[2 * (uint64_t)cap + d];"),
    # Delivery: reads ORDERLINE col 8 instead of col 3 (op_4_1)
    "delivery_read_column": (r"get\(tbl, o, OL, 3, od\) : 0; /\* op_4_1 \*/",
                             "get(tbl, o, OL, 8, od) : 0; /* op_4_1 */"),
}


def _mutant(tmp_path, name):
    pattern, repl = MISTAKES[name]
    src = open(os.path.join(ORACLE_SRC, "ltpg_oracle_tpcc.c")).read()
    new, n = re.subn(pattern, lambda m: repl, src, count=1)
    assert n == 1, f"mutation {name} did not apply"
    d = tmp_path / name
    d.mkdir()
    for f in ("ltpg_oracle.h", "ltpg_oracle_kv.c"):
        shutil.copy(os.path.join(ORACLE_SRC, f), d / f)
    (d / "ltpg_oracle_tpcc.c").write_text(new)
    so = d / "liboracle.so"
    subprocess.run(["gcc", "-O1", "-std=c99", "-fPIC", "-shared", "-o", str(so), str(d / "ltpg_oracle_kv.c"),
                    str(d / "ltpg_oracle_tpcc.c")], check=True)
    return oracle.load_variant(str(so))


@pytest.mark.parametrize("mistake", sorted(MISTAKES))
def test_pins_catch_planted_mistakes(tmp_path, mistake):
    """Each planted mistake fails a known answer or the pairwise /
    serializability pins, so those pins would catch it in the oracle."""
    lib = _mutant(tmp_path, mistake)
    caught = False
    for case in KAT["cases"]:
        ora = oracle.TpccOracle(case["W"], lib=lib)
        if not _check_kat(case, lambda tables, b, r, v, ro, dl: ora.epoch(tables, b, r, v, ro=ro, dl=dl)):
            caught = True
            break
    rng = np.random.default_rng(99)
    for i in range(0 if caught else 60):
        W = 1
        b = _random_batch(rng, W, int(rng.integers(1, 8)), int(rng.integers(1, 8)))
        tables0 = _tables(W)
        tables = [t.copy() for t in tables0]
        dl = _random_delivery(rng, W, int(rng.integers(1, 5)), np.arange(1, 60, 2)) if i % 3 == 2 else None
        ep = oracle.TpccOracle(W, lib=lib).epoch(tables, b, i % 2, 1, dl=dl)
        ok, res = _pins_hold(W, tables0, b, ep, i % 2) if dl is None else \
            _pins_hold_dl(W, tables0, b, None, dl, ep, i % 2)
        if not ok or not all(np.array_equal(tables[t], res[t]) for t in range(9)):
            caught = True
            break
    assert caught, f"planted mistake {mistake!r} passed every pin"
