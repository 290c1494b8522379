"""Shared helpers for order-entry tests: hand-built batches, fixture
decoding, and the independent checks that pin the oracle.

The independent checks deliberately do not reuse the oracle's registers:
  * op_sets lists each transaction's registrations and check slots straight
    from the reference's dispatch tables (TPCC-allopt-final/
    Execute_neworder.h, Execute_payment.h), and pairwise_conflicts counts a
    conflict for a check slot when some higher-priority transaction registered
    the consulted register on the same row -- by comparing transactions
    pairwise, not through minimum registers;
  * serial_check orders the committed transactions by their precedence
    constraints (a reader before every writer of a row it read; two writers
    of one row in priority order), checks the order exists (Aria's
    serializability guarantee), and executes them one at a time on a copy of
    the tables with the order-entry value rules, checking each transaction's
    reads against the delivered read values.
"""
import json
import os

import numpy as np

import ltpg_amd as L

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
WH, DI, CU, HI, NO, OR, OL, ST, IT = range(9)
COLS = [9, 11, 21, 5, 3, 8, 10, 17, 5]
NO_FIELDS = ["W_ID", "D_ID", "C_ID", "O_ID", "N_O_ID", "O_OL_CNT", "O_OL_ID"]
PAY_FIELDS = ["W_ID", "D_ID", "C_ID", "C_LAST", "isName", "C_D_ID", "C_W_ID", "H_AMOUNT", "H_ID"]
PAY_CUST_COLS = [3, 4, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20]


def rows_of(W, t):
    return 100000 if t == IT else [1, 10, 30000, 30000, 30000, 30000, 450000, 100000][t] * W


def make_batch(neworders, payments):
    """TpccBatch from dicts: NewOrder {W_ID, D_ID, C_ID, O_ID, N_O_ID, O_OL_ID,
    TID, lines: [(I_ID, SUPPLY_W_ID, QTY), ...]}, Payment {W_ID, D_ID, C_ID,
    C_LAST, isName, C_D_ID, C_W_ID, H_AMOUNT, H_ID, TID}; absent fields 0."""
    b = L.TpccBatch.empty(len(neworders), len(payments))
    for q, t in enumerate(neworders):
        for f in NO_FIELDS:
            b.nf(f)[q] = t.get(f, 0)
        lines = t.get("lines", [])
        b.nf("O_OL_CNT")[q] = t.get("O_OL_CNT", len(lines))
        for i, (item, sw, qty) in enumerate(lines):
            b.nf("OL_I_ID", i)[q] = item
            b.nf("OL_SUPPLY_W_ID", i)[q] = sw
            b.nf("OL_QUANTITY", i)[q] = qty
        b.nf("TID")[q] = t["TID"]
    for p, t in enumerate(payments):
        for f in PAY_FIELDS:
            b.pf(f)[p] = t.get(f, 0)
        b.pf("TID")[p] = t["TID"]
    return b


def kat_delivery(case):
    """DeliveryBatch of a case's 'delivery' list, or None."""
    if "delivery" not in case:
        return None
    dl = L.DeliveryBatch(len(case["delivery"]))
    for i, t in enumerate(case["delivery"]):
        for f in L.DELIVERY_FIELDS:
            getattr(dl, f)[i] = t.get(f, 0)
    return dl


def load_kat():
    with open(os.path.join(GOLDEN, "tpcc_kat.json")) as f:
        return json.load(f)


def bits(bitmap, n):
    return [int((int(bitmap[i // 64]) >> (i % 64)) & 1) for i in range(n)]


def unpack(conf):
    c = np.asarray(conf, dtype=np.uint32)
    return np.stack([c & 0xFF, (c >> 8) & 0xFF, (c >> 16) & 0xFF], axis=1)


# ---- op tables of the reference's dispatch --------------------------------

def _nlines(b, q):
    return min(int(b.nf("O_OL_CNT")[q]), 15)


def neworder_ops(b, q, W, rules):
    """(registrations, checks) of NewOrder q.  A registration is
    (table, row, 'R'|'W'); a check is (table, row, consulted 'R'|'W',
    counter 'raw'|'war'|'waw').  Slots whose row is outside its table are
    dropped (and counted as range errors)."""
    w, d = int(b.nf("W_ID")[q]), int(b.nf("D_ID")[q])
    c, noid, oid = int(b.nf("C_ID")[q]), int(b.nf("N_O_ID")[q]), int(b.nf("O_ID")[q])
    regs, checks, bad = [], [], 0
    slots = [(WH, w, "R", [("W", "raw")]),                                      # op_0_0 / check_0_0
             (DI, w * 10 + d, "R", [("W", "raw")]),                             # op_1_0 / check_1_0
             (CU, w * 30000 + d * 3000 + c, "R", [("W", "raw")]),               # op_2_0 / check_2_0
             (NO, w * 30000 + noid, "R" if rules == 0 else "W",                 # op_3_0
              ([] if rules == 0 else [("R", "war")]) + [("W", "waw")]),        # check_3_0 / 3_1
             (OR, w * 30000 + oid, "W", [("R", "war"), ("W", "waw")])]          # op_4_0 / check_4_*
    for i in range(_nlines(b, q)):
        item, sw = int(b.nf("OL_I_ID", i)[q]), int(b.nf("OL_SUPPLY_W_ID", i)[q])
        slots += [(IT, item, "R", [("W", "raw")]),                              # op_5_0 / check_5_0
                  (ST, sw * 100000 + item, "W", [("R", "war"), ("W", "waw")]),  # op_6_0 / check_6_*
                  (OL, w * 450000 + int(b.nf("O_OL_ID")[q]) + i, "W", [("R", "war"), ("W", "waw")])]
    for t, row, kind, cks in slots:
        if row >= rows_of(W, t):
            bad += 1
            continue
        regs.append((t, row, kind))
        checks += [(t, row, k, ctr) for k, ctr in cks]
    return regs, checks, bad


def payment_ops(b, p, W, cid):
    w, d = int(b.pf("W_ID")[p]), int(b.pf("D_ID")[p])
    cu = int(b.pf("C_W_ID")[p]) * 30000 + int(b.pf("C_D_ID")[p]) * 3000 + int(cid)
    hi = w * 30000 + int(b.pf("H_ID")[p])
    regs, checks, bad = [], [], 0
    slots = [(WH, w, "R", [("W", "raw")]),                      # op_0_0 / check_0_0
             (WH, w, None, []),                                 # op_1_0: isFrequent
             (DI, w * 10 + d, "R", [("W", "raw")]),             # op_2_0 / check_2_0
             (DI, w * 10 + d, None, []),                        # op_3_0: isFrequent
             (CU, cu, "R", [("W", "raw")]),                     # op_4_0 / check_4_0
             (CU, cu, "W", [("R", "war"), ("W", "waw")]),       # op_5_0 / check_5_*
             (HI, hi, "W", [("R", "war"), ("W", "waw")])]       # op_6_0 / check_6_*
    for t, row, kind, cks in slots:
        if row >= rows_of(W, t):
            bad += 1
            continue
        if kind:
            regs.append((t, row, kind))
        checks += [(t, row, k, ctr) for k, ctr in cks]
    return regs, checks, bad


def readonly_rows(ro, typ, i):
    """Rows read by read-only transaction i: Order-Status (TPCC-alltxn/
    Execute.h:1594-1660) CUSTOMER, ORDER, ORDERLINE; Stock-Level (:2157-2171)
    STOCK, with the column each reads."""
    if typ == "os":
        w = int(ro.os["W_ID"][i])
        return [(CU, w * 30000 + int(ro.os["D_ID"][i]) * 3000 + int(ro.os["C_ID"][i]), 2),
                (OR, w * 30000 + int(ro.os["O_ID"][i]), 0),
                (OL, w * 450000 + int(ro.os["OL_ID"][i]), 0)]
    return [(ST, int(ro.sl["W_ID"][i]) * 100000 + int(ro.sl["S_ID"][i]), 2)]


def readonly_ops(ro, typ, i, W):
    """(registrations, checks, range errors): R on each row read, raw
    against its W (check_0 .. check_2)."""
    regs, checks, bad = [], [], 0
    for t, row, _ in readonly_rows(ro, typ, i):
        if row >= rows_of(W, t):
            bad += 1
            continue
        regs.append((t, row, "R"))
        checks.append((t, row, "W", "raw"))
    return regs, checks, bad


def delivery_rows(dl, d):
    """(NEWORDER row, O_ID) of Delivery d: op_0_0 (TPCC-alltxn/Execute.h:
    1780-1784) and the warehouse-less O_ID row of op_1_0 .. op_5_0."""
    return int(dl.NO_W_ID[d]) * 30000 + int(dl.NO_O_ID[d]), int(dl.O_ID[d])


def delivery_ops(dl, d, W):
    """(registrations, checks, range errors) of Delivery d, slot by slot as
    its execute() chain registers (:1929-1977) and its check() chain reaches
    (:2108-2149; check_4_0 and check_5_0 sit behind repeated == 5 / == 6
    tests and never run)."""
    nr, o = delivery_rows(dl, d)
    slots = [(NO, nr, "W", [("R", "raw"), ("W", "raw")]),     # op_0_0 / check_0_0, check_0_1
             (OR, o, "R", [("W", "war")]),                     # op_1_0 / check_1_0
             (OR, o, "W", [("R", "raw"), ("W", "waw")]),       # op_2_0 / check_2_0, check_2_1
             (OL, o, "W", [("R", "raw"), ("W", "waw")]),       # op_3_0 / check_3_0, check_3_1
             (OL, o, "R", []),                                 # op_4_0 / (check_4_0 unreachable)
             (CU, o, "W", [("W", "waw")])]                     # op_5_0 / check_5_1 (check_5_0 unreachable)
    regs, checks, bad = [], [], 0
    for t, row, kind, cks in slots:
        if row >= rows_of(W, t):
            bad += 1
            continue
        regs.append((t, row, kind))
        checks += [(t, row, k, ctr) for k, ctr in cks]
    return regs, checks, bad


def all_ops(b, W, cid, rules, ro=None, dl=None):
    """[(priority, type, index, regs, checks)] and the range-error count."""
    out, bad = [], 0
    for q in range(b.n_no):
        r, c, e = neworder_ops(b, q, W, rules)
        out.append(((int(b.nf("TID")[q]), q), "no", q, r, c))
        bad += e
    for p in range(b.n_pay):
        r, c, e = payment_ops(b, p, W, cid[p])
        out.append(((int(b.pf("TID")[p]), b.n_no + p), "pay", p, r, c))
        bad += e
    if ro is not None:
        base = b.n_no + b.n_pay
        for typ, n, tids, off in (("os", ro.n_os, ro.os["TID"], base), ("sl", ro.n_sl, ro.sl["TID"], base + ro.n_os)):
            for i in range(n):
                r, c, e = readonly_ops(ro, typ, i, W)
                out.append(((int(tids[i]), off + i), typ, i, r, c))
                bad += e
    if dl is not None:
        base = b.n_no + b.n_pay + (ro.n_os + ro.n_sl if ro is not None else 0)
        for d in range(dl.n):
            r, c, e = delivery_ops(dl, d, W)
            out.append(((int(dl.TID[d]), base + d), "dl", d, r, c))
            bad += e
    return out, bad


def pairwise_conflicts(b, W, cid, rules, ro=None, dl=None):
    """raw/war/waw per transaction by pairwise comparison: a check slot of t
    on (table, row, kind) counts iff some transaction of higher priority
    registered `kind` on that row.  With read-only transactions `ro` also
    returns their counters: (no, pay, bad, os, sl); with Deliveries `dl`
    (no, pay, bad, os, sl, dl)."""
    txns, bad = all_ops(b, W, cid, rules, ro, dl)
    holders = {}
    for prio, _, _, regs, _ in txns:
        for t, row, kind in regs:
            holders.setdefault((t, row, kind), []).append(prio)
    conf = {"no": np.zeros((b.n_no, 3), np.uint32), "pay": np.zeros((b.n_pay, 3), np.uint32),
            "os": np.zeros((ro.n_os if ro else 0, 3), np.uint32), "sl": np.zeros((ro.n_sl if ro else 0, 3), np.uint32),
            "dl": np.zeros((dl.n if dl is not None else 0, 3), np.uint32)}
    col = {"raw": 0, "war": 1, "waw": 2}
    for prio, typ, i, _, checks in txns:
        for t, row, kind, ctr in checks:
            if any(h < prio for h in holders.get((t, row, kind), [])):
                conf[typ][i, col[ctr]] += 1
    if dl is not None:
        return conf["no"], conf["pay"], bad, conf["os"], conf["sl"], conf["dl"]
    if ro is None:
        return conf["no"], conf["pay"], bad
    return conf["no"], conf["pay"], bad, conf["os"], conf["sl"]


def aria(conf):
    raw, war, waw = conf[:, 0], conf[:, 1], conf[:, 2]
    return ((waw == 0) & ((raw == 0) | (war == 0))).astype(int)


def cid_closed_form(b, W):
    """C_IDs on the loader's data: customer row r of a district has C_LAST
    r % 20 (Database.h:180), so the k-th holder of name n is C_ID n + 20 k
    (k < 150); none (C_ID 0) otherwise or when the district does not exist."""
    out = b.pf("C_ID").astype(np.int64).copy()
    for p in range(b.n_pay):
        if b.pf("isName")[p]:
            cl = int(b.pf("C_LAST")[p])
            n, k = cl >> 8, cl & 0xFF
            dist = int(b.pf("C_W_ID")[p]) * 10 + int(b.pf("C_D_ID")[p])
            out[p] = n + 20 * k if (n < 20 and k < 150 and dist < 10 * W) else 0
    return out.astype(np.uint32)


# ---- serial execution ------------------------------------------------------

class Tables:
    """Column-major tables with (table, row, col) access."""

    def __init__(self, W, arrays):
        self.W = W
        self.a = [np.array(x, dtype=np.uint32, copy=True) for x in arrays]

    def get(self, t, row, c):
        return int(self.a[t][c * rows_of(self.W, t) + row])

    def set(self, t, row, c, v):
        self.a[t][c * rows_of(self.W, t) + row] = np.uint32(v & 0xFFFFFFFF)

    def ok(self, t, row):
        return row < rows_of(self.W, t)


def _neworder_effects(tb, b, q, rmw):
    """Execute NewOrder q on tb: returns its read values (51)."""
    w, d, c = int(b.nf("W_ID")[q]), int(b.nf("D_ID")[q]), int(b.nf("C_ID")[q])
    reads = [0] * 51
    if tb.ok(WH, w):
        reads[0] = tb.get(WH, w, 1)
    if tb.ok(DI, w * 10 + d):
        reads[1], reads[2] = tb.get(DI, w * 10 + d, 2), tb.get(DI, w * 10 + d, 4)
    cu = w * 30000 + d * 3000 + c
    if tb.ok(CU, cu):
        reads[3:6] = [tb.get(CU, cu, 4), tb.get(CU, cu, 10), tb.get(CU, cu, 13)]
    L_ = _nlines(b, q)
    for i in range(L_):
        item = int(b.nf("OL_I_ID", i)[q])
        if tb.ok(IT, item):
            reads[6 + 3 * i:9 + 3 * i] = [tb.get(IT, item, 2), tb.get(IT, item, 3), tb.get(IT, item, 4)]
    if not rmw:
        return reads
    oid = int(b.nf("O_ID")[q])
    nr, od = w * 30000 + int(b.nf("N_O_ID")[q]), w * 30000 + oid
    if tb.ok(NO, nr):
        for k, v in enumerate((oid, d, w)):
            tb.set(NO, nr, k, v)
    if tb.ok(OR, od):
        for k, v in enumerate((oid, d, w)):
            tb.set(OR, od, k, v)
    for i in range(L_):
        item, sw, qty = (int(b.nf(f, i)[q]) for f in ("OL_I_ID", "OL_SUPPLY_W_ID", "OL_QUANTITY"))
        sr, lr = sw * 100000 + item, w * 450000 + int(b.nf("O_OL_ID")[q]) + i
        if tb.ok(ST, sr):
            s_q = tb.get(ST, sr, 2)
            tb.set(ST, sr, 2, s_q - qty if s_q >= qty + 10 else s_q - qty + 91)
            tb.set(ST, sr, 3, tb.get(ST, sr, 3) + qty)
            tb.set(ST, sr, 4, tb.get(ST, sr, 4) + 1)
        if tb.ok(OL, lr):
            price = tb.get(IT, item, 2) if tb.ok(IT, item) else 0
            data = tb.get(IT, item, 3) if tb.ok(IT, item) else 0
            for k, v in enumerate((oid, d, w, i, item, sw, 0, qty, qty * price, data)):
                tb.set(OL, lr, k, v)
    return reads


def _payment_effects(tb, b, p, cid, rmw):
    w, d = int(b.pf("W_ID")[p]), int(b.pf("D_ID")[p])
    cu = int(b.pf("C_W_ID")[p]) * 30000 + int(b.pf("C_D_ID")[p]) * 3000 + int(cid)
    reads = [0] * 24
    if tb.ok(WH, w):
        reads[0:6] = [tb.get(WH, w, c) for c in range(3, 9)]
    if tb.ok(DI, w * 10 + d):
        reads[6:12] = [tb.get(DI, w * 10 + d, c) for c in range(5, 11)]
    if tb.ok(CU, cu):
        reads[12:24] = [tb.get(CU, cu, c) for c in PAY_CUST_COLS]
    if not rmw:
        return reads
    amt = int(b.pf("H_AMOUNT")[p])
    if tb.ok(CU, cu):
        tb.set(CU, cu, 5, tb.get(CU, cu, 5) - amt)
        tb.set(CU, cu, 9, tb.get(CU, cu, 9) + amt)
    hi = w * 30000 + int(b.pf("H_ID")[p])
    if tb.ok(HI, hi):
        for k, v in enumerate((int(cid), int(b.pf("C_D_ID")[p]), int(b.pf("C_W_ID")[p]), d, w)):
            tb.set(HI, hi, k, v)
    if tb.ok(WH, w):
        tb.set(WH, w, 2, tb.get(WH, w, 2) + amt)
    if tb.ok(DI, w * 10 + d):
        tb.set(DI, w * 10 + d, 3, tb.get(DI, w * 10 + d, 3) + amt)
    return reads


def _delivery_effects(tb, dl, d, rmw):
    """Execute Delivery d on tb: returns its read values (ORDER col 3,
    ORDERLINE col 3).  Writers of one NEWORDER row run in priority order,
    so the last of them leaves the row."""
    nr, o = delivery_rows(dl, d)
    reads = [tb.get(OR, o, 3) if tb.ok(OR, o) else 0, tb.get(OL, o, 3) if tb.ok(OL, o) else 0]
    if not rmw:
        return reads
    tid = int(dl.TID[d])
    if tb.ok(NO, nr):
        for k in range(3):
            tb.set(NO, nr, k, 0)
    if tb.ok(OR, o):
        tb.set(OR, o, 5, tid % 10 + 1)
    if tb.ok(OL, o):
        tb.set(OL, o, 6, tid)
    if tb.ok(CU, o):
        tb.set(CU, o, 5, tb.get(CU, o, 5) + (tb.get(OL, o, 8) if tb.ok(OL, o) else 0))
        tb.set(CU, o, 8, tb.get(CU, o, 8) + 1)
    return reads


def _rw_sets(b, W, cid, ro=None, dl=None):
    """Row-level read and write sets of every transaction (the YTD columns
    are commutative adds and do not order transactions)."""
    out = []
    if ro is not None:
        base = b.n_no + b.n_pay
        for typ, n, tids, off in (("os", ro.n_os, ro.os["TID"], base), ("sl", ro.n_sl, ro.sl["TID"], base + ro.n_os)):
            for i in range(n):
                out.append(((int(tids[i]), off + i), typ, i, {(t, r) for t, r, _ in readonly_rows(ro, typ, i)}, set()))
    for q in range(b.n_no):
        w, d, c = int(b.nf("W_ID")[q]), int(b.nf("D_ID")[q]), int(b.nf("C_ID")[q])
        rd = {(WH, w), (DI, w * 10 + d), (CU, w * 30000 + d * 3000 + c)}
        wr = {(NO, w * 30000 + int(b.nf("N_O_ID")[q])), (OR, w * 30000 + int(b.nf("O_ID")[q]))}
        for i in range(_nlines(b, q)):
            item, sw = int(b.nf("OL_I_ID", i)[q]), int(b.nf("OL_SUPPLY_W_ID", i)[q])
            rd.add((IT, item))
            wr.add((ST, sw * 100000 + item))
            wr.add((OL, w * 450000 + int(b.nf("O_OL_ID")[q]) + i))
        out.append(((int(b.nf("TID")[q]), q), "no", q, rd, wr))
    for p in range(b.n_pay):
        w, d = int(b.pf("W_ID")[p]), int(b.pf("D_ID")[p])
        cu = int(b.pf("C_W_ID")[p]) * 30000 + int(b.pf("C_D_ID")[p]) * 3000 + int(cid[p])
        rd = {(CU, cu)}
        wr = {(CU, cu), (HI, w * 30000 + int(b.pf("H_ID")[p]))}
        out.append(((int(b.pf("TID")[p]), b.n_no + p), "pay", p, rd, wr))
    if dl is not None:
        base = b.n_no + b.n_pay + (ro.n_os + ro.n_sl if ro is not None else 0)
        for d in range(dl.n):
            nr, o = delivery_rows(dl, d)
            out.append(((int(dl.TID[d]), base + d), "dl", d, {(OR, o), (OL, o), (CU, o)},
                        {(NO, nr), (OR, o), (OL, o), (CU, o)}))
    return out


def serial_check(W, tables0, b, cid, no_commit, pay_commit, rmw=True, ro=None, ro_commit=None, dl=None,
                 dl_commit=None):
    """(ok, message, final tables, reads) of a serial execution of the
    committed transactions in an order that honours every precedence
    constraint; reads are {('no', q): [..51], ('pay', p): [..24],
    ('os', o): [..3], ('sl', k): [..1], ('dl', d): [..2]}.  ro_commit:
    {'os': [...], 'sl': [...]} commit lists of the read-only transactions
    `ro`; dl_commit the commit list of the Deliveries `dl`."""
    commits = {"no": no_commit, "pay": pay_commit}
    if ro is not None:
        commits.update(ro_commit)
    if dl is not None:
        commits["dl"] = dl_commit
    txns = [t for t in _rw_sets(b, W, cid, ro, dl) if commits[t[1]][t[2]]]
    n = len(txns)
    succ = [set() for _ in range(n)]
    indeg = [0] * n
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            pi, _, _, ri, wi = txns[i]
            pj, _, _, rj, wj = txns[j]
            # snapshot reads: a reader precedes a writer of the row it read
            # (its own write of a row it read is not a constraint)
            before = bool(ri & wj)
            # two writers of one row: the later in priority order writes last
            before |= bool(wi & wj) and pi < pj
            if before and j not in succ[i]:
                succ[i].add(j)
                indeg[j] += 1
    order, ready = [], sorted(range(n), key=lambda k: txns[k][0])
    ready = [k for k in ready if indeg[k] == 0]
    while ready:
        ready.sort(key=lambda k: txns[k][0])
        k = ready.pop(0)
        order.append(k)
        for j in succ[k]:
            indeg[j] -= 1
            if indeg[j] == 0:
                ready.append(j)
    if len(order) != n:
        return False, "committed set has a precedence cycle (not serializable)", None, None
    tb = Tables(W, tables0)
    reads = {}
    for k in order:
        _, typ, i, _, _ = txns[k]
        if typ == "no":
            reads[(typ, i)] = _neworder_effects(tb, b, i, rmw)
        elif typ == "pay":
            reads[(typ, i)] = _payment_effects(tb, b, i, cid[i], rmw)
        elif typ == "dl":
            reads[(typ, i)] = _delivery_effects(tb, dl, i, rmw)
        else:
            reads[(typ, i)] = [tb.get(t, r, c) if tb.ok(t, r) else 0 for t, r, c in readonly_rows(ro, typ, i)]
    return True, "", tb.a, reads


def committed_reads(reads_flat, n, nv, commit):
    """{index: list} of the committed transactions' value-major read values."""
    r = np.asarray(reads_flat).reshape(nv, max(n, 1))[:, :n] if n else np.zeros((nv, 0), np.uint32)
    return {i: r[:, i].tolist() for i in range(n) if commit[i]}
