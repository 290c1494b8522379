"""Shared helpers for key-value tests: fixture decoding, batch construction,
and the independent checks that pin the oracle.

The independent checks deliberately do NOT reuse the oracle's register
algorithm:
  * pairwise_conflicts counts conflicts by comparing every pair of
    transactions directly (existence of a lower-TID conflicting op), instead
    of through min-TID registers;
  * serial_check rebuilds the committed set's precedence graph, checks it is
    acyclic (Aria's serializability guarantee) and replays the committed
    transactions one at a time in a topological order to get the final table.
"""
import json
import os

import numpy as np

KIND = {"n": 0, "r": 1, "w": 2, "s": 3}
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_kat():
    with open(os.path.join(GOLDEN, "kv_kat.json")) as f:
        return json.load(f)


def initial_table(rows, fields, fw):
    r = np.arange(rows, dtype=np.uint32)[:, None, None]
    f = np.arange(fields, dtype=np.uint32)[None, :, None]
    w = np.arange(fw, dtype=np.uint32)[None, None, :]
    return (1000 + 100 * r + 10 * f + w).astype(np.uint32).reshape(-1)


def case_arrays(case):
    """Slot-major keys/meta and tids of a fixture case."""
    txns = case["txns"]
    B = len(txns)
    S = max(len(t["ops"]) for t in txns)
    keys = np.zeros(S * B, dtype=np.uint64)
    meta = np.zeros(S * B, dtype=np.uint32)
    tids = np.array([t["tid"] for t in txns], dtype=np.uint32)
    for t, txn in enumerate(txns):
        for s, op in enumerate(txn["ops"]):
            kind, row, field = op[:3]
            keys[s * B + t] = row
            meta[s * B + t] = KIND[kind] | (field << 8)
            if kind == "s":  # ["s", row, field, length]
                meta[s * B + t] |= (op[3] - 1) << 24
    return keys, meta, tids, S


def expected_table(case):
    rows, fields, fw = case["rows"], case["fields"], case["field_words"]
    tab = initial_table(rows, fields, fw).reshape(rows, fields, fw)
    for r, f, w, v in case["expect"]["changed"]:
        tab[r, f, w] = v
    return tab.reshape(-1)


def bitmap_to_list(bitmap, B):
    return [int((int(bitmap[t // 64]) >> (t % 64)) & 1) for t in range(B)]


def unpack_conflicts(c):
    c = np.asarray(c, dtype=np.uint32)
    return np.stack([c & 0xFF, (c >> 8) & 0xFF, (c >> 16) & 0xFF], axis=1)


def ops_of(keys, meta, B, S, rows, fields, max_scan=0):
    """List per txn of (slot, kind, row, field) for valid ops; a scan of
    length n <= max_scan is one read (kind 1) per row of key .. key + n - 1
    inside the table."""
    out = [[] for _ in range(B)]
    for s in range(S):
        for t in range(B):
            m = int(meta[s * B + t])
            kind, field, row = m & 3, (m >> 8) & 0xFFFF, int(keys[s * B + t])
            if kind in (1, 2) and row < rows and field < fields:
                out[t].append((s, kind, row, field))
            elif kind == 3 and row < rows and field < fields and (m >> 24) + 1 <= max_scan:
                for r in range(row, min(row + (m >> 24) + 1, rows)):
                    out[t].append((s, 1, r, field))
    return out


def pairwise_conflicts(keys, meta, tids, B, S, rows, fields, max_scan=0):
    """raw/war/waw by direct pairwise comparison (no registers).  Priority
    is (TID, batch position): lower wins."""
    ops = ops_of(keys, meta, B, S, rows, fields, max_scan)
    readers, writers = {}, {}
    for t in range(B):
        for _, kind, row, _ in ops[t]:
            (readers if kind == 1 else writers).setdefault(row, []).append((int(tids[t]), t))
    res = np.zeros((B, 3), dtype=np.uint32)
    for t in range(B):
        me = (int(tids[t]), t)
        for _, kind, row, _ in ops[t]:
            lower_w = any(u < me for u in writers.get(row, []))
            lower_r = any(u < me for u in readers.get(row, []))
            if kind == 1:
                res[t, 0] += lower_w
            else:
                res[t, 1] += lower_r
                res[t, 2] += lower_w
    return res


def aria_commit(conf):
    raw, war, waw = conf[:, 0], conf[:, 1], conf[:, 2]
    return ((waw == 0) & ((raw == 0) | (war == 0))).astype(int)


def serial_check(table0, keys, meta, tids, B, S, rows, fields, fw, committed, values=1, max_scan=0):
    """Check the committed set is conflict-serializable and return the table
    obtained by running the committed transactions one by one.

    Every committed read saw the snapshot, so a committed reader of row x
    must come before any committed writer of x; two committed writers of one
    row are forbidden.  Returns (ok, message, final_table)."""
    ops = ops_of(keys, meta, B, S, rows, fields, max_scan)
    C = [t for t in range(B) if committed[t]]
    writers = {}
    for t in C:
        for _, kind, row, _ in ops[t]:
            if kind == 2:
                writers.setdefault(row, set()).add(t)
    for row, ws in writers.items():
        if len(ws) > 1:
            return False, f"row {row} has committed writers {sorted(ws)}", None
    succ = {t: set() for t in C}
    for t in C:
        for _, kind, row, _ in ops[t]:
            if kind == 1:
                for w in writers.get(row, ()):
                    if w != t:
                        succ[t].add(w)  # reader t before writer w
    # Kahn's algorithm; ties broken by TID for a fixed order
    indeg = {t: 0 for t in C}
    for t in C:
        for u in succ[t]:
            indeg[u] += 1
    ready = sorted((t for t in C if indeg[t] == 0), key=lambda t: int(tids[t]))
    order = []
    while ready:
        t = ready.pop(0)
        order.append(t)
        for u in succ[t]:
            indeg[u] -= 1
            if indeg[u] == 0:
                ready.append(u)
        ready.sort(key=lambda t: int(tids[t]))
    if len(order) != len(C):
        return False, "precedence graph of committed transactions has a cycle", None
    tab = np.array(table0, dtype=np.uint32).reshape(rows, fields, fw).copy()
    snap = tab.copy()
    for t in order:
        # Aria's execution model: a transaction reads the epoch snapshot and
        # buffers its writes, so its reads come before its own writes.
        for _, kind, row, field in ops[t]:
            if kind == 1 and not np.array_equal(tab[row, field], snap[row, field]):
                return False, f"txn {t} would read a value changed earlier in the serial order", None
        for _, kind, row, field in ops[t]:
            if kind == 2 and values == 1:
                tab[row, field] = tab[row, field] + np.uint32(int(tids[t]) | 1)
    return True, "", tab.reshape(-1)


def random_batch(rng, B, S, rows, fields, p_read=0.5, p_empty=0.1, distinct_tids=True, p_scan=0.0, max_len=1):
    keys = rng.integers(0, rows, size=S * B, dtype=np.uint64)
    kinds = np.where(rng.random(S * B) < p_read, 1, 2)
    kinds = np.where((kinds == 1) & (rng.random(S * B) < p_scan), 3, kinds)
    kinds = np.where(rng.random(S * B) < p_empty, 0, kinds)
    fld = rng.integers(0, fields, size=S * B)
    lens = rng.integers(1, max_len + 1, size=S * B)
    meta = (kinds | (fld << 8) | np.where(kinds == 3, (lens - 1) << 24, 0)).astype(np.uint32)
    if distinct_tids:
        tids = rng.permutation(np.arange(1, 4 * B + 1, dtype=np.uint32))[:B].astype(np.uint32)
    else:
        tids = rng.integers(0, 2**32 - 1, size=B, dtype=np.uint64).astype(np.uint32)
    return keys, meta, tids
