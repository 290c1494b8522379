"""Pins the key-value oracle (CPU only).

The reference ships no tests, fixtures or golden vectors for this path, so
the oracle is pinned by (1) hand-derived known-answer epochs
(tests/golden/kv_kat.json), (2) an independent pairwise computation of the
conflict counters, (3) Aria's serializability guarantee checked on the
committed set with a serial replay that also fixes the final table, and
(4) brute force over every TID permutation of tiny batches.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle
from tests import kvcases as kc

KAT = kc.load_kat()


@pytest.mark.parametrize("case", KAT["cases"], ids=[c["name"] for c in KAT["cases"]])
def test_oracle_known_answers(case):
    rows, fields, fw = case["rows"], case["fields"], case["field_words"]
    keys, meta, tids, S = kc.case_arrays(case)
    values = 1 if case["values"] == "rmw" else 0
    max_scan = case.get("max_scan", 0)
    ep = oracle.kv_epoch(kc.initial_table(rows, fields, fw), rows, fields, fw, keys, meta, tids, S, values,
                         hot_row=case.get("hot_row", oracle.NO_HOT_ROW), max_scan=max_scan)
    B = len(tids)
    L = max(1, max_scan)
    exp = case["expect"]
    assert kc.bitmap_to_list(ep.bitmap, B) == exp["commit"]
    assert kc.unpack_conflicts(ep.conflicts).tolist() == exp["conflicts"]
    assert ep.range_errors == exp["range_errors"]
    assert ep.commit_count == sum(exp["commit"])
    np.testing.assert_array_equal(ep.table, kc.expected_table(case))
    for t, s, words in exp["reads"]:
        i = (s * B + t) * L
        assert ep.read_values[i * fw:(i + 1) * fw].tolist() == words
    for t, s, j, words in exp.get("scan_reads", []):
        i = (s * B + t) * L + j
        assert ep.read_values[i * fw:(i + 1) * fw].tolist() == words


def _run(rows, fields, fw, keys, meta, tids, S, values=1):
    return oracle.kv_epoch(kc.initial_table(rows, fields, fw), rows, fields, fw, keys, meta, tids, S, values)


@pytest.mark.parametrize("seed", range(12))
def test_oracle_explicit_deltas(seed):
    """The deltas variant (served-form RMW deltas from 64-bit arrival
    numbers): deltas = tids | 1 reproduces the plain epoch exactly; other
    deltas leave every decision and read unchanged and add, per committed
    write op, its transaction's delta to each word of its field (a closed
    form summed here with numpy, independent of the oracle)."""
    rng = np.random.default_rng(100 + seed)
    B, S, rows, fields, fw = int(rng.integers(1, 40)), int(rng.integers(1, 6)), int(rng.integers(1, 12)), 3, 2
    keys, meta, tids = kc.random_batch(rng, B, S, rows, fields)
    t0 = kc.initial_table(rows, fields, fw)
    a = oracle.kv_epoch(t0, rows, fields, fw, keys, meta, tids, S)
    b = oracle.kv_epoch(t0, rows, fields, fw, keys, meta, tids, S, deltas=tids | 1)
    np.testing.assert_array_equal(a.table, b.table)
    deltas = rng.integers(0, 1 << 32, B, dtype=np.uint64).astype(np.uint32)
    c = oracle.kv_epoch(t0, rows, fields, fw, keys, meta, tids, S, deltas=deltas)
    np.testing.assert_array_equal(a.bitmap, c.bitmap)
    np.testing.assert_array_equal(a.read_values, c.read_values)
    want = t0.reshape(rows, fields, fw).astype(np.uint64)
    committed = kc.bitmap_to_list(a.bitmap, B)
    for s_ in range(S):
        for t in range(B):
            m = int(meta[s_ * B + t])
            if m & 3 == 2 and committed[t] and int(keys[s_ * B + t]) < rows:
                want[int(keys[s_ * B + t]), (m >> 8) & 0xFFFF] += int(deltas[t])
    np.testing.assert_array_equal(c.table, (want & 0xFFFFFFFF).astype(np.uint32).reshape(-1))


@pytest.mark.parametrize("seed", range(40))
def test_oracle_matches_pairwise_and_is_serializable(seed):
    rng = np.random.default_rng(seed)
    B = int(rng.integers(1, 40))
    S = int(rng.integers(1, 6))
    rows = int(rng.integers(1, 12))
    fields, fw = 3, int(rng.integers(1, 4))
    keys, meta, tids = kc.random_batch(rng, B, S, rows, fields)
    ep = _run(rows, fields, fw, keys, meta, tids, S)
    conf = kc.unpack_conflicts(ep.conflicts)
    np.testing.assert_array_equal(conf, kc.pairwise_conflicts(keys, meta, tids, B, S, rows, fields))
    committed = kc.bitmap_to_list(ep.bitmap, B)
    assert committed == kc.aria_commit(conf).tolist()
    ok, msg, final = kc.serial_check(kc.initial_table(rows, fields, fw), keys, meta, tids, B, S, rows,
                                     fields, fw, committed)
    assert ok, msg
    np.testing.assert_array_equal(ep.table, final)
    # committed reads return the snapshot
    snap = kc.initial_table(rows, fields, fw).reshape(rows, fields, fw)
    for s in range(S):
        for t in range(B):
            m = int(meta[s * B + t])
            if m & 3 == 1 and committed[t]:
                i = s * B + t
                np.testing.assert_array_equal(ep.read_values[i * fw:(i + 1) * fw],
                                              snap[int(keys[i]), (m >> 8) & 0xFFFF])


def test_oracle_bruteforce_all_tid_permutations():
    """Every TID order of small access patterns (<= 6 txns on <= 4 rows)."""
    rng = np.random.default_rng(1234)
    checked = 0
    for pattern in range(12):
        B = int(rng.integers(2, 7))
        S = int(rng.integers(1, 4))
        rows, fields, fw = int(rng.integers(1, 5)), 2, 1
        keys, meta, _ = kc.random_batch(rng, B, S, rows, fields, p_empty=0.0)
        for perm in itertools.permutations(range(1, B + 1)):
            tids = np.array(perm, dtype=np.uint32) * 3
            ep = _run(rows, fields, fw, keys, meta, tids, S)
            conf = kc.unpack_conflicts(ep.conflicts)
            np.testing.assert_array_equal(conf, kc.pairwise_conflicts(keys, meta, tids, B, S, rows, fields))
            committed = kc.bitmap_to_list(ep.bitmap, B)
            ok, msg, final = kc.serial_check(kc.initial_table(rows, fields, fw), keys, meta, tids, B, S,
                                             rows, fields, fw, committed)
            assert ok, msg
            np.testing.assert_array_equal(ep.table, final)
            # the lowest TID always commits
            assert committed[int(np.argmin(tids))] == 1
            checked += 1
    assert checked > 1000


@pytest.mark.parametrize("seed", range(10))
def test_oracle_equal_tids_break_ties_by_position(seed):
    """Equal TIDs are outside the reference's contract; the build orders
    them by batch position, and the pins still hold."""
    rng = np.random.default_rng(1000 + seed)
    B, S, rows, fields, fw = 30, 3, 6, 2, 1
    keys, meta, _ = kc.random_batch(rng, B, S, rows, fields)
    tids = rng.integers(0, 4, size=B).astype(np.uint32)  # many ties
    ep = _run(rows, fields, fw, keys, meta, tids, S)
    conf = kc.unpack_conflicts(ep.conflicts)
    np.testing.assert_array_equal(conf, kc.pairwise_conflicts(keys, meta, tids, B, S, rows, fields))
    ok, msg, final = kc.serial_check(kc.initial_table(rows, fields, fw), keys, meta, tids, B, S, rows, fields,
                                     fw, kc.bitmap_to_list(ep.bitmap, B))
    assert ok, msg
    np.testing.assert_array_equal(ep.table, final)


def test_oracle_copyback_never_changes_table():
    rng = np.random.default_rng(7)
    rows, fields, fw, B, S = 16, 4, 3, 64, 6
    keys, meta, tids = kc.random_batch(rng, B, S, rows, fields)
    ep = _run(rows, fields, fw, keys, meta, tids, S, values=0)
    np.testing.assert_array_equal(ep.table, kc.initial_table(rows, fields, fw))
    ep1 = _run(rows, fields, fw, keys, meta, tids, S, values=1)
    np.testing.assert_array_equal(ep.bitmap, ep1.bitmap)


@pytest.mark.parametrize("seed", range(30))
def test_oracle_scans_match_pairwise_and_are_serializable(seed):
    """Scans are reads of every row of their (clipped) range: conflicts match
    the pairwise computation over the expanded rows, the committed set is
    serializable, and every delivered row is the snapshot's."""
    rng = np.random.default_rng(1000 + seed)
    B, S = int(rng.integers(1, 30)), int(rng.integers(1, 5))
    rows, fields, fw, max_scan = int(rng.integers(2, 14)), 2, int(rng.integers(1, 3)), int(rng.integers(1, 6))
    keys, meta, tids = kc.random_batch(rng, B, S, rows, fields, p_scan=0.6, max_len=max_scan + 1)
    ep = oracle.kv_epoch(kc.initial_table(rows, fields, fw), rows, fields, fw, keys, meta, tids, S, 1,
                         max_scan=max_scan)
    conf = kc.unpack_conflicts(ep.conflicts)
    np.testing.assert_array_equal(conf, kc.pairwise_conflicts(keys, meta, tids, B, S, rows, fields, max_scan))
    committed = kc.bitmap_to_list(ep.bitmap, B)
    assert committed == kc.aria_commit(conf).tolist()
    ok, msg, final = kc.serial_check(kc.initial_table(rows, fields, fw), keys, meta, tids, B, S, rows, fields,
                                     fw, committed, max_scan=max_scan)
    assert ok, msg
    np.testing.assert_array_equal(ep.table, final)
    too_long = ((meta & 3) == 3) & ((meta >> 24) + 1 > max_scan)
    assert ep.range_errors == int(too_long.sum())
    snap = kc.initial_table(rows, fields, fw).reshape(rows, fields, fw)
    L = max(1, max_scan)
    for s in range(S):
        for t in range(B):
            i = s * B + t
            m = int(meta[i])
            if not committed[t] or m & 3 not in (1, 3) or too_long[i]:
                continue
            n = 1 if m & 3 == 1 else min((m >> 24) + 1, rows - int(keys[i]))
            for j in range(n):
                k = i * L + j
                np.testing.assert_array_equal(ep.read_values[k * fw:(k + 1) * fw],
                                              snap[int(keys[i]) + j, (m >> 8) & 0xFFFF])


def _rules_with_mistake(keys, meta, tids, B, S, rows, fields, fw, table0, mistake):
    """A CPU copy of the rules with one planted mistake (test-only)."""
    ops = kc.ops_of(keys, meta, B, S, rows, fields)
    R, W = {}, {}
    NONE = (1 << 64,)
    for t in range(B):
        for _, kind, row, _ in ops[t]:
            reg = R if (kind == 1) != (mistake == "swap_registers") else W
            reg[row] = min(reg.get(row, NONE), (int(tids[t]), t))
    conf = np.zeros((B, 3), dtype=np.uint32)
    gt = (lambda a, b: a >= b) if mistake == "ge" else (lambda a, b: a > b)
    for t in range(B):
        me = (int(tids[t]), t)
        for _, kind, row, _ in ops[t]:
            if kind == 1:
                conf[t, 0] += gt(me, W.get(row, NONE))
            else:
                conf[t, 1] += gt(me, (W if mistake == "war_from_w" else R).get(row, NONE))
                conf[t, 2] += gt(me, W.get(row, NONE))
    raw, war, waw = conf[:, 0], conf[:, 1], conf[:, 2]
    if mistake == "waw_only":
        commit = waw == 0
    elif mistake == "or_rule":
        commit = (waw == 0) | (raw == 0) | (war == 0)
    else:
        commit = (waw == 0) & ((raw == 0) | (war == 0))
    tab = np.array(table0, dtype=np.uint32).reshape(rows, fields, fw).copy()
    for t in range(B):
        if commit[t]:
            seen = set()
            for _, kind, row, field in ops[t]:
                if kind == 2:
                    if mistake == "apply_once" and (row, field) in seen:
                        continue
                    seen.add((row, field))
                    tab[row, field] += np.uint32(int(tids[t]) | 1)
    return commit.astype(int).tolist(), conf, tab.reshape(-1)


@pytest.mark.parametrize("mistake", ["swap_registers", "ge", "war_from_w", "waw_only", "or_rule", "apply_once"])
def test_pins_catch_planted_mistakes(mistake):
    """Each plausible mistake fails a known answer or the serializability
    replay, so the pins above would catch it in the oracle."""
    caught = False
    for case in KAT["cases"]:
        if case["values"] != "rmw" or case["expect"]["range_errors"] or "hot_row" in case or "max_scan" in case:
            continue
        rows, fields, fw = case["rows"], case["fields"], case["field_words"]
        keys, meta, tids, S = kc.case_arrays(case)
        commit, conf, tab = _rules_with_mistake(keys, meta, tids, len(tids), S, rows, fields, fw,
                                                kc.initial_table(rows, fields, fw), mistake)
        if (commit != case["expect"]["commit"] or conf.tolist() != case["expect"]["conflicts"]
                or not np.array_equal(tab, kc.expected_table(case))):
            caught = True
            break
    rng = np.random.default_rng(5)
    for _ in range(300 if not caught else 0):
        rows, fields, fw, B, S = 5, 2, 1, 12, 3
        keys, meta, tids = kc.random_batch(rng, B, S, rows, fields, p_empty=0.0)
        commit, _, tab = _rules_with_mistake(keys, meta, tids, B, S, rows, fields, fw,
                                             kc.initial_table(rows, fields, fw), mistake)
        ok, _, final = kc.serial_check(kc.initial_table(rows, fields, fw), keys, meta, tids, B, S, rows,
                                       fields, fw, commit)
        if not ok or not np.array_equal(final, tab):
            caught = True
            break
    assert caught, f"planted mistake {mistake!r} passed every pin"
