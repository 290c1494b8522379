"""The batched and served forms at their full configured sizes.

Batched (256 batches of 1e3-1e6 transactions, key-value on 10M rows x 10 x
100 B at Zipf 0.9 and order-entry at W=256) and served (Poisson-free waves
of 50/50 key-value requests on 100M rows x 10 x 100 B at Zipf 0.99 and
order-entry requests at W=1024).  Replaying every batch or epoch on
the single-threaded oracle would take hours, so:

- a seeded sample of batches / epochs is replayed on the oracle.  Commit
  decisions depend only on which ops touch which rows, not on the values,
  so a replay renumbers the rows (key-value) or warehouses (order-entry)
  the batch touches into a small table -- equality of rows, the only thing
  the rules look at, is kept;
- the final key-value table is checked in closed form on every row any
  committed write touched (and on untouched rows): a write adds its
  transaction's delta to each word of its field, and modular adds commute,
  so the final table is the initial one plus the sum of every committed
  write's delta, whatever the order.
"""
import time

import numpy as np
import pytest

import ltpg_amd as L
from oracle import oracle

SEED = 0x4C545047


def _renumbered_kv_epoch(keys_sm, meta_sm, tids, S, fields, deltas=None):
    """Oracle commit bitmap of one key-value batch (slot-major keys/meta)
    with its rows renumbered into a small table."""
    uniq, inv = np.unique(keys_sm, return_inverse=True)
    rows = len(uniq)
    table = np.zeros(rows * fields, dtype=np.uint32)
    ep = oracle.kv_epoch(table, rows, fields, 1, inv.astype(np.uint64), meta_sm, tids, S, want_reads=False,
                         inplace=True, deltas=deltas)
    return ep.bitmap


def _bits(bitmap, n):
    return np.unpackbits(np.asarray(bitmap).view(np.uint8), bitorder="little")[:n].astype(bool)


def _renumber_tpcc(b):
    """The batch with its warehouses renumbered 0 .. k-1, and k."""
    ws = [b.nf("W_ID"), b.pf("W_ID"), b.pf("C_W_ID")]
    lines = b.nf("OL_SUPPLY_W_ID")
    cnt = b.nf("O_OL_CNT").astype(np.int64)
    used = [lines[l][cnt > l] for l in range(15)]
    uniq = np.unique(np.concatenate([np.asarray(x, dtype=np.int64) for x in ws + used]))
    remap = {int(w): i for i, w in enumerate(uniq)}
    out = L.TpccBatch.empty(b.n_no, b.n_pay)
    for f in ("W_ID", "D_ID", "C_ID", "O_ID", "N_O_ID", "O_OL_CNT", "O_OL_ID", "TID"):
        out.nf(f)[:] = b.nf(f)
    for f in ("OL_I_ID", "OL_QUANTITY"):
        out.nf(f)[:, :] = b.nf(f)
    out.nf("OL_SUPPLY_W_ID")[:, :] = b.nf("OL_SUPPLY_W_ID")
    for f in ("W_ID", "D_ID", "C_ID", "C_LAST", "isName", "C_D_ID", "C_W_ID", "H_AMOUNT", "H_ID", "TID"):
        out.pf(f)[:] = b.pf(f)
    vmap = np.vectorize(lambda w: remap.get(int(w), 0), otypes=[np.uint32])
    out.nf("W_ID")[:] = vmap(b.nf("W_ID"))
    for l in range(15):
        live = cnt > l
        col = out.nf("OL_SUPPLY_W_ID", l)
        col[live] = vmap(b.nf("OL_SUPPLY_W_ID", l)[live])
    out.pf("W_ID")[:] = vmap(b.pf("W_ID"))
    out.pf("C_W_ID")[:] = vmap(b.pf("C_W_ID"))
    return out, len(uniq)


def _tpcc_oracle_bitmaps(b):
    """Oracle NewOrder / Payment bitmaps of an order-entry batch, replayed at
    its renumbered warehouse count."""
    rb, k = _renumber_tpcc(b)
    tables = oracle.tpcc_populate(k, SEED)
    ep = oracle.TpccOracle(k).epoch(tables, rb, 0, 1)
    return _bits(ep.no_bitmap, b.n_no), _bits(ep.pay_bitmap, b.n_pay)


def test_renumbering_keeps_commit_decisions():
    """The replay shortcut itself, on the CPU: renumbered key-value rows and
    order-entry warehouses give the oracle's decisions on the original ids."""
    rows, fields, S, B = 5000, 4, 8, 600
    keys, meta, tids = L.kv_generate(seed=3, rows=rows, fields=fields, n_txns=B, n_slots=S, min_ops=1, max_ops=S,
                                     theta=0.9)
    full = oracle.kv_epoch(np.zeros(rows * fields, np.uint32), rows, fields, 1, keys, meta, tids, S,
                           want_reads=False)
    np.testing.assert_array_equal(_renumbered_kv_epoch(keys, meta, tids, S, fields), full.bitmap)
    W = 6
    b = L.TpccGenerator(warehouses=W, seed=5).next(800)
    ep = oracle.TpccOracle(W).epoch(oracle.tpcc_populate(W, SEED), b, 0, 1)
    no, pay = _tpcc_oracle_bitmaps(b)
    np.testing.assert_array_equal(no, _bits(ep.no_bitmap, b.n_no))
    np.testing.assert_array_equal(pay, _bits(ep.pay_bitmap, b.n_pay))


@pytest.mark.gpu
def test_batched_form_at_full_size(gpu):
    """256 batches, sizes log-uniform in [1e3, 1e6], key-value (1-64 ops,
    Zipf 0.9, 10M rows x 10 x 100 B) or order-entry (W=256), run as the
    batched form's graphs in RMW mode."""
    rng = np.random.default_rng(SEED)
    nb, rows, fields, fw, S, W = 256, 10_000_000, 10, 25, 64, 256
    sizes = np.floor(np.exp(rng.uniform(np.log(1e3), np.log(1e6), nb))).astype(np.int64)
    is_kv = rng.random(nb) < 0.5
    kv = L.Context(workload=L.WORKLOAD_KV, kv_rows=rows, kv_fields=fields, kv_field_bytes=4 * fw,
                   max_txns=int(sizes.max()), kv_max_slots=S, values=L.VALUES_RMW)
    kv.kv_fill_reference()
    tp = L.Context(workload=L.WORKLOAD_TPCC, tpcc_warehouses=W, max_txns=int(sizes.max()), values=L.VALUES_RMW)
    tp.tpcc_populate(SEED)
    gen = L.TpccGenerator(warehouses=W, seed=SEED)
    descs, keep, tpcc_batches, tid = [], [], {}, 0

    def kv_batch(i):
        return L.kv_generate(seed=SEED + i, rows=rows, fields=fields, n_txns=int(sizes[i]), n_slots=S, min_ops=1,
                             max_ops=S, theta=0.9, tid_base=int(tid_base[i]))

    tid_base = np.r_[0, np.cumsum(sizes)[:-1]]
    for i in range(nb):
        if is_kv[i]:
            k, m, t = kv_batch(i)
            db = L.DeviceKvBatch(kv, k, m, t, S)
            keep.append(db)
            descs.append(L.kv_desc(db))
        else:
            b = gen.next(int(sizes[i]), tid_base=int(tid_base[i]))
            dev, bufs = L.upload_tpcc_batch(tp, b)
            keep.append(bufs)
            descs.append(L.tpcc_desc(dev))
            tpcc_batches[i] = b
    runner = L.Batched(kv, tp, descs)
    runner.autotune(reps=1)
    runner.run(L.RUN_GRAPH)
    got = {i: runner.bitmap(i, L.TYPE_KV) for i in range(nb) if is_kv[i]}
    got_no = {i: runner.bitmap(i, L.TYPE_NEWORDER) for i in tpcc_batches}
    got_pay = {i: runner.bitmap(i, L.TYPE_PAYMENT) for i in tpcc_batches}
    # sampled key-value batches on the oracle: the smallest, the largest and 4 seeded others
    kv_idx = np.nonzero(is_kv)[0]
    pick = {int(kv_idx[np.argmin(sizes[kv_idx])]), int(kv_idx[np.argmax(sizes[kv_idx])])}
    pick |= set(int(x) for x in np.random.default_rng(7).choice(kv_idx, 4, replace=False))
    # final table: every committed write adds (TID | 1) to each word of its field
    touched, deltas = [], []
    for i in kv_idx:
        k, m, t = kv_batch(int(i))
        n = int(sizes[i])
        c = _bits(got[int(i)], n)
        if int(i) in pick:
            np.testing.assert_array_equal(_bits(_renumbered_kv_epoch(k, m, t, S, fields), n), c,
                                          err_msg=f"key-value batch {i} ({n} txns)")
        # the minimum-priority transaction of a batch always commits
        assert c[np.argmin(t)], f"batch {i}: the minimum-TID transaction aborted"
        wr = ((m & 3) == 2).reshape(S, n) & c[None, :]
        touched.append(k.reshape(S, n)[wr].astype(np.int64) * fields + ((m.reshape(S, n)[wr] >> 8) & 0xFFFF))
        deltas.append(np.broadcast_to((t | 1).astype(np.uint64), (S, n))[wr])
    cells = np.concatenate(touched)
    uniq, inv = np.unique(cells, return_inverse=True)
    acc = np.zeros(len(uniq), dtype=np.uint64)
    np.add.at(acc, inv, np.concatenate(deltas))
    rows_t = np.unique(uniq // fields)
    sample = rows_t[np.random.default_rng(9).choice(len(rows_t), min(len(rows_t), 200_000), replace=False)]
    sample.sort()
    got_rows = kv.kv_table_read_rows(sample)
    want = np.broadcast_to(np.arange(fields, dtype=np.uint64)[None, :], (len(sample), fields)).copy()
    cell_of = sample[:, None].astype(np.int64) * fields + np.arange(fields)[None, :]
    pos = np.searchsorted(uniq, cell_of)
    hit = (pos < len(uniq)) & (uniq[np.minimum(pos, len(uniq) - 1)] == cell_of)
    want[hit] += acc[pos[hit]]
    np.testing.assert_array_equal(got_rows, np.broadcast_to(((want & 0xFFFFFFFF).astype(np.uint32))[:, :, None],
                                                            got_rows.shape))
    untouched = np.setdiff1d(np.random.default_rng(11).integers(0, rows, 5000), rows_t)
    np.testing.assert_array_equal(kv.kv_table_read_rows(untouched),
                                  np.broadcast_to(np.arange(fields, dtype=np.uint32)[None, :, None],
                                                  (len(untouched), fields, fw)))
    # sampled order-entry batches on the oracle: the smallest, the median, the largest
    oe = sorted(tpcc_batches, key=lambda i: sizes[i])
    for i in {oe[0], oe[len(oe) // 2], oe[-1]}:
        b = tpcc_batches[i]
        no, pay = _tpcc_oracle_bitmaps(b)
        np.testing.assert_array_equal(_bits(got_no[i], b.n_no), no, err_msg=f"order-entry batch {i} NewOrders")
        np.testing.assert_array_equal(_bits(got_pay[i], b.n_pay), pay, err_msg=f"order-entry batch {i} Payments")
    runner.close()
    for x in keep:
        for buf in (x if isinstance(x, list) else [x]):
            buf.free()
    kv.close()
    tp.close()


@pytest.mark.gpu
def test_served_form_at_full_size(gpu):
    """The served form on 100M rows x 10 x 100 B (Zipf 0.99, 1-64 ops) and
    W=1024, 50/50 key-value and order-entry requests, key-value epochs in the
    device loop: every request commits; a seeded sample of key-value epochs
    and every small order-entry epoch replay on the oracle; the touched rows
    end as the closed form says."""
    rows, fields, fw, S, W = 100_000_000, 10, 25, 64, 1024
    n_kv, n_oe = 60_000, 60_000
    kv = L.Context(workload=L.WORKLOAD_KV, kv_rows=rows, kv_fields=fields, kv_field_bytes=4 * fw,
                   max_txns=1 << 16, kv_max_slots=S, values=L.VALUES_RMW)
    kv.kv_fill_reference()
    tp = L.Context(workload=L.WORKLOAD_TPCC, tpcc_warehouses=W, max_txns=1 << 16, values=L.VALUES_RMW)
    tp.tpcc_populate(SEED)
    keys, meta, _ = L.kv_generate(seed=SEED + 5, rows=rows, fields=fields, n_txns=n_kv, n_slots=S, min_ops=1,
                                  max_ops=S, theta=0.99)
    keys = keys.reshape(S, n_kv).T.copy()
    meta = meta.reshape(S, n_kv).T.copy()
    gen = L.TpccGenerator(warehouses=W, seed=SEED)
    oe_batches = [gen.next(40) for _ in range(25)] + [gen.next((n_oe - 1000) // 10) for _ in range(10)]
    sv = L.Server(kv, tp, kv_ops=S, slots=2, record_epochs=1, epoch_floor=16)
    kv_ids, oe_index = [], {}
    waves = np.array_split(np.arange(n_kv), len(oe_batches))
    for w, b in zip(waves, oe_batches):
        first = sv.submit_kv(keys[w], meta[w])
        kv_ids.extend(range(first, first + len(w)))
        first = sv.submit_tpcc(b)
        for q in range(b.n_no):
            oe_index[first + q] = (b, "no", q)
        for p in range(b.n_pay):
            oe_index[first + b.n_no + p] = (b, "pay", p)
        time.sleep(0.002)
    sv.drain()
    done = {d["id"]: d for d in sv.poll(1 << 22)}
    st = sv.stats()
    kv_epochs = [sv.epoch(L.WORKLOAD_KV, e) for e in range(st["epochs"][0])]
    oe_epochs = [sv.epoch(L.WORKLOAD_TPCC, e) for e in range(st["epochs"][1])]
    dev = sv.device_stats()
    sv.stop()
    n_all = n_kv + sum(b.n_no + b.n_pay for b in oe_batches)
    assert len(done) == n_all, "every request commits"
    assert dev["kv_loop"] and dev["launches"] >= 1
    kv_pos = {rid: j for j, rid in enumerate(kv_ids)}
    # sampled key-value epochs on the oracle
    for e in np.random.default_rng(3).choice(len(kv_epochs), min(300, len(kv_epochs)), replace=False):
        ids = kv_epochs[int(e)].astype(np.int64)
        rq = np.array([kv_pos[int(i)] for i in ids])
        B = len(rq)
        bm = _renumbered_kv_epoch(keys[rq].T.reshape(-1).copy(), meta[rq].T.reshape(-1).copy(),
                                  np.arange(B, dtype=np.uint32), S, fields)
        want = np.array([done[int(i)]["epoch"] == e for i in ids])
        np.testing.assert_array_equal(_bits(bm, B), want, err_msg=f"key-value epoch {e}")
    # every order-entry epoch of at most 64 requests on the oracle (warehouses renumbered)
    replayed = 0
    for e, ids in enumerate(oe_epochs):
        if len(ids) > 64:
            continue
        nos = [int(i) for i in ids if oe_index[int(i)][1] == "no"]
        pays = [int(i) for i in ids if oe_index[int(i)][1] == "pay"]
        eb = L.TpccBatch.empty(len(nos), len(pays))
        for q, rid in enumerate(nos):
            b, _, pos = oe_index[rid]
            for f in ("W_ID", "D_ID", "C_ID", "O_ID", "N_O_ID", "O_OL_CNT", "O_OL_ID"):
                eb.nf(f)[q] = b.nf(f)[pos]
            for f in ("OL_I_ID", "OL_SUPPLY_W_ID", "OL_QUANTITY"):
                eb.nf(f)[:, q] = b.nf(f)[:, pos]
            eb.nf("TID")[q] = done[rid]["tid"]
        for p, rid in enumerate(pays):
            b, _, pos = oe_index[rid]
            for f in ("W_ID", "D_ID", "C_ID", "C_LAST", "isName", "C_D_ID", "C_W_ID", "H_AMOUNT", "H_ID"):
                eb.pf(f)[p] = b.pf(f)[pos]
            eb.pf("TID")[p] = done[rid]["tid"]
        no, pay = _tpcc_oracle_bitmaps(eb)
        np.testing.assert_array_equal(no, np.array([done[r]["epoch"] == e for r in nos], dtype=bool))
        np.testing.assert_array_equal(pay, np.array([done[r]["epoch"] == e for r in pays], dtype=bool))
        replayed += 1
    assert replayed >= 5
    # the key-value table: every request committed once, so the touched
    # fields hold the initial value plus every write's delta
    seqs = np.array([done[i]["seq"] for i in kv_ids], dtype=np.uint64)
    wr = (meta & 3) == 2
    cells = keys[wr].astype(np.int64) * fields + ((meta[wr] >> 8) & 0xFFFF)
    uniq, inv = np.unique(cells, return_inverse=True)
    acc = np.zeros(len(uniq), dtype=np.uint64)
    np.add.at(acc, inv, np.broadcast_to(((seqs & 0xFFFFFFFF) | 1)[:, None], wr.shape)[wr])
    rows_t = np.unique(uniq // fields)
    sample = np.sort(rows_t[np.random.default_rng(4).choice(len(rows_t), min(len(rows_t), 100_000), replace=False)])
    got_rows = kv.kv_table_read_rows(sample)
    want = np.broadcast_to(np.arange(fields, dtype=np.uint64)[None, :], (len(sample), fields)).copy()
    cell_of = sample[:, None].astype(np.int64) * fields + np.arange(fields)[None, :]
    pos = np.searchsorted(uniq, cell_of)
    hit = (pos < len(uniq)) & (uniq[np.minimum(pos, len(uniq) - 1)] == cell_of)
    want[hit] += acc[pos[hit]]
    np.testing.assert_array_equal(got_rows, np.broadcast_to(((want & 0xFFFFFFFF).astype(np.uint32))[:, :, None],
                                                            got_rows.shape))
    kv.close()
    tp.close()
