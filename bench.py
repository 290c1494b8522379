"""Benchmark of the deterministic concurrency-control hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload all|served-c5|kv-c2|tpcc-c3|batched-c4|...]

The headline is the served form on the served-form configuration: Poisson
arrivals of a 50/50 mix of key-value requests (1-64 ops, Zipf 0.99, 100M
rows x 10 fields x 100 B) and order-entry requests (W = 1024, 50/50
NewOrder / Payment) at --c5-rate M txn/s.  A step is one 1 ms window of
arrivals; the timed span runs from the first arrival to the last commit,
and ms_per_step = span / steps.  The line carries the plain per-batch form
on the same arrivals, p50/p99 commit latency of both (at the headline rate
and below capacity), abort rate and attempts per commit, the roofline of
the dominant kernel (the device loop, per epoch) and the CPU baseline (the
oracle replaying the served key-value lane on one host core).

Compact variants on the same line: the key-value plain form (10M rows x 10
x 100 B, 2^20 txns x 16 ops, Zipf 0.6 and 0.99; and the YCSB/1 mix of 9
reads + 1 RMW on uniform keys, where CHECK and WRITEBACK carry volume), the
order-entry plain form (W = 64, 2^18 txns) and the batched form (256 mixed
batches of 1e3-1e6 txns) against the plain loop, each with its roofline.
Every figure behind the line goes to --detail (gpurun_out/bench_detail.json).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
METRIC = "committed txns/sec at 1 GPU + abort rate; p99 commit latency vs plain form"
T_START = time.perf_counter()


def pinned_empty(nbytes):
    """(pointer, bytes buffer) of nbytes of pinned host memory (hipHostMalloc)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    p = ctypes.c_void_p()
    if hip.hipHostMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes), 0) != 0:
        raise RuntimeError("hipHostMalloc failed")
    return p, (ctypes.c_uint8 * nbytes).from_address(p.value)


def pinned_free(buf):
    import ctypes
    ctypes.CDLL("libamdhip64.so").hipHostFree(buf[0])


def log(msg):
    print(f"[bench {time.perf_counter() - T_START:7.1f}s] {msg}", file=sys.stderr, flush=True)


def kv_workload(name, theta):
    if name == "kv-c2":
        theta = 0.6 if theta is None else theta
        return dict(rows=10_000_000, fields=10, fw=25, B=1 << 20, S=16, theta=theta, read_pct=50,
                    desc=f"key-value plain form: 10M rows x 10 fields x 100 B, 2^20 txns x 16 ops, "
                         f"50/50 read/RMW, Zipf {theta}")
    if name == "kv-y1":
        return dict(rows=10_000_000, fields=10, fw=25, B=1 << 20, S=10, theta=0.0, read_pct=90, kinds=[1] * 9 + [2],
                    desc="key-value plain form, the YCSB/1 op mix (9 reads + 1 RMW per txn, uniform keys) at "
                         "10M rows x 10 fields x 100 B and 2^20 txns: low contention, so most transactions "
                         "commit and CHECK / WRITEBACK move their full volume")
    if name == "kv-c1":
        return dict(rows=1_000_000, fields=10, fw=1, B=65536, S=10, theta=0.8, read_pct=50,
                    desc="key-value plain form: 1M rows x 10 u32 fields, 65,536 txns x 10 ops, "
                         "50/50 read/RMW, Zipf 0.8")
    if name == "kv-scan":
        return dict(rows=1_000_000, fields=10, fw=1, B=10000, S=10, theta=0.0, read_pct=90, scan_len=10,
                    desc="key-value scans (the YCSB/2 configuration): 1M rows x 10 u32 fields, batches of "
                         "10,000 txns x 10 ops, 90 % scans of 10 rows / 10 % RMW, uniform keys")
    raise SystemExit(f"unknown workload {name}")


KERNELS = ["execute", "check", "writeback"]  # the fast CHECK includes the read-out


def kv_kernel_bytes(keys, meta, tids, bitmap, B, S, fw, rows=None, max_scan=0):
    """Algorithmic bytes each key-value phase must move for one batch (fast
    form): every byte the algorithm has to read or write, once.
    EXECUTE (write registration + read marking): the batch (tids, keys and
    meta of every slot), one 8-byte W probe per valid op, 8 bytes per W
    register written, one flag byte per doomed transaction and per raw / war
    flag set.  CHECK (commit + read-out): doomed, raw and war flags, the
    bitmap, then each committed transaction's ops and its read fields
    (snapshot gather + read-value store).  WRITEBACK: the bitmap, each
    committed transaction's tid and ops, and each committed write's field
    (load + store)."""
    F = 4 * fw
    kind = (meta & 3).reshape(S, B)
    rows_ = keys.reshape(S, B).astype(np.int64)
    nrows = int(rows_.max()) + 1 if rows_.size else 1
    rows = nrows if rows is None else rows
    is_w = kind == 2
    # rows read per slot: 1 for a read, the clipped range for a scan
    slen = (meta.reshape(S, B) >> 24).astype(np.int64) + 1
    nread = np.where(kind == 1, 1, np.where((kind == 3) & (slen <= max_scan), np.minimum(slen, rows - rows_), 0))
    prio = (tids.astype(np.int64) << 24) | np.arange(B, dtype=np.int64)
    big = np.iinfo(np.int64).max
    Wmin = np.full(max(nrows, rows), big, dtype=np.int64)
    pr = np.broadcast_to(prio, (S, B))
    np.minimum.at(Wmin, rows_[is_w], pr[is_w])
    waw = (is_w & (Wmin[rows_] < pr)).any(axis=0)
    # every row read, as (row, txn, priority)
    cnt = nread.reshape(-1)
    rr = np.repeat(rows_.reshape(-1), cnt) + (np.arange(int(cnt.sum())) - np.repeat(np.cumsum(cnt) - cnt, cnt))
    rt = np.repeat(np.broadcast_to(np.arange(B), (S, B)).reshape(-1), cnt)
    rp = prio[rt]
    raw = np.zeros(B, dtype=bool)
    raw[rt[Wmin[rr] < rp]] = True
    owner = Wmin[rr] & ((1 << 24) - 1)
    war_owners = np.unique(owner[(Wmin[rr] < big) & (Wmin[rr] > rp)])
    bits = np.unpackbits(bitmap.view(np.uint8), bitorder="little")[:B].astype(bool)
    committed = int(bits.sum())
    reg_ops = int(is_w.sum()) + int(cnt.sum())
    w_regs = int((Wmin < big).sum())
    execute = 4 * B + 12 * S * B + 8 * reg_ops + 8 * w_regs + int(waw.sum()) + int(raw.sum()) + war_owners.size
    check = 3 * B + B // 8 + committed * 12 * S + int((nread * bits[None, :]).sum()) * 2 * F
    writeback = B // 8 + committed * (4 + 12 * S) + int((is_w & bits[None, :]).sum()) * 2 * F
    return {"execute": execute, "check": check, "writeback": writeback}, committed


def choose_kv_shape(ctx, batch, reps=3):
    """Time the distinct key-value launch shapes (ltpg_amd.kv_shapes(): the
    LTPG_TUNE_* sets of per-op execute with block-local row folding, split
    register arrays, 64-transaction chunks, per-lane field copies and more
    items per lane, and the row-partitioned EXECUTE with the ones it uses)
    on one batch, one warm-up and `reps` epochs each, and keep the fastest for
    the run.  The timed loop is the plain per-batch loop either way."""
    import ltpg_amd as L

    times = {}
    for flags in L.kv_shapes():
        ctx.kv_set_tuning(flags)
        ctx.kv_submit_device(batch)
        ctx.epoch_run()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.kv_submit_device(batch)
            ctx.epoch_run()
        ctx.sync()
        times[flags] = (time.perf_counter() - t0) * 1e3 / reps
    best = min(times, key=times.get)
    ctx.kv_set_tuning(best)
    return best, {str(k): v for k, v in times.items()}


def cpu_baseline_kv(w, keys, meta, tids, budget_s=10.0):
    """Time the oracle (single-threaded C replay of the same rules, read
    staging included) on a bounded sample of the same workload: prefixes of
    the batch over the full-size table, growing until about budget_s of CPU
    work (then whole-batch epochs repeat until the budget is spent)."""
    from oracle import oracle

    rows, fields, fw, B, S = w["rows"], w["fields"], w["fw"], w["B"], w["S"]
    table = np.empty(rows * fields * fw, dtype=np.uint32)
    table.reshape(rows, fields, fw)[:] = np.arange(fields, dtype=np.uint32)[None, :, None]
    n = min(B, 1 << 14)
    total_s, total_commits, total_txns, epochs = 0.0, 0, 0, 0
    while total_s < budget_s:
        k = keys.reshape(S, B)[:, :n].copy()
        m = meta.reshape(S, B)[:, :n].copy()
        t = tids[:n].copy()
        t0 = time.perf_counter()
        ep = oracle.kv_epoch(table, rows, fields, fw, k, m, t, S, values=1, want_reads=True, inplace=True,
                             max_scan=w.get("scan_len", 0))
        dt = time.perf_counter() - t0
        total_s, total_commits, total_txns, epochs = (total_s + dt, total_commits + ep.commit_count,
                                                      total_txns + n, epochs + 1)
        if n < B:
            n = min(B, int(n * max(2.0, min(8.0, (budget_s - total_s) / max(dt, 1e-3)))))
    return {"value": total_commits / total_s, "unit": "txn/s", "cores": 1, "kind": "port",
            "attempted_txn_per_s": total_txns / total_s,
            "sample": f"oracle (single-threaded C replay, read staging included) over the full {rows}-row "
                      f"table: {epochs} epochs on prefixes of the batch, {total_txns} txns x {S} ops in "
                      f"{total_s:.1f} s"}


def reference_convention(ctx, submit, n_epochs, types, phases):
    """The reference's own measurement (TPCC-allopt-final/tp.cu:141-291,
    Predefine.h:15): epochs run one at a time, each phase followed by a
    synchronisation, then the commit counts read back; the first and last
    10 % of epochs are dropped; TPS = mean commits per epoch / mean epoch
    time (host clock).  Batches are already resident, so there is no H2D."""
    t_ep, c_ep = [], []
    for k in range(n_epochs):
        t1 = time.perf_counter()
        submit(k)
        for ph in phases:
            ctx.epoch_run(ph)
            ctx.sync()
        c = sum(ctx.commit_count(t) for t in types)
        t_ep.append(time.perf_counter() - t1)
        c_ep.append(c)
    lo, hi = n_epochs // 10, n_epochs - n_epochs // 10
    t_ep, c_ep = np.array(t_ep[lo:hi]), np.array(c_ep[lo:hi])
    return {"value": float(c_ep.mean() / t_ep.mean()), "unit": "txn/s", "epochs_kept": int(hi - lo),
            "ms_per_epoch": float(t_ep.mean() * 1e3),
            "what": "the reference's convention: per epoch submit, each phase + sync, commit counts read back; "
                    "first and last 10% of epochs dropped; mean commits / mean epoch time"}


def run_kv(args, theta, cpu_baseline=True, name=None, hot_row=None):
    """hot_row = 0: row 0 is the hot key, as in the reference's ZIPFIAN mode
    (its byte accounting still counts row-0 writes as registrations)."""
    import ltpg_amd as L

    w = kv_workload(name or args.workload, theta)
    if hot_row is not None:
        w["desc"] += f", row {hot_row} the hot key (the reference's ZIPFIAN mode)"
    rows, fields, fw, B, S = w["rows"], w["fields"], w["fw"], w["B"], w["S"]
    P = args.batches
    log(f"generating {P} batches: {w['desc']}")
    scan = w.get("scan_len", 0)
    gen = [L.kv_generate(seed=0x4C545047 + i, rows=rows, fields=fields, n_txns=B, n_slots=S,
                         read_pct=w["read_pct"], theta=w["theta"], tid_base=0, scan_len=scan) for i in range(P)]
    if "kinds" in w:
        # a fixed op kind per slot (the YCSB/1 mix: slots 0-8 read, slot 9 writes)
        kinds = np.repeat(np.asarray(w["kinds"], dtype=np.uint32), B)
        gen = [(k_, (m_ & ~np.uint32(3)) | kinds, t_) for (k_, m_, t_) in gen]
    ctx = L.Context(workload=L.WORKLOAD_KV, kv_rows=rows, kv_fields=fields, kv_field_bytes=4 * fw,
                    max_txns=B, kv_max_slots=S, values=L.VALUES_RMW, kv_max_scan=scan,
                    kv_hot_row=L.NO_HOT_ROW if hot_row is None else hot_row)
    ctx.kv_fill_reference()
    dev = [L.DeviceKvBatch(ctx, k, m, t, S) for (k, m, t) in gen]
    log("batches resident in HBM")

    # one pass per batch: commit sets depend only on the batch, which fixes
    # the committed-write bytes of every later run of it
    per_batch = []
    for i in range(P):
        ctx.kv_submit_device(dev[i])
        ctx.epoch_run()
        ctx.sync()
        per_batch.append(kv_kernel_bytes(gen[i][0], gen[i][1], gen[i][2], ctx.commit_bitmap(), B, S, fw, rows,
                                         scan))

    shape, shape_ms = choose_kv_shape(ctx, dev[0]) if not args.default_shape else (0, {})
    log(f"launch shape flags {shape}")

    for k in range(args.warmup):
        ctx.kv_submit_device(dev[k % P])
        ctx.epoch_run()
    ctx.sync()

    log(f"warmup done; timing {args.steps} steps")
    ctx.commit_totals(reset=True)
    ctx.set_timing(True)
    ctx.phase_times()  # start a fresh measurement
    t0 = time.perf_counter()
    for k in range(args.steps):
        ctx.kv_submit_device(dev[k % P])
        ctx.epoch_run()
    ctx.sync()
    elapsed = time.perf_counter() - t0
    phase_ms, n_timed = ctx.phase_times()
    ctx.set_timing(False)
    commits, attempts = ctx.commit_totals()

    log(f"timed region: {elapsed:.3f} s; latency loop")
    # reference-style synchronous tp loop: submit, phases, commit readback
    lat = []
    for k in range(max(5, args.steps // 2)):
        t1 = time.perf_counter()
        ctx.kv_submit_device(dev[k % P])
        ctx.epoch_run()
        ctx.commit_count()
        lat.append((time.perf_counter() - t1) * 1e3)

    ref_conv = reference_convention(ctx, lambda k: ctx.kv_submit_device(dev[k % P]), max(10, args.steps),
                                    [L.TYPE_KV], [L.PHASE_EXECUTE, L.PHASE_CHECK, L.PHASE_WRITEBACK])

    # the same epochs with the batch handed over in host memory: the submit
    # copies keys / meta / TIDs over PCIe into one of two device slots while
    # the previous epoch runs (pipelined), from pageable numpy arrays and
    # from pinned buffers; and one epoch at a time with the commit count read
    # back after each (the serial form)
    def pipelined(srcs, n_ep=6):
        ctx.sync()
        ctx.commit_totals(reset=True)
        t1 = time.perf_counter()
        for k in range(n_ep):
            ctx.kv_submit(*srcs[k % P], S)
            ctx.epoch_run()
        ctx.sync()
        dt = time.perf_counter() - t1
        c, _ = ctx.commit_totals()
        return {"value": c / dt, "unit": "txn/s", "ms_per_epoch": dt * 1e3 / n_ep}

    pinned, pin_bufs = [], []
    for (k_, m_, t_) in gen:
        arrs = []
        for a in (k_, m_, t_):
            buf = pinned_empty(a.nbytes)
            pin_bufs.append(buf)
            view = np.frombuffer(buf[1], dtype=a.dtype, count=a.size)
            view[:] = a
            arrs.append(view)
        pinned.append(tuple(arrs))
    pcie = {"pageable_pipelined": pipelined(gen), "pinned_pipelined": pipelined(pinned),
            "batch_bytes": int(gen[0][0].nbytes + gen[0][1].nbytes + gen[0][2].nbytes)}
    for buf in pin_bufs:
        pinned_free(buf)
    serial_ms, serial_commits = [], 0
    for k in range(3):
        t1 = time.perf_counter()
        ctx.kv_submit(gen[k % P][0], gen[k % P][1], gen[k % P][2], S)
        ctx.epoch_run()
        serial_commits += ctx.commit_count()
        serial_ms.append((time.perf_counter() - t1) * 1e3)
    pcie["serial"] = {"value": serial_commits / (sum(serial_ms) * 1e-3), "unit": "txn/s",
                      "ms_per_epoch": float(np.mean(serial_ms))}
    pcie["what"] = ("batch in host memory: ltpg_kv_submit(LTPG_MEM_HOST) copies it into one of two device slots on a "
                    "copy stream while the previous epoch runs (pipelined: batches submitted back to back; pageable "
                    "numpy arrays or hipHostMalloc'd buffers); serial: each epoch's commit count read back before "
                    "the next submit.  Never the headline value")

    avg_ms = {L.TIMING_NAMES[i]: phase_ms[i] / max(n_timed, 1) for i in range(L.TIMING_SLOTS)}
    bytes_avg = {k: sum(per_batch[j % P][0][k] for j in range(args.steps)) / args.steps for k in KERNELS}
    dom = max(KERNELS, key=lambda k: avg_ms[k])
    achieved = bytes_avg[dom] / (avg_ms[dom] * 1e-3) / 1e9
    kernel_gbs = {k: bytes_avg[k] / (avg_ms[k] * 1e-3) / 1e9 for k in KERNELS if avg_ms[k] > 0}
    # counted HBM bytes exist for the workloads profiled with --pmc
    traffic_src = {("kv-c2", 0.6): "kv_c2_traffic.json", ("kv-y1", 0.0): "kv_y1_traffic.json"}.get(
        (name or args.workload, w["theta"]))
    counted = {k: measured_traffic(k, source=traffic_src, partitioned=bool(shape & L.TUNE_PARTITIONED))
               for k in KERNELS}
    # EXECUTE is bound by random register accesses, not bytes: one probe +
    # atomicMin per write op and one W load per read op, against the rates
    # this device sustains for them over an array of the registers' size
    load_gops, atomic_gops = L.random_access_rates(ctx, rows * 16, 1 << 24)
    def rows_read(meta_):
        k_ = meta_ & 3
        return int((k_ == 1).sum()) + int(((meta_ >> 24).astype(np.int64) + 1)[k_ == 3].sum())

    n_w = sum(int(((gen[j % P][1] & 3) == 2).sum()) for j in range(args.steps)) / args.steps
    n_r = sum(rows_read(gen[j % P][1]) for j in range(args.steps)) / args.steps
    bound_ms = n_w / (atomic_gops * 1e6) + n_r / (load_gops * 1e6)
    # the read-out and write-back copy 100-byte fields between random rows
    # and dense arrays: their ceiling is this device's random-gather rate
    # over an array of the table's size, measured live for as many fields
    # as CHECK reads per epoch
    gather = None
    if 4 * fw <= 128:
        fsw = (fw + 15) // 16 * 16 if fw >= 16 else fw  # the table's field stride (words)
        n_fields = max(1, int(bytes_avg["check"] // (2 * 4 * fw)))
        g = L.random_gather_rate(ctx, rows * fields * 4 * fsw, 4 * fw, 4 * fsw, n_fields)
        gather = {"gbs": g, "fields": n_fields,
                  "check_frac": kernel_gbs.get("check", 0.0) / g if g else None,
                  "writeback_frac": kernel_gbs.get("writeback", 0.0) / g if g else None,
                  "what": "random field copies (table-sized array, field stride of the table) into a dense array, "
                          "read + written bytes per second (ltpg_random_gather_rate); fracs are the phases' "
                          "algorithmic GB/s over it"}
    random_access = {"kernel": "kv_execute", "write_ops": n_w, "read_ops": n_r,
                     "probe_atomic_gops": atomic_gops, "load_gops": load_gops,
                     "bound_ms": bound_ms, "measured_ms": avg_ms["execute"],
                     "frac": bound_ms / avg_ms["execute"] if avg_ms["execute"] > 0 else None,
                     "what": "write ops / probe+atomicMin rate + read ops / load rate, rates measured live "
                             "over an array the size of the registers (ltpg_random_access_rates)"}

    result = {
        "metric": METRIC,
        "value": commits / elapsed,
        "unit": "txn/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": w["desc"], "rows": rows, "fields": fields, "field_bytes": 4 * fw,
                   "batch_txns": B, "ops_per_txn": S, "zipf_theta": w["theta"], "form": "plain per-batch",
                   "inputs": "resident in HBM", "distinct_batches": P,
                   "launch_shape": {"flags": shape, "epoch_ms_by_flags": shape_ms,
                                    "chosen_by": "fastest of the 40 distinct LTPG_TUNE_* shapes on the first batch, "
                                                 "before the timed region"}},
        "abort_rate": 1.0 - commits / max(attempts, 1),
        "attempted_txns": attempts,
        "committed_txns": commits,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "what": "submit -> commit count on host, one synchronous epoch at a time"},
        "pcie_inclusive": pcie,
        "reference_convention": ref_conv,
        "kernel_ms": avg_ms,
        "kernel_algorithmic_gbs": kernel_gbs,
        "kernel_counted_bytes": counted,
        "random_gather_ceiling": gather,
        "kernel_counted_gbs": {k: counted[k] / (avg_ms[k] * 1e-3) / 1e9 for k in KERNELS
                               if counted[k] and avg_ms[k] > 0},
        "kernel_algorithmic_bytes": bytes_avg,
        "roofline": {"kernel": f"kv_{dom}", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": counted[dom],
                     "algorithmic_bytes_per_launch": bytes_avg[dom], "random_access": random_access},
    }
    for d in dev:
        d.free()
    ctx.close()
    if cpu_baseline and not args.no_cpu_baseline:
        log("cpu baseline (oracle)")
        k0, m0, t0_ = gen[0]
        result["cpu_baseline"] = cpu_baseline_kv(w, k0, m0, t0_)
    return result


TPCC_KERNELS = ["cid", "execute", "check", "writeback"]


def tpcc_kernel_bytes(b, no_commit, pay_commit, rules):
    """Algorithmic bytes of each order-entry kernel for one batch: every
    query field, register, table word, read value and flag the kernel has
    to read or write, once.  A u64 register update counts 16 B (8 read + 8
    written); a bitmap bit 1/8 B."""
    n_no, n_pay = b.n_no, b.n_pay
    L_ = np.minimum(b.nf("O_OL_CNT").astype(np.int64), 15)
    cno = np.asarray(no_commit, dtype=bool)
    cpay = np.asarray(pay_commit, dtype=bool)
    by_name = b.pf("isName").astype(bool)
    lines = int(L_.sum())
    lines_c = int(L_[cno].sum())
    nc, pc = int(cno.sum()), int(cpay.sum())
    reg = 16
    cid = n_pay * 12 + int(by_name.sum()) * (12 + 12 * 8 + 4) + n_pay * 4
    execute = (n_no * (8 * 4 + reg * (2 if rules else 1)) + lines * (2 * 4 + 2 * reg) + n_no
               + n_pay * (7 * 4 + 2 * reg) + n_pay)
    check = (n_no + n_pay) * (1 + 1 / 8) + nc * (4 * 4 + 6 * 4 + 51 * 4 + reg) + lines_c * (4 + 3 * 4) \
        + pc * (7 * 4 + 24 * 4 + 24 * 4)
    writeback = (n_no + n_pay) / 8 + nc * (7 * 4 + 8 + 2 * 12) + lines_c * (3 * 4 + 2 * 4 + 2 * 12 + 40) \
        + pc * (8 * 4 + 2 * 8 + 20 + 2 * 8)
    return {"cid": cid, "execute": execute, "check": check, "writeback": writeback}


def device_tpcc_batch(ctx, b):
    """Upload a TpccBatch; (NeworderBatch, PaymentBatch) of device pointers
    and the buffers that hold them."""
    import ltpg_amd as L

    nob, pab = L.DeviceBuffer(ctx, max(b.no.nbytes, 16)), L.DeviceBuffer(ctx, max(b.pay.nbytes, 16))
    nob.upload(b.no)
    pab.upload(b.pay)
    hn = L.neworder_struct(b.no, b.no_cap, b.n_no)
    hp = L.payment_struct(b.pay, b.pay_cap, b.n_pay)
    nb, pb = L.NeworderBatch(), L.PaymentBatch()
    nb.n, nb.line_stride, pb.n = hn.n, hn.line_stride, hp.n
    for name, _ in L.NeworderBatch._fields_[2:]:
        setattr(nb, name, nob.ptr + (getattr(hn, name) - b.no.ctypes.data))
    for name, _ in L.PaymentBatch._fields_[2:]:
        setattr(pb, name, pab.ptr + (getattr(hp, name) - b.pay.ctypes.data))
    return (nb, pb), (nob, pab)


def cpu_baseline_tpcc(W, batches, seed, budget_s=10.0, ros=None):
    """The oracle (single-threaded C replay of the reference's rules, every
    register of every slot, read staging and RMW write-back) on whole batches
    of the same workload over the full-size tables, until about budget_s."""
    from oracle import oracle

    tables = oracle.tpcc_populate(W, seed)
    ora = oracle.TpccOracle(W)
    total_s, commits, txns, epochs = 0.0, 0, 0, 0
    while total_s < budget_s:
        b = batches[epochs % len(batches)]
        ro = ros[epochs % len(ros)] if ros else None
        t0 = time.perf_counter()
        ep = ora.epoch(tables, b, 0, 1, ro=ro)
        total_s += time.perf_counter() - t0
        commits += ep.no_commits + ep.pay_commits + (ep.ro["os_commits"] + ep.ro["sl_commits"] if ro else 0)
        txns += b.n_no + b.n_pay + (ro.n_os + ro.n_sl if ro else 0)
        epochs += 1
    return {"value": commits / total_s, "unit": "txn/s", "cores": 1, "kind": "port",
            "attempted_txn_per_s": txns / total_s,
            "sample": f"oracle (single-threaded C replay) over the full W={W} tables: {epochs} whole "
                      f"batches, {txns} txns in {total_s:.1f} s"}


def run_tpcc(args, cpu_baseline=True, mix=False):
    """C3 (mix=False), or with mix=True the same tables and batch size with
    45 % NewOrder, 43 % Payment, 6 % Order-Status and 6 % Stock-Level (half
    of the read-only transactions on rows the batch's NewOrders write)."""
    import ltpg_amd as L

    W, n, P = 64, 1 << 18, args.batches
    seed = 0x4C545047
    n_ro = n * 6 // 100 if mix else 0
    if mix:
        desc = (f"order-entry mix, plain form: W={W}, 2^18 txns: 45% NewOrder, 43% Payment, 6% Order-Status, "
                f"6% Stock-Level (half on rows the NewOrders write), NURand items, 1% remote supply, "
                f"15% remote customer, 60% by name")
    else:
        desc = (f"order-entry plain form: W={W}, 2^18 txns, 50/50 NewOrder/Payment, NURand items, 1% remote "
                f"supply, 15% remote customer, 60% by name")
    log(f"generating {P} batches: {desc}")
    gen = L.TpccGenerator(warehouses=W, seed=seed, neworder_pct=51 if mix else 50)
    n_rw = n - 2 * n_ro
    batches = [gen.next(n_rw, tid_base=i * n) for i in range(P)]
    ros = [L.readonly_batch(batches[i], W, n_ro, n_ro, seed=seed + i, tid_lo=i * n, tid_hi=i * n + n_rw)
           for i in range(P)] if mix else None
    ctx = L.Context(workload=L.WORKLOAD_TPCC, tpcc_warehouses=W, max_txns=n, values=L.VALUES_RMW)
    ctx.tpcc_populate(seed)
    dev = [device_tpcc_batch(ctx, b) for b in batches]
    rdev = [L.upload_readonly_batch(ctx, r) for r in ros] if mix else []

    def submit(i):
        ctx.tpcc_submit(dev[i][0], mem=L.MEM_DEVICE)
        if mix:
            ctx.tpcc_submit_readonly(rdev[i][0], mem=L.MEM_DEVICE)
    log("order-entry batches resident in HBM")

    per_batch = []
    for i in range(P):
        submit(i)
        ctx.epoch_run()
        ctx.sync()
        nc = np.unpackbits(ctx.commit_bitmap(L.TYPE_NEWORDER, n_txns=batches[i].n_no).view(np.uint8),
                           bitorder="little")[:batches[i].n_no]
        pc = np.unpackbits(ctx.commit_bitmap(L.TYPE_PAYMENT, n_txns=batches[i].n_pay).view(np.uint8),
                           bitorder="little")[:batches[i].n_pay]
        per_batch.append(tpcc_kernel_bytes(batches[i], nc, pc, 0))
        if mix:
            # read-only transactions (in EXECUTE): their fields, one W
            # register per row read, the table word and its read value,
            # the commit bit
            per_batch[-1]["execute"] += n_ro * (6 * 4 + 3 * 8 + 3 * 8 + 1 / 8) + n_ro * (3 * 4 + 8 + 8 + 1 / 8)

    for k in range(args.warmup):
        submit(k % P)
        ctx.epoch_run()
    ctx.sync()
    log(f"warmup done; timing {args.steps} steps")
    types = (L.TYPE_NEWORDER, L.TYPE_PAYMENT) + ((L.TYPE_ORDERSTATUS, L.TYPE_STOCKLEVEL) if mix else ())
    for t in types:
        ctx.commit_totals(t, reset=True)
    ctx.set_timing(True)
    ctx.phase_times()
    t0 = time.perf_counter()
    for k in range(args.steps):
        submit(k % P)
        ctx.epoch_run()
    ctx.sync()
    elapsed = time.perf_counter() - t0
    phase_ms, n_timed = ctx.phase_times()
    ctx.set_timing(False)
    (cn, an), (cp, ap_) = ctx.commit_totals(L.TYPE_NEWORDER), ctx.commit_totals(L.TYPE_PAYMENT)
    c_ro = a_ro = 0
    for t in types[2:]:
        c_, a_ = ctx.commit_totals(t)
        c_ro, a_ro = c_ro + c_, a_ro + a_

    lat = []
    for k in range(max(5, args.steps // 2)):
        t1 = time.perf_counter()
        submit(k % P)
        ctx.epoch_run()
        ctx.commit_count(L.TYPE_PAYMENT)
        lat.append((time.perf_counter() - t1) * 1e3)

    ref_conv = reference_convention(ctx, lambda k: submit(k % P), max(10, args.steps), list(types),
                                    [L.PHASE_CID, L.PHASE_EXECUTE, L.PHASE_CHECK, L.PHASE_WRITEBACK])
    avg_ms = {L.TIMING_NAMES[i]: phase_ms[i] / max(n_timed, 1) for i in range(L.TIMING_SLOTS)}
    bytes_avg = {k: sum(per_batch[j % P][k] for j in range(args.steps)) / args.steps for k in TPCC_KERNELS}
    dom = max(TPCC_KERNELS, key=lambda k: avg_ms[k])
    achieved = bytes_avg[dom] / (avg_ms[dom] * 1e-3) / 1e9
    result = {
        "metric": METRIC, "value": (cn + cp + c_ro) / elapsed, "unit": "txn/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": desc, "warehouses": W, "batch_txns": n, "neworder_pct": 45 if mix else 50,
                   "payment_pct": 43 if mix else 50, "orderstatus_pct": 6 if mix else 0,
                   "stocklevel_pct": 6 if mix else 0,
                   "form": "plain per-batch", "inputs": "resident in HBM", "distinct_batches": P},
        "abort_rate": 1.0 - (cn + cp + c_ro) / max(an + ap_ + a_ro, 1),
        "abort_rate_by_type": {"neworder": 1.0 - cn / max(an, 1), "payment": 1.0 - cp / max(ap_, 1)},
        "committed_txns": cn + cp + c_ro, "attempted_txns": an + ap_ + a_ro,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "what": "submit -> commit count on host, one synchronous epoch at a time"},
        "reference_convention": ref_conv,
        "kernel_ms": avg_ms,
        "kernel_algorithmic_gbs": {k: bytes_avg[k] / (avg_ms[k] * 1e-3) / 1e9 for k in TPCC_KERNELS
                                   if avg_ms.get(k, 0) > 0},
        "kernel_algorithmic_bytes": bytes_avg,
        "roofline": {"kernel": f"tpcc_{dom}", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom, "tpcc"),
                     "algorithmic_bytes_per_launch": bytes_avg[dom]},
    }
    for (_, bufs) in dev + rdev:
        for x in bufs:
            x.free()
    ctx.close()
    if cpu_baseline and not args.no_cpu_baseline:
        log("cpu baseline (order-entry oracle)")
        result["cpu_baseline"] = cpu_baseline_tpcc(W, batches, seed, ros=ros)
    return result


def run_batched_c4(args):
    """Batched form against the plain per-batch loop on the same mixed batch
    list: 256 batches, sizes floor(exp(U[ln 1e3, ln 1e6])) transactions, each
    key-value (1-64 ops, Zipf 0.9, the 10M x 10 x 100 B table) or order-entry
    (W = 256) by a seeded coin.  Commit sets of the two forms must match."""
    import ltpg_amd as L

    seed = 0x4C545047
    nb = args.c4_batches
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(1e3), np.log(1e6), nb))).astype(np.int64)
    is_kv = rng.random(nb) < 0.5
    desc = (f"batched form vs plain loop: {nb} batches, sizes log-uniform in [1e3, 1e6] txns, mixed key-value "
            f"(1-64 ops, Zipf 0.9, 10M x 10 x 100 B) and order-entry (W=256, 50/50 NewOrder/Payment)")
    log(f"C4: {desc}; {int(sizes.sum())} txns ({int(sizes[is_kv].sum())} key-value)")
    rows, fields, fw, S = 10_000_000, 10, 25, 64
    kv = L.Context(workload=L.WORKLOAD_KV, kv_rows=rows, kv_fields=fields, kv_field_bytes=4 * fw,
                   max_txns=int(max(sizes.max(), 1)), kv_max_slots=S, values=L.VALUES_RMW)
    kv.kv_fill_reference()
    tp = L.Context(workload=L.WORKLOAD_TPCC, tpcc_warehouses=256, max_txns=int(max(sizes.max(), 1)),
                   values=L.VALUES_RMW)
    tp.tpcc_populate(seed)
    gen = L.TpccGenerator(warehouses=256, seed=seed)
    descs, keep, kinds = [], [], []
    kv_ops = {"batch_bytes": 0, "valid": 0, "writes": 0, "reads": 0}
    tid = 0
    for i in range(nb):
        n = int(sizes[i])
        if is_kv[i]:
            k, m, t = L.kv_generate(seed=seed + i, rows=rows, fields=fields, n_txns=n, n_slots=S, min_ops=1,
                                    max_ops=S, theta=0.9, read_pct=50, tid_base=tid)
            kinds_ = m & 3
            kv_ops["batch_bytes"] += 4 * n + 12 * S * n
            kv_ops["valid"] += int((kinds_ != 0).sum())
            kv_ops["writes"] += int((kinds_ == 2).sum())
            kv_ops["reads"] += int((kinds_ == 1).sum())
            db = L.DeviceKvBatch(kv, k, m, t, S)
            keep.append(db)
            descs.append(L.kv_desc(db))
            kinds.append("kv")
        else:
            b = gen.next(n, tid_base=tid)
            dev, bufs = L.upload_tpcc_batch(tp, b)
            keep.append(bufs)
            descs.append(L.tpcc_desc(dev))
            kinds.append("tpcc")
        tid += n
        if (i + 1) % 32 == 0:
            log(f"C4: {i + 1}/{nb} batches resident in HBM")
    runner = L.Batched(kv, tp, descs)

    def bitmaps():
        out = []
        for i, kind in enumerate(kinds):
            if kind == "kv":
                out.append(runner.bitmap(i, L.TYPE_KV))
            else:
                out.append(np.concatenate([runner.bitmap(i, L.TYPE_NEWORDER), runner.bitmap(i, L.TYPE_PAYMENT)]))
        return out

    phases = {}

    def timed(mode, label=None):
        # warm-up pass; for the plain loop it also yields the per-phase
        # kernel times (HIP events), kept out of the wall-clock pass
        if label:
            for c in (kv, tp):
                c.set_timing(True)
                c.phase_times()
        runner.run(mode)
        if label:
            phases[label] = {}
            for name, c in (("key_value", kv), ("order_entry", tp)):
                ms, _ = c.phase_times()
                c.set_timing(False)
                phases[label][name] = {L.TIMING_NAMES[i]: float(ms[i]) for i in range(L.TIMING_SLOTS)}
        t0 = time.perf_counter()
        runner.run(mode)
        return time.perf_counter() - t0, runner.times(), bitmaps()

    log("C4: plain per-batch loop (default launch shapes)")
    plain_s, plain_t, plain_bm = timed(L.RUN_PLAIN, "plain_default_shape")
    log("C4: autotuning launch shapes")
    tuning = runner.autotune(reps=3)
    log("C4: plain per-batch loop with the tuned launch shapes")
    tuned_s, _, _ = timed(L.RUN_PLAIN, "plain_tuned_shapes")
    if args.c4_plain_only:  # for kernel traces: the profiler cannot follow graph launches
        runner.close()
        return {"plain_ms": plain_s * 1e3, "plain_tuned_ms": tuned_s * 1e3, "phase_ms_per_pass": phases,
                "autotuner": tuning}
    log("C4: batched form (graphs, two streams, tuned shapes)")
    graph_s, graph_t, graph_bm = timed(L.RUN_GRAPH)
    same = all(np.array_equal(a, b) for a, b in zip(plain_bm, graph_bm))
    commits = int(sum(int(np.unpackbits(b.view(np.uint8)).sum()) for b in graph_bm))
    total = int(sizes.sum())
    # roofline of the dominant kernel, key-value EXECUTE, per pass over the
    # key-value batches (tuned plain leg, HIP events): algorithmic bytes are
    # the batches (tids, keys, metas) and one 8-byte register probe per valid
    # op (a lower bound: register writes and flag bytes left out); its
    # random-access bound is write ops / probe+atomic rate + read ops / load
    # rate over an array the size of the registers
    ex_ms = phases["plain_tuned_shapes"]["key_value"]["execute"]
    ex_bytes = kv_ops["batch_bytes"] + 8 * kv_ops["valid"]
    load_gops, atomic_gops = L.random_access_rates(kv, rows * 16, 1 << 24)
    bound_ms = kv_ops["writes"] / (atomic_gops * 1e6) + kv_ops["reads"] / (load_gops * 1e6)
    ex_gbs = ex_bytes / (ex_ms * 1e-3) / 1e9 if ex_ms > 0 else None
    roof = {"kernel": "kv_execute (all key-value batches of a pass)", "bound": "hbm", "achieved": ex_gbs,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ex_gbs / HBM_PEAK_GBS if ex_gbs else None, "traffic": None,
            "algorithmic_bytes_per_pass": ex_bytes, "ms_per_pass": ex_ms,
            "random_access": {"write_ops": kv_ops["writes"], "read_ops": kv_ops["reads"],
                              "probe_atomic_gops": atomic_gops, "load_gops": load_gops, "bound_ms": bound_ms,
                              "frac": bound_ms / ex_ms if ex_ms > 0 else None}}
    # a persistent kernel would replace each kernel boundary inside an
    # epoch's graph by a grid-wide barrier: both measured live, at the grid
    # sizes of the small size classes (2^10-2^14 transactions) and of the chip
    sync = {str(b): L.sync_costs(kv, b, 1000) for b in (64, 256, 1024)}
    runner.close()
    for x in keep:
        for buf in (x if isinstance(x, list) else [x]):
            buf.free()
    kv.close()
    tp.close()
    return {
        "sync_costs_us": sync,
        "value": commits / graph_s, "unit": "txn/s", "ms_per_step": graph_s * 1e3,
        "plain_value": commits / plain_s, "plain_ms": plain_s * 1e3, "speedup_vs_plain": plain_s / graph_s,
        "plain_tuned_value": commits / tuned_s, "plain_tuned_ms": tuned_s * 1e3,
        "commit_sets_identical": bool(same), "committed_txns": commits, "attempted_txns": total,
        "abort_rate": 1.0 - commits / max(total, 1),
        "latency_ms": {
            "what": "completion time of each batch from the start of the run (all batches arrive at once)",
            "plain_p50": float(np.percentile(plain_t, 50)), "plain_p99": float(np.percentile(plain_t, 99)),
            "batched_p50": float(np.percentile(graph_t, 50)), "batched_p99": float(np.percentile(graph_t, 99))},
        "phase_ms_per_pass": phases,
        "roofline": roof,
        "autotuner": tuning,
        "config": {"workload": desc, "batches": nb, "kv_batches": int(is_kv.sum()), "seed": seed,
                   "form": "batched (hipGraph per workload, concurrent streams, autotuned) vs plain loop",
                   "inputs": "resident in HBM"},
    }


def _c5_rate(args, L, kv, tp, S, rate, n, keys, meta, tb, at, kind, plain, profile=False):
    """One offered rate of the served-form workload through a fresh server
    (plain = 1: the same admission run as the plain per-batch form).
    Requests are submitted in 1 ms windows as their scheduled arrival time
    comes; latency is commit time minus the scheduled arrival."""
    oe_tid = np.r_[tb.nf("TID"), tb.pf("TID")]
    oe_order = np.argsort(oe_tid, kind="stable")  # request k of the batch's own arrival order
    arrival_of = {}  # request id -> arrival index
    if profile:
        os.environ["LTPG_SERVE_PROFILE"] = "1"
    sv = L.Server(kv, tp, kv_ops=S, slots=2, hbm_budget=args.c5_budget_mb << 20, plain=plain,
                  epoch_floor=args.c5_epoch_floor, ring_requests=1 << 20)
    os.environ.pop("LTPG_SERVE_PROFILE", None)
    # every window's submissions prepared before the clock starts
    window = 1e-3
    edges = np.searchsorted(at, np.arange(window, at[-1] + window, window), side="right")
    ki = oi = prev = 0
    plan = []  # (due time, key-value slice, its arrivals, order-entry batch, its arrivals)
    for e in edges:
        if e <= prev:
            continue
        sel = np.arange(prev, e)
        kv_arr = sel[kind[prev:e] == 0]
        oe_arr = sel[kind[prev:e] == 1]
        kv_part = (keys[ki:ki + len(kv_arr)], meta[ki:ki + len(kv_arr)]) if len(kv_arr) else None
        ki += len(kv_arr)
        oe_part = None
        if len(oe_arr):
            part, perm = _tpcc_select(tb, oe_order[oi:oi + len(oe_arr)])
            oe_part, oe_arr = part, oe_arr[perm]
            oi += len(oe_arr)
        plan.append((at[e - 1] * 1e9, kv_part, kv_arr, oe_part, oe_arr))
        prev = e
    kv_ids, oe_ids = [], []
    t0_ns = time.monotonic_ns()
    for due, kv_part, kv_arr, oe_part, oe_arr in plan:
        while time.monotonic_ns() - t0_ns < due:
            pass
        if kv_part is not None:
            kv_ids.append((sv.submit_kv(*kv_part), kv_arr))
        if oe_part is not None:
            oe_ids.append((sv.submit_tpcc(oe_part), oe_arr))
    submit_s = (time.monotonic_ns() - t0_ns) * 1e-9
    done = []
    deadline = time.monotonic() + args.c5_drain_s
    form = "plain" if plain else "served"
    while len(done) < n and time.monotonic() < deadline:
        time.sleep(0.02)
        done += sv.poll(1 << 22)
    log(f"C5 {form}: {rate}M txn/s: {len(done)}/{n} committed")
    st = sv.stats()
    dev = sv.device_stats()
    dropped = sv.cancel()
    for first, arr in kv_ids + oe_ids:
        for j, a_i in enumerate(arr):
            arrival_of[first + j] = a_i
    sched = np.array([t0_ns + at[arrival_of[d["id"]]] * 1e9 for d in done])
    commit = np.array([d["commit_ns"] for d in done], dtype=np.float64)
    kinds = np.array([d["type"] for d in done])
    lat = (commit - sched) * 1e-6
    span_s = (commit.max() - t0_ns) * 1e-9
    att = sum(st["attempts"])
    out = {"offered_txn_per_s": rate * 1e6, "requests": n, "arrival_window_s": float(at[-1]),
           "committed": len(done), "dropped_at_cap": dropped, "completed_fraction": len(done) / n,
           "drain_cap_s": args.c5_drain_s, "committed_txn_per_s": len(done) / span_s,
           "submit_s": submit_s, "offered_effective_txn_per_s": n / max(submit_s, at[-1]), "span_s": span_s,
           "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                          "max": float(lat.max())},
           "latency_ms_kv": {"p50": float(np.percentile(lat[kinds == 0], 50)),
                             "p99": float(np.percentile(lat[kinds == 0], 99))} if (kinds == 0).any() else None,
           "epochs": {"kv": st["epochs"][0], "tpcc": st["epochs"][1]},
           "max_epoch_txns": {"kv": st["max_epoch_txns"][0], "tpcc": st["max_epoch_txns"][1]},
           "attempts_per_commit": att / max(sum(st["committed"]), 1), "attempts_kv": st["attempts"][0],
           "abort_rate": 1.0 - sum(st["committed"]) / max(att, 1),
           "abort_rate_by_type": {name: 1.0 - st["committed"][i] / st["attempts"][i]
                                  for i, name in enumerate(["kv", "neworder", "payment"]) if st["attempts"][i]}}
    if dev["kv_loop"]:
        out["device_loop"] = dev
    return out


def run_kv_small(args):
    """Small epochs in the plain form: the reference's loop (submit, every
    phase, synchronise, read the commit count) over 2,000 epochs of 32
    transactions x 1-64 ops (Zipf 0.99) on the 10M-row x 10 x 100 B table,
    once with the separate phase kernels (the served form's shape for larger
    epochs) and once with TUNE_SMALL_EPOCH (one single-workgroup kernel per
    epoch).  Both runs start from the same table and commit the same
    transactions."""
    import ltpg_amd as L

    rows, fields, fw, B, S, n_ep = 10_000_000, 10, 25, 32, 64, 2000
    keys, meta, tids = L.kv_generate(seed=0x4C545047, rows=rows, fields=fields, n_txns=B * n_ep, n_slots=S,
                                     min_ops=1, max_ops=S, theta=0.99)
    k3, m3, t2 = keys.reshape(S, n_ep, B), meta.reshape(S, n_ep, B), tids.reshape(n_ep, B)
    desc = (f"key-value plain form, small epochs: {n_ep} epochs of {B} txns x 1-{S} ops, Zipf 0.99, "
            f"{rows // 1_000_000}M rows x {fields} x 100 B; a synchronisation and the commit count after each")
    log(f"small epochs: {desc}")
    res = {}
    for name, flags in (("separate_kernels", L.TUNE_EXEC_PER_OP | L.TUNE_SMALL_CHUNKS | L.TUNE_LANE_COPY),
                        ("single_workgroup", L.TUNE_SMALL_EPOCH)):
        ctx = L.Context(workload=L.WORKLOAD_KV, kv_rows=rows, kv_fields=fields, kv_field_bytes=4 * fw, max_txns=B,
                        kv_max_slots=S, values=L.VALUES_RMW)
        ctx.kv_fill_reference()
        ctx.kv_set_tuning(flags)
        batches = [L.DeviceKvBatch(ctx, np.ascontiguousarray(k3[:, e, :]).reshape(-1),
                                   np.ascontiguousarray(m3[:, e, :]).reshape(-1), t2[e], S) for e in range(n_ep)]
        ctx.sync()
        committed, t0 = 0, time.perf_counter()
        for b in batches:
            ctx.kv_submit_device(b)
            ctx.epoch_run(L.PHASE_ALL)
            ctx.sync()
            committed += ctx.commit_count()
        dt = time.perf_counter() - t0
        res[name] = {"ms_per_epoch": dt * 1e3 / n_ep, "committed_txn_per_s": committed / dt, "committed": committed}
        log(f"small epochs, {name}: {dt * 1e6 / n_ep:.1f} us per epoch, {committed / dt / 1e3:.1f}K committed/s")
        for b in batches:
            b.free()
        ctx.close()
    assert res["separate_kernels"]["committed"] == res["single_workgroup"]["committed"]
    v = res["single_workgroup"]
    return {"metric": METRIC, "value": v["committed_txn_per_s"], "unit": "txn/s", "n_gpus": 1, "steps": n_ep,
            "warmup": 0, "ms_per_step": v["ms_per_epoch"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u32", "data": "synthetic", "forms": res,
            "speedup_vs_separate_kernels": v["committed_txn_per_s"] / res["separate_kernels"]["committed_txn_per_s"],
            "config": {"workload": desc}}


def cpu_baseline_c5(keys, meta, S, fields, fw, floor, growth, cap_max, budget_s=10.0, max_requests=200_000):
    """The oracle replaying the served form's key-value lane on one host
    core: the C5 key-value requests, formed into epochs by the device loop's
    rule (the aborts of the last epoch head the next one, arrivals fill it up
    to the cap max(floor, growth x last commits), at most cap_max), each
    epoch one oracle epoch with rank priorities and the arrival deltas.  The
    rows an epoch touches are renumbered into a small table for the call
    (the rules see only row equality; values move in and out of a compact
    copy of the touched rows, which starts as the reference fill); only the
    oracle calls are timed.  Runs until budget_s of oracle time or
    max_requests have committed."""
    from oracle import oracle

    n = min(len(keys), max_requests)
    uniq, inv = np.unique(keys[:n].reshape(-1), return_inverse=True)
    inv = inv.reshape(n, S)
    g = np.empty((len(uniq), fields, fw), dtype=np.uint32)
    g[:] = np.arange(fields, dtype=np.uint32)[None, :, None]
    carry, nxt, last, commits, epochs, t_or = [], 0, None, 0, 0, 0.0
    while (carry or nxt < n) and t_or < budget_s:
        if growth == 0:
            cap = cap_max
        elif last is None:
            cap = min(cap_max, floor)
        else:
            cap = min(cap_max, max(floor, growth * last))
        take = max(0, cap - len(carry))
        members = carry + list(range(nxt, min(n, nxt + take)))
        nxt = min(n, nxt + take)
        B = len(members)
        rows_e, loc = np.unique(inv[members].reshape(-1), return_inverse=True)
        table = np.ascontiguousarray(g[rows_e]).reshape(-1)
        k = loc.reshape(B, S).T.reshape(-1).astype(np.uint64)
        m = meta[members].T.reshape(-1).copy()
        tids = np.arange(B, dtype=np.uint32)
        deltas = ((np.asarray(members, dtype=np.uint64) & 0xFFFFFFFF) | 1).astype(np.uint32)
        t1 = time.perf_counter()
        ep = oracle.kv_epoch(table, len(rows_e), fields, fw, k, m, tids, S, values=1, want_reads=False,
                             inplace=True, deltas=deltas)
        t_or += time.perf_counter() - t1
        g[rows_e] = table.reshape(len(rows_e), fields, fw)
        ok = [ep.committed(t) for t in range(B)]
        carry = [r for r, c in zip(members, ok) if not c]
        last = sum(ok)
        commits += last
        epochs += 1
    return {"value": commits / t_or, "unit": "txn/s", "cores": 1, "kind": "port",
            "sample": f"oracle replay of the served key-value lane (device-loop epoch rule, cap {cap_max}, floor "
                      f"{floor}, growth {growth}) over the first {n} C5 key-value requests: {commits} committed in "
                      f"{epochs} epochs, {t_or:.1f} s of oracle time; order-entry requests not in the sample"}


def host_cpu():
    """Model and core count of the host this runs on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(),
            "usable_cores": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def c5_inputs(L, rate, seconds, rows, fields, S, gen, rng, seed, min_requests=20000):
    """Arrivals of one offered rate: at least min_requests (a longer window
    at low rates), half key-value (1-64 ops, Zipf 0.99), half order-entry."""
    n = max(min_requests, int(rate * 1e6 * seconds))
    n_kv = n // 2
    keys, meta, _ = L.kv_generate(seed=seed + int(rate * 1000), rows=rows, fields=fields, n_txns=n_kv,
                                  n_slots=S, min_ops=1, max_ops=S, theta=0.99)
    keys = keys.reshape(S, n_kv).T.copy()
    meta = meta.reshape(S, n_kv).T.copy()
    tb = gen.next(n - n_kv)
    at = np.cumsum(rng.exponential(1.0 / (rate * 1e6), n))
    kind = rng.permutation(np.r_[np.zeros(n_kv, np.int8), np.ones(n - n_kv, np.int8)])
    return n, keys, meta, tb, at, kind


# The device loop's epoch (at most 64 requests, no kept reads, admitted by
# the epoch before it) as a chain of steps, counted from kv_serve_loop_kernel
# along its longest lane: 3 LDS-only workgroup barriers (after registration,
# marking and requeue); 18 dependent LDS accesses (registration: op words 1;
# marking: a row's minimum and the flag store, for two op slots, 3; CHECK:
# flags and the wave's commit arrays 2; next admission: control words and
# op writes 2; write-back: commit list, then op and sequence number, 3 rounds
# of ~2,900 field words over 960 lanes at C5, 6; requeue: destination, op,
# its write 3; the record's destination 1); 2 dependent LDS atomics (the
# registration's compare-and-swap and minimum); one release publish (wave 0,
# while the other waves issue the write-back).  The ring's PCIe reads are
# prefetched an epoch ahead and the write-back's HBM atomics are never waited
# for, so neither is on the chain.
CHAIN_STEPS = {"barrier_ns": 3, "lds_load_ns": 18, "lds_atomic_ns": 2, "publish_ns": 1}


def chain_bound(lat, epoch_us):
    """Dependent-chain roofline of one device-loop epoch: the measured step
    latencies (ltpg_latency_costs) times the steps the epoch chains."""
    bound_us = sum(lat[k] * c for k, c in CHAIN_STEPS.items()) / 1e3
    return {"steps": CHAIN_STEPS, "step_ns": {k: round(v, 1) for k, v in lat.items()}, "bound_us": bound_us,
            "epoch_us": epoch_us, "frac": bound_us / epoch_us if epoch_us else None,
            "what": "sum of step latencies along the epoch's dependent chain, measured live in one workgroup; "
                    "frac = bound / measured epoch time (the rest is ALU work, bank conflicts and same-address "
                    "serialization inside each step)"}


def run_served_c5(args):
    """Served form under Poisson arrivals: a 50/50 mix of key-value requests
    (1-64 ops, Zipf 0.99, 100M rows x 10 x 100 B) and order-entry requests
    (W = 1024, 50/50 NewOrder / Payment) offered at each rate in --c5-rates,
    the plain per-batch form on the same arrivals at the rates in
    --c5-plain-rates.  Requests are submitted in 1 ms windows as their
    scheduled arrival time comes; latency is commit time minus the scheduled
    arrival (so a submitter that falls behind the schedule counts as
    queueing).  Aborted requests re-queue with their priority.  The drain is
    capped at --c5-drain-s: requests still uncommitted then are dropped and
    counted (the completed fraction), and latency covers the committed ones.

    The headline is the served form at --c5-rate: its arrival window is
    --steps windows of 1 ms (after --warmup windows of arrivals served and
    drained untimed on the same tables), and ms_per_step is the span from
    the first arrival to the last commit divided by steps."""
    import ltpg_amd as L

    seed = 0x4C545047
    rows, fields, fw, S, W = args.c5_rows, 10, 25, 64, 1024
    desc = (f"served form (C5): Poisson arrivals, 50/50 key-value (1-64 ops, Zipf 0.99, {rows // 1_000_000}M rows x "
            f"10 x 100 B) and order-entry (W={W}, 50/50 NewOrder/Payment) requests")
    log(f"C5: {desc}")
    kv = L.Context(workload=L.WORKLOAD_KV, kv_rows=rows, kv_fields=fields, kv_field_bytes=4 * fw,
                   max_txns=1 << 20, kv_max_slots=S, values=L.VALUES_RMW)
    kv.kv_fill_reference()
    # host-formed epochs (the plain form's) of up to 2048 op slots run as one
    # single-workgroup kernel; larger ones per-op registration with row
    # folding, 64-transaction chunks and per-lane field copies
    kv.kv_set_tuning(L.TUNE_EXEC_PER_OP | L.TUNE_SMALL_CHUNKS | L.TUNE_LANE_COPY)
    tp = L.Context(workload=L.WORKLOAD_TPCC, tpcc_warehouses=W, max_txns=1 << 20, values=L.VALUES_RMW)
    tp.tpcc_populate(seed)
    log("C5: tables resident")
    rng = np.random.default_rng(seed)
    gen = L.TpccGenerator(warehouses=W, seed=seed)
    head_rate = args.c5_rate
    if args.warmup > 0:
        n, keys, meta, tb, at, kind = c5_inputs(L, head_rate, args.warmup * 1e-3, rows, fields, S, gen, rng,
                                                seed + 17, min_requests=1000)
        _c5_rate(args, L, kv, tp, S, head_rate, n, keys, meta, tb, at, kind, 0)
        log("C5: warm-up served and drained")
    rates = sorted(set(args.c5_rates) | {head_rate})
    served, plain = {}, {}
    inputs = {}
    for rate in rates:
        secs = args.steps * 1e-3 if rate == head_rate else args.c5_seconds
        inputs[rate] = c5_inputs(L, rate, secs, rows, fields, S, gen, rng, seed,
                                 min_requests=1000 if rate == head_rate else 20000)
        n, keys, meta, tb, at, kind = inputs[rate]
        forms = [("served", served, 0)] + ([("plain", plain, 1)] if rate in args.c5_plain_rates else [])
        for form, dest, pl in forms:
            r = _c5_rate(args, L, kv, tp, S, rate, n, keys, meta, tb, at, kind, pl)
            dest[rate] = r
            log(f"C5 {form}: offered {rate}M txn/s -> {r['committed_txn_per_s'] / 1e6:.3f}M committed/s, "
                f"p50 {r['latency_ms']['p50']:.2f} ms, p99 {r['latency_ms']['p99']:.2f} ms, "
                f"{r['completed_fraction'] * 100:.0f}% in the drain cap, submit {r['submit_s']:.3f} s for "
                f"{r['arrival_window_s']:.3f} s of arrivals")
    # the per-part chain of the device loop's epochs: one profiled pass at the
    # headline rate (its clock stamps add to each epoch, so it is not timed)
    n, keys, meta, tb, at, kind = inputs[head_rate]
    prof = _c5_rate(args, L, kv, tp, S, head_rate, n, keys, meta, tb, at, kind, 0, profile=True)
    # the step latencies the epoch chains (barrier, LDS load / atomic, PCIe
    # round trip, release publish), measured in one 1024-thread workgroup
    lat = L.latency_costs(kv)
    # the reference's ZIPFIAN mode: row 0 as the hot key (its writes reduce
    # commutatively and never conflict), on a fresh table
    hot = {}
    kv.close()
    if args.c5_hot_rates:
        kv = L.Context(workload=L.WORKLOAD_KV, kv_rows=rows, kv_fields=fields, kv_field_bytes=4 * fw,
                       max_txns=1 << 20, kv_max_slots=S, values=L.VALUES_RMW, kv_hot_row=0)
        kv.kv_fill_reference()
        kv.kv_set_tuning(L.TUNE_EXEC_PER_OP | L.TUNE_SMALL_CHUNKS | L.TUNE_LANE_COPY)
        for rate in args.c5_hot_rates:
            n, keys, meta, tb, at, kind = inputs.get(rate) or c5_inputs(L, rate, args.c5_seconds, rows, fields, S,
                                                                        gen, rng, seed)
            hot[rate] = {form: _c5_rate(args, L, kv, tp, S, rate, n, keys, meta, tb, at, kind, pl)
                         for form, pl in (("served", 0), ("plain", 1))}
            log(f"C5 hot row 0 at {rate}M/s: served {hot[rate]['served']['committed_txn_per_s'] / 1e6:.3f}M, "
                f"plain {hot[rate]['plain']['committed_txn_per_s'] / 1e6:.3f}M committed/s")
        kv.close()
    tp.close()
    h = served[head_rate]
    # roofline of the dominant kernel, the device loop, per epoch: HBM bytes
    # are the committed writes' read-modify-writes of their fields (every
    # request commits once; a request's write ops are its WRITE slots)
    n, keys, meta, tb, at, kind = inputs[head_rate]
    writes = int(((meta & 3) == 2).sum())
    dl = h.get("device_loop", {})
    epochs = max(dl.get("epochs", 0), 1)
    epoch_us = dl.get("busy_ns", 0) / epochs / 1e3
    hbm_per_epoch = writes * 2 * 4 * fw / epochs
    achieved = hbm_per_epoch / (epoch_us * 1e-6) / 1e9 if epoch_us > 0 else None
    pdl = prof.get("device_loop", {})
    parts = {k: v / max(pdl.get("profile_epochs", 0), 1) / 1e3 for k, v in pdl.get("profile_ns", {}).items()
             if k != "unused"}
    roof = {"kernel": "kv_serve_loop_kernel", "unit_of_work": "one epoch", "bound": "hbm",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": None,
            "algorithmic_bytes_per_epoch": hbm_per_epoch,
            # request ring reads (12 B per padded op slot) and done-ring records
            "pcie_bytes_per_epoch": (12 * S * (n // 2) + 8 * (epochs + h["attempts_kv"])) / epochs,
            "epoch_us": epoch_us, "epochs": dl.get("epochs", 0), "kernel_busy_ms": dl.get("busy_ns", 0) / 1e6,
            "latency_chain": {"what": "the epoch is a dependent chain of LDS phases separated by workgroup barriers "
                                      "plus one PCIe-published record; per-part device time of a profiled pass "
                                      "(thread 0's clock stamps; the profiled epoch is longer than the plain one "
                                      "by the stamps' own cost)", "parts_us": parts,
                              "profiled_epoch_us": sum(parts.values()) if parts else None},
            "chain_bound": chain_bound(lat, epoch_us)}
    res = {"value": h["committed_txn_per_s"], "unit": "txn/s", "span_s": h["span_s"], "served": served,
           "plain": plain, "profiled_pass": prof, "hot_row_0": hot, "roofline": roof,
           "config": {"workload": desc, "offered_txn_per_s": head_rate * 1e6, "rates_mtxn_s": rates,
                      "plain_rates_mtxn_s": sorted(args.c5_plain_rates), "hbm_budget_mb_per_slot": args.c5_budget_mb,
                      "kv_form": "device loop: one persistent single-workgroup kernel forms and runs the key-value "
                                 "epochs (requests through a pinned ring, aborts kept on the device)",
                      "tpcc_form": "host-formed epochs, 2 in flight", "epoch_growth": 8,
                      "epoch_floor": args.c5_epoch_floor, "inputs": "arrive over time; tables resident in HBM"}}
    if not args.no_cpu_baseline:
        log("C5: cpu baseline (oracle replay of the served key-value lane)")
        floor = args.c5_epoch_floor
        res["cpu_baseline"] = cpu_baseline_c5(keys, meta, S, fields, fw, floor, 8, 2048 // S)
        res["cpu_baseline"]["host"] = host_cpu()
    return res


def _tpcc_select(b, idx):
    """Order-entry requests idx (indices into NewOrders then Payments of b)
    as a new TpccBatch, and the permutation of idx in the order the server
    numbers them (the batch's NewOrders first, then its Payments)."""
    import ltpg_amd as L

    idx = np.asarray(idx)
    is_no = idx < b.n_no
    nos, pays = idx[is_no], idx[~is_no] - b.n_no
    out = L.TpccBatch.empty(len(nos), len(pays))
    for f in ("W_ID", "D_ID", "C_ID", "O_ID", "N_O_ID", "O_OL_CNT", "O_OL_ID", "TID"):
        out.nf(f)[:] = b.nf(f)[nos]
    for f in ("OL_I_ID", "OL_SUPPLY_W_ID", "OL_QUANTITY"):
        out.nf(f)[:, :] = b.nf(f)[:, nos]
    for f in ("W_ID", "D_ID", "C_ID", "C_LAST", "isName", "C_D_ID", "C_W_ID", "H_AMOUNT", "H_ID", "TID"):
        out.pf(f)[:] = b.pf(f)[pays]
    perm = np.r_[np.nonzero(is_no)[0], np.nonzero(~is_no)[0]]
    return out, perm


def measured_traffic(phase, family="kv", source=None, partitioned=False):
    """HBM bytes per launch of a kernel from a committed rocprofv3 --pmc
    summary of the same workload (profiles/<source>), or None when none
    exists for it.  The row-partitioned EXECUTE has its own entry
    (<family>_execute_partitioned)."""
    if source is None:
        source = "tpcc_c3_traffic.json" if family == "tpcc" else None
    if source is None:
        return None
    key = f"{family}_{phase}" + ("_partitioned" if partitioned and phase == "execute" else "")
    try:
        with open(os.path.join(ROOT, "profiles", source)) as f:
            return json.load(f).get(key)
    except (OSError, ValueError):
        return None


def copy_calibration(gib):
    """Copy `gib` GiB device -> device five times with the library's copy
    kernel: the known-size traffic that calibrates FETCH_SIZE / WRITE_SIZE."""
    import ltpg_amd as L

    ctx = L.Context(workload=L.WORKLOAD_KV, kv_rows=64, kv_fields=1, kv_field_bytes=4, max_txns=64,
                    kv_max_slots=1)
    n = int(gib * (1 << 30))
    a, b = L.DeviceBuffer(ctx, n), L.DeviceBuffer(ctx, n)
    t0 = time.perf_counter()
    for _ in range(5):
        b.copy_from(a)
    ctx.sync()
    dt = time.perf_counter() - t0
    print(json.dumps({"copy_bytes_per_launch": n, "launches": 5, "gbs": 5 * 2 * n / dt / 1e9}))
    a.free()
    b.free()
    ctx.close()


ROOF_KEYS = ("kernel", "bound", "achieved", "peak", "unit", "frac", "traffic")


def compact(v, name):
    """The few figures of a variant that go on the bench line."""
    out = {"workload": name, "value": v.get("value"), "unit": v.get("unit", "txn/s")}
    for k in ("ms_per_step", "abort_rate", "speedup_vs_plain", "plain_value", "commit_sets_identical"):
        if k in v:
            out[k] = v[k]
    if "latency_ms" in v:
        out["latency_ms"] = {k: v["latency_ms"][k] for k in v["latency_ms"] if k != "what"}
    shape = v.get("config", {}).get("launch_shape")
    if shape:
        out["launch_shape_flags"] = shape["flags"]
    if "roofline" in v:
        out["roofline"] = {k: v["roofline"].get(k) for k in ROOF_KEYS}
        ra = v["roofline"].get("random_access")
        if ra:
            out["roofline"]["random_access_frac"] = ra.get("frac")
    if "kernel_algorithmic_gbs" in v:
        out["phase_hbm_frac"] = {k: round(g / HBM_PEAK_GBS, 4) for k, g in v["kernel_algorithmic_gbs"].items()}
    if v.get("random_gather_ceiling"):
        out["random_gather_ceiling_gbs"] = round(v["random_gather_ceiling"]["gbs"], 1)
    if v.get("kernel_counted_gbs"):
        out["phase_counted_hbm_frac"] = {k: round(g / HBM_PEAK_GBS, 4) for k, g in v["kernel_counted_gbs"].items()}
    if "sync_costs_us" in v:
        out["sync_costs_us"] = {b: {k: round(x, 2) for k, x in c.items()} for b, c in v["sync_costs_us"].items()}
    if "reference_convention" in v:
        out["reference_convention_txn_per_s"] = v["reference_convention"]["value"]
    if "cpu_baseline" in v:
        out["cpu_baseline"] = {k: v["cpu_baseline"][k] for k in ("value", "unit", "cores", "kind")}
    return out


def write_detail(path, detail):
    if not path:
        return None
    try:
        os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
        with open(path, "w") as f:
            json.dump(detail, f, indent=1, default=float)
        return path
    except OSError as e:
        log(f"detail file not written: {e}")
        return None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="all",
                    choices=["all", "served-c5", "kv-c2", "kv-y1", "kv-c1", "kv-scan", "kv-small", "tpcc-c3", "tpcc-mix",
                             "batched-c4"])
    ap.add_argument("--c5-rate", type=float, default=5.0,
                    help="offered load of the headline served run, M txn/s (its window is --steps ms)")
    ap.add_argument("--c5-rates", type=float, nargs="+", default=[0.02, 0.1, 1, 5, 25, 50],
                    help="offered loads of the served form, M txn/s")
    ap.add_argument("--c5-plain-rates", type=float, nargs="*", default=[0.02, 0.1, 1, 5],
                    help="offered loads (M txn/s) also run in the plain per-batch form")
    ap.add_argument("--c5-seconds", type=float, default=0.02, help="arrival window per offered load")
    ap.add_argument("--c5-drain-s", type=float, default=8.0, help="cap on the drain after the arrivals, per load")
    ap.add_argument("--c5-hot-rates", type=float, nargs="*", default=[5],
                    help="offered rates (M txn/s) also run with row 0 as the hot key (the reference's ZIPFIAN mode)")
    ap.add_argument("--c5-rows", type=int, default=100_000_000, help="key-value rows of the served form")
    ap.add_argument("--c5-budget-mb", type=int, default=4096, help="HBM budget per in-flight epoch, MiB")
    ap.add_argument("--c5-epoch-floor", type=int, default=16,
                    help="smallest epoch the commit-based epoch cap allows (ltpg_serve_config.epoch_floor)")
    ap.add_argument("--c4-batches", type=int, default=256, help="batches of the batched-form comparison")
    ap.add_argument("--c4-plain-only", action="store_true",
                    help="C4: run the plain legs only (kernel traces of the plain loop)")
    ap.add_argument("--theta", type=float, default=None, help="Zipf theta of kv-c2 (default 0.6)")
    ap.add_argument("--batches", type=int, default=2, help="distinct device-resident batches cycled")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--default-shape", action="store_true",
                    help="key-value plain runs keep the default launch shape instead of the fastest distinct one")
    ap.add_argument("--detail", default=os.path.join("gpurun_out", "bench_detail.json"),
                    help="file for every figure behind the line (empty: none)")
    ap.add_argument("--copy-calibration", type=float, default=None, metavar="GIB",
                    help="only run the known-size copy used to calibrate traffic counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the sharding design is written up "
              "beside the kernels' description).", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1:
        raise SystemExit("--steps must be >= 1")
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    base = {"metric": METRIC, "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u32", "data": "synthetic"}
    detail = {}
    if args.workload in ("all", "served-c5"):
        c5 = run_served_c5(args)
        detail["served_c5"] = c5
        h = c5["served"][args.c5_rate]
        p = c5["plain"].get(args.c5_rate)
        low = min(c5["served"])
        result = dict(base, value=c5["value"], unit="txn/s", steps=args.steps, warmup=args.warmup,
                      ms_per_step=c5["span_s"] * 1e3 / args.steps)
        result["config"] = dict(c5["config"], timed="first arrival to last commit of --steps 1 ms arrival windows")
        result["abort_rate"] = h["abort_rate"]
        result["abort_rate_by_type"] = h["abort_rate_by_type"]
        result["attempts_per_commit"] = h["attempts_per_commit"]
        result["latency_ms"] = {k: h["latency_ms"][k] for k in ("p50", "p99")}
        if p:
            result["plain_form"] = {"value": p["committed_txn_per_s"],
                                    "latency_ms": {k: p["latency_ms"][k] for k in ("p50", "p99")},
                                    "completed_fraction": p["completed_fraction"]}
            result["speedup_vs_plain"] = c5["value"] / max(p["committed_txn_per_s"], 1e-9)
        result["below_capacity"] = {
            "offered_txn_per_s": low * 1e6, "served_txn_per_s": c5["served"][low]["committed_txn_per_s"],
            "served_latency_ms": {k: c5["served"][low]["latency_ms"][k] for k in ("p50", "p99")},
            **({"plain_txn_per_s": c5["plain"][low]["committed_txn_per_s"],
                "plain_latency_ms": {k: c5["plain"][low]["latency_ms"][k] for k in ("p50", "p99")}}
               if low in c5["plain"] else {})}
        result["by_rate"] = [{"offered": r, "served_txn_per_s": round(c5["served"][r]["committed_txn_per_s"]),
                              "served_p99_ms": round(c5["served"][r]["latency_ms"]["p99"], 3),
                              **({"plain_txn_per_s": round(c5["plain"][r]["committed_txn_per_s"]),
                                  "plain_p99_ms": round(c5["plain"][r]["latency_ms"]["p99"], 3),
                                  "plain_completed": round(c5["plain"][r]["completed_fraction"], 3)}
                                 if r in c5["plain"] else {})} for r in sorted(c5["served"])]
        result["roofline"] = {k: c5["roofline"].get(k) for k in ROOF_KEYS}
        result["roofline"].update({"unit_of_work": "epoch", "epoch_us": c5["roofline"]["epoch_us"],
                                   "epochs": c5["roofline"]["epochs"],
                                   "chain_parts_us": c5["roofline"]["latency_chain"]["parts_us"],
                                   "chain_bound_us": c5["roofline"]["chain_bound"]["bound_us"],
                                   "chain_frac": c5["roofline"]["chain_bound"]["frac"],
                                   "chain_step_ns": c5["roofline"]["chain_bound"]["step_ns"]})
        if "cpu_baseline" in c5:
            result["cpu_baseline"] = c5["cpu_baseline"]
        variants = []
        if args.workload == "all":
            for th in (0.6, 0.99):
                v = run_kv(args, th, cpu_baseline=th == 0.6, name="kv-c2")
                detail[f"kv_c2_theta{th}"] = v
                variants.append(compact(v, f"C2 key-value plain form, Zipf {th}"))
            v = run_kv(args, None, cpu_baseline=False, name="kv-y1")
            detail["kv_y1_mix"] = v
            variants.append(compact(v, "key-value plain form, YCSB/1 mix (9 reads + 1 RMW, uniform), low contention"))
            v = run_tpcc(args)
            detail["tpcc_c3"] = v
            variants.append(compact(v, "C3 order-entry plain form"))
            v = run_batched_c4(args)
            detail["batched_c4"] = v
            variants.append(compact(v, "C4 batched form vs plain loop"))
            result["variants"] = variants
    else:
        if args.workload == "tpcc-c3":
            v = run_tpcc(args)
        elif args.workload == "tpcc-mix":
            v = run_tpcc(args, mix=True)
        elif args.workload == "kv-small":
            v = run_kv_small(args)
        elif args.workload == "batched-c4":
            v = dict(base, steps=1, warmup=1, **run_batched_c4(args))
        else:
            v = run_kv(args, args.theta if args.workload == "kv-c2" else None)
        detail[args.workload] = v
        result = dict(base)
        result.update({k: v[k] for k in v if k in ("value", "unit", "steps", "warmup", "ms_per_step", "config",
                                                   "abort_rate", "latency_ms", "roofline", "cpu_baseline",
                                                   "speedup_vs_plain", "plain_value", "commit_sets_identical")})
    where = write_detail(args.detail, detail)
    result["detail"] = where
    print(json.dumps(result, default=float))


if __name__ == "__main__":
    main()
