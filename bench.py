"""Benchmark of the deterministic concurrency-control hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload kv-c2] [--theta 0.99]

A step is one epoch of the plain per-batch form (EXECUTE, CHECK+COMMIT,
WRITEBACK) over one batch whose inputs are already resident in HBM.  The
default workload is the key-value configuration the metric is quoted on:
10M rows x 10 fields x 100 B (row-major 1,000 B records), batches of 2^20
transactions x 16 ops, 50/50 read / read-modify-write, Zipf theta 0.99.

Prints one JSON line: committed txns/s (whole job), abort rate, p50/p99
commit latency of the reference-style synchronous loop, per-phase device
time, the roofline of the dominant kernel and the CPU baseline (the oracle
timed on this host's cores on a bounded sample of the same workload).
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


def log(msg):
    print(f"[bench {time.perf_counter() - T_START:7.1f}s] {msg}", file=sys.stderr, flush=True)


def kv_workload(name, theta):
    if name == "kv-c2":
        theta = 0.6 if theta is None else theta
        return dict(rows=10_000_000, fields=10, fw=25, B=1 << 20, S=16, theta=theta, read_pct=50,
                    desc=f"key-value plain form: 10M rows x 10 fields x 100 B, 2^20 txns x 16 ops, "
                         f"50/50 read/RMW, Zipf {theta}")
    if name == "kv-c1":
        return dict(rows=1_000_000, fields=10, fw=1, B=65536, S=10, theta=0.8, read_pct=50,
                    desc="key-value plain form: 1M rows x 10 u32 fields, 65,536 txns x 10 ops, "
                         "50/50 read/RMW, Zipf 0.8")
    raise SystemExit(f"unknown workload {name}")


KERNELS = ["execute", "check", "readout", "writeback"]


def kv_kernel_bytes(keys, meta, tids, bitmap, B, S, fw):
    """Algorithmic bytes each key-value kernel must move for one batch: every
    byte the algorithm has to read or write, once ('Algorithmic
    bytes').  The fast CHECK's early exit is simulated exactly: a surviving
    transaction visits slots until it has both a raw and a war conflict."""
    F = 4 * fw
    kind = (meta & 3).reshape(S, B)
    rows_ = keys.reshape(S, B).astype(np.int64)
    is_r, is_w = kind == 1, kind == 2
    prio = (tids.astype(np.int64) << 24) | np.arange(B, dtype=np.int64)
    big = np.iinfo(np.int64).max
    nrows = int(rows_.max()) + 1 if rows_.size else 1
    Rmin = np.full(nrows, big, dtype=np.int64)
    Wmin = np.full(nrows, big, dtype=np.int64)
    pr = np.broadcast_to(prio, (S, B))
    np.minimum.at(Rmin, rows_[is_r], pr[is_r])
    np.minimum.at(Wmin, rows_[is_w], pr[is_w])
    waw = (is_w & (Wmin[rows_] < pr)).any(axis=0)
    raw_op = is_r & (Wmin[rows_] < pr)
    war_op = is_w & (Rmin[rows_] < pr)
    done = np.logical_or.accumulate(raw_op, axis=0) & np.logical_or.accumulate(war_op, axis=0)
    before_done = np.vstack([np.zeros((1, B), bool), done[:-1]])  # slot s visited iff not done before s
    visited = (~before_done) & ~waw[None, :]
    bits = np.unpackbits(bitmap.view(np.uint8), bitorder="little")[:B].astype(bool)
    committed = int(bits.sum())
    valid = int((is_r | is_w).sum())
    regs_written = int((Rmin < big).sum() + (Wmin < big).sum())
    execute = 4 * B + 12 * S * B + 8 * valid + 8 * regs_written + int(waw.sum())
    check = 5 * B + B // 8 + 20 * int((visited & (is_r | is_w)).sum())
    readout = B // 8 + committed * 12 * S + int((is_r & bits[None, :]).sum()) * 2 * F
    writeback = B // 8 + committed * (4 + 12 * S) + int((is_w & bits[None, :]).sum()) * 2 * F
    return {"execute": execute, "check": check, "readout": readout, "writeback": writeback}, committed


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
        ep = oracle.kv_epoch(table, rows, fields, fw, k, m, t, S, values=1, want_reads=True, inplace=True)
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


def run_kv(args, theta, cpu_baseline=True):
    import ltpg_amd as L

    w = kv_workload(args.workload, theta)
    rows, fields, fw, B, S = w["rows"], w["fields"], w["fw"], w["B"], w["S"]
    P = args.batches
    log(f"generating {P} batches: {w['desc']}")
    gen = [L.kv_generate(seed=0x4C545047 + i, rows=rows, fields=fields, n_txns=B, n_slots=S,
                         read_pct=w["read_pct"], theta=w["theta"], tid_base=0) for i in range(P)]
    ctx = L.Context(workload=L.WORKLOAD_KV, kv_rows=rows, kv_fields=fields, kv_field_bytes=4 * fw,
                    max_txns=B, kv_max_slots=S, values=L.VALUES_RMW)
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
        per_batch.append(kv_kernel_bytes(gen[i][0], gen[i][1], gen[i][2], ctx.commit_bitmap(), B, S, fw))

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

    avg_ms = {L.TIMING_NAMES[i]: phase_ms[i] / max(n_timed, 1) for i in range(L.TIMING_SLOTS)}
    bytes_avg = {k: sum(per_batch[j % P][0][k] for j in range(args.steps)) / args.steps for k in KERNELS}
    dom = max(KERNELS, key=lambda k: avg_ms[k])
    achieved = bytes_avg[dom] / (avg_ms[dom] * 1e-3) / 1e9
    kernel_gbs = {k: bytes_avg[k] / (avg_ms[k] * 1e-3) / 1e9 for k in KERNELS if avg_ms[k] > 0}

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
                   "inputs": "resident in HBM", "distinct_batches": P},
        "abort_rate": 1.0 - commits / max(attempts, 1),
        "attempted_txns": attempts,
        "committed_txns": commits,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "what": "submit -> commit count on host, one synchronous epoch at a time"},
        "kernel_ms": avg_ms,
        "kernel_algorithmic_gbs": kernel_gbs,
        "kernel_algorithmic_bytes": bytes_avg,
        "roofline": {"kernel": f"kv_{dom}", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom),
                     "algorithmic_bytes_per_launch": bytes_avg[dom]},
    }
    for d in dev:
        d.free()
    ctx.close()
    if cpu_baseline and not args.no_cpu_baseline:
        log("cpu baseline (oracle)")
        k0, m0, t0_ = gen[0]
        result["cpu_baseline"] = cpu_baseline_kv(w, k0, m0, t0_)
    return result


def measured_traffic(phase):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if
    one exists for this kernel (profiles/kv_c2_traffic.json)."""
    p = os.path.join(ROOT, "profiles", "kv_c2_traffic.json")
    try:
        with open(p) as f:
            return json.load(f).get(f"kv_{phase}")
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


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="kv-c2")
    ap.add_argument("--theta", type=float, default=None,
                    help="one Zipf theta (default for kv-c2: 0.6 headline plus 0.99 variant)")
    ap.add_argument("--batches", type=int, default=2, help="distinct device-resident batches cycled")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=float, default=None, metavar="GIB",
                    help="only run the known-size copy used to calibrate traffic counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (see DESIGN.md, 'Sharding').",
              file=sys.stderr)
        sys.exit(2)
    if args.steps < 1:
        raise SystemExit("--steps must be >= 1")
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    thetas = [args.theta] if args.theta is not None else ([0.6, 0.99] if args.workload == "kv-c2" else [None])
    result = run_kv(args, thetas[0])
    variants = []
    for th in thetas[1:]:
        v = run_kv(args, th, cpu_baseline=False)
        variants.append({k: v[k] for k in ("value", "unit", "ms_per_step", "abort_rate", "latency_ms",
                                           "kernel_ms", "kernel_algorithmic_gbs", "roofline")}
                        | {"config": v["config"]})
    if variants:
        result["variants"] = variants
    print(json.dumps(result))


if __name__ == "__main__":
    main()
