"""A compiled C++ caller of include/ltpg.h (tests/abi/integration_loops.cpp),
built with g++ and linked against libltpg.so as an application would link
it.  The CPU test builds it and runs its host-only calls; the GPU test runs
the key-value loop (YCSB/0/ycsb.cu:78-152), the order-entry loop
(TPCC-allopt-final/tp.cu:136-261) and a served segment, and replays every
epoch it dumped on the oracle.  Struct layouts are therefore exercised as a
C++ compiler lays them out, not as ctypes declares them."""
import os
import subprocess

import numpy as np
import pytest

import ltpg_amd as L
from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "ltpg_amd")
SRC = os.path.join(ROOT, "tests", "abi", "integration_loops.cpp")

# the driver's workload sizes (integration_loops.cpp)
KV_ROWS, KV_FIELDS, KV_S, KV_B, KV_EPOCHS = 50000, 10, 8, 6000, 3
TPCC_W, TPCC_EPOCHS, SEED = 2, 3, 0x4C545047
SV_ROWS, SV_FIELDS, SV_N, SV_OPS = 4096, 4, 2000, 4


def _build(tmp_path):
    assert os.path.exists(os.path.join(LIBDIR, "libltpg.so")), "libltpg.so not built"
    exe = str(tmp_path / "integration_loops")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"), SRC,
           "-L" + LIBDIR, "-lltpg", "-Wl,-rpath," + LIBDIR, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_compiled_caller_builds_and_runs_host_calls(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe, "--host-only"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host-only ok" in r.stdout


def _load(d, name, dtype):
    return np.fromfile(os.path.join(d, name), dtype=dtype)


@pytest.mark.gpu
def test_compiled_caller_loops_match_oracle(gpu, tmp_path):
    exe = _build(tmp_path)
    out = tmp_path / "dump"
    out.mkdir()
    print(f"running {exe}", flush=True)
    r = subprocess.run([exe, str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    d = str(out)

    # key-value loop: column-major snapshot -> the oracle's row-major table
    table = np.ascontiguousarray(_load(d, "kv_init.u32", np.uint32).reshape(KV_FIELDS, KV_ROWS).T)
    total = 0
    for e in range(KV_EPOCHS):
        keys = _load(d, f"kv_keys{e}.u64", np.uint64)
        meta = _load(d, f"kv_meta{e}.u32", np.uint32)
        tids = _load(d, f"kv_tids{e}.u32", np.uint32)
        ep = oracle.kv_epoch(table, KV_ROWS, KV_FIELDS, 1, keys, meta, tids, KV_S, values=L.VALUES_RMW, inplace=True)
        bitmap = _load(d, f"kv_bitmap{e}.u64", np.uint64)
        np.testing.assert_array_equal(bitmap, ep.bitmap)
        assert int(_load(d, f"kv_commits{e}.u32", np.uint32)[0]) == ep.commit_count
        assert 0 < ep.commit_count < KV_B
        committed = np.unpackbits(ep.bitmap.view(np.uint8), bitorder="little")[:KV_B].astype(bool)
        mask = (((meta.reshape(KV_S, KV_B) & 3) == L.KV_READ) & committed[None, :]).reshape(-1)
        reads = _load(d, f"kv_reads{e}.u32", np.uint32)
        np.testing.assert_array_equal(reads[mask], ep.read_values[mask])
        total += ep.commit_count
    assert list(_load(d, "kv_totals.u64", np.uint64)) == [total, KV_EPOCHS * KV_B]
    final = _load(d, "kv_final.u32", np.uint32).reshape(KV_FIELDS, KV_ROWS).T
    np.testing.assert_array_equal(final, table.reshape(KV_ROWS, KV_FIELDS))

    # order-entry loop
    tables = oracle.tpcc_populate(TPCC_W, SEED)
    ora = oracle.TpccOracle(TPCC_W)
    for e in range(TPCC_EPOCHS):
        n_no, n_pay = (int(x) for x in _load(d, f"tpcc_n{e}.u32", np.uint32))
        no, pay = _load(d, f"no_block{e}.u32", np.uint32), _load(d, f"pay_block{e}.u32", np.uint32)
        cap = no.size // 53
        ep = ora.epoch(tables, L.TpccBatch(no, cap, n_no, pay, pay.size // 10, n_pay), rules=L.RULES_REFERENCE,
                       values=L.VALUES_RMW)
        assert list(_load(d, f"tpcc_commits{e}.u32", np.uint32)) == [ep.no_commits, ep.pay_commits]
        np.testing.assert_array_equal(_load(d, f"no_bitmap{e}.u64", np.uint64), ep.no_bitmap)
        np.testing.assert_array_equal(_load(d, f"pay_bitmap{e}.u64", np.uint64), ep.pay_bitmap)
    for t in range(9):
        np.testing.assert_array_equal(_load(d, f"tpcc_table{t}.u32", np.uint32), tables[t], err_msg=f"table {t}")
    ora.close()

    # served segment: every request committed once, so each written word
    # holds the sum of its writers' deltas ((u32)seq | 1), seq = id here
    rec = _load(d, "serve_done.u64", np.uint64).reshape(-1, 4)
    assert sorted(rec[:, 0].tolist()) == list(range(SV_N))
    np.testing.assert_array_equal(rec[:, 1], rec[:, 0])
    keys = _load(d, "serve_keys.u64", np.uint64).reshape(SV_N, SV_OPS)
    meta = _load(d, "serve_meta.u32", np.uint32).reshape(SV_N, SV_OPS)
    want = np.zeros((SV_ROWS, SV_FIELDS), np.uint64)
    for i in range(SV_N):
        for k in range(SV_OPS):
            if meta[i, k] & 3 == L.KV_WRITE:
                want[int(keys[i, k]), int(meta[i, k] >> 8)] += (i & 0xFFFFFFFF) | 1
    got = _load(d, "serve_final.u32", np.uint32).reshape(SV_ROWS, SV_FIELDS)
    np.testing.assert_array_equal(got, (want & 0xFFFFFFFF).astype(np.uint32))
    assert int(rec[:, 3].max()) > 1, "the 64-row segment should re-queue some requests"
