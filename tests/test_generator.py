"""CPU checks of the seeded key-value generator (host code in libltpg)."""
import numpy as np
import pytest

import ltpg_amd as L


def test_same_seed_same_batch_and_thread_count_invariance():
    a = L.kv_generate(seed=5, rows=10_000, n_txns=9000, n_slots=12, theta=0.9, threads=1)
    b = L.kv_generate(seed=5, rows=10_000, n_txns=9000, n_slots=12, theta=0.9, threads=7)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    c = L.kv_generate(seed=6, rows=10_000, n_txns=9000, n_slots=12, theta=0.9)
    assert not np.array_equal(a[0], c[0])


def test_counter_based_stream_offsets():
    """Transaction t of a batch starting at txn_base depends only on
    (seed, txn_base + t, slot)."""
    B, S = 300, 5
    k, m, _ = L.kv_generate(seed=1, rows=1000, n_txns=B, n_slots=S, theta=0.7)
    k2, m2, t2 = L.kv_generate(seed=1, rows=1000, n_txns=100, n_slots=S, theta=0.7, txn_base=200,
                               tid_base=200)
    k = k.reshape(S, B)[:, 200:]
    m = m.reshape(S, B)[:, 200:]
    np.testing.assert_array_equal(k, k2.reshape(S, 100))
    np.testing.assert_array_equal(m, m2.reshape(S, 100))
    np.testing.assert_array_equal(t2, np.arange(200, 300))


def test_shape_kinds_fields_and_distinct_keys():
    B, S, rows, fields = 20_000, 16, 5000, 10
    keys, meta, tids = L.kv_generate(seed=2, rows=rows, fields=fields, n_txns=B, n_slots=S, theta=0.99,
                                     read_pct=50)
    assert keys.max() < rows
    kinds = meta & 3
    assert set(np.unique(kinds)) == {1, 2}
    assert abs((kinds == 1).mean() - 0.5) < 0.01
    fld = (meta >> 8) & 0xFFFF
    assert fld.max() == fields - 1 and fld.min() == 0
    per_txn = np.sort(keys.reshape(S, B).T, axis=1)
    assert (np.diff(per_txn.astype(np.int64), axis=1) != 0).all()
    np.testing.assert_array_equal(tids, np.arange(B))


def test_ragged_ops_use_empty_slots():
    B, S = 5000, 64
    keys, meta, _ = L.kv_generate(seed=3, rows=100_000, n_txns=B, n_slots=S, min_ops=1, max_ops=64,
                                  theta=0.9)
    counts = ((meta.reshape(S, B) & 3) != 0).sum(axis=0)
    assert counts.min() == 1 and counts.max() == 64
    assert abs(counts.mean() - 32.5) < 1.0
    # empty slots come after the used ones
    used = (meta.reshape(S, B) & 3) != 0
    assert (used == (np.arange(S)[:, None] < counts[None, :])).all()


@pytest.mark.parametrize("theta", [0.6, 0.8, 0.99])
def test_zipf_head_matches_closed_form(theta):
    """Gray et al.'s method gives rank 0 (row 0) with probability 1/zeta(n)
    and rank 1 with probability 2^-theta/zeta(n)."""
    rows, B, S = 1000, 50_000, 4
    keys, _, _ = L.kv_generate(seed=11, rows=rows, n_txns=B, n_slots=S, theta=theta, distinct_keys=False)
    zeta = (np.arange(1, rows + 1, dtype=np.float64) ** -theta).sum()
    n = keys.size
    p0 = 1.0 / zeta
    f0 = (keys == 0).mean()
    assert abs(f0 - p0) < 5 * np.sqrt(p0 * (1 - p0) / n)
    # rank 1 maps to row M mod rows with M = (2^61 - 1) mod rows
    r1 = ((1 << 61) - 1) % rows
    p1 = 2.0 ** -theta / zeta
    f1 = (keys == r1).mean()
    assert abs(f1 - p1) < 5 * np.sqrt(p1 * (1 - p1) / n)


def test_uniform_keys_cover_table():
    keys, _, _ = L.kv_generate(seed=4, rows=64, n_txns=20_000, n_slots=4, theta=0.0, distinct_keys=False)
    counts = np.bincount(keys.astype(np.int64), minlength=64)
    assert counts.min() > 0.8 * counts.mean() and counts.max() < 1.2 * counts.mean()


def test_invalid_generator_parameters():
    with pytest.raises(L.LtpgError):
        L.kv_generate(rows=4, n_txns=10, n_slots=8)  # distinct keys impossible
    with pytest.raises(L.LtpgError):
        L.kv_generate(rows=100, n_txns=10, n_slots=65)
    with pytest.raises(L.LtpgError):
        L.kv_generate(rows=100, n_txns=10, n_slots=4, theta=1.0)
