"""GPU parity of the EM table queries and Philox draws (c8mi_em.h) against the
oracle (oracle/native_oracle.c, itself pinned bit-for-bit to the reference's
Philox.hpp and ProposalNativeQueries.hpp compiled in place).

* Philox: the four words of every key are integer-identical; the uniforms
  (word + 0.5) / 2^32 are bit-identical.
* Native queries on synthetic format-v7 tables (every query kind, every
  kinematic and direct-rate model): the status of every query is identical.
  Values carry the device's log/exp/pow, which may differ from the host's by an
  ulp, so the bar is 1e-9 relative (the per-step physics bar is 1e-6); Newton
  and bisection counts are identical for at least 99.9 % of the root solves
  (an ulp can move a stopping test by one iteration).
"""
import numpy as np
import pytest

import c8mi
import oracle_lib
from c8mi import tables

pytestmark = pytest.mark.gpu


def test_random_words_match_oracle():
    rng = np.random.default_rng(11)
    n = 4096
    keys = np.zeros(n, dtype=c8mi.RANDOM_KEY_DTYPE)
    for name in ("seed", "shower_id", "history_id", "step_id", "draw_id"):
        keys[name] = rng.integers(0, 2**63, n, dtype=np.uint64)
    keys["process_id"] = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    keys[:4]["history_id"] = [0, 1, 2**32, 2**64 - 1]
    words, uniforms = c8mi.random_draws(keys)
    for i in range(0, n, 7):
        k = keys[i]
        ident = [int(k[f]) for f in ("seed", "shower_id", "history_id", "step_id", "process_id", "draw_id")]
        assert list(words[i]) == oracle_lib.random_words(*ident), i
        lane = i % 4
        assert uniforms[i, lane] == oracle_lib.uniform_open01(*ident, lane)
    assert np.array_equal(uniforms, (words.astype(np.float64) + 0.5) / 2.0**32)
    assert uniforms.min() > 0.0 and uniforms.max() < 1.0


def _compare(seed, noise=0.0, wiggle=0.0, count=20000, polynomial_every=3):
    t = tables.random_table_set(seed, polynomial_every=polynomial_every, noise=noise, wiggle=wiggle)
    q = tables.random_queries(t, count, seed=seed + 100)
    want = oracle_lib.native_queries(t, q)
    dev = c8mi.EmTables(t)
    got = dev.query(q)
    dev.close()
    bad = np.nonzero(got["status"] != want["status"])[0]
    assert bad.size == 0, (
        f"status differs at {bad[:8]}: kinds {q['kind'][bad[:8]]}, "
        f"device {got['status'][bad[:8]]}, oracle {want['status'][bad[:8]]}")
    ok = want["status"] == 0
    a, b = got["value"][ok], want["value"][ok]
    assert np.array_equal(np.isnan(a), np.isnan(b))
    finite = np.isfinite(b)
    err = np.abs(a[finite] - b[finite]) / np.maximum(np.abs(b[finite]), 1e-300)
    worst = int(np.argmax(err)) if err.size else 0
    assert err.size == 0 or err.max() <= 1e-9, f"max rel err {err.max():.3e} at success #{worst}"
    solves = (want["newton_iterations"] + want["bisection_iterations"]) > 0
    same = ((got["newton_iterations"] == want["newton_iterations"])
            & (got["bisection_iterations"] == want["bisection_iterations"]))
    assert same[solves].mean() >= 0.999
    return got, want, q


@pytest.mark.parametrize("seed,noise,wiggle", [(1, 0.0, 0.0), (2, 0.0, 0.0), (3, 0.05, 0.0),
                                               (4, 0.2, 0.0), (5, 0.0, 0.4)])
def test_native_queries_match_oracle(seed, noise, wiggle):
    got, want, q = _compare(seed, noise, wiggle)
    # the batch reached the root solvers and the error statuses
    assert {0, 2, 3, 4, 5, 6}.issubset(set(int(s) for s in want["status"]))
    lf = q["kind"] == tables.QUERY["loss_fraction"]
    assert (got["newton_iterations"][lf] > 0).any()


def test_native_queries_without_polynomial_cells():
    # every bicubic column evaluated by the 16-term Hermite sum
    _compare(9, polynomial_every=0, count=8000)


def test_native_queries_empty_batch():
    t = tables.random_table_set(1)
    dev = c8mi.EmTables(t)
    assert dev.query(np.zeros(0, dtype=tables.QUERY_DTYPE)).size == 0
    assert dev.bytes() > 0
