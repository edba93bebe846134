"""Pins the EM oracle's Philox and native-table restatements (oracle/native_oracle.c).

* Philox4x32-10: the reference's known-answer vector
  (tests/kokkos/em/stochastic/testStochasticContracts.cpp:57-59) and, on random
  counters/keys and random draw identities, the reference's own Philox.hpp
  compiled in place (oracle/_ref).
* Native table queries: every query kind of ProposalNativeQueries.inl on
  synthetic format-v7 tables, compared with the reference's own header compiled
  in place: status, Newton and bisection counts and value must be identical
  (both are host FP64 builds without contraction over the same libm).
* Closed forms the queries must honour: cubic/bicubic interpolation is exact at
  the nodes, physical_loss and transformed_loss are inverses, and
  energy_after_continuous_loss(E, range(E) - range(E')) returns E'.
"""
import math

import numpy as np
import pytest

import oracle_lib
from c8mi import tables

REF = oracle_lib.load_ref()
needs_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref not built (no reference tree at build time)")


def test_philox_known_answer():
    assert oracle_lib.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]


@needs_ref
def test_philox_matches_reference_code():
    rng = np.random.default_rng(5)
    for _ in range(2000):
        c = [int(x) for x in rng.integers(0, 2**32, 4, dtype=np.uint64)]
        k = [int(x) for x in rng.integers(0, 2**32, 2, dtype=np.uint64)]
        assert oracle_lib.philox(c, k) == oracle_lib.philox(c, k, lib=REF)
    for _ in range(2000):
        ident = [int(x) for x in rng.integers(0, 2**63, 6, dtype=np.uint64)]
        ident[4] &= 0xFFFFFFFF
        assert oracle_lib.random_words(*ident) == oracle_lib.random_words(*ident, lib=REF)
        lane = int(rng.integers(0, 8))
        assert oracle_lib.uniform_open01(*ident, lane) == REF.ref_uniform_open01(*ident, lane)


def test_uniform_is_open_interval():
    # (word + 0.5) / 2^32 never reaches 0 or 1
    assert 0.0 < (0 + 0.5) / 2**32 and (2**32 - 1 + 0.5) / 2**32 < 1.0
    u = [oracle_lib.uniform_open01(1, 1, h, 0, 0x454d0001, 0) for h in range(1000)]
    assert 0.0 < min(u) and max(u) < 1.0
    assert abs(np.mean(u) - 0.5) < 0.05


@needs_ref
@pytest.mark.parametrize("seed,noise,wiggle", [(1, 0.0, 0.0), (2, 0.0, 0.0), (3, 0.05, 0.0),
                                               (4, 0.2, 0.0), (5, 0.0, 0.4)])
def test_native_queries_match_reference_code(seed, noise, wiggle):
    t = tables.random_table_set(seed, noise=noise, wiggle=wiggle)
    q = tables.random_queries(t, 6000, seed=seed + 10)
    mine = oracle_lib.native_queries(t, q)
    ref = oracle_lib.native_queries(t, q, reference=True)
    assert np.array_equal(mine["status"], ref["status"])
    assert np.array_equal(mine["newton_iterations"], ref["newton_iterations"])
    assert np.array_equal(mine["bisection_iterations"], ref["bisection_iterations"])
    assert np.array_equal(mine["value"].view(np.uint64), ref["value"].view(np.uint64))
    # the queries reach the interesting branches
    statuses = set(int(s) for s in mine["status"])
    assert {0, 2, 3, 4, 5, 6}.issubset(statuses)
    ok = mine["status"] == 0
    lf = ok & (q["kind"] == tables.QUERY["loss_fraction"])
    assert (mine["newton_iterations"][lf] > 0).any()
    assert (mine["bisection_iterations"][lf] > 0).any()
    if wiggle:
        # non-bracketing endpoints: PROPOSAL's fallback bisection reports no root
        assert (mine["status"][q["kind"] == tables.QUERY["loss_fraction"]] == 8).any()


def test_interpolation_exact_at_nodes():
    t = tables.random_table_set(7)
    c = t.totals[0]
    ax = c["spline"]["axis"]
    nodes = tables.back_transform(ax, np.arange(2, int(ax["nodes"]) - 1))
    q = np.zeros(nodes.size, dtype=tables.QUERY_DTYPE)
    q["kind"] = tables.QUERY["total_rate"]
    q["pdg_id"] = c["pdg_id"]
    q["energy_mev"] = nodes
    out = oracle_lib.native_queries(t, q)
    off = int(c["spline"]["coefficient_offset"])
    stored = t.cubic_values[off + 2: off + int(ax["nodes"]) - 1]
    expected = np.maximum(stored, 0.0)
    assert np.allclose(out["value"], expected, rtol=1e-12, atol=0)


def test_energy_after_loss_inverts_range():
    t = tables.random_table_set(8)
    rng = np.random.default_rng(3)
    for pdg in (11, -11, 22):
        u = t.utility[t.utility["pdg_id"] == pdg][0]
        lo = float(u["lower_energy_limit_mev"])
        e0 = 10 ** rng.uniform(math.log10(lo) + 0.5, 6.5, 200)
        e1 = np.maximum(e0 * rng.uniform(0.2, 0.99, 200), 1.05 * lo)
        q = np.zeros(600, dtype=tables.QUERY_DTYPE)
        q["pdg_id"] = pdg
        q["kind"][:200] = tables.QUERY["range"]
        q["energy_mev"][:200] = e0
        q["kind"][200:400] = tables.QUERY["range"]
        q["energy_mev"][200:400] = e1
        r = oracle_lib.native_queries(t, q[:400])
        assert (r["status"] == 0).all()
        q2 = np.zeros(200, dtype=tables.QUERY_DTYPE)
        q2["kind"] = tables.QUERY["energy_after_loss"]
        q2["pdg_id"] = pdg
        q2["energy_mev"] = e0
        q2["argument"] = r["value"][:200] - r["value"][200:400]
        out = oracle_lib.native_queries(t, q2)
        assert (out["status"] == 0).all()
        # 20-bit Newton stop: relative accuracy about 2^-19 of the bracket
        assert np.allclose(out["value"], e1, rtol=2e-5)
