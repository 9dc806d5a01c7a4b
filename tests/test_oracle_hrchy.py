"""Pins the oracle's hierarchy (alloc_hrchy_space, hrchy_space.cc:137-231).

Checks: the reference's own src/hrchy_space.cc compiled into oracle/_ref
(identical breadth-first order and +-1 tables), the closed-form count
C(K+L, L), and the table invariants (uniqueness, tiers, ptr_m1[ptr_p1] = id,
sentinel row).
"""
import itertools
from math import comb

import numpy as np
import pytest

from libheom_amd import configs
from oracle import oracle as O

CASES = [(1, 0), (1, 5), (2, 10), (3, 4), (4, 3), (5, 5), (8, 3), (21, 2), (16, 3), (7, 4)]


def oracle_tables(K, L):
    m = configs.random_model(2, K, 1, L, seed=0, dense_v=False)
    o = O.Oracle(O.HILB, 2, L).load(m)
    return o.tables()


@pytest.mark.parametrize("K,L", CASES)
def test_count_closed_form(K, L):
    n, _, _ = oracle_tables(K, L)
    assert n.shape[0] == comb(K + L, L) == O.hrchy_count(L, K)


@pytest.mark.parametrize("K,L", CASES)
def test_matches_reference_hrchy_space(K, L):
    ref = O.ref_hrchy(K, L)
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    for a, b in zip(oracle_tables(K, L), ref):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("K,L", CASES)
def test_table_invariants(K, L):
    n, p1, m1 = oracle_tables(K, L)
    nh = n.shape[0]
    assert len({tuple(r) for r in n}) == nh
    tiers = n.sum(axis=1)
    assert np.all(np.diff(tiers) >= 0)  # breadth-first: tiers are contiguous
    assert tiers.max(initial=0) <= L
    assert np.all(p1[nh] == nh) and np.all(m1[nh] == nh)
    for l in range(nh):
        for k in range(K):
            if p1[l, k] != nh:
                e = n[l].copy()
                e[k] += 1
                assert np.array_equal(n[p1[l, k]], e)
                assert m1[p1[l, k], k] == l
            else:
                assert tiers[l] == L
            if m1[l, k] != nh:
                e = n[l].copy()
                e[k] -= 1
                assert np.array_equal(n[m1[l, k]], e)
            else:
                assert n[l, k] == 0


def test_bfs_order_small_case_by_hand():
    # K=2, L=2 visiting rule: [0,0] then +e0, +e1; pop [1,0] -> [2,0],[1,1];
    # pop [0,1] -> only k >= 1: [0,2]
    n, _, _ = oracle_tables(2, 2)
    assert [tuple(r) for r in n] == [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2)]


def test_every_multi_index_present():
    K, L = 4, 3
    n, _, _ = oracle_tables(K, L)
    want = {c for c in itertools.product(range(L + 1), repeat=K) if sum(c) <= L}
    assert {tuple(r) for r in n} == want
