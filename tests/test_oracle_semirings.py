"""Pins the oracle's CountingTropical{Float64,Float64} and TropicalAndOr modes on the CPU
with answers fixed by combinatorics: counting maximum independent sets (the Petersen graph
has independence number 4 and exactly 5 maximum independent sets; a 6-cycle has 3 and 2),
brute-force enumeration of independent sets, and boolean matrix products and reachability
against integer matrix products.  No GPU.
"""
import itertools

import numpy as np

from oracle import loop
from omeinsum_mi355x import networks


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


PETERSEN = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (0, 5), (1, 6), (2, 7), (3, 8), (4, 9),
            (5, 7), (7, 9), (9, 6), (6, 8), (8, 5)]


def mis_network(n, edges):
    """Independent-set network in the counting semiring: a vertex tensor [(0, 1), (1, 1)]
    per vertex (weight 1 when chosen) and an edge tensor that forbids choosing both ends."""
    ct = lambda v, c: complex(v, c)  # noqa: E731
    zero = ct(-np.inf, 0.0)
    vert = np.array([ct(0, 1), ct(1, 1)])
    edge = np.array([[ct(0, 1), ct(0, 1)], [ct(0, 1), zero]])
    ixs = [[v] for v in range(n)] + [list(e) for e in edges]
    xs = [vert] * n + [edge] * len(edges)
    return ixs, xs


def brute_mis(n, edges):
    """(independence number, number of maximum independent sets) by enumerating subsets."""
    masks = np.arange(1 << n, dtype=np.int64)
    ok = np.ones(1 << n, dtype=bool)
    for a, b in edges:
        ok &= ~(((masks >> a) & 1).astype(bool) & ((masks >> b) & 1).astype(bool))
    size = np.zeros(1 << n, dtype=np.int64)
    for v in range(n):
        size += (masks >> v) & 1
    size = np.where(ok, size, -1)
    best = int(size.max())
    return best, int(np.sum(size == best))


def test_counting_tropical_known_answers():
    for n, edges, want in [(10, PETERSEN, (4, 5)),
                           (6, [(i, (i + 1) % 6) for i in range(6)], (3, 2))]:
        ixs, xs = mis_network(n, edges)
        y = loop.loop_einsum(ixs, [], xs, {v: 2 for v in range(n)}, semiring="counting")
        assert (y.real, y.imag) == want
        assert brute_mis(n, edges) == want


def test_counting_tropical_random_graphs_vs_enumeration():
    for seed in range(3):
        n = 12
        edges = networks.random_regular_graph(n, 3, seed=seed)
        ixs, xs = mis_network(n, edges)
        y = loop.loop_einsum(ixs, [], xs, {v: 2 for v in range(n)}, semiring="counting")
        assert (y.real, y.imag) == brute_mis(n, edges)


def test_counting_tropical_semiring_rules():
    # (+) keeps the larger n with its count and adds counts of equal n; (*) adds n and
    # multiplies counts: check one product-sum against a direct evaluation
    r = rng(7201)
    a = r.integers(-3, 4, (4, 5)) + 1j * r.integers(1, 4, (4, 5))
    b = r.integers(-3, 4, (5, 3)) + 1j * r.integers(1, 4, (5, 3))
    y = loop.loop_einsum([["i", "j"], ["j", "k"]], ["i", "k"], [a, b],
                         {"i": 4, "j": 5, "k": 3}, semiring="counting")
    for i in range(4):
        for k in range(3):
            terms = [(a[i, j].real + b[j, k].real, a[i, j].imag * b[j, k].imag) for j in range(5)]
            best = max(t[0] for t in terms)
            assert y[i, k] == complex(best, sum(t[1] for t in terms if t[0] == best))


def test_and_or_products_and_reachability():
    r = rng(7202)
    a = r.random((6, 7)) < 0.3
    b = r.random((7, 5)) < 0.3
    y = loop.loop_einsum([["i", "j"], ["j", "k"]], ["i", "k"], [a, b],
                         {"i": 6, "j": 7, "k": 5}, semiring="andor")
    assert y.dtype == np.bool_
    assert np.array_equal(y, (a.astype(np.int64) @ b.astype(np.int64)) > 0)
    # reachability in a directed 8-cycle by repeated squaring of (I or A)
    n = 8
    adj = np.zeros((n, n), dtype=bool)
    for i in range(n):
        adj[i, (i + 1) % n] = True
        adj[i, i] = True
    m = adj
    for _ in range(3):
        m = loop.loop_einsum([["i", "j"], ["j", "k"]], ["i", "k"], [m, m],
                             {"i": n, "j": n, "k": n}, semiring="andor")
    assert m.all()
    # scaling: y <- false*y | true*x keeps x; y <- true*y | true*x is or
    y0 = r.random((6, 5)) < 0.5
    z = loop.loop_einsum([["i", "j"], ["j", "k"]], ["i", "k"], [a, b], {"i": 6, "j": 7, "k": 5},
                         y=y0.copy(), sx=True, sy=True, semiring="andor")
    assert np.array_equal(z, y | y0)


def test_brute_mis_itself():
    # the enumeration helper on graphs with closed forms: empty graph of n vertices (n, 1),
    # complete graph K_n (1, n)
    assert brute_mis(5, []) == (5, 1)
    assert brute_mis(5, list(itertools.combinations(range(5), 2))) == (1, 5)
