"""Pins the oracle (CPU) against the reference's known answers, closed forms, numpy.einsum
and brute force.  No GPU.  A planted-mistake check at the end shows the checks bite."""
import itertools

import numpy as np
import pytest

from oracle import default_backend as db
from oracle import loop
from omeinsum_mi355x import codes, contractionorder as co, networks


def rng(seed=0):
    return np.random.Generator(np.random.PCG64(seed))


def randn(r, shape):
    return np.asarray(r.normal(size=tuple(shape)) if len(shape) else r.normal())


# ---- loop_einsum: closed forms (test/einsum.jl:103-212) -------------------------------


def test_loop_closed_forms():
    r = rng(1)
    a, b, c = r.normal(size=(2, 2)), r.random((2, 2)), r.random((2, 2))
    v = r.random(2)
    t = r.normal(size=(2, 2, 2, 2))
    le = loop.loop_einsum
    assert np.allclose(le([[1, 2], [2, 3], [3, 4]], [1, 4], [a, b, c]), a @ b @ c)
    assert np.allclose(le([[1, 2], [2, 3], [3, 4]], [4, 1], [a, b, c]), (a @ b @ c).T)
    assert np.allclose(le([[1, 2], [2]], [1], [a, v]), a @ v)
    assert np.allclose(le([[1, 2], [1, 2]], [], [a, a]), np.sum(a * a))
    assert np.allclose(le([[1, 1]], [], [a]), np.trace(a))
    aa = r.random((2, 4, 4, 2))
    assert np.allclose(le([[1, 2, 2, 1]], [], [aa]),
                       sum(aa[i, j, j, i] for i in range(2) for j in range(4)))
    assert np.allclose(le([[1, 2, 2, 3]], [1, 3], [aa]), sum(aa[:, i, i, :] for i in range(4)))
    assert np.allclose(le([[1, 2, 2, 3]], [3, 1], [aa]), sum(aa[:, i, i, :] for i in range(4)).T)
    assert np.allclose(le([[1, 2, 2, 3]], [1, 2, 3], [aa]),
                       np.stack([aa[:, i, i, :] for i in range(4)], axis=1))
    assert np.allclose(le([[1, 2, 3, 4]], [2, 3, 1, 4], [t]), np.transpose(t, (1, 2, 0, 3)))
    ta = np.zeros((2, 2))
    for i, j, k, l in itertools.product(range(2), repeat=4):
        ta[i, l] += t[i, j, k, l] * a[j, k]
    assert np.allclose(le([[1, 2, 3, 4], [2, 3]], [1, 4], [t, a]), ta)
    aaa = np.zeros((2, 2, 2))
    for i, j, k, l in itertools.product(range(2), repeat=4):
        aaa[j, k, l] += a[i, j] * a[i, k] * a[i, l]
    assert np.allclose(le([[1, 2], [1, 3], [1, 4]], [2, 3, 4], [a, a, a]), aaa)
    a3 = r.random((2, 2, 5))
    assert np.allclose(le([[1, 2, 3]], [1, 2], [a3]), a3.sum(axis=2))
    assert np.allclose(le([[1, 2], [3, 4]], [1, 2, 3, 4], [a, b]),
                       a[:, :, None, None] * b[None, None, :, :])
    # projecting to diag: ii -> ii
    d = le([[1, 1]], [1, 1], [a])
    assert np.array_equal(d, np.diag(np.diag(a)))


def test_loop_trivial_int_and_broadcast():
    # ein"->ii"(asarray(1); size_info i=5) == I(5)  (test/einsum.jl:191-194)
    one = np.array(1, dtype=np.int64)
    out = loop.loop_einsum([[]], ["i", "i"], [one], {"i": 5})
    assert np.array_equal(out, np.eye(5, dtype=np.int64))
    # ein",->"(3, 4) === 12  (test/einsum.jl:198-200)
    out = loop.loop_einsum([[], []], [], [np.array(3), np.array(4)])
    assert out.dtype == np.int64 and out[()] == 12


def test_issue_136():
    # test/einsum.jl:306-309
    out = loop.loop_einsum([[1, 2, 3], [2]], [1, 3], [np.ones((2, 2, 1)), np.ones(2)])
    assert np.array_equal(out, np.array([[2.0], [2.0]]))
    out = loop.loop_einsum([[1, 2, 3], [2]], [1, 3], [np.ones((2, 2, 0)), np.ones(2)])
    assert out.shape == (2, 0)


def _s_tensor():
    s = np.zeros((3, 3, 3), dtype=np.int64)
    for i, j, k in itertools.product(range(3), repeat=3):
        s[i, j, k] = int(len({i, j, k}) == 3)
    return s


def test_coloring_known_answers():
    # docs/src/applications.md:20-34
    s = _s_tensor()
    assert loop.loop_einsum([list("ijk"), list("ijk")], [], [s, s])[()] == 6
    code = "afl,bhn,cjf,dlh,enj,ago,big,cki,dmk,eom"
    ixs = [list(w) for w in code.split(",")]
    assert loop.loop_einsum(ixs, [], [s] * 10)[()] == 0


def test_loop_vs_numpy_einsum_random_codes():
    r = rng(7)
    letters = "abcdefg"
    for trial in range(60):
        nin = int(r.integers(1, 4))
        ixs = []
        for _ in range(nin):
            k = int(r.integers(0, 4))
            ixs.append(list(r.choice(list(letters), size=k, replace=False)))
        alll = sorted(set(l for ix in ixs for l in ix))
        ny = int(r.integers(0, len(alll) + 1))
        iy = list(r.choice(alll, size=ny, replace=False)) if ny else []
        sd = {l: int(r.integers(1, 4)) for l in letters}
        for dt in (np.float64, np.complex128, np.int64):
            xs = []
            for ix in ixs:
                shape = [sd[l] for l in ix]
                if dt is np.int64:
                    xs.append(np.asarray(r.integers(-8, 9, size=tuple(shape)), dtype=np.int64))
                elif dt is np.complex128:
                    xs.append(randn(r, shape) + 1j * randn(r, shape))
                else:
                    xs.append(randn(r, shape))
            spec = ",".join("".join(ix) for ix in ixs) + "->" + "".join(iy)
            ref = np.einsum(spec, *xs) if xs else None
            got = loop.loop_einsum(ixs, iy, xs, sd)
            if dt is np.int64:
                assert np.array_equal(got, ref), spec
            else:
                assert np.allclose(got, ref, rtol=1e-12, atol=1e-12), spec


def test_loop_scaling_semantics():
    # loop_einsum!(..., 4.0, 2.0) == 4*a*a + 2*y; with a repeated output label only the
    # diagonal receives the product (test/roceinsum.jl:10-21)
    r = rng(3)
    a = r.normal(size=(3, 3))
    y = np.ones((3, 3))
    out = loop.loop_einsum([[1, 2], [2, 3]], [1, 3], [a, a], y=y.copy(), sx=4.0, sy=2.0)
    assert np.allclose(out, 4 * a @ a + 2)
    out3 = loop.loop_einsum([[1, 2], [2, 3]], [1, 3, 3], [a, a], y=np.ones((3, 3, 3)), sx=4.0, sy=2.0)
    ref = 2 * np.ones((3, 3, 3))
    res = 4 * a @ a
    for k in range(3):
        ref[:, k, k] += res[:, k]
    assert np.allclose(out3, ref)
    # sy == 0 does not read y (NaN-safe overwrite)
    out = loop.loop_einsum([[1, 2], [2, 3]], [1, 3], [a, a], y=np.full((3, 3), np.nan), sx=1.0, sy=0.0)
    assert np.allclose(out, a @ a)


def test_tropical_loop_brute_force():
    r = rng(11)
    a = r.integers(-5, 6, size=(3, 4)).astype(np.float64)
    b = r.integers(-5, 6, size=(4, 2)).astype(np.float64)
    got = loop.loop_einsum([[1, 2], [2, 3]], [1, 3], [a, b], tropical=True)
    ref = np.max(a[:, :, None] + b[None, :, :], axis=1)
    assert np.array_equal(got, ref)


# ---- analyze_binary fixture (test/binaryrules.jl:5-16) ------------------------------


def test_analyze_binary_fixture():
    sd = {1: 1, 2: 2, 3: 3, 4: 4, 6: 6, 7: 7, 8: 8}
    c1, c2, cy, s1, s2, s3, i1, i2, iyb = db.analyze_binary(
        [1, 2, 3, 4, 8], [2, 6, 6, 8, 4, 2], [7, 2, 1, 2, 2, 6], sd)
    assert c1 == [1, 4, 8, 2] and c2 == [4, 8, 6, 2] and cy == [1, 6, 2]
    assert s1 == [1, 32, 2] and s2 == [32, 6, 2]
    assert i1 == ["i", "j", "l"] and i2 == ["j", "k", "l"] and iyb == ["i", "k", "l"]


# ---- the DefaultBackend restatement vs the brute force ------------------------------

BINARY_SWEEP_LETTERS = [(), ("i",), ("j",), ("i", "j"), ("j", "i")]


def binary_sweep_codes():
    """The 5x5x5x2 label patterns of test/binaryrules.jl:18-45."""
    out = []
    for has_batch in (True, False):
        for i1 in [(), ("i",), ("j",), ("i", "j"), ("j", "i")]:
            for i2 in [(), ("k",), ("j",), ("k", "j"), ("j", "k")]:
                for i3 in [(), ("i",), ("k",), ("i", "k"), ("k", "i")]:
                    b = ("l",) if has_batch else ()
                    out.append((list(i1 + b), list(i2 + b), list(i3 + b)))
    return out


def test_default_backend_binary_sweep():
    r = rng(5)
    sd = {l: 5 for l in "ijkl"}
    for ix1, ix2, iy in binary_sweep_codes():
        a = randn(r, [5] * len(ix1))
        b = randn(r, [5] * len(ix2))
        got = db.einsum([ix1, ix2], iy, [a, b], sd)
        ref = loop.loop_einsum([ix1, ix2], iy, [a, b], sd)
        assert np.allclose(got, ref), (ix1, ix2, iy)


COMPOSITE = ["ijl,jl->il", "ab,bc->ac", "abb,bc->ac", "ab,bc->acc", "ab,bce->ac",
             "ab,bc->ace", "bal,bcl->lcae", "bal,bcl->ca", "bal,bc->lca", "ddebal,bcf->lcac",
             "c,c->cc", "c,cc->c", "cc,c->cc", "cc,cc->cc", "cc,cb->bc", "cb,bc->cc",
             "ac,cc->ac"]
UNARY = ["ij->", "ij->j", "ii->", "ii->i", "ijk->kij", "a->aaaa", "ac->acc", "->ii", "i->ik",
         "->ik", "illljkk->kijjcc"]


def _parse(spec):
    l, r_ = spec.split("->")
    return [list(w) for w in l.split(",")], list(r_)


@pytest.mark.parametrize("spec", COMPOSITE + UNARY)
def test_default_backend_codes(spec):
    r = rng(9)
    ixs, iy = _parse(spec)
    sd = {l: 3 for l in "abcdefijkl"}
    for dt in (np.complex128, np.int64):
        xs = []
        for ix in ixs:
            shape = [3] * len(ix)
            if dt is np.int64:
                xs.append(np.asarray(r.integers(-8, 9, size=tuple(shape)), dtype=np.int64))
            else:
                xs.append(randn(r, shape) + 1j * randn(r, shape))
        got = db.einsum(ixs, iy, xs, sd)
        ref = loop.loop_einsum(ixs, iy, xs, sd)
        if dt is np.int64:
            assert np.array_equal(got, ref), spec
        else:
            assert np.allclose(got, ref), spec


def test_default_backend_scaling_and_regression():
    r = rng(13)
    sd = {1: 3, 2: 3, 3: 3, 4: 4, 5: 5}
    ix, iy, iz = [1, 2, 3, 3, 4], [3, 5, 1, 1, 2, 5], [1, 2, 3, 4, 5, 5]
    x = r.normal(size=(3, 3, 3, 3, 4))
    y = r.normal(size=(3, 5, 3, 3, 3, 5))
    z = r.normal(size=(3, 3, 3, 4, 5, 5))
    for sx, sy in ((1.0, 0.0), (5.0, 3.0), (5.0, 1.0)):
        got = db.einsum_inplace([ix, iy], iz, [x, y], np.asfortranarray(z.copy()), sx, sy, sd)
        ref = loop.loop_einsum([ix, iy], iz, [x, y], sd, y=z.copy(), sx=sx, sy=sy)
        assert np.allclose(got, ref)
    # regression code of test/binaryrules.jl:77-81
    a, b = r.normal(size=[2] * 8), r.normal(size=[2] * 5)
    ixs = [[10, 5, 4, 2, 6, 6, 7, 9], [1, 9, 8, 2, 10]]
    assert np.allclose(db.einsum(ixs, [4, 10, 9, 7, 10], [a, b]),
                       loop.loop_einsum(ixs, [4, 10, 9, 7, 10], [a, b]))


# ---- trees, slicing, tropical fullerene ----------------------------------------------


def test_nested_and_sliced():
    r = rng(17)
    ne = codes.ein("(ij,jk),(kl,lm)->im")
    xs = [r.normal(size=s) for s in ((2, 3), (3, 4), (4, 5), (5, 6))]
    sd = codes.get_size_dict(ne.getixsv(), xs)
    flat = loop.loop_einsum(ne.getixsv(), ne.getiyv(), xs, sd)
    nested = db.nested_einsum(ne.to_plain(), xs, sd)
    assert np.allclose(nested, flat)
    sliced = db.sliced_einsum(["i", "l"], ne.to_plain(), xs, sd)
    assert np.allclose(sliced, flat)


def test_fullerene_tropical_66():
    # test/contractionorder.jl:31-70: TropicalF64([-1 1; 1 -1]) on 90 edges, [0,0] on 60
    # vertices contracts to 66, unsliced and sliced
    edges = networks.fullerene_edges()
    assert len(edges) == 90
    ixs = [list(e) for e in edges] + [[i] for i in range(60)]
    code = codes.EinCode(ixs, [])
    sd = {i: 2 for i in range(60)}
    xs = [np.array([[-1.0, 1.0], [1.0, -1.0]])] * 90 + [np.zeros(2)] * 60
    tree = co.optimize_code(code, sd)
    assert tree.flatten() == code
    val = db.nested_einsum(tree.to_plain(), xs, sd, semiring="max_plus")
    assert val[()] == 66.0
    sliced = co.slice_code(tree, sd, sc_target=8)
    assert sliced.slicing
    assert co.contraction_complexity(sliced, sd)[1] <= 8
    val = db.sliced_einsum(sliced.slicing, tree.to_plain(), xs, sd, semiring="max_plus")
    assert val[()] == 66.0


def test_complexity_fixtures():
    ne = codes.ein("(ij,jkc),klc->il")
    sd = {l: 10 for l in "ijklc"}
    tc, sc, rw = co.contraction_complexity(ne, sd)
    assert np.isclose(tc, np.log2(20000)) and np.isclose(sc, np.log2(1000))
    assert np.isclose(rw, np.log2(100 + 1000 + 1000 + 1000 + 1000 + 100))
    assert co.peak_memory(ne, sd) == 2100
    assert co.flop(codes.ein("i->"), {"i": 4}) == 4


# ---- gradients ----------------------------------------------------------------------


def test_einsum_grad_closed_form():
    # src/autodiff.jl:14-21: einsum_grad(ij,jk; 1) == c * transpose(b)
    r = rng(19)
    a, b = r.random((2, 2)), r.random((2, 2))
    c = a @ b
    sd = {"i": 2, "j": 2, "k": 2}
    g = db.einsum_grad([["i", "j"], ["j", "k"]], [a, b], ["i", "k"], sd, c, 0)
    assert np.allclose(g, c @ b.T)


def test_cost_and_gradient_vs_finite_differences():
    r = rng(23)
    ne = codes.ein("(ij,jk),ki->")
    xs = [r.normal(size=s) + 1j * r.normal(size=s) for s in ((2, 3), (3, 4), (4, 2))]
    cost, grads = db.cost_and_gradient(ne.to_plain(), xs)
    assert np.isclose(cost[()], np.einsum("ij,jk,ki->", *xs))
    # d cost / d conj(x) convention of the reference: grad = conj(d f / d x) for holomorphic f
    eps = 1e-6
    for k in range(3):
        for idx in itertools.product(*[range(s) for s in xs[k].shape]):
            xp = [x.copy() for x in xs]
            xp[k][idx] += eps
            fd = (np.einsum("ij,jk,ki->", *xp) - cost[()]) / eps
            assert np.isclose(np.conj(grads[k][idx]), fd, rtol=1e-4, atol=1e-6)


def test_petersen_vertex_gradient_sum_zero():
    # docs/src/applications.md:50-55: relaxing one vertex still gives no colouring
    s = _s_tensor()
    code = codes.ein("afl,bhn,cjf,dlh,enj,ago,big,cki,dmk,eom->")
    sd = codes.uniformsize(code, 3)
    tree = co.optimize_code(code, sd)
    cost, grads = db.cost_and_gradient(tree.to_plain(), [s] * 10)
    assert cost[()] == 0
    assert np.sum(grads[0]) == 0


# ---- the checks bite: a planted mistake in the oracle is caught ---------------------


def test_planted_mistake_is_caught(monkeypatch):
    """Plant a plausible mistake in the restatement (the inverse permutation in
    tensorpermute!) and check that the comparison against the brute force and the closed
    form catches it."""
    orig = db.tensorpermute

    def broken(C, A, perm, sx, sy, sr):
        inv = [perm.index(p) for p in range(len(perm))]
        return orig(C, A, inv, sx, sy, sr)
    monkeypatch.setattr(db, "tensorpermute", broken)
    r = rng(29)
    a = r.normal(size=(3, 4, 5))
    got = None
    try:
        got = db.einsum([["i", "j", "k"]], ["j", "k", "i"], [a])
    except (ValueError, IndexError):
        return  # caught as a shape error
    ref = loop.loop_einsum([["i", "j", "k"]], ["j", "k", "i"], [a])
    assert got.shape != ref.shape or not np.allclose(got, ref)
