"""CPU-only checks: the C ABI library exports every declared symbol; host-side parsing,
trees and planners behave like the reference's."""
import ctypes
import os
import re

import numpy as np
import pytest

from omeinsum_mi355x import codes, contractionorder as co, networks
from omeinsum_mi355x.codes import ArgumentError, DimensionMismatch, EinCode

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "omeinsum_mi355x.h")
LIB = os.path.join(ROOT, "omeinsum.jl_amd", "omeinsum_mi355x", "libomeinsum_mi355x.so")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(ome_\w+)\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("ome_init", "ome_einsum", "ome_loop_einsum", "ome_batched_gemm", "ome_tree_create",
              "ome_tree_execute", "ome_tree_gradient", "ome_server_submit", "ome_permute"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the library first (__graft_entry__.build())"
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert lib.ome_abi_version() == 1


def test_python_binding_covers_header():
    from omeinsum_mi355x import _lib
    assert set(declared_symbols()) == set(_lib.SIGNATURES)
    _lib.load()


def test_init_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from omeinsum_mi355x import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.ome_init(0, 0, None, ctypes.byref(h))
    assert rc != 0
    with pytest.raises(RuntimeError):
        _lib.context()


# ---- parser and codes (src/interfaces.jl, src/einsequence.jl) --------------------------


def test_ein_parser():
    c = codes.ein("ij,jk -> ik")
    assert isinstance(c, EinCode) and c.ixs == [["i", "j"], ["j", "k"]] and c.iy == ["i", "k"]
    c = codes.ein(",->")
    assert c.ixs == [[], []] and c.iy == []
    c = codes.ein("αβ,βγ->αγ")
    assert c.ixs == [["α", "β"], ["β", "γ"]]
    with pytest.raises(ArgumentError):
        codes.ein("ij,j1->i")
    with pytest.raises(ArgumentError):
        codes.ein("ij,jk")


def test_nested_parser():
    # test/einsequence.jl:5-48
    with pytest.raises(ArgumentError):
        codes.parse_nested("((ij,jk),km")
    e1 = codes.ein("((i,ij),i),j->i")
    assert e1.getiyv() == ["i"]
    assert e1.rootcode() == EinCode([["i", "j"], ["j"]], ["i"])
    lf = e1.siblings()[1]
    assert lf.isleaf() and lf.tensorindex == 3
    with pytest.raises(RuntimeError):
        lf.getiyv()
    assert codes.ein("((i,ij),i),j->ik").getixsv() == [["i"], ["i", "j"], ["i"], ["j"]]
    ne = codes.ein("(ij,j),k->")
    assert ne.flatten() == codes.ein("ij,j,k->")
    ne2 = codes.ein("(ij, jk), kl->il").replace({"i": "a", "j": "b", "k": "c", "l": "d"})
    assert ne2 == codes.ein("(ab, bc), cd->ad")


def test_get_size_dict_errors():
    a = np.zeros((3, 3))
    with pytest.raises(ArgumentError):
        codes.get_size_dict([[1, 2], [2, 3]], (a, a, a))
    with pytest.raises(ArgumentError):
        codes.get_size_dict([[1, 2], [2, 3]], (np.zeros((3, 3, 3)), a))
    b = np.zeros((3, 2))
    sd = codes.get_size_dict([["i", "j"], ["k", "j"]], (b, b))
    assert (sd["i"], sd["j"], sd["k"]) == (3, 2, 3)
    with pytest.raises(DimensionMismatch):
        codes.get_size_dict([["i", "j"], ["j", "k"]], (b, b))


def test_sliced_einsum_api():
    # test/slicing.jl:13-32
    se = codes.SlicedEinsum(["i", "l"], codes.ein("(ij,jk),(kl,lm)->im"))
    assert se.flatten() == se.eins.flatten()
    xs = (np.zeros((2, 3)), np.zeros((3, 4)), np.zeros((4, 5)), np.zeros((5, 6)))
    assert codes.get_size_dict(se.getixsv(), xs) == {"i": 2, "j": 3, "k": 4, "l": 5, "m": 6}
    assert se.getixsv() == [["i", "j"], ["j", "k"], ["k", "l"], ["l", "m"]]
    assert se.getiyv() == ["i", "m"]
    assert se.uniquelabels() == ["i", "j", "k", "l", "m"]


def test_greedy_and_writejson(tmp_path):
    code = EinCode([["a", "b"], ["b", "c"], ["c", "d"]], ["a", "d"])
    tree = co.optimize_code(code, codes.uniformsize(code, 2))
    assert tree.flatten() == code
    for obj in (tree, co.slice_code(tree, codes.uniformsize(code, 2), sc_target=1)):
        f = str(tmp_path / "t.json")
        co.writejson(f, obj)
        assert co.readjson(f) == obj


def test_random_regular_graph():
    e = networks.random_regular_graph(100, 3, seed=2000)
    assert len(e) == 150
    deg = np.zeros(100, int)
    for u, v in e:
        assert u < v
        deg[u] += 1
        deg[v] += 1
    assert np.all(deg == 3)
    assert len(set(e)) == 150
    assert e == networks.random_regular_graph(100, 3, seed=2000)
