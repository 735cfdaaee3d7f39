"""The device against the committed golden fixtures (tests/golden/): every flat einsum case
through both backends (integer and max-plus bit-exact, floating point within the tolerances
of helpers.F_TOL), and the 100-vertex network's stored tree against its stored FP64 value."""
import json
import os

import numpy as np
import pytest

from helpers import assert_close

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    yield
    import omeinsum_mi355x as om
    om.set_einsum_backend(om.DefaultBackend())


def test_golden_einsum_cases():
    import omeinsum_mi355x as om
    meta = json.load(open(os.path.join(HERE, "einsum_cases.json")))
    arr = np.load(os.path.join(HERE, "einsum_cases.npz"))
    for backend in (om.DefaultBackend(), om.MI355XBackend(), om.MI355XBackend(precision="bf16x3")):
        om.set_einsum_backend(backend)
        for c, case in enumerate(meta):
            ixs = [list(w) for w in case["spec"].split("->")[0].split(",")]
            iy = list(case["spec"].split("->")[1])
            xs = [torch.from_numpy(np.ascontiguousarray(arr["c%d_x%d" % (c, k)])).cuda()
                  for k in range(len(ixs))]
            if case["tropical"]:
                xs = [om.Tropical(x) for x in xs]
            y = om.einsum(om.EinCode(ixs, iy), xs, case["sizes"])
            got = om.to_numpy(y)
            ref = arr["c%d_y" % c]
            if case["tropical"] or case["dtype"] == "int64":
                assert np.array_equal(got, ref), (backend, case["spec"])
            else:
                k = int(np.prod([case["sizes"][l] for l in set(l for ix in ixs for l in ix) - set(iy)] or [1]))
                # a sum to one scalar is judged elementwise only: its norm-wise error is
                # the cancellation of the random-sign terms, not the arithmetic's
                assert_close(got, ref, np.dtype(case["dtype"]), k, case["spec"],
                             norm=ref.size > 1)


def test_golden_network_value():
    import omeinsum_mi355x as om
    from omeinsum_mi355x import contractionorder as co, networks
    fx = json.load(open(os.path.join(HERE, "network_100.json")))
    tmp = "/tmp/ome_golden_tree_gpu.json"
    with open(tmp, "w") as f:
        json.dump(fx["tree"], f)
    tree = co.readjson(tmp)
    edges = [tuple(e) for e in fx["edges"]]
    code = networks.network_code(fx["n"], edges)
    sd = {e: 2 for e in range(len(edges))}
    xs = networks.gaussian_tensors(code, sd, seed=fx["tensor_seed"], dtype=np.complex64)
    ref = complex(*fx["value_fp64"])
    for backend in (om.MI355XBackend(precision="fp32"), om.DefaultBackend()):
        om.set_einsum_backend(backend)
        got = complex(om.to_numpy(tree(*[torch.from_numpy(x).cuda() for x in xs]))[()])
        assert abs(got - ref) <= 1e-4 * abs(ref), (backend, got, ref)
