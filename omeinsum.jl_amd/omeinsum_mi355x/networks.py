"""Tensor-network workload generators (host side, deterministic from a seed).

* ``random_regular_graph``: uniform-pairing model with rejection of loops and multi-edges;
  numpy PCG64 stream, so a seed fixes the graph on every machine.
* ``fullerene``: the C60 construction of test/contractionorder.jl:32-48.
* ``network_code``: the contraction code of a graph (one tensor per vertex carrying the
  labels of its incident edges, contracted to a scalar), as the 3-regular benchmark
  networks and the fullerene test use it.
"""
import numpy as np

from .codes import EinCode


def random_regular_graph(n, d=3, seed=0, max_tries=10000):
    """Edges (u, v), u < v, of a simple d-regular graph on n vertices (n*d even)."""
    if (n * d) % 2:
        raise ValueError("n*d must be even for a d-regular graph")
    rng = np.random.Generator(np.random.PCG64(seed))
    for _ in range(max_tries):
        stubs = np.repeat(np.arange(n), d)
        rng.shuffle(stubs)
        pairs = stubs.reshape(-1, 2)
        u = np.minimum(pairs[:, 0], pairs[:, 1])
        v = np.maximum(pairs[:, 0], pairs[:, 1])
        if np.any(u == v):
            continue
        keys = u * n + v
        if len(np.unique(keys)) != len(keys):
            continue
        order = np.lexsort((v, u))
        return [(int(a), int(b)) for a, b in zip(u[order], v[order])]
    raise RuntimeError("no simple regular graph found")


def fullerene():
    """The 60 vertex coordinates of C60 (test/contractionorder.jl:32-46)."""
    phi = (1 + 5 ** 0.5) / 2
    res = []
    for (x, y, z) in ((0.0, 1.0, 3 * phi), (1.0, 2 + phi, 2 * phi), (phi, 2.0, 2 * phi + 1.0)):
        for (a, b, c) in ((x, y, z), (y, z, x), (z, x, y)):
            for loc in ((a, b, c), (a, b, -c), (a, -b, c), (a, -b, -c), (-a, b, c),
                        (-a, b, -c), (-a, -b, c), (-a, -b, -c)):
                if loc not in res:
                    res.append(loc)
    return res


def fullerene_edges():
    """Pairs (i, j), i < j, 0-based, closer than sqrt(5) (test/contractionorder.jl:48),
    enumerated in the reference's comprehension order (j outer, i inner)."""
    xyz = fullerene()
    edges = []
    for j, (i1, j1, k1) in enumerate(xyz):
        for i, (i2, j2, k2) in enumerate(xyz):
            if i < j and (i2 - i1) ** 2 + (j2 - j1) ** 2 + (k2 - k1) ** 2 < 5.0:
                edges.append((i, j))
    return edges


def network_code(n, edges):
    """EinCode of a graph: vertex v's tensor carries its incident edge labels (edge index e)
    in edge order; output is a scalar."""
    inc = [[] for _ in range(n)]
    for e, (u, v) in enumerate(edges):
        inc[u].append(e)
        inc[v].append(e)
    return EinCode(inc, [])


def gaussian_tensors(code, size_dict, seed, dtype=np.complex64):
    """i.i.d. complex Gaussian entries with sigma = 1/sqrt(2) per component (or real N(0,1)),
    axis p of tensor k carrying label code.ixs[k][p]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    xs = []
    for ix in code.ixs:
        shape = [size_dict[l] for l in ix]
        if np.dtype(dtype).kind == "c":
            s = 1 / np.sqrt(2)
            x = rng.normal(0, s, shape) + 1j * rng.normal(0, s, shape)
        else:
            x = rng.normal(0, 1, shape)
        xs.append(np.asarray(x, dtype=dtype))
    return xs


def request_stream(nreq=48, seed=5000):
    """The serving workload's request list: request q is a random 3-regular graph on an even
    number of vertices n uniform in [40, 140], with its own graph seed and tensor seed, all
    drawn from one PCG64 stream in the order n, graph seed, tensor seed.  Returns dicts with
    keys n, graph_seed, tensor_seed, edges."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(nreq):
        n = 2 * int(rng.integers(20, 71))
        gseed = int(rng.integers(1 << 30))
        tseed = int(rng.integers(1 << 30))
        out.append({"n": n, "graph_seed": gseed, "tensor_seed": tseed,
                    "edges": random_regular_graph(n, 3, seed=gseed)})
    return out
