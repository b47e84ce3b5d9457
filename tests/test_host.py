"""Host side of the product, no GPU: the C-ABI library loads and exports every
declared symbol; its index layer is bit-exact with the oracle and the hand-derived
fixture; construction errors follow the reference's ArgumentError cases."""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest

from conftest import ROOT, oracle_for
from test_oracle import _complete4, _random_mixed

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _declared_functions():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(nd_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(ND):
    lib = C.CDLL(ND._abi.LIB_PATH)
    names = _declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes prototypes cover the same set
    assert set(ND._abi.PROTOTYPES) >= set(names) - {"nd_plan_create_noalloc"}


def test_layout_matches_fixture(ND):
    fx = json.load(open(os.path.join(GOLDEN, "complete4_layout.json")))
    nw = _complete4(ND)
    assert (nw.dim, nw.pdim) == (fx["dim"], fx["pdim"])
    v, e = nw.ranges()
    for j, name in enumerate(["v_data", "v_out", "v_para", "v_aggr"]):
        assert v[:, j, :].tolist() == fx[name], name
    for j, name in enumerate(["e_data", "e_out_src", "e_out_dst", "e_para", "e_gbuf_src", "e_gbuf_dst"]):
        assert e[:, j, :].tolist() == fx[name], name
    rng_, amap, gmap = nw.maps()
    assert list(rng_) == fx["aggregation_range"]
    assert amap.tolist() == fx["aggregation_map"]
    assert gmap.tolist() == fx["gather_map"]
    rp, cols = nw.csr()
    # rows are aggregation indices; the inverse map lists (aggr index, sources) in vertex order
    byrow = {a + 1: cols[rp[a]:rp[a + 1]].tolist() for a in range(len(rp) - 1)}
    assert [[d, byrow[d]] for d, _ in fx["inverse_map"]] == fx["inverse_map"]


@pytest.mark.parametrize("seed,directed", [(1, True), (2, False), (3, True)])
def test_layout_bit_exact_with_oracle(ND, seed, directed):
    nw, _, _ = _random_mixed(ND, n=400, seed=seed, directed=directed)
    on = oracle_for(nw)
    v, e = nw.ranges()
    ov, oe = on.ranges()
    assert np.array_equal(v, ov) and np.array_equal(e, oe)
    r, a, g = nw.maps()
    oa, og = on.maps()
    assert (r[0], r[1]) == (on.agg_first, on.agg_last)
    assert np.array_equal(a, oa) and np.array_equal(g, og)
    # CSR rows == inverse map lists (no self-loops: same per-destination order)
    rp, cols = nw.csr()
    d, ptr, s = on.inverse_map()
    for w in range(len(d)):
        row = d[w] - 1
        assert np.array_equal(cols[rp[row]:rp[row + 1]], s[ptr[w]:ptr[w + 1]])
    for is_edge, n in ((False, on.nvb), (True, on.neb)):
        bs = nw.batches(is_edge)
        assert len(bs) == n
        for b in range(n):
            ob = on.batch(is_edge, b)
            for key in ("count", "state_first", "param_first", "out_first", "in_first"):
                assert bs[b][key] == ob[key], key
            assert np.array_equal(bs[b]["indices"], ob["indices"])


def test_construction_errors(ND):
    g = ND.complete_graph(3)
    L = ND.Lib
    with pytest.raises(ND.ArgumentError):   # construction.jl:99-101
        ND.Network(g, [L.kuramoto_second(), ND.VertexModel(dim=2, g=[1, 2]), L.kuramoto_second()],
                   L.kuramoto_edge())
    with pytest.raises(ND.ArgumentError):   # construction.jl:50
        ND.Network(g, [L.kuramoto_second()] * 2, L.kuramoto_edge())
    with pytest.raises(ND.UnsupportedError):
        ND.VertexModel(f=lambda dv, v, acc, p, t: None, dim=1)
    with pytest.raises(ND.ArgumentError):
        ND.Network(g, L.kuramoto_second(), L.kuramoto_edge(), execution=ND.ExecutionStyle(True))


def test_generators(ND):
    g = ND.barabasi_albert(5000, 2, seed=3)
    assert g.ne() == 2 * (5000 - 2)
    assert np.all(g.src < g.dst)
    key = g.src * 10 ** 6 + g.dst
    assert np.all(np.diff(key) > 0)          # sorted, no duplicates
    assert ND.barabasi_albert(5000, 2, seed=3).src.tolist() == g.src.tolist()   # reproducible
    er = ND.erdos_renyi(1000, 4000, seed=1)
    assert er.ne() == 4000 and np.all(er.src < er.dst)
    rg = ND.random_geometric(20000, np.sqrt(4 / (np.pi * 20000)), seed=1)
    assert 3.5 < 2 * rg.ne() / 20000 < 4.1
    ws = ND.watts_strogatz(100, 4, 0.5, seed=1)
    assert ws.ne() == 200


def test_graph_edge_order_like_graphs_jl(ND):
    g = ND.SimpleGraph(4, [3, 2, 4, 1], [1, 3, 2, 2])
    assert g.edges() == [(1, 2), (1, 3), (2, 3), (2, 4)]
    d = ND.SimpleDiGraph(3, [3, 1, 2], [1, 2, 1])
    assert d.edges() == [(1, 2), (2, 1), (3, 1)]


def test_default_parameters(ND):
    nw = ND.Network(ND.complete_graph(3), ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    p = nw.default_p()
    assert p[:9].tolist() == [1.0, 0.1, 1.0] * 3
    assert np.isnan(p[9:]).all()
