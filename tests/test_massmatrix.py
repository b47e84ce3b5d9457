"""Mass matrices (construct_mass_matrix, src/construction.jl:409-453): the product's C
implementation and the oracle's numpy restatement against the known answers of the
reference's construction_test.jl:57-86 and massmatrix_test.jl:20-47, then against each
other on random networks.  Host only, no GPU."""
import json
import os

import numpy as np
import pytest

from conftest import oracle_for

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _user_mass(spec):
    kind = spec[0]
    if kind == "I":
        return None
    if kind == "uniform":
        return float(spec[1])
    if kind == "diag":
        return [float(x) for x in spec[1]]
    return np.array(spec[1], dtype=np.float64)


def _oracle_mass(spec):
    return tuple(spec) if spec[0] != "I" else ("I",)


class _Builder:
    """Components of a fixture case; one device function per distinct function name."""

    def __init__(self, ND):
        self.ND, self.fns = ND, {}

    def fn(self, name, role):
        key = (name, role)
        if key not in self.fns:
            A = self.ND._abi
            model = A.ND_VM_NOOP if role == "vertex_f" else A.ND_EM_NOOP
            self.fns[key] = self.ND.DeviceFunction(name, role, model, None, 0)
        return self.fns[key]

    def vertex(self, s):
        ND = self.ND
        return ND.VertexModel(f=self.fn(s["fn"], "vertex_f"), dim=s["dim"], g=ND.StateMask(s["g"]),
                              mass_matrix=_user_mass(s["mass"]))

    def edge(self, s):
        ND = self.ND
        g = s["g"]
        wrap = dict(AntiSymmetric=ND.AntiSymmetric, Symmetric=ND.Symmetric, Directed=ND.Directed)
        if g[0] == "Fiducial":
            return ND.EdgeModel(f=self.fn(s["fn"], "edge_f"), g=ND.Fiducial(g[1], g[2]), dim=s["dim"],
                                mass_matrix=_user_mass(s["mass"]))
        if g[1] == "function":
            return ND.EdgeModel(g=wrap[g[0]](self.fn(s["fn"], "edge_g")), outdim=1,
                                mass_matrix=_user_mass(s["mass"]))
        return ND.EdgeModel(f=self.fn(s["fn"], "edge_f"), g=wrap[g[0]](g[1]), dim=s["dim"],
                            mass_matrix=_user_mass(s["mass"]))


def _expand(fx, case, key, n):
    x = case[key]
    if isinstance(x, list):
        return x
    if x in fx["lists"]:
        return [fx["models"][m] for m in fx["lists"][x]]
    name, count = x.split("*")
    assert int(count) == n
    return [fx["models"][name]] * n


def _case_network(ND, fx, case):
    b = _Builder(ND)
    G = ND.SimpleDiGraph if case["directed"] else ND.SimpleGraph
    g = G(case["nv"], case["src"], case["dst"])
    vspecs = _expand(fx, case, "vertices", g.nv())
    especs = _expand(fx, case, "edges", g.ne())
    # one model object per distinct spec, like the reference's shared model instances
    vobj, eobj = {}, {}
    vm = [vobj.setdefault(json.dumps(s, sort_keys=True), b.vertex(s)) for s in vspecs]
    em = [eobj.setdefault(json.dumps(s, sort_keys=True), b.edge(s)) for s in especs]
    return ND.Network(g, vm, em), vspecs, especs


def _expected_matrix(exp, dim):
    M = np.diag(np.asarray(exp["diag"], np.float64))
    assert M.shape == (dim, dim)
    for r, c, v in exp["entries"]:
        M[r - 1, c - 1] = v
    return M


def _cases():
    fx = json.load(open(os.path.join(GOLDEN, "massmatrix_cases.json")))
    return fx, [c["name"] for c in fx["cases"]]


@pytest.mark.parametrize("name", _cases()[1])
def test_mass_matrix_known_answers(ND, name):
    fx, _ = _cases()
    case = next(c for c in fx["cases"] if c["name"] == name)
    nw, vspecs, especs = _case_network(ND, fx, case)
    exp = case["expected"]
    got = nw.mass_matrix
    shape, o = oracle_for(nw).mass_matrix([_oracle_mass(s["mass"]) for s in vspecs],
                                          [_oracle_mass(s["mass"]) for s in especs])
    assert shape == exp["shape"]
    if exp["shape"] == "identity":
        assert isinstance(got, ND.UniformScaling) and got == ND.I
        assert o is None
    elif exp["shape"] == "diagonal":
        assert isinstance(got, ND.Diagonal)
        assert np.array_equal(got.diag, np.asarray(exp["diag"], np.float64))
        assert np.array_equal(o, got.diag)
    else:
        M = _expected_matrix(exp, nw.dim)
        assert isinstance(got, np.ndarray)
        assert np.array_equal(got, M)
        assert np.array_equal(o, M)


def _random_mass(rng, dim):
    r = rng.integers(0, 6)
    if r == 0 or dim == 0:
        return ("I",) if r % 2 == 0 else ("uniform", float(rng.integers(0, 3)))
    if r == 1:
        return ("uniform", float(rng.integers(0, 3)))
    if r == 2:
        return ("diag", rng.integers(0, 3, dim).astype(float).tolist())
    if r == 3:
        return ("diag", [1.0] * dim)              # equals I: skipped
    if r == 4:
        return ("dense", np.eye(dim).tolist())    # equals I: skipped
    m = rng.integers(0, 3, (dim, dim)).astype(float)
    return ("dense", m.tolist())


@pytest.mark.parametrize("seed", range(6))
def test_mass_matrix_random_networks_match_oracle(ND, seed):
    rng = np.random.default_rng(100 + seed)
    b = _Builder(ND)
    n = int(rng.integers(3, 12))
    g = ND.erdos_renyi(n, int(rng.integers(n, 2 * n)), seed=seed)
    vdims = [1, 2, 3]
    vm, vspec = [], []
    for i in range(n):
        d = vdims[int(rng.integers(0, 3))]
        mm = _random_mass(rng, d)
        s = dict(fn=f"v{d}_{rng.integers(0, 2)}", dim=d, g=[1], mass=list(mm))
        vspec.append(s)
        vm.append(b.vertex(s))
    em, espec = [], []
    for e in range(g.ne()):
        d = int(rng.integers(0, 3))
        mm = _random_mass(rng, d)
        if d == 0:
            s = dict(fn="e0", dim=0, g=["Directed", "function"], mass=list(mm))
        else:
            s = dict(fn=f"e{d}", dim=d, g=["Fiducial", [1], [d]], mass=list(mm))
        espec.append(s)
        em.append(b.edge(s))
    nw = ND.Network(g, vm, em)
    shape, o = oracle_for(nw).mass_matrix([_oracle_mass(s["mass"]) for s in vspec],
                                          [_oracle_mass(s["mass"]) for s in espec])
    got = nw.mass_matrix
    if shape == "identity":
        assert got == ND.I
    elif shape == "diagonal":
        assert isinstance(got, ND.Diagonal) and np.array_equal(got.diag, o)
    else:
        assert isinstance(got, np.ndarray) and np.array_equal(got, o)


def test_mass_matrix_argument_errors(ND):
    f = ND.DeviceFunction("fv", "vertex_f", ND._abi.ND_VM_NOOP, None, 0)
    with pytest.raises(ND.ArgumentError):
        ND.VertexModel(f=f, dim=2, g=1, mass_matrix=[1.0, 0.0, 1.0])
    with pytest.raises(ND.ArgumentError):
        ND.VertexModel(f=f, dim=2, g=1, mass_matrix=np.eye(3))
    with pytest.raises(ND.ArgumentError):
        ND.VertexModel(f=f, dim=2, g=1, mass_matrix="zero")
    v = ND.VertexModel(f=f, dim=2, g=1, mass_matrix=0)
    assert v.mass_matrix == ND.UniformScaling(0)
    assert ND.VertexModel(f=f, dim=2, g=1, mass_matrix=[1, 0]).mass_matrix == ND.Diagonal([1.0, 0.0])


def test_explicit_integrators_reject_mass_matrices(ND):
    """Algebraic states need an implicit method; the explicit steppers refuse them before
    touching the device, as explicit OrdinaryDiffEq methods refuse a mass matrix."""
    g = ND.path_graph(4)
    alg = ND.VertexModel(f=ND.components.load_vertex_f, sym=["theta"], psym=["tau", "P"], g=1, mass_matrix=0)
    nw = ND.Network(g, [ND.Lib.kuramoto_second(), alg, alg, ND.Lib.kuramoto_second()], ND.Lib.kuramoto_edge())
    assert isinstance(nw.mass_matrix, ND.Diagonal)
    assert np.array_equal(nw.mass_matrix.diag, [1, 1, 1, 1, 0, 0])    # swing batch first, then the loads
    with pytest.raises(ND.UnsupportedError):
        nw.rk4(None, None, 0.0, 0.01, 10)
    with pytest.raises(ND.UnsupportedError):
        ND.Ensemble([g], alg, ND.Lib.kuramoto_edge())


def test_host_only_kinds_have_no_plan(ND):
    """Edges whose outputs are wrapped StateMasks, or with states but no state function,
    lay out like the reference's but have no HIP kernel: plan creation refuses them
    (checked before any device call)."""
    fx, _ = _cases()
    nw, _, _ = _case_network(ND, fx, fx["cases"][0])
    with pytest.raises(ND.UnsupportedError):
        nw.plan
