"""Component models: VertexModel / EdgeModel, output wrappers and feed-forward traits.

Mirrors src/component_functions.jl of NetworkDynamics.jl: the constructors
(`VertexModel` :251-303, `EdgeModel` :306-362), the output wrappers
(`StateMask` :81-99, `AntiSymmetric` :117-127, `Symmetric` :142-152,
`Directed` :167-175, `Fiducial` :190-203) and the feed-forward traits (:6-64).

The functions `f` and `g` cannot be arbitrary Python callables: they run inside
fused HIP kernels, so they are chosen from a fixed set of device functions
(`DeviceFunction` instances defined in `library`).  Anything else raises
`UnsupportedError` at construction time.
"""
import numpy as np

from . import _abi as A


class FeedForwardType:
    _id = None

    def __repr__(self):
        return type(self).__name__ + "()"

    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self))


class PureFeedForward(FeedForwardType):
    """g(outs..., ins..., p, t)"""
    _id = A.ND_FF_PUREFEEDFORWARD


class FeedForward(FeedForwardType):
    """g(outs..., x, ins..., p, t)"""
    _id = A.ND_FF_FEEDFORWARD


class NoFeedForward(FeedForwardType):
    """g(outs..., x, p, t)"""
    _id = A.ND_FF_NOFEEDFORWARD


class PureStateMap(FeedForwardType):
    """g(outs..., x)"""
    _id = A.ND_FF_PURESTATEMAP


def hasff(x):
    return isinstance(fftype(x), (PureFeedForward, FeedForward))


class DeviceFunction:
    """A component function implemented by the HIP kernels.

    role: "vertex_f" (f of a vertex), "edge_g" (single-sided edge output, used
    inside AntiSymmetric/Symmetric/Directed), "edge_g2" (two-sided edge output
    g(osrc, odst, v_s, v_d, p, t)), or "edge_f" (state function of an ODE edge).
    """

    def __init__(self, name, role, model, dim=None, pdim=0, ref="", ext=0, ff=None):
        self.name, self.role, self.model, self.dim, self.pdim, self.ref = name, role, model, dim, pdim, ref
        self.ext = ext          # external inputs the function reads
        self.ff = ff            # role "vertex_g": the output's feed-forward trait

    def __repr__(self):
        return f"<device function {self.name}>"


class StateMask:
    """Output picked from the state: out[i] = x[idxs[i]] (1-based)."""

    def __init__(self, idxs):
        if isinstance(idxs, int):
            idxs = [idxs]
        elif isinstance(idxs, range):
            idxs = list(idxs)
        self.idxs = tuple(int(i) for i in idxs)
        if not self.idxs or min(self.idxs) < 1:
            raise A.ArgumentError("StateMask indices are 1-based and non-empty")

    def __len__(self):
        return len(self.idxs)

    def __repr__(self):
        return f"StateMask({list(self.idxs)})"


def _as_mask_or_fn(g):
    if isinstance(g, (int, range, list, tuple)):
        return StateMask(g)
    return g


class AntiSymmetric:
    """y_dst = g(...); y_src = -y_dst"""
    wrapper = A.ND_WRAP_ANTISYMMETRIC

    def __init__(self, g):
        self.g = _as_mask_or_fn(g)


class Symmetric:
    """y_dst = g(...); y_src = y_dst"""
    wrapper = A.ND_WRAP_SYMMETRIC

    def __init__(self, g):
        self.g = _as_mask_or_fn(g)


class Directed:
    """y_dst = g(...); no src output"""
    wrapper = A.ND_WRAP_DIRECTED

    def __init__(self, g):
        self.g = _as_mask_or_fn(g)


class Fiducial:
    """Independent src and dst outputs: Fiducial(src, dst) or Fiducial(src=..., dst=...)."""
    wrapper = A.ND_WRAP_FIDUCIAL

    def __init__(self, src=None, dst=None):
        if src is None or dst is None:
            raise A.ArgumentError("Fiducial needs src and dst")
        self.src = _as_mask_or_fn(src)
        self.dst = _as_mask_or_fn(dst)


class VIndex:
    """VIndex(component, sym): a vertex state or output by vertex index (1-based) or unique
    name, and symbol (or 1-based state index) -- the references an `extin` list holds."""

    kind = "V"

    def __init__(self, comp, sym):
        self.comp, self.sym = comp, sym

    def __repr__(self):
        return f"{type(self).__name__}({self.comp!r}, {self.sym!r})"


class EIndex(VIndex):
    """EIndex(component, sym): the edge counterpart of VIndex."""

    kind = "E"


def _insym(insym, indim, is_edge):
    """Input symbols (component_functions.jl:864-911): insym as given, or from indim the
    names i / i1, i2, ...; an edge's plain list becomes src/dst names with the prefixes
    "src₊" / "dst₊".  None: no input symbols (inputs cannot be named)."""
    if insym is None and indim is not None:
        insym = ["i"] if indim == 1 else [f"i{k + 1}" for k in range(indim)]
    if insym is None:
        return None
    if is_edge:
        if isinstance(insym, dict):
            return dict(src=[str(x) for x in insym.get("src", [])], dst=[str(x) for x in insym.get("dst", [])])
        return dict(src=["src₊" + str(x) for x in insym], dst=["dst₊" + str(x) for x in insym])
    return [str(x) for x in insym]


def _extin(extin, f):
    if extin is None:
        refs = []
    else:
        refs = list(extin)
        for r in refs:
            if not isinstance(r, VIndex):
                raise A.ArgumentError(f"external inputs are VIndex / EIndex references, got {r!r}")
    need = getattr(f, "ext", 0) if f is not None else 0
    if f is not None and f.model not in (A.ND_VM_NOOP, A.ND_EM_NOOP) and len(refs) != need:
        raise A.ArgumentError(f"{f.name} reads {need} external input(s), extin has {len(refs)}")
    if len(refs) > A.ND_MAX_EXT:
        raise A.UnsupportedError(f"at most {A.ND_MAX_EXT} external inputs per component")
    return refs


class UniformScaling:
    """LinearAlgebra.UniformScaling: s times the identity of any size."""

    def __init__(self, s=1.0):
        self.s = float(s)

    def __eq__(self, other):
        if isinstance(other, UniformScaling):
            return self.s == other.s
        return NotImplemented

    def __hash__(self):
        return hash(("UniformScaling", self.s))

    def __call__(self, n):
        return Diagonal(np.full(int(n), self.s))

    def __repr__(self):
        return "I" if self.s == 1.0 else f"UniformScaling({self.s:g})"


I = UniformScaling(1.0)


class Diagonal:
    """LinearAlgebra.Diagonal: a square matrix given by its diagonal."""

    def __init__(self, d):
        self.diag = np.asarray(d, dtype=np.float64).reshape(-1)

    def toarray(self):
        return np.diag(self.diag)

    def __eq__(self, other):
        if isinstance(other, Diagonal):
            return np.array_equal(self.diag, other.diag)
        if isinstance(other, UniformScaling):
            return bool(np.all(self.diag == other.s))
        if isinstance(other, np.ndarray):
            return other.shape == (len(self.diag),) * 2 and np.array_equal(self.toarray(), other)
        return NotImplemented

    __hash__ = None

    def __repr__(self):
        return f"Diagonal({self.diag.tolist()})"


def _normalize_mass(mm, dim):
    """The user's mass matrix as VertexModel/EdgeModel keep it (component_functions.jl:948-963):
    nothing -> I, a number -> UniformScaling, a vector -> Diagonal (length dim), a matrix
    (dim x dim) kept; anything else is an ArgumentError."""
    if mm is None:
        return I
    if isinstance(mm, (UniformScaling, Diagonal)):
        if isinstance(mm, Diagonal) and len(mm.diag) != dim:
            raise A.ArgumentError("Size of mass matrix must match dimension of component.")
        return mm
    if isinstance(mm, (int, float, np.integer, np.floating)) and not isinstance(mm, bool):
        return UniformScaling(mm)
    try:
        a = np.asarray(mm, dtype=np.float64)
    except (TypeError, ValueError):
        a = None
    if a is None or a.ndim not in (1, 2):
        raise A.ArgumentError(f"Mass matrix must be a vector, square matrix, a uniform scaling, or scalar. Got {mm!r}.")
    if a.ndim == 1:
        if len(a) != dim:
            raise A.ArgumentError("If given as a vector, mass matrix must have length equal to dimension of component.")
        return Diagonal(a)
    if a.shape != (dim, dim):
        raise A.ArgumentError("Size of mass matrix must match dimension of component.")
    return a.copy()


def is_identity_mass(mm):
    """mm == I, comparing as LinearAlgebra does (a Diagonal or matrix entry by entry)."""
    if isinstance(mm, UniformScaling):
        return mm.s == 1.0
    if isinstance(mm, Diagonal):
        return bool(np.all(mm.diag == 1.0))
    return bool(np.array_equal(mm, np.eye(len(mm))))


def _mass_struct(mm):
    """(nd_mass fields, keep-alive buffer) of a normalised mass matrix."""
    if isinstance(mm, UniformScaling):
        if mm.s == 1.0:
            return (A.ND_MASS_IDENTITY, None), None
        buf = np.array([mm.s])
        return (A.ND_MASS_UNIFORM, buf), buf
    if isinstance(mm, Diagonal):
        buf = np.ascontiguousarray(mm.diag)
        return (A.ND_MASS_DIAGONAL, buf), buf
    buf = np.asfortranarray(mm).ravel(order="F").copy()
    return (A.ND_MASS_DENSE, buf), buf


def fftype(x):
    if isinstance(x, (VertexModel, EdgeModel)):
        return x.ff
    if isinstance(x, StateMask):
        return PureStateMap()
    raise A.ArgumentError(f"no feed-forward trait for {x!r}")


def _split_sym(syms):
    """sym / psym entries may be `name` or `(name, default)` (the reference's `:x => 1`)."""
    names, defaults = [], []
    for s in syms:
        if isinstance(s, tuple):
            names.append(str(s[0]))
            defaults.append(float(s[1]))
        else:
            names.append(str(s))
            defaults.append(None)
    return names, defaults


class ComponentModel:
    is_edge = False

    def kind(self):
        """The nd_kind description handed to the C ABI."""
        return dict(is_edge=int(self.is_edge), model=self.model, wrapper=self.wrapper, ff=self.ff._id,
                    dim=self.dim, pdim=self.pdim, outdim_src=self.outdim_src, outdim_dst=self.outdim_dst,
                    mask_src=tuple(self.mask_src), mask_dst=tuple(self.mask_dst), extdim=len(self.extin))

    def component_hash(self):
        """What batches components together (_component_hash, construction.jl:245-256):
        the kind plus the identity of the state and output functions, so that two
        distinct device functions of one equation still form two batches."""
        k = self.kind()
        return tuple(sorted(k.items())) + (("f", id(self.f)), ("g", id(getattr(self, "_gfn", None))))

    def copy(self):
        import copy as _c
        return _c.copy(self)

    @property
    def outdim(self):
        return self.outdim_dst if not self.is_edge else dict(src=self.outdim_src, dst=self.outdim_dst)


class VertexModel(ComponentModel):
    """VertexModel(f=..., g=StateMask(...), dim/sym, pdim/psym, name) -- component_functions.jl:251-303."""

    is_edge = False

    def __init__(self, f=None, g=None, dim=None, sym=None, pdim=None, psym=None, outdim=None, name=None,
                 mass_matrix=None, extin=None, insym=None, indim=None):
        self.insym = _insym(insym, indim, False)
        if isinstance(g, DeviceFunction) and g.role == "vertex_g":
            self._init_output_model(f, g, dim, sym, pdim, psym, outdim, name, mass_matrix, extin)
            return
        if f is None:
            f = noop_vertex
        if not isinstance(f, DeviceFunction) or f.role != "vertex_f":
            raise A.UnsupportedError(f"vertex f must be a vertex device function of the library, got {f!r}")
        self.f = f
        self._gfn = None
        self.model = f.model
        if sym is not None:
            self.sym, self.sym_default = _split_sym(sym)
        else:
            d = dim if dim is not None else f.dim
            if d is None:
                raise A.ArgumentError("VertexModel needs dim or sym")
            self.sym = [f"x{i + 1}" for i in range(d)] if d != 1 else ["x"]
            self.sym_default = [None] * d
        self.dim = len(self.sym)
        if dim is not None and dim != self.dim:
            raise A.ArgumentError("dim and sym disagree")
        if f.dim is not None and f.dim != self.dim:
            raise A.ArgumentError(f"{f.name} needs dim={f.dim}")
        if psym is not None:
            self.psym, self.psym_default = _split_sym(psym)
        else:
            pd = pdim if pdim is not None else f.pdim
            self.psym = [f"p{i + 1}" for i in range(pd)] if pd != 1 else ["p"]
            self.psym_default = [None] * pd
        self.pdim = len(self.psym)
        if pdim is not None and pdim != self.pdim:
            raise A.ArgumentError("pdim and psym disagree")
        if f.model != A.ND_VM_NOOP and self.pdim != f.pdim:
            raise A.ArgumentError(f"{f.name} needs pdim={f.pdim}")
        if g is None:
            g = StateMask(1)
        g = _as_mask_or_fn(g)
        if not isinstance(g, StateMask):
            raise A.UnsupportedError("vertex outputs must be StateMasks (feed-forward vertices have no HIP kernel)")
        if max(g.idxs) > self.dim:
            raise A.ArgumentError("StateMask index beyond dim")
        if outdim is not None and outdim != len(g):
            raise A.ArgumentError("outdim and g disagree")
        self.g = g
        self.ff = PureStateMap()
        self.wrapper = A.ND_WRAP_NONE
        self.outdim_src, self.outdim_dst = 0, len(g)
        self.mask_src, self.mask_dst = (), g.idxs
        self.name = name or f.name
        self.mass_matrix = _normalize_mass(mass_matrix, self.dim)
        self.extin = _extin(extin, f)

    def _init_output_model(self, f, g, dim, sym, pdim, psym, outdim, name, mass_matrix, extin):
        """A vertex whose output is computed by a library g without states: NoFeedForward
        g(out, p, t) or an injector's PureFeedForward g(out, in, p, t)."""
        if f is not None:
            raise A.UnsupportedError(f"{g.name} has no state function in the library")
        if (dim or 0) != 0 or sym:
            raise A.ArgumentError(f"{g.name} has no states")
        od = getattr(g, "outdim", 1)
        if outdim is not None and outdim != od:
            raise A.ArgumentError("outdim and g disagree")
        self.f, self._gfn, self.model = None, g, g.model
        self.sym, self.sym_default, self.dim = [], [], 0
        if psym is not None:
            self.psym, self.psym_default = _split_sym(psym)
        else:
            pd = pdim if pdim is not None else g.pdim
            self.psym = [f"p{i + 1}" for i in range(pd)] if pd != 1 else ["p"]
            self.psym_default = [None] * pd
        self.pdim = len(self.psym)
        if self.pdim != g.pdim:
            raise A.ArgumentError(f"{g.name} needs pdim={g.pdim}")
        self.g = g
        self.ff = g.ff
        self.wrapper = A.ND_WRAP_NONE
        self.outdim_src, self.outdim_dst = 0, od
        self.mask_src, self.mask_dst = (), ()
        self.name = name or g.name
        self.mass_matrix = _normalize_mass(mass_matrix, 0)
        self.extin = _extin(extin, None)

    def __repr__(self):
        fn = self.f.name if self.f is not None else self.g.name
        return f"VertexModel :{self.name} (dim={self.dim}, pdim={self.pdim}, {'f' if self.f is not None else 'g'}={fn})"


class EdgeModel(ComponentModel):
    """EdgeModel(f=..., g=AntiSymmetric(...)/Fiducial(...)/..., dim, pdim, outdim) -- component_functions.jl:306-362."""

    is_edge = True

    def __init__(self, f=None, g=None, dim=None, sym=None, pdim=None, psym=None, outdim=None, outsym=None,
                 name=None, mass_matrix=None, extin=None, insym=None, indim=None):
        self.insym = _insym(insym, indim, True)
        if g is None:
            raise A.ArgumentError("EdgeModel needs g")
        if f is not None and (not isinstance(f, DeviceFunction) or f.role != "edge_f"):
            raise A.UnsupportedError(f"edge f must be an edge device function of the library, got {f!r}")
        self.f = f
        if sym is not None:
            self.sym, self.sym_default = _split_sym(sym)
        else:
            d = dim if dim is not None else (f.dim if f is not None else 0)
            self.sym = [f"e{i + 1}" for i in range(d)]
            self.sym_default = [None] * d
        self.dim = len(self.sym)
        if dim is not None and dim != self.dim:
            raise A.ArgumentError("dim and sym disagree")
        od = outdim
        if outsym is not None:
            od = len(outsym)
        fn = None
        if isinstance(g, (AntiSymmetric, Symmetric, Directed)):
            inner = g.g
            self.wrapper = g.wrapper
            if isinstance(inner, StateMask):
                # host layout only: nd_plan_create rejects it (no HIP kernel)
                if f is None:
                    raise A.ArgumentError("a StateMask output needs an edge state (f)")
                od = len(inner) if od is None else od
                if od != len(inner):
                    raise A.ArgumentError("outdim and g disagree")
                if max(inner.idxs) > self.dim:
                    raise A.ArgumentError("StateMask index beyond dim")
                self.outdim_dst = od
                self.outdim_src = 0 if isinstance(g, Directed) else od
                self.mask_dst = inner.idxs
                self.mask_src = () if isinstance(g, Directed) else inner.idxs
                self.ff = PureStateMap()
            elif not isinstance(inner, DeviceFunction) or inner.role != "edge_g":
                raise A.UnsupportedError(f"edge g must be a library device function, got {inner!r}")
            else:
                fn = inner
                od = 1 if od is None else od
                self.outdim_dst = od
                self.outdim_src = 0 if isinstance(g, Directed) else od
                self.mask_src, self.mask_dst = (), ()
                self.ff = PureFeedForward()
        elif isinstance(g, Fiducial):
            self.wrapper = A.ND_WRAP_FIDUCIAL
            if not (isinstance(g.src, StateMask) and isinstance(g.dst, StateMask)):
                raise A.UnsupportedError("Fiducial outputs in the HIP kernels are StateMask pairs")
            if f is None:
                raise A.ArgumentError("a Fiducial StateMask output needs an edge state (f)")
            self.outdim_src, self.outdim_dst = len(g.src), len(g.dst)
            self.mask_src, self.mask_dst = g.src.idxs, g.dst.idxs
            if max(self.mask_src + self.mask_dst) > self.dim:
                raise A.ArgumentError("StateMask index beyond dim")
            self.ff = PureStateMap()
        elif isinstance(g, DeviceFunction) and g.role == "edge_g2":
            self.wrapper = A.ND_WRAP_NONE
            fn = g
            od = 1 if od is None else od
            self.outdim_src = self.outdim_dst = od
            self.mask_src, self.mask_dst = (), ()
            self.ff = PureFeedForward()
        else:
            raise A.UnsupportedError(f"edge output {g!r} has no HIP kernel")
        self.g = g
        if fn is not None and f is not None:
            raise A.UnsupportedError("edges with both a state function and a feed-forward output have no HIP kernel")
        self.model = fn.model if fn is not None else f.model
        natural_p = (fn or f).pdim
        if psym is not None:
            self.psym, self.psym_default = _split_sym(psym)
        else:
            pd = pdim if pdim is not None else natural_p
            self.psym = [f"p{i + 1}" for i in range(pd)] if pd != 1 else ["p"]
            self.psym_default = [None] * pd
        self.pdim = len(self.psym)
        if pdim is not None and pdim != self.pdim:
            raise A.ArgumentError("pdim and psym disagree")
        if self.model != A.ND_EM_NOOP and self.pdim != natural_p:
            raise A.ArgumentError(f"{(fn or f).name} needs pdim={natural_p}")
        self.name = name or (fn or f).name
        self._gfn = fn
        self.outsym = [str(x) for x in outsym] if outsym is not None else None
        self.mass_matrix = _normalize_mass(mass_matrix, self.dim)
        self.extin = _extin(extin, f)

    def __repr__(self):
        return f"EdgeModel :{self.name} (dim={self.dim}, pdim={self.pdim}, outdim={self.outdim})"


# ---- device functions (the kernels' equations) -------------------------------------

noop_vertex = DeviceFunction("noop_vertex", "vertex_f", A.ND_VM_NOOP, None, 0,
                             "test/aggregators_test.jl:15")
diffusionvertex = DeviceFunction("diffusionvertex!", "vertex_f", A.ND_VM_DIFFUSION, 1, 0,
                                 "test/ComponentLibrary.jl:42-45")
# dv .= esum with any number of states (one per output component)
diffusionvertex_vec = DeviceFunction("diffusionvertex_f!", "vertex_f", A.ND_VM_DIFFUSION, None, 0,
                                     "docs/examples/getting_started_with_network_dynamics.jl:129-132")
kuramoto_vertex = DeviceFunction("kuramoto_vertex!", "vertex_f", A.ND_VM_KURAMOTO, 1, 1,
                                 "test/ComponentLibrary.jl:69-71")
kuramoto_inertia = DeviceFunction("kuramoto_inertia!", "vertex_f", A.ND_VM_SWING, 2, 3,
                                  "test/ComponentLibrary.jl:59-63")
kuramoto_inertia_unit_damping = DeviceFunction("kuramoto_inertia_unit_damping!", "vertex_f",
                                               A.ND_VM_SWING_UNIT_DAMPING, 2, 1,
                                               "benchmark/benchmark_models.jl:37-41")
load_vertex_f = DeviceFunction("load_vertex!", "vertex_f", A.ND_VM_LOAD, 1, 2, "build-defined")

noop_edge = DeviceFunction("noop_edge", "edge_g", A.ND_EM_NOOP, 0, 0, "test/aggregators_test.jl:20")
noop_edge2 = DeviceFunction("noop_edge2", "edge_g2", A.ND_EM_NOOP, 0, 0, "test/aggregators_test.jl:27-28")
noop_edge_f = DeviceFunction("noop_edge_f", "edge_f", A.ND_EM_NOOP, None, 0, "test/massmatrix_test.jl:13-17")
diffusionedge = DeviceFunction("diffusionedge!", "edge_g", A.ND_EM_DIFFUSION, 0, 0,
                               "test/diffusion_test.jl:57-60")
diffusionedge_weighted = DeviceFunction("diffusionedge_weighted!", "edge_g", A.ND_EM_DIFFUSION_WEIGHTED, 0, 1,
                                        "test/ComponentLibrary.jl:8-10")
kuramoto_edge_g = DeviceFunction("kuramoto_edge!", "edge_g", A.ND_EM_KURAMOTO, 0, 1,
                                 "test/ComponentLibrary.jl:51-53")
static_line_g = DeviceFunction("static_line!", "edge_g", A.ND_EM_STATIC_LINE, 0, 2, "build-defined")
diffusionedge_fid = DeviceFunction("diffusionedge_fid!", "edge_g2", A.ND_EM_DIFFUSION_FIDUCIAL, 0, 1,
                                   "test/ComponentLibrary.jl:22-25")
diffusion_dedge = DeviceFunction("diffusion_dedge!", "edge_f", A.ND_EM_ODE_SIN_DIFFUSION, 2, 1,
                                 "test/ComponentLibrary.jl:30-34")
diffusion_dedge_fast = DeviceFunction("diffusion_dedge_fast!", "edge_f", A.ND_EM_ODE_SIN_DIFFUSION_FAST, 2, 0,
                                      "benchmark/benchmark_models.jl:11-15")
real_ode_edge = DeviceFunction("real_ode_edge!", "edge_f", A.ND_EM_ODE_LINEAR, 2, 0,
                               "test/diffusion_test.jl:102-106")
# one-state line lag de = 1/tau (sin(v_s - v_d) - e) (the dst half of ComponentLibrary.jl:30-34),
# output through a wrapped StateMask (AntiSymmetric(1) gives the src side -e)
ode_lag_f = DeviceFunction("ode_lag!", "edge_f", A.ND_EM_ODE_LAG, 1, 1, "test/ComponentLibrary.jl:30-32")
# de[1] = ext[1]: the edge of test/external_inputs_test.jl:45-47
feext = DeviceFunction("feext", "edge_f", A.ND_EM_EXT_DRIVEN, 1, 0, "test/external_inputs_test.jl:45-47", ext=1)
resistor_g = DeviceFunction("resistor_g", "edge_g", A.ND_EM_RESISTOR, 0, 1, "docs/src/external_inputs.md:66-72")
f_load = DeviceFunction("f_load", "vertex_f", A.ND_VM_CAPACITOR_SINK, 1, 1, "docs/src/external_inputs.md:76-82")
f_source = DeviceFunction("f_source", "vertex_f", A.ND_VM_PI_SOURCE, 2, 4, "docs/src/external_inputs.md:86-99",
                          ext=1)
# docs/examples/stress_on_truss.jl:45-84: the truss's free and fixed nodes (two outputs: the
# position) and its beams (two outputs: the force)
truss_free_f = DeviceFunction("free_f", "vertex_f", A.ND_VM_TRUSS_FREE, 4, 3, "docs/examples/stress_on_truss.jl:55-61")
truss_fixed_g = DeviceFunction("fixed_g", "vertex_g", A.ND_VM_TRUSS_FIXED, 0, 2, "docs/examples/stress_on_truss.jl:45-48",
                               ff=NoFeedForward())
truss_fixed_g.outdim = 2
beam_g = DeviceFunction("edge_g!", "edge_g", A.ND_EM_BEAM, 0, 2, "docs/examples/stress_on_truss.jl:77-84")

# docs/examples/directed_and_weighted_graphs.jl:89-101: FitzHugh-Nagumo neurons coupled by
# directed, weighted electrical gap junctions
fhn_electrical_vertex = DeviceFunction("fhn_electrical_vertex!", "vertex_f", A.ND_VM_FHN, 2, 2,
                                       "docs/examples/directed_and_weighted_graphs.jl:89-93")
electrical_edge_g = DeviceFunction("electrical_edge!", "edge_g", A.ND_EM_ELECTRICAL, 0, 2,
                                   "docs/examples/directed_and_weighted_graphs.jl:97-100")
# test/loopback_test.jl:19-131: the components of the injector network and of its merged twin
capacitor_f = DeviceFunction("capacitor", "vertex_f", A.ND_VM_CAPACITOR, 1, 1, "test/loopback_test.jl:53-63")
voltage_source_g = DeviceFunction("voltage_source", "vertex_g", A.ND_VM_VOLTAGE_SOURCE, 0, 1,
                                  "test/loopback_test.jl:19-34", ff=NoFeedForward())
resistor_injector_g = DeviceFunction("resistor_injector", "vertex_g", A.ND_VM_RESISTOR_INJECTOR, 0, 1,
                                     "test/loopback_test.jl:66-78", ff=PureFeedForward())
inductor_injector_f = DeviceFunction("inductor_injector", "vertex_f", A.ND_VM_INDUCTOR_INJECTOR, 1, 1,
                                     "test/loopback_test.jl:80-92")
crl_f = DeviceFunction("crl", "vertex_f", A.ND_VM_CRL, 2, 3, "test/loopback_test.jl:113-131")
loopback_g = DeviceFunction("LOOPBACK_G", "edge_g", A.ND_EM_LOOPBACK, 0, 0, "src/post_utils.jl:105-108")
