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

    def __init__(self, name, role, model, dim=None, pdim=0, ref=""):
        self.name, self.role, self.model, self.dim, self.pdim, self.ref = name, role, model, dim, pdim, ref

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
                    mask_src=tuple(self.mask_src), mask_dst=tuple(self.mask_dst))

    def component_hash(self):
        """What batches components together (_component_hash, construction.jl:245-256)."""
        k = self.kind()
        return tuple(sorted(k.items()))

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
                 mass_matrix=None, **kwargs):
        if kwargs.get("extin") is not None:
            raise A.UnsupportedError("external inputs (extin) have no HIP kernel")
        if f is None:
            f = noop_vertex
        if not isinstance(f, DeviceFunction) or f.role != "vertex_f":
            raise A.UnsupportedError(f"vertex f must be a vertex device function of the library, got {f!r}")
        self.f = f
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
        self.mass_matrix = mass_matrix

    def __repr__(self):
        return f"VertexModel :{self.name} (dim={self.dim}, pdim={self.pdim}, f={self.f.name})"


class EdgeModel(ComponentModel):
    """EdgeModel(f=..., g=AntiSymmetric(...)/Fiducial(...)/..., dim, pdim, outdim) -- component_functions.jl:306-362."""

    is_edge = True

    def __init__(self, f=None, g=None, dim=None, sym=None, pdim=None, psym=None, outdim=None, outsym=None,
                 name=None, mass_matrix=None, **kwargs):
        if kwargs.get("extin") is not None:
            raise A.UnsupportedError("external inputs (extin) have no HIP kernel")
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
                raise A.UnsupportedError("StateMask edge outputs need Fiducial in the HIP kernels")
            if not isinstance(inner, DeviceFunction) or inner.role != "edge_g":
                raise A.UnsupportedError(f"edge g must be a library device function, got {inner!r}")
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
        self.mass_matrix = mass_matrix

    def __repr__(self):
        return f"EdgeModel :{self.name} (dim={self.dim}, pdim={self.pdim}, outdim={self.outdim})"


# ---- device functions (the kernels' equations) -------------------------------------

noop_vertex = DeviceFunction("noop_vertex", "vertex_f", A.ND_VM_NOOP, None, 0,
                             "test/aggregators_test.jl:15")
diffusionvertex = DeviceFunction("diffusionvertex!", "vertex_f", A.ND_VM_DIFFUSION, 1, 0,
                                 "test/ComponentLibrary.jl:42-45")
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
