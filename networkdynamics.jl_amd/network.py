"""Network: the reference's `Network(g, vertexm, edgem; execution, aggregator)`
constructor and its `nw(du, u, p, t)` call, executed by libnd_hip.so.

Reference: src/construction.jl:31-236 (constructor), src/coreloop.jl:1-109
(call, get_buffers), src/executionstyles.jl (execution styles),
src/aggregators.jl (aggregators).  Device arrays are torch CUDA (HIP) tensors;
work is queued on torch's current stream.
"""
import ctypes as C
import operator

import numpy as np

from . import _abi as A
from . import components as comp
from . import graphs


class ExecutionStyle:
    """ExecutionStyle{buffered} (executionstyles.jl:13)."""

    def __init__(self, buffered=True):
        self.buffered = bool(buffered)

    def __repr__(self):
        return f"{type(self).__name__}{{{str(self.buffered).lower()}}}()"


class HIPExecution(ExecutionStyle):
    """Fused HIP kernels on the current device.  `buffered` is kept for API
    parity with the reference's styles; both values run the same fused kernel
    (the gather is fused into the edge evaluation)."""


def usebuffer(ex):
    return ex.buffered


class Aggregator:
    def __init__(self, f=operator.add):
        if f not in (operator.add, "+", sum):
            raise A.ArgumentError(f"{type(self).__name__} only supports + as reducer")
        self.f = operator.add


class SegmentedAggregator(Aggregator):
    """Deterministic destination-segmented reduction (no float atomics).  Segments
    up to `hub_degree` entries are summed in SequentialAggregator's order
    (aggregators.jl:140-151); longer ones by a fixed wave-wide tree."""


def _as_list(models, n, name):
    if isinstance(models, comp.ComponentModel):
        return [models] * n, True
    models = list(models)
    if len(models) != n:
        raise A.ArgumentError(f"expected {n} {name} models, got {len(models)}")
    return models, False


def _edge_outsym_flat(m):
    """An edge's output symbols, src side then dst side (outsym_flat with the naming of
    _symvec_to_sym_tup, component_functions.jl:1040-1075): a single-sided wrapper over
    names s gives AntiSymmetric src "₋s" / dst s, Symmetric s / s and Directed dst s
    only, s being the outsym given or, for a StateMask, the masked state names.  For
    other edges the outsym list names the dst-side outputs."""
    w = getattr(m, "wrapper", None)
    if w in (A.ND_WRAP_ANTISYMMETRIC, A.ND_WRAP_SYMMETRIC, A.ND_WRAP_DIRECTED):
        base = list(m.outsym) if m.outsym else [m.sym[j - 1] for j in m.mask_dst]
        if w == A.ND_WRAP_DIRECTED:
            return base
        return (["₋" + s for s in base] if w == A.ND_WRAP_ANTISYMMETRIC else list(base)) + base
    return [None] * m.outdim_src + list(m.outsym or [])


def _tensor_ptr(x, n, what, torch):
    if not isinstance(x, torch.Tensor):
        raise A.ArgumentError(f"{what} must be a torch tensor on the HIP device")
    if x.dtype != torch.float64:
        # network_structure.jl:146-151: caches are Float64
        raise A.ArgumentError(f"Network caches are initialized with Float64 but is used for {x.dtype} data!")
    if x.device.type != "cuda":
        raise A.ArgumentError(f"{what} must live on the HIP device, got {x.device}")
    if x.dim() != 1 or not x.is_contiguous():
        raise A.ArgumentError(f"{what} must be a contiguous vector")
    return x.data_ptr(), x.numel()


class Network:
    """Network(g, vertexm, edgem; execution=HIPExecution(True), aggregator=SegmentedAggregator(+))."""

    def __init__(self, g, vertexm, edgem, execution=None, aggregator=None, check_graphelement=True,
                 dealias=False, verbose=False):
        if execution is None:
            execution = HIPExecution(True)
        if aggregator is None:
            aggregator = SegmentedAggregator(operator.add)
        if isinstance(aggregator, type):
            aggregator = aggregator(operator.add)
        if not isinstance(execution, HIPExecution):
            raise A.ArgumentError(f"Execution type {execution} not supported (choose HIPExecution)")
        if not isinstance(aggregator, SegmentedAggregator):
            raise A.ArgumentError(f"Aggregator {aggregator} not supported (choose SegmentedAggregator)")
        if not isinstance(g, graphs.SimpleGraph):
            raise A.ArgumentError("g must be a SimpleGraph / SimpleDiGraph")
        self.g = g
        self.execution = execution
        self.aggregator = aggregator
        nv, ne = g.nv(), g.ne()
        vm, _ = _as_list(vertexm, nv, "vertex")
        em, _ = _as_list(edgem, ne, "edge")
        for m in vm:
            if not isinstance(m, comp.VertexModel):
                raise A.ArgumentError(f"Expected VertexModels, got {type(m)}")
        for m in em:
            if not isinstance(m, comp.EdgeModel):
                raise A.ArgumentError(f"Expected EdgeModels, got {type(m)}")
        if dealias:
            vm = [m.copy() for m in vm]
            em = [m.copy() for m in em]
        self.vertexm, self.edgem = vm, em
        # kind table: one entry per distinct component hash, in first-appearance order
        kinds, key_to_kind = [], {}
        vkind = np.empty(nv, np.int32)
        ekind = np.empty(max(ne, 1), np.int32)
        cache = {}
        for arr, models in ((vkind, vm), (ekind, em)):
            for i, m in enumerate(models):
                k = cache.get(id(m))
                if k is None:
                    h = m.component_hash()
                    k = key_to_kind.get(h)
                    if k is None:
                        k = len(kinds)
                        key_to_kind[h] = k
                        kinds.append(m.kind())
                    cache[id(m)] = k
                arr[i] = k
        self.kinds = kinds
        self.vertex_kind, self.edge_kind = vkind, ekind[:ne]
        karr = (A.NDKind * len(kinds))()
        for a, k in zip(karr, kinds):
            for f in ("is_edge", "model", "wrapper", "ff", "dim", "pdim", "outdim_src", "outdim_dst", "extdim"):
                setattr(a, f, int(k[f]))
            for j, x in enumerate(k["mask_src"]):
                a.mask_src[j] = x
            for j, x in enumerate(k["mask_dst"]):
                a.mask_dst[j] = x
        L = A.lib()
        h = C.c_void_p()
        A.check(L.nd_network_create(nv, ne, g.src.ctypes.data, g.dst.ctypes.data, len(kinds),
                                    C.cast(karr, C.c_void_p), vkind.ctypes.data, ekind.ctypes.data, C.byref(h)))
        self._net = h
        info = A.NDNetworkInfo()
        A.check(L.nd_network_get_info(h, C.byref(info)))
        self.info = info
        self._plan = None
        if info.lastidx_extbuf > 0:
            m = self.extmap()
            A.check(L.nd_network_set_extmap(h, len(m), m.ctypes.data))
        if verbose:
            print(f"Create dynamic network with {nv} vertices and {ne} edges:")
            for b in self.batches(False):
                print(f" - VertexBatch: dim={kinds[b['kind']]['dim']}, pdim={kinds[b['kind']]['pdim']}, length={b['count']}")
            for b in self.batches(True):
                print(f" - EdgeBatch: dim={kinds[b['kind']]['dim']}, pdim={kinds[b['kind']]['pdim']}, length={b['count']}")

    def __del__(self):
        L = getattr(A, "_lib", None) if A is not None else None
        if L is None:
            return
        if getattr(self, "_plan", None):
            L.nd_plan_destroy(self._plan)
            self._plan = None
        if getattr(self, "_net", None):
            L.nd_network_destroy(self._net)
            self._net = None

    # ---- sizes (network_structure.jl:119-135) --------------------------------------------
    @property
    def dim(self):
        return int(self.info.dim)

    @property
    def pdim(self):
        return int(self.info.pdim)

    def nv(self):
        return int(self.info.nv)

    def ne(self):
        return int(self.info.ne)

    # ---- layout exports (integer, bit-exact with the reference's IndexManager) ------------
    def batches(self, is_edge):
        out = []
        n = self.info.nebatches if is_edge else self.info.nvbatches
        for b in range(n):
            x = A.NDBatch()
            A.check(A.lib().nd_network_get_batch(self._net, int(is_edge), b, C.byref(x)))
            idx = np.ctypeslib.as_array(x.indices, shape=(x.count,)).copy() if x.count else np.zeros(0, np.int64)
            out.append(dict(kind=x.kind, count=x.count, indices=idx, state_first=x.state_first,
                            param_first=x.param_first, out_first=x.out_first, in_first=x.in_first))
        return out

    def ranges(self):
        """(v_ranges[nv,4,2], e_ranges[ne,6,2]) 1-based (first, last): v_data, v_out, v_para, v_aggr /
        e_data, e_out.src, e_out.dst, e_para, e_gbufr.src, e_gbufr.dst."""
        v = np.zeros((self.nv(), 4, 2), np.int64)
        e = np.zeros((max(self.ne(), 1), 6, 2), np.int64)
        A.check(A.lib().nd_network_export_ranges(self._net, v.ctypes.data, e.ctypes.data))
        return v, e[: self.ne()]

    def maps(self):
        """(AggregationMap.range, AggregationMap.map, EagerGBufProvider.map), 1-based."""
        i = self.info
        n = max(i.agg_last - i.agg_first + 1, 0)
        a = np.zeros(max(n, 1), np.int64)
        gm = np.zeros(max(i.lastidx_gbuf, 1), np.int64)
        A.check(A.lib().nd_network_export_maps(self._net, a.ctypes.data, gm.ctypes.data))
        return (int(i.agg_first), int(i.agg_last)), a[:n], gm[: i.lastidx_gbuf]

    def csr(self):
        nnz = C.c_int64()
        A.check(A.lib().nd_network_csr_size(self._net, C.byref(nnz)))
        rp = np.zeros(self.info.lastidx_aggr + 1, np.int64)
        cols = np.zeros(max(nnz.value, 1), np.int64)
        A.check(A.lib().nd_network_export_csr(self._net, rp.ctypes.data, cols.ctypes.data))
        return rp, cols[: nnz.value]

    # ---- external inputs (external_inputs.jl) ------------------------------------------------
    def ext_ranges(self):
        """(v_ext[nv, 2], e_ext[ne, 2]): 1-based (first, last) external-buffer ranges."""
        v = np.zeros((self.nv(), 2), np.int64)
        e = np.zeros((max(self.ne(), 1), 2), np.int64)
        A.check(A.lib().nd_network_export_ext_ranges(self._net, v.ctypes.data, e.ctypes.data))
        return v, e[: self.ne()]

    def _component(self, ref):
        models = self.vertexm if ref.kind == "V" else self.edgem
        c = ref.comp
        if isinstance(c, str):
            hits = [i for i, m in enumerate(models) if m.name == c]
            if len(hits) != 1:
                raise A.ArgumentError(f"Cannot resolve external input {ref!r}: {len(hits)} components named {c!r}")
            return hits[0]
        i = int(c) - 1
        if not 0 <= i < len(models):
            raise A.ArgumentError(f"Cannot resolve external input {ref!r}: no such component")
        return i

    def _symidx_to_extidx(self, ref):
        """external_inputs.jl:33-46: +u-index (StateBufIdx) or -output-index (OutBufIdx)."""
        i = self._component(ref)
        v, e = self.ranges()
        if ref.kind == "V":
            m, first = self.vertexm[i], v[i, 0, 0]
        else:
            m, first = self.edgem[i], e[i, 0, 0]
        if isinstance(ref.sym, (int, np.integer)):
            if 1 <= int(ref.sym) <= m.dim:
                return int(first + int(ref.sym) - 1)
        elif ref.sym in m.sym:
            return int(first + m.sym.index(ref.sym))
        elif ref.kind == "E" and ref.sym in _edge_outsym_flat(m):
            # the edge's outputs: src side, then dst side from e[i, 2, 0]
            return -int(e[i, 2, 0] - m.outdim_src + _edge_outsym_flat(m).index(ref.sym))
        raise A.ArgumentError(f"Cannot resolve external input {ref!r}! External inputs musst be states or "
                              "outputs of non-feed-forward components.")

    def extmap(self):
        """ExtMap (external_inputs.jl:10-31) as +u / -o 1-based indices per external-buffer entry."""
        n = int(self.info.lastidx_extbuf)
        if n == 0:
            raise A.ArgumentError("There are no external inputs.")
        out = np.zeros(n, np.int64)
        ve, ee = self.ext_ranges()
        for rng, models in ((ve, self.vertexm), (ee, self.edgem)):
            for i, m in enumerate(models):
                for j, ref in enumerate(m.extin):
                    out[rng[i, 0] - 1 + j] = self._symidx_to_extidx(ref)
        return out

    def collect_externals(self, extbuf, u, o):
        """collect_externals!(map, extbuf, u, o) (external_inputs.jl:49-77) on the device."""
        import torch
        eptr, n = _tensor_ptr(extbuf, self.info.lastidx_extbuf, "extbuf", torch)
        uptr, _ = _tensor_ptr(u, self.dim, "u", torch)
        optr, _ = _tensor_ptr(o, self.info.lastidx_out, "o", torch)
        if n != self.info.lastidx_extbuf or u.numel() != self.dim or o.numel() != self.info.lastidx_out:
            raise A.ArgumentError("extbuf / u / o sizes do not match the network")
        stream = torch.cuda.current_stream(u.device).cuda_stream
        A.check(A.lib().nd_collect_externals(self.plan, eptr, uptr, optr, stream))

    # ---- mass matrix (construction.jl:409-453) ----------------------------------------------
    @property
    def mass_matrix(self):
        """construct_mass_matrix: `I` when every component's mass matrix equals the
        identity, a `Diagonal` when all others are diagonal, else a dense (dim, dim) array."""
        if getattr(self, "_mass", None) is None:
            keep = []

            def arr(models):
                n = len(models)
                a = (A.NDMass * max(n, 1))()
                for x, m in zip(a, models):
                    (form, buf), k = comp._mass_struct(m.mass_matrix)
                    x.form = form
                    x.m = buf.ctypes.data if buf is not None else None
                    keep.append(k)
                return a

            va, ea = arr(self.vertexm), arr(self.edgem)
            L, shape = A.lib(), C.c_int32()
            A.check(L.nd_network_mass_matrix(self._net, C.cast(va, C.c_void_p), C.cast(ea, C.c_void_p),
                                             C.byref(shape), None))
            if shape.value == A.ND_MASS_IDENTITY:
                self._mass = comp.I
            elif shape.value == A.ND_MASS_DIAGONAL:
                d = np.empty(self.dim)
                A.check(L.nd_network_mass_matrix(self._net, C.cast(va, C.c_void_p), C.cast(ea, C.c_void_p),
                                                 C.byref(shape), d.ctypes.data))
                self._mass = comp.Diagonal(d)
            else:
                m = np.empty((self.dim, self.dim), order="F")
                A.check(L.nd_network_mass_matrix(self._net, C.cast(va, C.c_void_p), C.cast(ea, C.c_void_p),
                                                 C.byref(shape), m.ctypes.data))
                self._mass = np.ascontiguousarray(m)
        return self._mass

    def _require_identity_mass(self, what):
        if not comp.is_identity_mass(self.mass_matrix):
            raise A.UnsupportedError(f"{what} is an explicit method and cannot use a mass matrix "
                                     "(the network has algebraic or scaled states)")

    def variable_symbols(self):
        """The state symbols in flat state order: ("V", i, sym) / ("E", i, sym) with 1-based
        component indices, the VIndex / EIndex entries of SII.variable_symbols(nw)."""
        out = [None] * self.dim
        v, e = self.ranges()
        for i, m in enumerate(self.vertexm):
            for j, s in enumerate(m.sym):
                out[v[i, 0, 0] - 1 + j] = ("V", i + 1, s)
        for i, m in enumerate(self.edgem):
            for j, s in enumerate(m.sym):
                out[e[i, 0, 0] - 1 + j] = ("E", i + 1, s)
        return out

    def parameter_symbols(self):
        """The parameter symbols in flat parameter order, like variable_symbols."""
        out = [None] * self.pdim
        v, e = self.ranges()
        for i, m in enumerate(self.vertexm):
            for j, s in enumerate(m.psym):
                out[v[i, 2, 0] - 1 + j] = ("V", i + 1, s)
        for i, m in enumerate(self.edgem):
            for j, s in enumerate(m.psym):
                out[e[i, 3, 0] - 1 + j] = ("E", i + 1, s)
        return out

    # ---- parameters / states ----------------------------------------------------------------
    def default_p(self):
        """Flat parameter vector from the models' defaults (NaN where none)."""
        p = np.full(self.pdim, np.nan)
        v, e = self.ranges()
        for i, m in enumerate(self.vertexm):
            for j, d in enumerate(m.psym_default):
                if d is not None:
                    p[v[i, 2, 0] - 1 + j] = d
        for i, m in enumerate(self.edgem):
            for j, d in enumerate(m.psym_default):
                if d is not None:
                    p[e[i, 3, 0] - 1 + j] = d
        return p

    # ---- device plan ---------------------------------------------------------------------
    @property
    def plan(self):
        if self._plan is None:
            import torch  # noqa: F401  (initialises the HIP runtime on torch's device)
            h = C.c_void_p()
            A.check(A.lib().nd_plan_create(self._net, C.byref(h)))
            self._plan = h
        return self._plan

    def tune(self, hub_degree=0, block_size=0, big_hub_degree=0, rk4_path=0, gather_mode=0):
        """0 keeps a setting.  rk4_path: 1 automatic (one workgroup for small single-batch
        networks), 2 four launches per RK4 step.  gather_mode: 1 automatic, 2 neighbour
        outputs read from u, 3 compact output copy plus slot-ordered parameter copy."""
        t = A.NDPlanTuning(int(hub_degree), int(block_size), int(big_hub_degree), int(rk4_path), int(gather_mode))
        A.check(A.lib().nd_plan_tune(self.plan, C.byref(t)))

    def tuning(self):
        t = A.NDPlanTuning()
        A.check(A.lib().nd_plan_get_tuning(self.plan, C.byref(t)))
        return dict(hub_degree=t.hub_degree, block_size=t.block_size, big_hub_degree=t.big_hub_degree,
                    rk4_path=t.rk4_path, gather_mode=t.gather_mode)

    def __call__(self, du, u, p, t=0.0, perturb=None, perturb_maps=None):
        """nw(du, u, p, t; perturb, perturb_maps): coreloop.jl:1-102.  du is overwritten;
        stream-ordered on torch's current stream.  perturb: channel values added at the
        buffer positions of perturb_maps (vo_map / vi_map / eo_map / ei_map, 1-based
        (buffer index, channel) pairs, as linear_analysis.perturbation_channels builds
        them); a parameter map is the caller's to apply to p."""
        import torch
        if perturb is not None:
            self.perturbed_jvp(du, None, u, None, p, t, perturb, None, perturb_maps)
            return None
        dptr, ndu = _tensor_ptr(du, self.dim, "du", torch)
        uptr, nu = _tensor_ptr(u, self.dim, "u", torch)
        if self.pdim > 0:
            pptr, np_ = _tensor_ptr(p, self.pdim, "p", torch)
        else:
            pptr, np_ = (p.data_ptr(), p.numel()) if isinstance(p, torch.Tensor) else (None, 0)
        stream = torch.cuda.current_stream(du.device).cuda_stream
        A.check(A.lib().nd_rhs(self.plan, dptr, uptr, pptr, float(t), ndu, nu, np_, stream))
        return None

    def jvp(self, du, jv, u, v, p, t=0.0):
        """du = f(u) and jv = J(u) v: nw evaluated on dual numbers u + eps*v, the
        Jacobian-vector product ForwardDiff forms in linearize_network
        (linear_analysis.jl:74-85).  Stream-ordered on torch's current stream."""
        import torch
        dptr, ndu = _tensor_ptr(du, self.dim, "du", torch)
        jptr, njv = _tensor_ptr(jv, self.dim, "jv", torch)
        uptr, nu = _tensor_ptr(u, self.dim, "u", torch)
        vptr, nv = _tensor_ptr(v, self.dim, "v", torch)
        if njv != ndu or nv != nu:      # the tangent vectors have the state's layout
            raise A.ArgumentError(f"v and jv must have the length of u and du ({self.dim})")
        if self.pdim > 0:
            pptr, np_ = _tensor_ptr(p, self.pdim, "p", torch)
        else:
            pptr, np_ = (p.data_ptr(), p.numel()) if isinstance(p, torch.Tensor) else (None, 0)
        stream = torch.cuda.current_stream(du.device).cuda_stream
        A.check(A.lib().nd_rhs_jvp(self.plan, dptr, jptr, uptr, vptr, pptr, float(t), ndu, nu, np_, stream))

    def jvp_multi(self, du, jv, u, v, p, t=0.0):
        """du = f(u) and jv[j] = J(u) v[j] for the k = v.shape[0] <= 8 rows of v in one
        launch (nd_rhs_jvp_multi): the chunked duals ForwardDiff carries through the
        network's DiffCache buffers (construction.jl:205-208).  v, jv: (k, dim) float64
        device tensors, C-contiguous.  Raises UnsupportedError for networks whose vertex
        batches take the generic kernels (jvp serves them one tangent at a time)."""
        import torch
        dptr, ndu = _tensor_ptr(du, self.dim, "du", torch)
        uptr, nu = _tensor_ptr(u, self.dim, "u", torch)
        for name, x in (("v", v), ("jv", jv)):
            if not isinstance(x, torch.Tensor) or x.dtype != torch.float64 or x.device.type != "cuda" or \
                    x.dim() != 2 or x.shape[1] != self.dim or not x.is_contiguous():
                raise A.ArgumentError(f"{name} must be a contiguous float64 device tensor of shape (k, {self.dim})")
        k = v.shape[0]
        if jv.shape[0] != k or not 1 <= k <= 8:
            raise A.ArgumentError(f"v and jv need the same number of rows k, 1 <= k <= 8 (got {k}, {jv.shape[0]})")
        if self.pdim > 0:
            pptr, np_ = _tensor_ptr(p, self.pdim, "p", torch)
        else:
            pptr, np_ = (p.data_ptr(), p.numel()) if isinstance(p, torch.Tensor) else (None, 0)
        stream = torch.cuda.current_stream(du.device).cuda_stream
        A.check(A.lib().nd_rhs_jvp_multi(self.plan, dptr, jv.data_ptr(), uptr, v.data_ptr(), int(k), pptr, float(t),
                                         ndu, nu, np_, stream))

    def jacobian(self, u, p, t=0.0, colored=True):
        """Dense Jacobian d(du)/du at u (dim x dim, float64 device tensor): the matrix
        ForwardDiff.jacobian builds in linearize_network (linear_analysis.jl:84).  colored:
        one JVP per color of the structural pattern (sparsity.py), scattered into the dense
        matrix; otherwise one JVP per unit vector."""
        import torch
        if colored:
            from .sparsity import color_columns, jac_prototype, sparse_jacobian
            if getattr(self, "_jac_colors", None) is None:
                P = jac_prototype(self)
                self._jac_colors = (P, color_columns(P)[0])
            P, colors = self._jac_colors
            J = sparse_jacobian(self, u, p, t, P, colors).toarray()
            return torch.from_numpy(J).to(u.device)
        n = self.dim
        J = torch.empty((n, n), dtype=torch.float64, device=u.device)
        du = torch.empty(n, dtype=torch.float64, device=u.device)
        e = torch.zeros(n, dtype=torch.float64, device=u.device)
        col = torch.empty(n, dtype=torch.float64, device=u.device)
        for j in range(n):
            e.zero_()
            e[j] = 1.0
            self.jvp(du, col, u, e, p, t)
            J[:, j] = col
        return J

    def rk4(self, u, p, t0, dt, nsteps, work=None):
        """Fixed-step classic RK4 in place on the device vector u (nsteps steps of dt).
        Networks with a mass matrix other than I raise UnsupportedError, as explicit
        OrdinaryDiffEq methods do."""
        import torch
        self._require_identity_mass("rk4")
        uptr, nu = _tensor_ptr(u, self.dim, "u", torch)
        if self.pdim > 0:
            pptr, np_ = _tensor_ptr(p, self.pdim, "p", torch)
        else:
            pptr, np_ = None, 0
        wptr = None
        if work is not None:
            wptr, nw_ = _tensor_ptr(work, 3 * self.dim, "work", torch)
            if nw_ < 3 * self.dim:
                raise A.ArgumentError("work needs 3*dim entries")
        stream = torch.cuda.current_stream(u.device).cuda_stream
        A.check(A.lib().nd_rk4(self.plan, uptr, pptr, float(t0), float(dt), int(nsteps), wptr, nu, np_, stream))

    def tsit5(self, u, p, tspan, reltol=1e-8, abstol=1e-8, dt0=0.0, maxiters=10 ** 6):
        """Adaptive Tsit5 (Tsitouras 5(4) pair, OrdinaryDiffEq's PI controller) in place on
        the device vector u from tspan[0] to tspan[1]; returns t, naccept, nreject, retcode
        and the path taken.  retcode: 0 reached tspan[1], 1 maxiters, 2 step size below
        1e-14*max(1, |t1|), 3 NaN in the step size or the error estimate.

        Networks that nd_tsit5 runs in one workgroup (one vertex batch of Kuramoto or swing
        vertices, one batch of AntiSymmetric Kuramoto or static-line edges, up to 2048
        vertices) take every step in one launch; any other network steps on the host, one
        device RHS per stage, with the same tableau and controller."""
        import torch
        from .ensemble import STAT_DTYPE, tsit5_per_instance
        self._require_identity_mass("tsit5")
        uptr, nu = _tensor_ptr(u, self.dim, "u", torch)
        if self.pdim > 0:
            pptr, np_ = _tensor_ptr(p, self.pdim, "p", torch)
        else:
            pptr, np_ = None, 0
        stats = torch.zeros(STAT_DTYPE.itemsize // 8, dtype=torch.float64, device=u.device)
        stream = torch.cuda.current_stream(u.device).cuda_stream
        rc = A.lib().nd_tsit5(self.plan, uptr, pptr, float(tspan[0]), float(tspan[1]), float(reltol), float(abstol),
                              float(dt0), int(maxiters), stats.data_ptr(), nu, np_, stream)
        if rc == A.ND_ERR_UNSUPPORTED:
            info = tsit5_per_instance(self, u, p, tspan, reltol, abstol, dt0, maxiters)
            info.update(path="host steps")
            return info
        A.check(rc)
        st = stats.cpu().numpy().view(STAT_DTYPE)[0]
        return dict(t=float(st["t"]), naccept=int(st["naccept"]), nreject=int(st["nreject"]),
                    retcode=int(st["retcode"]), path="one workgroup")

    def get_buffers(self, u, p, t=0.0):
        """get_buffers(nw, u, p, t; initbufs=true): (o, aggbuf) device tensors, coreloop.jl:103-109."""
        import torch
        uptr, _ = _tensor_ptr(u, self.dim, "u", torch)
        pptr = p.data_ptr() if isinstance(p, torch.Tensor) and p.numel() else None
        o = torch.empty(self.info.lastidx_out, dtype=torch.float64, device=u.device)
        agg = torch.empty(self.info.lastidx_aggr, dtype=torch.float64, device=u.device)
        stream = torch.cuda.current_stream(u.device).cuda_stream
        A.check(A.lib().nd_get_buffers(self.plan, o.data_ptr(), agg.data_ptr() if agg.numel() else None,
                                       uptr, pptr, float(t), stream))
        return o, agg

    def _set_perturbation(self, ndelta, maps):
        key = (int(ndelta),) + tuple(tuple(map(tuple, maps.get(k, []) or [])) for k in
                                     ("vo_map", "vi_map", "eo_map", "ei_map"))
        if getattr(self, "_pert_key", None) == key:
            return
        arrs = [np.ascontiguousarray(np.asarray(maps.get(k, []) or [], np.int64).reshape(-1, 2))
                for k in ("vo_map", "vi_map", "eo_map", "ei_map")]
        args = []
        for a in arrs:
            args += [len(a), a.ctypes.data if len(a) else None]
        A.check(A.lib().nd_plan_set_perturbation(self.plan, int(ndelta), *args))
        self._pert_key = key

    def perturbed_jvp(self, du, jv, u, v, p, t, delta, ddelta, perturb_maps, buffers=False, dp=None):
        """nw(du, u + eps v, p + eps dp, t; perturb = delta + eps ddelta, perturb_maps) on
        dual numbers (the buffered loop of coreloop.jl:1-102 with the four perturbation
        points): du = f, jv = its directional derivative (None: not wanted).  v, dp,
        ddelta: None for zero tangents.  With buffers=True returns (o, do, aggbuf, dagg),
        the perturbed buffers and their tangents."""
        import torch
        dptr, _ = _tensor_ptr(du, self.dim, "du", torch)
        uptr, _ = _tensor_ptr(u, self.dim, "u", torch)
        jptr = _tensor_ptr(jv, self.dim, "jv", torch)[0] if jv is not None else None
        vptr = _tensor_ptr(v, self.dim, "v", torch)[0] if v is not None else None
        delta = torch.as_tensor(np.asarray(delta, np.float64) if not isinstance(delta, torch.Tensor) else delta,
                                dtype=torch.float64, device=u.device).contiguous()
        self._set_perturbation(delta.numel(), perturb_maps or {})
        if ddelta is not None:
            ddelta = torch.as_tensor(np.asarray(ddelta, np.float64) if not isinstance(ddelta, torch.Tensor)
                                     else ddelta, dtype=torch.float64, device=u.device).contiguous()
            if ddelta.numel() != delta.numel():
                raise A.ArgumentError("ddelta must have the length of delta")
        if self.pdim > 0:
            pptr, _ = _tensor_ptr(p, self.pdim, "p", torch)
        else:
            pptr = p.data_ptr() if isinstance(p, torch.Tensor) and p.numel() else None
        dpptr = _tensor_ptr(dp, self.pdim, "dp", torch)[0] if dp is not None else None
        bufs = None
        if buffers:
            bufs = [torch.empty(n, dtype=torch.float64, device=u.device) for n in
                    (self.info.lastidx_out, self.info.lastidx_out, self.info.lastidx_aggr, self.info.lastidx_aggr)]
        bp = [b.data_ptr() if b is not None and b.numel() else None for b in (bufs or [None] * 4)]
        stream = torch.cuda.current_stream(u.device).cuda_stream
        A.check(A.lib().nd_rhs_perturbed_jvp(self.plan, dptr, jptr, uptr, vptr, pptr, dpptr, float(t),
                                             delta.data_ptr() if delta.numel() else None,
                                             ddelta.data_ptr() if ddelta is not None and ddelta.numel() else None,
                                             *bp, stream))
        return tuple(bufs) if buffers else None

    def gather(self, gbuf, o):
        """gather!(EagerGBufProvider, gbuf, o): gbuf[k] = o[map[k]] (gbufs.jl:25)."""
        import torch
        stream = torch.cuda.current_stream(o.device).cuda_stream
        A.check(A.lib().nd_gather(self.plan, gbuf.data_ptr(), o.data_ptr(), stream))

    def aggregate(self, aggbuf, o):
        """aggregate!(aggregator, aggbuf, o): additive (aggregators_test.jl:69-79)."""
        import torch
        stream = torch.cuda.current_stream(o.device).cuda_stream
        A.check(A.lib().nd_aggregate(self.plan, aggbuf.data_ptr(), o.data_ptr(), stream))

    def __repr__(self):
        return (f"Network with {self.dim} states and {self.pdim} parameters\n"
                f" ├ {self.nv()} vertices ({self.info.nvbatches} batches)\n"
                f" └ {self.ne()} edges ({self.info.nebatches} batches)")
