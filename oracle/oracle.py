"""ctypes front end of the CPU oracle (nd_oracle.c).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker, never by the product.
It restates NetworkDynamics.jl's construction (src/construction.jl:31-256,
src/network_structure.jl:224-289), maps (src/aggregators.jl:54-86,203-235,
src/gbufs.jl:11-22) and the SequentialExecution coreloop (src/coreloop.jl:1-102);
see the header of nd_oracle.c for the per-function citations and pinning.
`OracleNet.mass_matrix` restates construct_mass_matrix (src/construction.jl:409-453)
in numpy; it is pinned by the known answers of test/construction_test.jl:57-86 and
test/massmatrix_test.jl:20-47.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libnd_oracle.so")
ND_MAX_OUT = 4


class NDKind(C.Structure):
    _fields_ = [("is_edge", C.c_int32), ("model", C.c_int32), ("wrapper", C.c_int32),
                ("ff", C.c_int32), ("dim", C.c_int32), ("pdim", C.c_int32),
                ("outdim_src", C.c_int32), ("outdim_dst", C.c_int32),
                ("mask_src", C.c_int32 * ND_MAX_OUT), ("mask_dst", C.c_int32 * ND_MAX_OUT),
                ("extdim", C.c_int32)]


def build():
    """Compile libnd_oracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "libnd_oracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
        L.ndo_net_create.restype = P
        L.ndo_net_create.argtypes = [C.c_int64, C.c_int64, i64p, i64p, C.c_int32, P, i32p, i32p,
                                     C.c_char_p, C.c_int64]
        L.ndo_net_destroy.argtypes = [P]
        L.ndo_net_info.argtypes = [P, i64p]
        L.ndo_net_ranges.argtypes = [P, i64p, i64p]
        L.ndo_net_maps.argtypes = [P, i64p, i64p]
        L.ndo_net_inv_nnz.restype = C.c_int64
        L.ndo_net_inv_nnz.argtypes = [P]
        L.ndo_net_inv.argtypes = [P, i64p, i64p, i64p]
        L.ndo_net_batch.argtypes = [P, C.c_int32, C.c_int32, i64p, P]
        L.ndo_net_extbuf_len.restype = C.c_int64
        L.ndo_net_extbuf_len.argtypes = [P]
        L.ndo_net_ext_ranges.argtypes = [P, i64p, i64p]
        L.ndo_net_set_extmap.restype = C.c_int
        L.ndo_net_set_extmap.argtypes = [P, C.c_int64, i64p]
        L.ndo_collect_externals.argtypes = [P, f64p, f64p, f64p]
        L.ndo_rhs.argtypes = [P, P, f64p, f64p, C.c_double, C.c_int32, P, P]
        L.ndo_rhs.restype = C.c_int
        L.ndo_rhs_perturbed.argtypes = [P, P, f64p, f64p, C.c_double, f64p] + [P, C.c_int64] * 4 + [P, P]
        L.ndo_rhs_perturbed.restype = C.c_int
        L.ndo_rhs_threaded.argtypes = [P, f64p, f64p, f64p, C.c_double, C.c_int32]
        L.ndo_rhs_scale.argtypes = [P, f64p, f64p, f64p, C.c_double]
        L.ndo_rk4.argtypes = [P, f64p, f64p, C.c_double, C.c_double, C.c_int64]
        L.ndo_tsit5.argtypes = [P, f64p, f64p] + [C.c_double] * 7 + [C.c_int64, P, C.c_int, C.c_double,
                                                                     C.c_double, C.c_double, C.c_int64, f64p]
        L.ndo_tsit5.restype = C.c_int
        L.ndo_tsit5_tableau.argtypes = [f64p, f64p, f64p]
        L.ndo_tsit5_dense.argtypes = [f64p]
        L.ndo_aggregate_seq.argtypes = [P, f64p, f64p]
        L.ndo_aggregate_inv.argtypes = [P, f64p, f64p, C.c_int32]
        L.ndo_aggregate_naive.argtypes = [P, f64p, f64p]
        L.ndo_gather.argtypes = [P, f64p, f64p]
        L.ndo_check_sizes.argtypes = [P, C.c_int64, C.c_int64, C.c_int64]
        L.ndo_check_sizes.restype = C.c_int
        L.ndo_batchstride_range.argtypes = [C.c_int64, C.c_int32, i32p, C.c_int64, C.c_int32,
                                            C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.ndo_batchstride_fullstride.argtypes = [C.c_int32, i32p]
        L.ndo_batchstride_fullstride.restype = C.c_int64
        _lib = L
    return _lib


def tsit5_tableau():
    """The oracle's Tsit5 coefficients: A (7x7, lower triangular), btilde (7), c (7)."""
    A6 = np.zeros(42)
    bt = np.zeros(7)
    c = np.zeros(7)
    lib().ndo_tsit5_tableau(A6, bt, c)
    A = np.zeros((7, 7))
    A[:, :6] = A6.reshape(7, 6)
    return A, bt, c


def tsit5_dense():
    """The oracle's dense-output table r (7x4): b_i(s) = sum_j r[i, j-1] s^j."""
    r = np.zeros(28)
    lib().ndo_tsit5_dense(r)
    return r.reshape(7, 4)


def make_kind_array(kinds):
    """kinds: sequence of dicts with the nd_kind field names."""
    arr = (NDKind * len(kinds))()
    for a, k in zip(arr, kinds):
        for f in ("is_edge", "model", "wrapper", "ff", "dim", "pdim", "outdim_src", "outdim_dst"):
            setattr(a, f, int(k[f]))
        for j, m in enumerate(k.get("mask_src", ())):
            a.mask_src[j] = int(m)
        for j, m in enumerate(k.get("mask_dst", ())):
            a.mask_dst[j] = int(m)
        a.extdim = int(k.get("extdim", 0))
    return arr


def batchstride_range(first, strides, i, j=0):
    """(first, last) of element i (j=0) or of its sub-range j (1-based), utils.jl:21-29."""
    s = np.asarray(strides, dtype=np.int32)
    a, b = C.c_int64(), C.c_int64()
    lib().ndo_batchstride_range(first, len(s), s, i, j, C.byref(a), C.byref(b))
    return a.value, b.value


def batchstride_fullstride(strides):
    s = np.asarray(strides, dtype=np.int32)
    return int(lib().ndo_batchstride_fullstride(len(s), s))


class OracleNet:
    """The reference's Network restated on the CPU (one instance owns its caches)."""

    def __init__(self, nv, src, dst, kinds, vkind, ekind):
        L = lib()
        self.nv = int(nv)
        self.src = np.ascontiguousarray(src, dtype=np.int64)
        self.dst = np.ascontiguousarray(dst, dtype=np.int64)
        self.ne = len(self.src)
        self._kinds = make_kind_array(kinds)
        vk = np.ascontiguousarray(vkind, dtype=np.int32)
        ek = np.ascontiguousarray(ekind, dtype=np.int32)
        err = C.create_string_buffer(256)
        h = L.ndo_net_create(self.nv, self.ne, self.src, self.dst, len(kinds),
                             C.cast(self._kinds, C.c_void_p), vk, ek, err, 256)
        if not h:
            raise ValueError(err.value.decode())
        self._h = h
        info = np.zeros(11, np.int64)
        L.ndo_net_info(h, info)
        (self.dim, self.pdim, self.lastidx_out, self.lastidx_aggr, self.lastidx_gbuf,
         self.agg_first, self.agg_last, self.vdepth, self.edepth, self.nvb, self.neb) = [int(x) for x in info]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().ndo_net_destroy(h)
            self._h = None

    def ranges(self):
        v = np.zeros((self.nv, 4, 2), np.int64)
        e = np.zeros((max(self.ne, 1), 6, 2), np.int64)
        lib().ndo_net_ranges(self._h, v.reshape(-1), e.reshape(-1))
        return v, e[: self.ne]

    def ext_ranges(self):
        v = np.zeros((self.nv, 2), np.int64)
        e = np.zeros((max(self.ne, 1), 2), np.int64)
        lib().ndo_net_ext_ranges(self._h, v.reshape(-1), e.reshape(-1))
        return v, e[: self.ne]

    @property
    def lastidx_extbuf(self):
        return int(lib().ndo_net_extbuf_len(self._h))

    def set_extmap(self, m):
        m = np.ascontiguousarray(m, np.int64)
        if lib().ndo_net_set_extmap(self._h, len(m), m):
            raise ValueError("ExtMap length differs from the external buffer")

    def collect_externals(self, u, o):
        ext = np.empty(max(self.lastidx_extbuf, 1))
        lib().ndo_collect_externals(self._h, ext, np.ascontiguousarray(u, np.float64), np.ascontiguousarray(o, np.float64))
        return ext[: self.lastidx_extbuf]

    def maps(self):
        n = max(self.agg_last - self.agg_first + 1, 0)
        a = np.zeros(max(n, 1), np.int64)
        g = np.zeros(max(self.lastidx_gbuf, 1), np.int64)
        lib().ndo_net_maps(self._h, a, g)
        return a[:n], g[: self.lastidx_gbuf]

    def inverse_map(self):
        nnz = lib().ndo_net_inv_nnz(self._h)
        d = np.zeros(max(self.lastidx_aggr, 1), np.int64)
        p = np.zeros(self.lastidx_aggr + 1, np.int64)
        s = np.zeros(max(nnz, 1), np.int64)
        lib().ndo_net_inv(self._h, d, p, s)
        return d[: self.lastidx_aggr], p, s[:nnz]

    def batch(self, is_edge, b):
        out = np.zeros(6, np.int64)
        lib().ndo_net_batch(self._h, int(is_edge), int(b), out, None)
        n = int(out[1])
        idx = np.zeros(max(n, 1), np.int64)
        lib().ndo_net_batch(self._h, int(is_edge), int(b), out, idx.ctypes.data_as(C.c_void_p))
        return dict(kind=int(out[0]), count=n, state_first=int(out[2]), param_first=int(out[3]),
                    out_first=int(out[4]), in_first=int(out[5]), indices=idx[:n])

    def check_sizes(self, len_du, len_u, len_p):
        return lib().ndo_check_sizes(self._h, len_du, len_u, len_p)

    def rhs(self, u, p, t=0.0, buffered=True):
        u = np.ascontiguousarray(u, np.float64)
        p = np.ascontiguousarray(p if len(p) else np.zeros(1), np.float64)
        du = np.empty(self.dim, np.float64)
        if self.check_sizes(len(du), len(u), self.pdim if self.pdim == 0 else len(p)):
            raise ValueError("du or u does not have expected size")
        lib().ndo_rhs(self._h, du.ctypes.data_as(C.c_void_p), u, p, float(t), int(buffered), None, None)
        return du

    def rhs_perturbed(self, u, p, delta, maps, t=0.0):
        """nw(du, u, p, t; perturb=delta, perturb_maps): maps = dict of vo_map, vi_map,
        eo_map, ei_map, each a list of 1-based (buffer index, channel) pairs.  Returns
        (du, o, aggbuf) of the perturbed loop."""
        u = np.ascontiguousarray(u, np.float64)
        p = np.ascontiguousarray(p if len(p) else np.zeros(1), np.float64)
        delta = np.ascontiguousarray(delta if len(delta) else np.zeros(1), np.float64)
        du = np.empty(self.dim, np.float64)
        o = np.empty(self.lastidx_out, np.float64)
        a = np.empty(max(self.lastidx_aggr, 1), np.float64)
        arrs = [np.ascontiguousarray(np.asarray(maps.get(k, []), np.int64).reshape(-1, 2))
                for k in ("vo_map", "vi_map", "eo_map", "ei_map")]
        args = []
        for m in arrs:
            args += [m.ctypes.data_as(C.c_void_p) if len(m) else None, len(m)]
        lib().ndo_rhs_perturbed(self._h, du.ctypes.data_as(C.c_void_p), u, p, float(t), delta, *args,
                                o.ctypes.data_as(C.c_void_p), a.ctypes.data_as(C.c_void_p))
        return du, o, a[: self.lastidx_aggr]

    def buffers(self, u, p, t=0.0):
        u = np.ascontiguousarray(u, np.float64)
        p = np.ascontiguousarray(p if len(p) else np.zeros(1), np.float64)
        o = np.empty(self.lastidx_out, np.float64)
        a = np.empty(max(self.lastidx_aggr, 1), np.float64)
        lib().ndo_rhs(self._h, None, u, p, float(t), 1, o.ctypes.data_as(C.c_void_p), a.ctypes.data_as(C.c_void_p))
        return o, a[: self.lastidx_aggr]

    def rhs_threaded(self, u, p, t=0.0, nthreads=1, du=None):
        u = np.ascontiguousarray(u, np.float64)
        p = np.ascontiguousarray(p if len(p) else np.zeros(1), np.float64)
        if du is None:
            du = np.empty(self.dim, np.float64)
        lib().ndo_rhs_threaded(self._h, du, u, p, float(t), int(nthreads))
        return du

    def scale(self, u, p, t=0.0):
        u = np.ascontiguousarray(u, np.float64)
        p = np.ascontiguousarray(p if len(p) else np.zeros(1), np.float64)
        S = np.empty(self.dim, np.float64)
        lib().ndo_rhs_scale(self._h, S, u, p, float(t))
        return S

    def rk4(self, u0, p, t0, dt, nsteps):
        u = np.array(u0, dtype=np.float64, copy=True)
        p = np.ascontiguousarray(p if len(p) else np.zeros(1), np.float64)
        lib().ndo_rk4(self._h, u, p, float(t0), float(dt), int(nsteps))
        return u

    def tsit5(self, u0, p, tspan, reltol=1e-8, abstol=1e-8, dt0=0.0, dtmax=0.0, dtmin=0.0, maxiters=0,
              line_limit=None, steady_state=(0.0, 0.0), preset_trip=(0.0, -1), line_events="crossing"):
        """Adaptive Tsit5 from tspan[0] to tspan[1] (ndo_tsit5).  Returns (u, p, stats):
        p with tripped lines' first parameter at 0; stats: t, dt, naccept, nreject, nf,
        ntrip and retcode (0 reached t1, 1 maxiters, 2 dt < dtmin, 3 NaN, 4 steady state).
        preset_trip: (time, 0-based edge), edge -1 for none.  line_events: "crossing"
        (each crossing of |P| - limit located on the dense output) or "step_end"
        (|P| > limit checked at the end of each accepted step)."""
        if line_events not in ("crossing", "step_end"):
            raise ValueError("line_events must be 'crossing' or 'step_end'")
        u = np.array(u0, dtype=np.float64, copy=True)
        if self.check_sizes(len(u), len(u), self.pdim if self.pdim == 0 else len(p)):
            raise ValueError("u or p does not have expected size")
        p = np.array(p if len(p) else np.zeros(1), dtype=np.float64, copy=True)
        lim = None
        if line_limit is not None:
            lim = np.ascontiguousarray(line_limit, np.float64)
            if len(lim) != self.ne:
                raise ValueError("line_limit needs one entry per edge")
        st = np.zeros(6)
        rc = lib().ndo_tsit5(self._h, u, p, float(tspan[0]), float(tspan[1]), float(reltol), float(abstol),
                             float(dt0), float(dtmax), float(dtmin), int(maxiters),
                             None if lim is None else lim.ctypes.data_as(C.c_void_p),
                             2 if line_events == "crossing" else 1, float(steady_state[0]), float(steady_state[1]), float(preset_trip[0]),
                             int(preset_trip[1]), st)
        stats = dict(t=st[0], dt=st[1], naccept=int(st[2]), nreject=int(st[3]), nf=int(st[4]), ntrip=int(st[5]),
                     retcode=int(rc))
        return u, p[: self.pdim], stats

    def aggregate(self, aggbuf, data, method="seq", nthreads=1):
        aggbuf = np.ascontiguousarray(aggbuf, np.float64)
        data = np.ascontiguousarray(data, np.float64)
        if method == "seq":
            lib().ndo_aggregate_seq(self._h, aggbuf, data)
        elif method == "inv":
            lib().ndo_aggregate_inv(self._h, aggbuf, data, int(nthreads))
        elif method == "naive":
            lib().ndo_aggregate_naive(self._h, aggbuf, data)
        else:
            raise ValueError(method)
        return aggbuf

    def gather(self, o):
        o = np.ascontiguousarray(o, np.float64)
        g = np.empty(max(self.lastidx_gbuf, 1), np.float64)
        lib().ndo_gather(self._h, g, o)
        return g[: self.lastidx_gbuf]

    def mass_matrix(self, vmass, emass):
        """construct_mass_matrix(im), src/construction.jl:409-453, over this net's own ranges.
        vmass[nv], emass[ne]: ("I",), ("uniform", s), ("diag", vector) or ("dense", rows).
        Returns ("identity", None), ("diagonal", vector) or ("dense", matrix)."""
        v, e = self.ranges()
        blocks = []                         # (first, last) state range and its dense block
        for rng, masses in ((v[:, 0, :], vmass), (e[:, 0, :] if self.ne else np.zeros((0, 2), np.int64), emass)):
            for (first, last), mm in zip(rng, masses):
                n = int(last - first + 1)
                kind = mm[0]
                if kind == "I":
                    continue
                if kind == "uniform":       # UniformScaling != I compares the scale only
                    if float(mm[1]) == 1.0:
                        continue
                    block = float(mm[1]) * np.eye(n)
                elif kind == "diag":
                    block = np.diag(np.asarray(mm[1], np.float64).reshape(n))
                elif kind == "dense":
                    block = np.asarray(mm[1], np.float64).reshape(n, n)
                else:
                    raise ValueError(kind)
                if kind != "uniform" and np.array_equal(block, np.eye(n)):
                    continue                # Diagonal / Matrix == I entry by entry
                blocks.append((int(first), n, block))
        if not blocks:
            return "identity", None
        M = np.eye(self.dim)
        for first, n, block in blocks:
            M[first - 1:first - 1 + n, first - 1:first - 1 + n] = block
        if all(np.count_nonzero(b - np.diag(np.diag(b))) == 0 for _, _, b in blocks):
            return "diagonal", np.diag(M).copy()
        return "dense", M
