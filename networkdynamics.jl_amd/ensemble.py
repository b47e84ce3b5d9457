"""Batched ensembles of independent networks (basin-stability style), and the
per-instance reference path they are compared with.

`Ensemble(graphs, vertexm, edgem)` -> one device integrator for all instances
(nd_ensemble_* in include/nd_hip.h).  `tsit5_per_instance` integrates one
network at a time with the network RHS kernel and host-side step control: the
plain per-instance GPU path.
"""
import ctypes as C

import numpy as np

from . import _abi as A
from . import components as comp

TSIT5, RK4 = 0, 1
NCLASS = 6
DEFAULT_CLASSES = [(128, 2), (256, 2), (512, 2), (512, 3), (512, 4), (512, 4)]   # (threads, vertices per thread)


def tsit5_tableau():
    """The Tsitouras (2011) 5(4) pair the device steppers use, read from the library
    (nd_tsit5_tableau): A (7x7, strictly lower triangular; row s = stage s+1), btilde =
    b - bhat (7; b is A's last row, first same as last) and the nodes c (7)."""
    a = np.zeros(42)
    bt = np.zeros(7)
    c = np.zeros(7)
    A.check(A.lib().nd_tsit5_tableau(a.ctypes.data, bt.ctypes.data, c.ctypes.data))
    tab = np.zeros((7, 7))
    tab[:, :6] = a.reshape(7, 6)
    return tab, bt, c


def tsit5_dense():
    """The dense output the ensemble kernel locates line crossings on (nd_tsit5_dense):
    r (7x4) with b_i(s) = sum_j r[i, j-1] s^j."""
    r = np.zeros(28)
    A.check(A.lib().nd_tsit5_dense(r.ctypes.data))
    return r.reshape(7, 4)


# when line limits are checked (nd_ensemble_opts.line_events)
LINE_EVENTS = {None: 0, "step_end": 1, "crossing": 2}


class _Opts(C.Structure):
    _fields_ = [("method", C.c_int32), ("line_events", C.c_int32), ("reltol", C.c_double), ("abstol", C.c_double),
                ("dt0", C.c_double), ("dtmax", C.c_double), ("dtmin", C.c_double), ("maxiters", C.c_int64),
                ("line_limit", C.c_void_p), ("ss_abstol", C.c_double), ("ss_reltol", C.c_double),
                ("event_t", C.c_void_p), ("event_e", C.c_void_p)]

# retcode of an instance stopped at a steady state
TERMINATED = 4
# DiffEqCallbacks' TerminateSteadyState defaults
STEADY_STATE_TOLS = (1e-8, 1e-6)


STAT_DTYPE = np.dtype([("t", "f8"), ("dt", "f8"), ("naccept", "i8"), ("nreject", "i8"), ("nf", "i8"),
                       ("retcode", "i4"), ("ntrip", "i4")])


def _kind_struct(m):
    k = A.NDKind()
    d = m.kind()
    for f in ("is_edge", "model", "wrapper", "ff", "dim", "pdim", "outdim_src", "outdim_dst", "extdim"):
        setattr(k, f, int(d[f]))
    for j, x in enumerate(d["mask_src"]):
        k.mask_src[j] = x
    for j, x in enumerate(d["mask_dst"]):
        k.mask_dst[j] = x
    return k


class Ensemble:
    """Independent networks sharing one vertex model and one edge model."""

    def __init__(self, graphs, vertexm, edgem):
        if not isinstance(vertexm, comp.VertexModel) or not isinstance(edgem, comp.EdgeModel):
            raise A.ArgumentError("Ensemble needs one VertexModel and one EdgeModel")
        if not (comp.is_identity_mass(vertexm.mass_matrix) and comp.is_identity_mass(edgem.mass_matrix)):
            raise A.UnsupportedError("the ensemble integrators are explicit methods and cannot use a mass matrix")
        self.graphs = list(graphs)
        self.vertexm, self.edgem = vertexm, edgem
        self.ninst = len(self.graphs)
        nv = np.array([g.nv() for g in self.graphs], np.int64)
        ne = np.array([g.ne() for g in self.graphs], np.int64)
        src = np.ascontiguousarray(np.concatenate([g.src for g in self.graphs]) if ne.sum() else np.zeros(1, np.int64))
        dst = np.ascontiguousarray(np.concatenate([g.dst for g in self.graphs]) if ne.sum() else np.zeros(1, np.int64))
        self.nv, self.ne = nv, ne
        self._vk, self._ek = _kind_struct(vertexm), _kind_struct(edgem)
        h = C.c_void_p()
        L = A.lib()
        A.check(L.nd_ensemble_create(self.ninst, nv.ctypes.data, ne.ctypes.data, src.ctypes.data, dst.ctypes.data,
                                     C.byref(self._vk), C.byref(self._ek), C.byref(h)))
        self._h = h
        ts, tp = C.c_int64(), C.c_int64()
        self.state_offsets = np.zeros(self.ninst, np.int64)
        self.param_offsets = np.zeros(self.ninst, np.int64)
        A.check(L.nd_ensemble_sizes(h, C.byref(ts), C.byref(tp), self.state_offsets.ctypes.data,
                                    self.param_offsets.ctypes.data))
        self.total_states, self.total_params = ts.value, tp.value
        self.classes = list(DEFAULT_CLASSES)

    def __del__(self):
        L = getattr(A, "_lib", None) if A is not None else None
        if L is not None and getattr(self, "_h", None):
            L.nd_ensemble_destroy(self._h)
            self._h = None

    def set_class(self, cls, threads, vper):
        A.check(A.lib().nd_ensemble_set_class(self._h, int(cls), int(threads), int(vper)))
        self.classes[int(cls)] = (int(threads), int(vper))

    def set_classes(self, classes):
        """classes: up to six (threads, vertices per thread) pairs; missing ones repeat the last."""
        cl = [tuple(map(int, c)) for c in classes]
        if not 1 <= len(cl) <= NCLASS:
            raise A.ArgumentError(f"need 1..{NCLASS} classes")
        cl += [cl[-1]] * (NCLASS - len(cl))
        for c, (t, v) in enumerate(cl):
            self.set_class(c, t, v)

    def state_slice(self, i):
        return slice(int(self.state_offsets[i]), int(self.state_offsets[i] + self.nv[i] * self.vertexm.dim))

    def param_slice(self, i):
        n = self.nv[i] * self.vertexm.pdim + self.ne[i] * self.edgem.pdim
        return slice(int(self.param_offsets[i]), int(self.param_offsets[i] + n))

    def solve(self, u, p, tspan, method=TSIT5, reltol=1e-8, abstol=1e-8, dt=0.0, dtmax=0.0, maxiters=0,
              line_limit=None, steady_state=None, preset_trip=None, line_events=None):
        """Integrates all instances in place on the device vector u; returns per-instance stats.
        line_limit: per-edge limits (device, instance order) of the cascading-failure
        callback (condition |P| - limit, affect K = 0, in p as well).  line_events:
        "crossing" locates every up- or down-crossing of the condition inside the step on
        Tsit5's dense output, as the reference's ContinuousComponentCallback does;
        "step_end" trips lines whose |P| exceeds the limit at the end of an accepted step;
        None: "crossing" for Tsit5, "step_end" for RK4.
        steady_state: True or (abstol, reltol) to stop each instance after the first
        accepted step where every |du_i| <= abstol or <= reltol*|u_i| (retcode TERMINATED,
        stats t = the stopping time), as a TerminateSteadyState callback does.
        preset_trip: (times, edges), one entry per instance (edge 0-based within the
        instance, -1 = none): a PresetTimeComponentCallback that sets the edge's K to 0
        at that time (docs/examples/cascading_failure.jl:125-131); Tsit5 steps to it
        exactly, RK4 needs it on the step grid."""
        import torch
        self._check(u, p)
        stats = torch.zeros(self.ninst * STAT_DTYPE.itemsize // 8, dtype=torch.float64, device=u.device)
        o = self._opts(method, reltol, abstol, dt, dtmax, maxiters, line_limit, steady_state, preset_trip,
                       tspan, u.device, line_events)
        stream = torch.cuda.current_stream(u.device).cuda_stream
        A.check(A.lib().nd_ensemble_solve(self._h, u.data_ptr(), p.data_ptr(), float(tspan[0]), float(tspan[1]),
                                          C.byref(o), stats.data_ptr(), stream))
        return stats

    def _opts(self, method, reltol, abstol, dt, dtmax, maxiters, line_limit=None, steady_state=None,
              preset_trip=None, tspan=None, device=None, line_events=None):
        if line_events not in LINE_EVENTS:
            raise A.ArgumentError("line_events must be None, 'step_end' or 'crossing'")
        ss = (0.0, 0.0)
        if steady_state is True:
            ss = STEADY_STATE_TOLS
        elif steady_state:
            ss = tuple(float(x) for x in steady_state)
            if len(ss) != 2 or min(ss) < 0:
                raise A.ArgumentError("steady_state must be True or (abstol, reltol) with tolerances >= 0")
        lim = None
        if line_limit is not None:
            import torch
            if line_limit.dtype != torch.float64 or line_limit.device.type != "cuda" or \
                    line_limit.numel() != int(self.ne.sum()):
                raise A.ArgumentError("line_limit must be a float64 device vector with one entry per edge")
            lim = line_limit.data_ptr()
        ev_t = ev_e = None
        if preset_trip is not None:
            import torch
            times = np.asarray(preset_trip[0], np.float64).reshape(-1)
            edges = np.asarray(preset_trip[1], np.int32).reshape(-1)
            if len(times) != self.ninst or len(edges) != self.ninst:
                raise A.ArgumentError("preset_trip needs one time and one edge per instance")
            if ((edges >= self.ne) | (edges < -1)).any():
                raise A.ArgumentError("preset_trip edge outside its instance")
            if method == RK4 and tspan is not None:
                k = np.rint((times - tspan[0]) / dt)
                on = np.abs(tspan[0] + k * dt - times) <= 1e-9 * np.maximum(1.0, np.abs(times))
                if not on[(edges >= 0) & (times > tspan[0]) & (times <= tspan[1])].all():
                    raise A.ArgumentError("RK4: preset trip times must lie on the step grid t0 + k*dt")
            self._events = (torch.from_numpy(times).to(device), torch.from_numpy(edges).to(device))
            ev_t, ev_e = self._events[0].data_ptr(), self._events[1].data_ptr()
        return _Opts(int(method), LINE_EVENTS[line_events], float(reltol), float(abstol), float(dt), float(dtmax), 0.0, int(maxiters), lim,
                     ss[0], ss[1], ev_t, ev_e)

    def _check(self, u, p):
        import torch
        if u.dtype != torch.float64 or p.dtype != torch.float64 or u.device.type != "cuda":
            raise A.ArgumentError("u and p must be float64 device tensors")
        if u.numel() != self.total_states or p.numel() != self.total_params:
            raise A.ArgumentError("u / p do not match the ensemble's sizes")

    def capture(self, u, p, tspan, method=TSIT5, reltol=1e-8, abstol=1e-8, dt=0.0, dtmax=0.0, maxiters=0,
                line_limit=None, steady_state=None, line_events=None):
        """Captures the solve of (u, p, tspan) into a hipGraph; `.launch()` replays it in place on u."""
        import torch
        self._check(u, p)
        stats = torch.zeros(self.ninst * STAT_DTYPE.itemsize // 8, dtype=torch.float64, device=u.device)
        o = self._opts(method, reltol, abstol, dt, dtmax, maxiters, line_limit, steady_state,
                       line_events=line_events)
        h = C.c_void_p()
        A.check(A.lib().nd_ensemble_graph_create(self._h, u.data_ptr(), p.data_ptr(), float(tspan[0]),
                                                 float(tspan[1]), C.byref(o), stats.data_ptr(), C.byref(h)))
        return EnsembleGraph(self, h, u, p, stats)

    @staticmethod
    def stats_to_numpy(stats):
        return stats.cpu().numpy().view(STAT_DTYPE)


class EnsembleGraph:
    """A captured ensemble solve.  Holds its ensemble and buffers so they outlive it."""

    def __init__(self, ens, h, u, p, stats):
        self.ens, self._h, self.u, self.p, self.stats = ens, h, u, p, stats

    def launch(self):
        import torch
        stream = torch.cuda.current_stream(self.u.device).cuda_stream
        A.check(A.lib().nd_ensemble_graph_launch(self._h, stream))
        return self.stats

    def __del__(self):
        L = getattr(A, "_lib", None) if A is not None else None
        if L is not None and getattr(self, "_h", None):
            L.nd_ensemble_graph_destroy(self._h)
            self._h = None


def tsit5_per_instance(nw, u, p, tspan, reltol=1e-8, abstol=1e-8, dt0=0.0, maxiters=10 ** 7, dtmin=0.0):
    """Adaptive Tsit5 for one network: device RHS per stage at t + c_s*h, host-side step
    control (the tableau, PI controller, initial-dt heuristic and retcodes of the ensemble
    kernel).  u: device vector, in place.  Returns t, naccept, nreject and retcode: 0
    reached tspan[1], 1 maxiters, 2 dt below dtmin (default 1e-14*max(1, |t1|)), 3 NaN
    in the error estimate or the step size."""
    import torch
    t0, t1 = float(tspan[0]), float(tspan[1])
    if not t1 >= t0:
        raise A.ArgumentError("tspan[1] < tspan[0]")
    if t1 == t0:
        return dict(t=t0, naccept=0, nreject=0, retcode=0)
    tab, btilde, c = tsit5_tableau()
    if not dtmin > 0:
        dtmin = 1e-14 * max(1.0, abs(t1))
    ks = [torch.empty_like(u) for _ in range(7)]
    ys = torch.empty_like(u)
    nw(ks[0], u, p, t0)
    n = u.numel()

    def norm(x, ref):
        return torch.sqrt(torch.sum((x / (abstol + ref.abs() * reltol)) ** 2) / n).item()

    dt = dt0
    if not dt > 0:
        d0, d1 = norm(u, u), norm(ks[0], u)
        h0 = 1e-6 if (d0 < 1e-5 or d1 < 1e-5) else 0.01 * d0 / d1
        h0 = min(h0, t1 - t0)
        torch.add(u, ks[0], alpha=h0, out=ys)
        nw(ks[1], ys, p, t0 + h0)
        d2 = norm(ks[1] - ks[0], u) / h0
        m = max(d1, d2)
        h1 = max(1e-6, h0 * 1e-3) if m <= 1e-15 else (0.01 / m) ** (1.0 / 6.0)
        dt = min(100 * h0, h1)
    dt = min(dt, t1 - t0)
    t, qold, naccept, nreject, it, retcode = t0, 1e-4, 0, 0, 0, 0
    while t < t1:
        if it >= maxiters:
            retcode = 1
            break
        if dt < dtmin:
            retcode = 2
            break
        it += 1
        last = t + dt >= t1
        h = t1 - t if last else dt
        for s in range(1, 7):
            ys.copy_(u)
            for j in range(s):
                ys.add_(ks[j], alpha=h * tab[s, j])
            nw(ks[s], ys, p, t + c[s] * h)
        err = sum(float(bt) * k for bt, k in zip(btilde, ks)) * h
        sc = abstol + torch.maximum(u.abs(), ys.abs()) * reltol
        eest = torch.sqrt(torch.sum((err / sc) ** 2) / n).item()
        if eest != eest:
            retcode = 3
            break
        q11 = eest ** 0.14
        if eest <= 1.0:
            qq = 0.1 if eest == 0 else max(0.1, min(5.0, q11 / qold ** 0.08 / 0.9))
            tn = t + h
            # a step ending within 100 ulp of t1 ends on it (fixed_t_for_floatingpoint_error!)
            t = t1 if last or abs(tn - t1) < 100.0 * 2.220446049250313e-16 * max(abs(t), abs(t1)) else tn
            u.copy_(ys)
            ks[0].copy_(ks[6])
            qold = max(eest, 1e-4)
            dt = h / qq
            naccept += 1
        else:
            dt = h / min(5.0, q11 / 0.9)
            nreject += 1
        if dt != dt:
            retcode = 3
            break
    return dict(t=t, naccept=naccept, nreject=nreject, retcode=retcode)
