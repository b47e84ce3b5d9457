"""ctypes binding of libnd_hip.so (declarations in include/nd_hip.h).

The library is built in-tree by build.py.  A missing library is an error:
there is no CPU fallback for the right-hand side.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnd_hip.so")

ND_OK, ND_ERR_ARGUMENT, ND_ERR_UNSUPPORTED, ND_ERR_DEVICE, ND_ERR_ALLOC, ND_ERR_NOT_CONVERGED = range(6)

ND_WRAP_NONE, ND_WRAP_ANTISYMMETRIC, ND_WRAP_SYMMETRIC, ND_WRAP_DIRECTED, ND_WRAP_FIDUCIAL = range(5)
ND_FF_PURESTATEMAP, ND_FF_PUREFEEDFORWARD, ND_FF_FEEDFORWARD, ND_FF_NOFEEDFORWARD = range(4)
(ND_VM_NOOP, ND_VM_DIFFUSION, ND_VM_KURAMOTO, ND_VM_SWING, ND_VM_SWING_UNIT_DAMPING, ND_VM_LOAD,
 ND_VM_CAPACITOR_SINK, ND_VM_PI_SOURCE, ND_VM_CAPACITOR, ND_VM_VOLTAGE_SOURCE, ND_VM_RESISTOR_INJECTOR,
 ND_VM_INDUCTOR_INJECTOR, ND_VM_CRL, ND_VM_FHN, ND_VM_TRUSS_FREE, ND_VM_TRUSS_FIXED) = range(16)
(ND_EM_NOOP, ND_EM_DIFFUSION, ND_EM_DIFFUSION_WEIGHTED, ND_EM_KURAMOTO, ND_EM_STATIC_LINE,
 ND_EM_DIFFUSION_FIDUCIAL, ND_EM_ODE_SIN_DIFFUSION, ND_EM_ODE_SIN_DIFFUSION_FAST, ND_EM_ODE_LINEAR,
 ND_EM_RESISTOR, ND_EM_LOOPBACK, ND_EM_ODE_LAG, ND_EM_EXT_DRIVEN, ND_EM_ELECTRICAL, ND_EM_BEAM) = range(15)
ND_MAX_OUT = 4
ND_MAX_EXT = 8
ND_MASS_IDENTITY, ND_MASS_UNIFORM, ND_MASS_DIAGONAL, ND_MASS_DENSE = range(4)


class ArgumentError(ValueError):
    """The reference's ArgumentError (e.g. src/coreloop.jl:2-7)."""


class UnsupportedError(NotImplementedError):
    """A component model or option without a HIP implementation."""


class DeviceError(RuntimeError):
    """A HIP runtime failure."""


class NotConvergedError(RuntimeError):
    """An iterative solver stopped before reaching its tolerance."""


class NDKind(C.Structure):
    _fields_ = [("is_edge", C.c_int32), ("model", C.c_int32), ("wrapper", C.c_int32),
                ("ff", C.c_int32), ("dim", C.c_int32), ("pdim", C.c_int32),
                ("outdim_src", C.c_int32), ("outdim_dst", C.c_int32),
                ("mask_src", C.c_int32 * ND_MAX_OUT), ("mask_dst", C.c_int32 * ND_MAX_OUT),
                ("extdim", C.c_int32)]


class NDNetworkInfo(C.Structure):
    _fields_ = [("nv", C.c_int64), ("ne", C.c_int64), ("dim", C.c_int64), ("pdim", C.c_int64),
                ("lastidx_out", C.c_int64), ("lastidx_aggr", C.c_int64), ("lastidx_gbuf", C.c_int64),
                ("agg_first", C.c_int64), ("agg_last", C.c_int64),
                ("vdepth", C.c_int32), ("edepth", C.c_int32),
                ("nvbatches", C.c_int32), ("nebatches", C.c_int32),
                ("lastidx_extbuf", C.c_int64)]


class NDBatch(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad_", C.c_int32), ("count", C.c_int64),
                ("indices", C.POINTER(C.c_int64)), ("state_first", C.c_int64),
                ("param_first", C.c_int64), ("out_first", C.c_int64), ("in_first", C.c_int64)]


class NDMass(C.Structure):
    _fields_ = [("form", C.c_int32), ("pad_", C.c_int32), ("m", C.c_void_p)]


class NDPlanTuning(C.Structure):
    _fields_ = [("hub_degree", C.c_int32), ("block_size", C.c_int32), ("big_hub_degree", C.c_int32),
                ("rk4_path", C.c_int32), ("gather_mode", C.c_int32)]


P = C.c_void_p
I64 = C.c_int64
I32 = C.c_int32
D = C.c_double

# name -> (restype, argtypes); mirrors include/nd_hip.h
PROTOTYPES = {
    "nd_network_create": (C.c_int, [I64, I64, P, P, I32, P, P, P, C.POINTER(P)]),
    "nd_network_destroy": (None, [P]),
    "nd_network_get_info": (C.c_int, [P, C.POINTER(NDNetworkInfo)]),
    "nd_network_get_batch": (C.c_int, [P, I32, I32, C.POINTER(NDBatch)]),
    "nd_network_export_ranges": (C.c_int, [P, P, P]),
    "nd_network_export_maps": (C.c_int, [P, P, P]),
    "nd_network_csr_size": (C.c_int, [P, C.POINTER(I64)]),
    "nd_network_export_csr": (C.c_int, [P, P, P]),
    "nd_network_mass_matrix": (C.c_int, [P, P, P, C.POINTER(I32), P]),
    "nd_network_export_ext_ranges": (C.c_int, [P, P, P]),
    "nd_network_set_extmap": (C.c_int, [P, I64, P]),
    "nd_collect_externals": (C.c_int, [P, P, P, P, P]),
    "nd_plan_create": (C.c_int, [P, C.POINTER(P)]),
    "nd_plan_destroy": (None, [P]),
    "nd_plan_tune": (C.c_int, [P, C.POINTER(NDPlanTuning)]),
    "nd_plan_get_tuning": (C.c_int, [P, C.POINTER(NDPlanTuning)]),
    "nd_rhs": (C.c_int, [P, P, P, P, D, I64, I64, I64, P]),
    "nd_get_buffers": (C.c_int, [P, P, P, P, P, D, P]),
    "nd_rk4": (C.c_int, [P, P, P, D, D, I64, P, I64, I64, P]),
    "nd_rk4_noalloc": (C.c_int, [P, P, P, D, D, I64, I64, I64, P]),
    "nd_tsit5": (C.c_int, [P, P, P, D, D, D, D, D, I64, P, I64, I64, P]),
    "nd_rhs_jvp": (C.c_int, [P, P, P, P, P, P, D, I64, I64, I64, P]),
    "nd_rhs_jvp_multi": (C.c_int, [P, P, P, P, P, C.c_int32, P, D, I64, I64, I64, P]),
    "nd_plan_set_perturbation": (C.c_int, [P, I64, I64, P, I64, P, I64, P, I64, P]),
    "nd_rhs_perturbed_jvp": (C.c_int, [P, P, P, P, P, P, P, D, P, P, P, P, P, P, P]),
    "nd_malloc": (C.c_int, [C.c_size_t, C.POINTER(P)]),
    "nd_free": (C.c_int, [P]),
    "nd_memcpy_h2d": (C.c_int, [P, P, C.c_size_t]),
    "nd_memcpy_d2h": (C.c_int, [P, P, C.c_size_t]),
    "nd_gather": (C.c_int, [P, P, P, P]),
    "nd_aggregate": (C.c_int, [P, P, P, P]),
    "nd_ensemble_create": (C.c_int, [I32, P, P, P, P, P, P, C.POINTER(P)]),
    "nd_ensemble_destroy": (None, [P]),
    "nd_ensemble_sizes": (C.c_int, [P, P, P, P, P]),
    "nd_ensemble_set_class": (C.c_int, [P, I32, I32, I32]),
    "nd_ensemble_solve": (C.c_int, [P, P, P, D, D, P, P, P]),
    "nd_tsit5_tableau": (C.c_int, [P, P, P]),
    "nd_tsit5_dense": (C.c_int, [P]),
    "nd_ensemble_graph_create": (C.c_int, [P, P, P, D, D, P, P, P]),
    "nd_ensemble_graph_launch": (C.c_int, [P, P]),
    "nd_ensemble_graph_destroy": (None, [P]),
    "nd_fixpoint_create": (C.c_int, [I32, P, P, P, P, C.POINTER(P)]),
    "nd_fixpoint_destroy": (None, [P]),
    "nd_fixpoint_solve": (C.c_int, [P, P, P, P, P, P, P, P]),
    "nd_mp_param_count": (I64, [I32]),
    "nd_mp_workspace_floats": (I64, [I64]),
    "nd_mp_forward": (C.c_int, [I64, P, P, P, P, I32, P, P, P, P]),
    "nd_mp_backward": (C.c_int, [I64, P, P, P, P, I32, P, P, P, P, P, P]),
    "nd_mp_adam": (C.c_int, [I64, P, P, P, P, C.c_float, C.c_float, C.c_float, C.c_float, I64, P]),
    "nd_ws_features": (C.c_int, [I64, P, P, P, P, P, P, P, P, P]),
    "nd_ws_residual": (C.c_int, [I64, P, P, P, P, P, P, P, P]),
    "nd_graph_erdos_renyi": (C.c_int, [I64, I64, C.c_uint64, C.POINTER(P)]),
    "nd_graph_barabasi_albert": (C.c_int, [I64, I64, C.c_uint64, C.POINTER(P)]),
    "nd_graph_random_geometric": (C.c_int, [I64, D, C.c_uint64, P, C.POINTER(P)]),
    "nd_graph_watts_strogatz": (C.c_int, [I64, I64, D, C.c_uint64, C.POINTER(P)]),
    "nd_edgelist_size": (I64, [P]),
    "nd_edgelist_copy": (C.c_int, [P, P, P]),
    "nd_edgelist_free": (None, [P]),
    "nd_last_error": (C.c_char_p, []),
    "nd_device_synchronize": (C.c_int, []),
}

_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libnd_hip.so not built ({LIB_PATH}); run networkdynamics.jl_amd/build.py "
                              "or __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    """Raise the Python twin of an ND_ERR_* status."""
    if rc == ND_OK:
        return
    msg = lib().nd_last_error().decode(errors="replace")
    if rc == ND_ERR_ARGUMENT:
        raise ArgumentError(msg)
    if rc == ND_ERR_UNSUPPORTED:
        raise UnsupportedError(msg)
    if rc == ND_ERR_NOT_CONVERGED:
        raise NotConvergedError(msg)
    if rc == ND_ERR_ALLOC:
        raise MemoryError(msg)
    raise DeviceError(msg)
