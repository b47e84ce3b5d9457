// Device equations of the component library, shared by the RHS plan kernels and the
// ensemble integrator.  Operation order follows the reference's model definitions.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/nd_hip.h"

namespace nd {

__host__ __device__ constexpr bool is_ode_model(int em) {
    return em == ND_EM_ODE_SIN_DIFFUSION || em == ND_EM_ODE_SIN_DIFFUSION_FAST || em == ND_EM_ODE_LINEAR;
}

// dst-side output of a pure feed-forward edge model g(v_s, v_d, p)
template <int EM>
__device__ __forceinline__ double g_dst(double vs, double vd, const double* __restrict__ pe) {
    if constexpr (EM == ND_EM_KURAMOTO) return pe[0] * sin(vs - vd);             // ComponentLibrary.jl:51-53
    else if constexpr (EM == ND_EM_STATIC_LINE) return pe[0] * sin(vs - vd - pe[1]);
    else if constexpr (EM == ND_EM_DIFFUSION) return vs - vd;                     // diffusion_test.jl:57-60
    else if constexpr (EM == ND_EM_DIFFUSION_WEIGHTED) return pe[0] * (vs - vd);  // ComponentLibrary.jl:8-10
    else if constexpr (EM == ND_EM_DIFFUSION_FIDUCIAL) return pe[0] * (vs - vd);  // ComponentLibrary.jl:22-25
    else return __builtin_nan("");
}

__device__ __forceinline__ double g_dst_rt(int em, double vs, double vd, const double* __restrict__ pe) {
    switch (em) {
    case ND_EM_KURAMOTO: return g_dst<ND_EM_KURAMOTO>(vs, vd, pe);
    case ND_EM_STATIC_LINE: return g_dst<ND_EM_STATIC_LINE>(vs, vd, pe);
    case ND_EM_DIFFUSION: return g_dst<ND_EM_DIFFUSION>(vs, vd, pe);
    case ND_EM_DIFFUSION_WEIGHTED: return g_dst<ND_EM_DIFFUSION_WEIGHTED>(vs, vd, pe);
    case ND_EM_DIFFUSION_FIDUCIAL: return g_dst<ND_EM_DIFFUSION_FIDUCIAL>(vs, vd, pe);
    default: return __builtin_nan("");
    }
}

template <int VM>
struct VModel { static constexpr int dim = -1; };
template <> struct VModel<ND_VM_DIFFUSION> { static constexpr int dim = 1; };
template <> struct VModel<ND_VM_KURAMOTO> { static constexpr int dim = 1; };
template <> struct VModel<ND_VM_SWING> { static constexpr int dim = 2; };
template <> struct VModel<ND_VM_SWING_UNIT_DAMPING> { static constexpr int dim = 2; };
template <> struct VModel<ND_VM_LOAD> { static constexpr int dim = 1; };

// vertex f (apply_compf, coreloop.jl:213-216); operation order as the reference writes it
template <int VM>
__device__ __forceinline__ void vertex_f(double* __restrict__ dx, const double* __restrict__ x, double acc,
                                         const double* __restrict__ pv, int dim) {
    if constexpr (VM == ND_VM_DIFFUSION) {                 // ComponentLibrary.jl:42-45
        dx[0] = acc;
    } else if constexpr (VM == ND_VM_KURAMOTO) {           // ComponentLibrary.jl:69-71
        dx[0] = pv[0] + acc;
    } else if constexpr (VM == ND_VM_SWING) {              // ComponentLibrary.jl:59-63
        const double w = x[1];
        dx[0] = w;
        dx[1] = 1.0 / pv[0] * (pv[2] - pv[1] * w + acc);
    } else if constexpr (VM == ND_VM_SWING_UNIT_DAMPING) { // benchmark_models.jl:37-41
        const double w = x[1];
        dx[0] = w;
        double d = pv[0] - 1.0 * w;
        d += acc;
        dx[1] = d;
    } else if constexpr (VM == ND_VM_LOAD) {
        dx[0] = 1.0 / pv[0] * (pv[1] + acc);
    } else {                                               // f does nothing: du keeps its zero fill
        for (int j = 0; j < dim; ++j) dx[j] = 0.0;
    }
}

// Parameters each edge model reads (p entries per edge).
template <int EM> struct EModel { static constexpr int np = 0; };
template <> struct EModel<ND_EM_KURAMOTO> { static constexpr int np = 1; };
template <> struct EModel<ND_EM_STATIC_LINE> { static constexpr int np = 2; };
template <> struct EModel<ND_EM_DIFFUSION_WEIGHTED> { static constexpr int np = 1; };
template <> struct EModel<ND_EM_DIFFUSION_FIDUCIAL> { static constexpr int np = 1; };

// parameters per vertex model
template <int VM> struct VPar { static constexpr int np = 0; };
template <> struct VPar<ND_VM_KURAMOTO> { static constexpr int np = 1; };
template <> struct VPar<ND_VM_SWING> { static constexpr int np = 3; };
template <> struct VPar<ND_VM_SWING_UNIT_DAMPING> { static constexpr int np = 1; };
template <> struct VPar<ND_VM_LOAD> { static constexpr int np = 2; };

}  // namespace nd
