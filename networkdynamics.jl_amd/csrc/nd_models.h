// Device equations of the component library, shared by the RHS plan kernels and the
// ensemble integrator.  Operation order follows the reference's model definitions.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/nd_hip.h"
#include "nd_math.h"

namespace nd {

using ::sqrt;   // the double overload stays visible next to sqrt(Dual)

// First-order dual number v + d*eps (one tangent direction), the forward-mode scalar
// the Jacobian-vector product evaluates the models with.  The model functions take the
// parameters as plain doubles, or as Duals where a derivative in a parameter is wanted
// (the parameter channels of linearize_network).
struct Dual {
    double v, d;
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator+(double c, Dual a) { return {c + a.v, a.d}; }
__device__ __forceinline__ Dual operator+(Dual a, double c) { return {a.v + c, a.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dual operator-(Dual a, double c) { return {a.v - c, a.d}; }
__device__ __forceinline__ Dual operator-(double c, Dual a) { return {c - a.v, -a.d}; }
__device__ __forceinline__ Dual operator*(double c, Dual a) { return {c * a.v, c * a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, double c) { return {a.v * c, a.d * c}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, double c) { return {a.v / c, a.d / c}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
__device__ __forceinline__ Dual operator/(double c, Dual b) { return {c / b.v, -c * b.d / (b.v * b.v)}; }
// the models' sine: nd_sin (nd_math.h) on doubles, and its derivative nd_cos on duals
__device__ __forceinline__ double msin(double x) { return nd_sin(x); }
__device__ __forceinline__ Dual msin(Dual a) { return {nd_sin(a.v), nd_cos(a.v) * a.d}; }
__device__ __forceinline__ Dual sqrt(Dual a) {
    const double r = ::sqrt(a.v);
    return {r, a.d / (2.0 * r)};
}
__device__ __forceinline__ Dual dual_nan() { return {__builtin_nan(""), __builtin_nan("")}; }

template <typename S> __device__ __forceinline__ S nan_of();
template <> __device__ __forceinline__ double nan_of<double>() { return __builtin_nan(""); }
template <> __device__ __forceinline__ Dual nan_of<Dual>() { return dual_nan(); }
template <typename S> __device__ __forceinline__ S zero_of();
template <> __device__ __forceinline__ double zero_of<double>() { return 0.0; }
template <> __device__ __forceinline__ Dual zero_of<Dual>() { return {0.0, 0.0}; }

// K tangent directions at once, the chunk of a ForwardDiff Dual{T,Float64,K}: value v
// and partials d[0..K).  Every operation applies the single-tangent rule above to each
// partial in the same operation order, so partial j equals the Dual along direction j
// bit for bit.
template <int K>
struct DualK {
    double v;
    double d[K];
};
#define ND_DUALK_OP(ret_v, part)                                    \
    DualK<K> r;                                                     \
    r.v = (ret_v);                                                  \
    _Pragma("unroll") for (int j = 0; j < K; ++j) r.d[j] = (part);  \
    return r;
template <int K> __device__ __forceinline__ DualK<K> operator+(DualK<K> a, DualK<K> b) { ND_DUALK_OP(a.v + b.v, a.d[j] + b.d[j]) }
template <int K> __device__ __forceinline__ DualK<K> operator+(double c, DualK<K> a) { ND_DUALK_OP(c + a.v, a.d[j]) }
template <int K> __device__ __forceinline__ DualK<K> operator+(DualK<K> a, double c) { ND_DUALK_OP(a.v + c, a.d[j]) }
template <int K> __device__ __forceinline__ DualK<K> operator-(DualK<K> a, DualK<K> b) { ND_DUALK_OP(a.v - b.v, a.d[j] - b.d[j]) }
template <int K> __device__ __forceinline__ DualK<K> operator-(DualK<K> a) { ND_DUALK_OP(-a.v, -a.d[j]) }
template <int K> __device__ __forceinline__ DualK<K> operator-(DualK<K> a, double c) { ND_DUALK_OP(a.v - c, a.d[j]) }
template <int K> __device__ __forceinline__ DualK<K> operator-(double c, DualK<K> a) { ND_DUALK_OP(c - a.v, -a.d[j]) }
template <int K> __device__ __forceinline__ DualK<K> operator*(double c, DualK<K> a) { ND_DUALK_OP(c * a.v, c * a.d[j]) }
template <int K> __device__ __forceinline__ DualK<K> operator*(DualK<K> a, double c) { ND_DUALK_OP(a.v * c, a.d[j] * c) }
template <int K> __device__ __forceinline__ DualK<K> operator*(DualK<K> a, DualK<K> b) {
    ND_DUALK_OP(a.v * b.v, a.d[j] * b.v + a.v * b.d[j])
}
template <int K> __device__ __forceinline__ DualK<K> operator/(DualK<K> a, double c) { ND_DUALK_OP(a.v / c, a.d[j] / c) }
template <int K> __device__ __forceinline__ DualK<K> operator/(DualK<K> a, DualK<K> b) {
    const double bb = b.v * b.v;
    ND_DUALK_OP(a.v / b.v, (a.d[j] * b.v - a.v * b.d[j]) / bb)
}
template <int K> __device__ __forceinline__ DualK<K> operator/(double c, DualK<K> b) {
    const double bb = b.v * b.v;
    ND_DUALK_OP(c / b.v, -c * b.d[j] / bb)
}
template <int K> __device__ __forceinline__ DualK<K> msin(DualK<K> a) {
    const double cs = nd_cos(a.v);
    ND_DUALK_OP(nd_sin(a.v), cs * a.d[j])
}
template <int K> __device__ __forceinline__ DualK<K> sqrt(DualK<K> a) {
    const double rt = ::sqrt(a.v);
    ND_DUALK_OP(rt, a.d[j] / (2.0 * rt))
}
#undef ND_DUALK_OP
template <int K> __device__ __forceinline__ DualK<K> dualk_const(double x) {
    DualK<K> r;
    r.v = x;
#pragma unroll
    for (int j = 0; j < K; ++j) r.d[j] = x == x ? 0.0 : x;
    return r;
}
template <> __device__ __forceinline__ DualK<4> nan_of<DualK<4>>() { return dualk_const<4>(__builtin_nan("")); }
template <> __device__ __forceinline__ DualK<8> nan_of<DualK<8>>() { return dualk_const<8>(__builtin_nan("")); }
template <> __device__ __forceinline__ DualK<4> zero_of<DualK<4>>() { return dualk_const<4>(0.0); }
template <> __device__ __forceinline__ DualK<8> zero_of<DualK<8>>() { return dualk_const<8>(0.0); }

__host__ __device__ constexpr bool is_ode_model(int em) {
    return em == ND_EM_ODE_SIN_DIFFUSION || em == ND_EM_ODE_SIN_DIFFUSION_FAST || em == ND_EM_ODE_LINEAR ||
           em == ND_EM_ODE_LAG || em == ND_EM_EXT_DRIVEN;
}
// state dimension of each ODE edge model, and the external inputs its f reads
__host__ __device__ constexpr int ode_dim(int em) {
    return (em == ND_EM_ODE_LAG || em == ND_EM_EXT_DRIVEN) ? 1 : 2;
}
__host__ __device__ constexpr int edge_ext_inputs(int em) { return em == ND_EM_EXT_DRIVEN ? 1 : 0; }

// f of an ODE edge: dx from its state x, the endpoint outputs and its external inputs;
// S = double or Dual
template <int EM, typename S, typename P = double>
__device__ __forceinline__ void edge_ode_f(S* dx, const S* x, S vs, S vd, const S* ext, const P* __restrict__ pe) {
    if constexpr (EM == ND_EM_ODE_SIN_DIFFUSION) {            // ComponentLibrary.jl:30-34
        dx[0] = 1.0 / pe[0] * (msin(vs - vd) - x[0]);
        dx[1] = 1.0 / pe[0] * (msin(vd - vs) - x[1]);
    } else if constexpr (EM == ND_EM_ODE_SIN_DIFFUSION_FAST) { // benchmark_models.jl:11-15
        dx[0] = 100.0 * (msin(vs - vd) - x[0]);
        dx[1] = 100.0 * (msin(vd - vs) - x[1]);
    } else if constexpr (EM == ND_EM_ODE_LAG) {               // ComponentLibrary.jl:30-32, one state
        dx[0] = 1.0 / pe[0] * (msin(vs - vd) - x[0]);
    } else if constexpr (EM == ND_EM_EXT_DRIVEN) {            // external_inputs_test.jl:45-47
        dx[0] = ext[0];
    } else {                                                   // diffusion_test.jl:102-106
        dx[0] = vs - vd - x[0];
        dx[1] = vd - vs - x[1];
    }
}
template <typename S, typename P = double>
__device__ __forceinline__ void edge_ode_f_rt(int em, S* dx, const S* x, S vs, S vd, const S* ext,
                                              const P* __restrict__ pe) {
    switch (em) {
    case ND_EM_ODE_SIN_DIFFUSION: edge_ode_f<ND_EM_ODE_SIN_DIFFUSION, S, P>(dx, x, vs, vd, ext, pe); break;
    case ND_EM_ODE_SIN_DIFFUSION_FAST: edge_ode_f<ND_EM_ODE_SIN_DIFFUSION_FAST, S, P>(dx, x, vs, vd, ext, pe); break;
    case ND_EM_ODE_LAG: edge_ode_f<ND_EM_ODE_LAG, S, P>(dx, x, vs, vd, ext, pe); break;
    case ND_EM_EXT_DRIVEN: edge_ode_f<ND_EM_EXT_DRIVEN, S, P>(dx, x, vs, vd, ext, pe); break;
    default: edge_ode_f<ND_EM_ODE_LINEAR, S, P>(dx, x, vs, vd, ext, pe); break;
    }
}

// dst-side output of a pure feed-forward edge model g(v_s, v_d, p); S = double or Dual
template <int EM, typename S = double, typename P = double>
__device__ __forceinline__ S g_dst(S vs, S vd, const P* __restrict__ pe) {
    if constexpr (EM == ND_EM_KURAMOTO) return pe[0] * msin(vs - vd);             // ComponentLibrary.jl:51-53
    else if constexpr (EM == ND_EM_STATIC_LINE) return pe[0] * msin(vs - vd - pe[1]);
    else if constexpr (EM == ND_EM_DIFFUSION) return vs - vd;                     // diffusion_test.jl:57-60
    else if constexpr (EM == ND_EM_DIFFUSION_WEIGHTED) return pe[0] * (vs - vd);  // ComponentLibrary.jl:8-10
    else if constexpr (EM == ND_EM_DIFFUSION_FIDUCIAL) return pe[0] * (vs - vd);  // ComponentLibrary.jl:22-25
    else if constexpr (EM == ND_EM_RESISTOR) return (vs - vd) / pe[0];            // external_inputs.md:66-68
    else if constexpr (EM == ND_EM_LOOPBACK) return -1.0 * vs;                     // post_utils.jl:105-108
    else if constexpr (EM == ND_EM_ELECTRICAL) return pe[0] * (vs - vd) * pe[1];  // directed_and_weighted_graphs.jl:97-100
    else return nan_of<S>();
}

template <typename S = double, typename P = double>
__device__ __forceinline__ S g_dst_rt(int em, S vs, S vd, const P* __restrict__ pe) {
    switch (em) {
    case ND_EM_KURAMOTO: return g_dst<ND_EM_KURAMOTO, S, P>(vs, vd, pe);
    case ND_EM_STATIC_LINE: return g_dst<ND_EM_STATIC_LINE, S, P>(vs, vd, pe);
    case ND_EM_DIFFUSION: return g_dst<ND_EM_DIFFUSION, S, P>(vs, vd, pe);
    case ND_EM_DIFFUSION_WEIGHTED: return g_dst<ND_EM_DIFFUSION_WEIGHTED, S, P>(vs, vd, pe);
    case ND_EM_DIFFUSION_FIDUCIAL: return g_dst<ND_EM_DIFFUSION_FIDUCIAL, S, P>(vs, vd, pe);
    case ND_EM_RESISTOR: return g_dst<ND_EM_RESISTOR, S, P>(vs, vd, pe);
    case ND_EM_LOOPBACK: return g_dst<ND_EM_LOOPBACK, S, P>(vs, vd, pe);
    case ND_EM_ELECTRICAL: return g_dst<ND_EM_ELECTRICAL, S, P>(vs, vd, pe);
    default: return nan_of<S>();
    }
}

template <int VM>
struct VModel { static constexpr int dim = -1; };
template <> struct VModel<ND_VM_DIFFUSION> { static constexpr int dim = 1; };
template <> struct VModel<ND_VM_FHN> { static constexpr int dim = 2; };
template <> struct VModel<ND_VM_TRUSS_FREE> { static constexpr int dim = 4; };
template <> struct VModel<ND_VM_TRUSS_FIXED> { static constexpr int dim = 0; };
template <> struct VModel<ND_VM_KURAMOTO> { static constexpr int dim = 1; };
template <> struct VModel<ND_VM_SWING> { static constexpr int dim = 2; };
template <> struct VModel<ND_VM_SWING_UNIT_DAMPING> { static constexpr int dim = 2; };
template <> struct VModel<ND_VM_LOAD> { static constexpr int dim = 1; };
template <> struct VModel<ND_VM_CAPACITOR_SINK> { static constexpr int dim = 1; };
template <> struct VModel<ND_VM_PI_SOURCE> { static constexpr int dim = 2; };
template <> struct VModel<ND_VM_CAPACITOR> { static constexpr int dim = 1; };
template <> struct VModel<ND_VM_VOLTAGE_SOURCE> { static constexpr int dim = 0; };
template <> struct VModel<ND_VM_RESISTOR_INJECTOR> { static constexpr int dim = 0; };
template <> struct VModel<ND_VM_INDUCTOR_INJECTOR> { static constexpr int dim = 1; };
template <> struct VModel<ND_VM_CRL> { static constexpr int dim = 2; };

// vertex f (apply_compf, coreloop.jl:213-216); operation order as the reference writes it
template <int VM, typename S = double, typename P = double>
__device__ __forceinline__ void vertex_f(S* __restrict__ dx, const S* __restrict__ x, S acc,
                                         const P* __restrict__ pv, int dim) {
    if constexpr (VM == ND_VM_DIFFUSION) {                 // ComponentLibrary.jl:42-45
        dx[0] = acc;
    } else if constexpr (VM == ND_VM_KURAMOTO) {           // ComponentLibrary.jl:69-71
        dx[0] = pv[0] + acc;
    } else if constexpr (VM == ND_VM_SWING) {              // ComponentLibrary.jl:59-63
        const S w = x[1];
        dx[0] = w;
        dx[1] = 1.0 / pv[0] * (pv[2] - pv[1] * w + acc);
    } else if constexpr (VM == ND_VM_SWING_UNIT_DAMPING) { // benchmark_models.jl:37-41
        const S w = x[1];
        dx[0] = w;
        S d = pv[0] - 1.0 * w;
        d = d + acc;
        dx[1] = d;
    } else if constexpr (VM == ND_VM_LOAD) {
        dx[0] = 1.0 / pv[0] * (pv[1] + acc);
    } else if constexpr (VM == ND_VM_FHN) {                // directed_and_weighted_graphs.jl:89-93
        const S a = x[0], b = x[1];
        dx[0] = a - a * a * a / 3.0 - b + acc;              // v[1]^3 is v[1]*v[1]*v[1] (literal power)
        dx[1] = (a - pv[0]) * pv[1];
    } else {                                               // f does nothing: du keeps its zero fill
        for (int j = 0; j < dim; ++j) dx[j] = zero_of<S>();
    }
}

// Vertex f with the arguments the fused kernels do not carry: external inputs ext
// (external_inputs.jl; apply_compf passes them after the aggregated input) and the time
// t.  The models without either reduce to vertex_f.
template <int VM, typename S = double, typename P = double>
__device__ __forceinline__ void vertex_f_ext(S* __restrict__ dx, const S* __restrict__ x, S acc,
                                             const P* __restrict__ pv, const S* __restrict__ ext, double t,
                                             int dim) {
    if constexpr (VM == ND_VM_CAPACITOR_SINK) {            // external_inputs.md:76-79
        dx[0] = 1.0 / pv[0] * (acc - (1.0 + 0.1 * msin(t)));
    } else if constexpr (VM == ND_VM_PI_SOURCE) {          // external_inputs.md:87-92
        const S d = pv[1] - ext[0];                        // tracking error Vref - V_remote
        dx[1] = d;                                         // integrator state
        const S inj = pv[3] * d + pv[2] * x[1];            // Kp*d + Ki*xi
        dx[0] = 1.0 / pv[0] * (acc + inj);
    } else if constexpr (VM == ND_VM_CAPACITOR) {          // loopback_test.jl:61: D(v) ~ i / C
        dx[0] = acc / pv[0];
    } else if constexpr (VM == ND_VM_INDUCTOR_INJECTOR) {  // loopback_test.jl:89: D(i) ~ v / L
        dx[0] = acc / pv[0];
    } else if constexpr (VM == ND_VM_CRL) {                // loopback_test.jl:113-131
        const S v = x[0];
        dx[0] = (acc - v / pv[1] - x[1]) / pv[0];          // cap.p.i = i - v/R - iL
        dx[1] = v / pv[2];
    } else {
        vertex_f<VM, S, P>(dx, x, acc, pv, dim);
    }
}

// Computed vertex outputs: NoFeedForward g(out, x, p, t) and the PureFeedForward g of an
// injector, g(out, in, p, t), whose input is its hub's output (the loopback copy).
template <typename S, typename P = double>
__device__ __forceinline__ S vertex_out_nff(int vm, const P* __restrict__ pv, int j = 0) {
    if (vm == ND_VM_VOLTAGE_SOURCE) return S{} + pv[0];    // loopback_test.jl:31: p.v ~ V
    if (vm == ND_VM_TRUSS_FIXED) return S{} + pv[j];       // stress_on_truss.jl:45-47: pos .= p
    return nan_of<S>();
}
template <typename S, typename P = double>
__device__ __forceinline__ S vertex_out_ff(int vm, S in, const P* __restrict__ pv) {
    if (vm == ND_VM_RESISTOR_INJECTOR) return in / pv[0];   // loopback_test.jl:75: p.i ~ p.v / R
    return nan_of<S>();
}

// Component j of a two-output vector edge model's dst output, from the source and
// destination outputs ps, pd (stress_on_truss.jl:77-84: the beam force
// F = K*(L - d)*(pd - ps)/d, d = |pd - ps|; x^2 is x*x)
template <typename S, typename P = double>
__device__ __forceinline__ S beam_dst(const S* ps, const S* pd, const P* __restrict__ pe, int j) {
    const S dx = pd[0] - ps[0], dy = pd[1] - ps[1];
    const S d = sqrt(dx * dx + dy * dy);
    const S f = pe[0] * (pe[1] - d);
    return j == 0 ? f * dx / d : f * dy / d;
}

// f of a free truss node (stress_on_truss.jl:55-61): states (vx, vy, x, y), inputs (Fx, Fy)
template <typename S, typename P = double>
__device__ __forceinline__ void truss_free_f(S* dx, const S* x, const S* acc, const P* __restrict__ pv) {
    dx[0] = (acc[0] - pv[1] * x[0]) / pv[0];
    dx[1] = (acc[1] - pv[1] * x[1]) / pv[0];
    dx[1] = dx[1] - pv[2];
    dx[2] = x[0];
    dx[3] = x[1];
}

// Parameters each edge model reads (p entries per edge).
template <int EM> struct EModel { static constexpr int np = 0; };
template <> struct EModel<ND_EM_KURAMOTO> { static constexpr int np = 1; };
template <> struct EModel<ND_EM_STATIC_LINE> { static constexpr int np = 2; };
template <> struct EModel<ND_EM_DIFFUSION_WEIGHTED> { static constexpr int np = 1; };
template <> struct EModel<ND_EM_DIFFUSION_FIDUCIAL> { static constexpr int np = 1; };
template <> struct EModel<ND_EM_RESISTOR> { static constexpr int np = 1; };

// parameters per vertex model
template <int VM> struct VPar { static constexpr int np = 0; };
template <> struct VPar<ND_VM_KURAMOTO> { static constexpr int np = 1; };
template <> struct VPar<ND_VM_SWING> { static constexpr int np = 3; };
template <> struct VPar<ND_VM_SWING_UNIT_DAMPING> { static constexpr int np = 1; };
template <> struct VPar<ND_VM_LOAD> { static constexpr int np = 2; };
template <> struct VPar<ND_VM_CAPACITOR_SINK> { static constexpr int np = 1; };
template <> struct VPar<ND_VM_PI_SOURCE> { static constexpr int np = 4; };
template <> struct VPar<ND_VM_CAPACITOR> { static constexpr int np = 1; };
template <> struct VPar<ND_VM_VOLTAGE_SOURCE> { static constexpr int np = 1; };
template <> struct VPar<ND_VM_RESISTOR_INJECTOR> { static constexpr int np = 1; };
template <> struct VPar<ND_VM_INDUCTOR_INJECTOR> { static constexpr int np = 1; };
template <> struct VPar<ND_VM_CRL> { static constexpr int np = 3; };
template <> struct VPar<ND_VM_FHN> { static constexpr int np = 2; };
template <> struct VPar<ND_VM_TRUSS_FREE> { static constexpr int np = 3; };
template <> struct VPar<ND_VM_TRUSS_FIXED> { static constexpr int np = 2; };

// external inputs each vertex model reads
template <int VM> struct VExt { static constexpr int n = 0; };
template <> struct VExt<ND_VM_PI_SOURCE> { static constexpr int n = 1; };

}  // namespace nd
