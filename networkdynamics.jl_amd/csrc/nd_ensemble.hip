// Batched ensemble integrator: thousands of independent (uneven) networks, one
// workgroup per instance, whole trajectory in one launch per size class.
//
// Per instance everything a stage touches is on-chip:
//   registers  per-vertex state y and the Runge-Kutta stage derivatives (up to V
//              vertices per thread);
//   LDS        stage vertex outputs theta[n], edge parameters, the destination
//              lists (neighbour | role bit, edge index) and their chunking.
// A stage: write theta, barrier, evaluate chunks (<= kChunk entries of one vertex's
// destination list each, so hubs are split over threads), barrier, each vertex sums
// its chunk partials in order and applies its model.  Vertices whose list fits one
// chunk are summed in the reference order (increasing output index).
//
// Methods: Tsit5 (Tsitouras 2011) with the PI step-size controller of
// OrdinaryDiffEq (beta1 = 7/50, beta2 = 2/25, qmin = 1/5, qmax = 10, gamma = 9/10,
// error norm sqrt(mean((err / (abstol + reltol*max(|y|,|ynew|)))^2))), and fixed-step
// classic RK4.  Step acceptance is block-uniform, so every wave leaves the loop
// together.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "nd_internal.h"
#include "nd_models.h"

namespace {

using namespace nd;
constexpr int kWave = 64;
constexpr int kChunk = 8;          // destination-list entries per chunk

struct InstDesc {
    int32_t n, ne, nslot, nchunk;
    int64_t uoff, poff, soff, csoff, cvoff, cpoff;
    int32_t lds_bytes, cls;
};

struct SolveArgs {
    const InstDesc* inst;
    const int32_t* order;
    const uint16_t* ends;         // per edge: local src, local dst (2 per edge, offset soff)
    const uint16_t* slot_e;       // local edge index | 0x8000 if this vertex is the edge's src
    const uint16_t* cs;           // chunk starts (nchunk + 1 per instance)
    const uint16_t* cv;           // chunk vertex
    const uint16_t* cptr;         // vertex -> chunk range (n + 1 per instance)
    double* u;
    double* p;                    // written only where a line trips (line_limit set)
    const double* limit;          // per edge (instance order), or null: no line tripping
    const double* event_t;        // per instance: preset trip time, or null
    const int32_t* event_e;       // per instance: the edge it trips (local index, -1 none)
    nd_ensemble_stat* stats;
    double t0, t1, dt0, reltol, abstol, dtmax, dtmin;
    double ss_abstol, ss_reltol;  // steady-state termination (ss != 0)
    int64_t maxiters;
    int32_t first, count;
    int32_t ss;
};


__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    const int t = static_cast<int>(threadIdx.x), nw = static_cast<int>(blockDim.x) / kWave;
    if ((t & (kWave - 1)) == 0) red[t / kWave] = v;
    __syncthreads();
    if (t == 0) {
        double s = red[0];
        for (int w = 1; w < nw; ++w) s += red[w];
        red[16] = s;
    }
    __syncthreads();
    const double r = red[16];
    __syncthreads();
    return r;
}

struct Lds {
    double* theta;
    double* epar;
    double* fval;                 // dst-side output of each edge at the current stage
    double* partial;
    double* red;
    uint16_t *ends, *eidx, *cs, *cv, *cptr;
};

// Vertex parameters stay in registers for the whole solve up to three vertices per
// thread; at four they are read from memory (L1/L2-resident) in each vertex update, which
// frees the registers whose spill would otherwise cost more
template <int V>
constexpr bool kParamsInMemory = V >= 4;

// out = f(ys) for the thread's vertices (three block barriers): each edge's output is
// evaluated once (the src side takes it negated, AntiSymmetric), then every chunk sums
// its entries in list order.  pv: the vertex parameters in registers, or pg: the
// instance's vertex parameters in memory (kParamsInMemory)
template <int VM, int EM, int V>
__device__ __forceinline__ void rhs(const InstDesc& I, const Lds& L, const double (&ys)[V][VModel<VM>::dim],
                                    const double (&pv)[V][VPar<VM>::np > 0 ? VPar<VM>::np : 1],
                                    const double* __restrict__ pg, double (&out)[V][VModel<VM>::dim]) {
    constexpr int D = VModel<VM>::dim;
    constexpr int NPE = EModel<EM>::np;
    const int T = static_cast<int>(blockDim.x), t = static_cast<int>(threadIdx.x);
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int v = t + k * T;
        if (v < I.n) L.theta[v] = ys[k][0];
    }
    __syncthreads();
    for (int e = t; e < I.ne; e += T)
        L.fval[e] = g_dst<EM>(L.theta[L.ends[2 * e]], L.theta[L.ends[2 * e + 1]], L.epar + e * (NPE > 0 ? NPE : 1));
    __syncthreads();
    for (int c = t; c < I.nchunk; c += T) {
        double acc = 0.0;
        const int s1 = L.cs[c + 1];
        for (int s = L.cs[c]; s < s1; ++s) {
            const int se = L.eidx[s];
            const double f = L.fval[se & 0x7fff];
            acc += (se & 0x8000) ? -f : f;
        }
        L.partial[c] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int v = t + k * T;
        if (v < I.n) {
            double acc = 0.0;
            const int c1 = L.cptr[v + 1];
            for (int c = L.cptr[v]; c < c1; ++c) acc += L.partial[c];
            if constexpr (kParamsInMemory<V>)
                vertex_f<VM>(out[k], ys[k], acc, pg + static_cast<int64_t>(v) * VPar<VM>::np, D);
            else
                vertex_f<VM>(out[k], ys[k], acc, pv[k], D);
        }
    }
}

// fval[e] = dst-side output of every edge at the state ys (one block barrier pair)
template <int VM, int EM, int V>
__device__ __forceinline__ void edge_flows(const InstDesc& I, const Lds& L, const double (&ys)[V][VModel<VM>::dim]) {
    constexpr int NPE = EModel<EM>::np;
    const int T = static_cast<int>(blockDim.x), t = static_cast<int>(threadIdx.x);
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int v = t + k * T;
        if (v < I.n) L.theta[v] = ys[k][0];
    }
    __syncthreads();
    for (int e = t; e < I.ne; e += T)
        L.fval[e] = g_dst<EM>(L.theta[L.ends[2 * e]], L.theta[L.ends[2 * e + 1]], L.epar + e * (NPE > 0 ? NPE : 1));
    __syncthreads();
}

// Line tripping (the component callback of the reference's cascading-failure example,
// docs/examples/cascading_failure.jl: condition |P_line| - limit, affect K = 0), checked
// at the end of every accepted step on the flows in fval: a line whose |flow| exceeds
// its limit gets its first parameter set to 0, in LDS and in p.  Returns the number of
// lines tripped in this check (block-uniform).
template <int VM, int EM>
__device__ __forceinline__ int trip_lines(const SolveArgs& a, const InstDesc& I, const Lds& L) {
    constexpr int NPE = EModel<EM>::np > 0 ? EModel<EM>::np : 1;
    const int T = static_cast<int>(blockDim.x), t = static_cast<int>(threadIdx.x);
    const double* lim = a.limit + I.soff / 2;
    double* pe_g = a.p + I.poff + static_cast<int64_t>(I.n) * VPar<VM>::np;
    double cnt = 0.0;
    for (int e = t; e < I.ne; e += T) {
        if (L.epar[e * NPE] != 0.0 && fabs(L.fval[e]) > lim[e]) {
            L.epar[e * NPE] = 0.0;
            pe_g[static_cast<int64_t>(e) * NPE] = 0.0;
            cnt += 1.0;
        }
    }
    return static_cast<int>(block_sum(cnt, L.red));
}

// The preset-time trip of one edge (PresetTimeComponentCallback with the same affect):
// its first parameter goes to 0 in LDS and in p.  Returns 1 if it was not 0 before
// (block-uniform).
template <int VM, int EM>
__device__ __forceinline__ int trip_edge(const SolveArgs& a, const InstDesc& I, const Lds& L, int e) {
    constexpr int NPE = EModel<EM>::np > 0 ? EModel<EM>::np : 1;
    const bool was = L.epar[e * NPE] != 0.0;
    __syncthreads();
    if (threadIdx.x == 0) {
        L.epar[e * NPE] = 0.0;
        a.p[I.poff + static_cast<int64_t>(I.n) * VPar<VM>::np + static_cast<int64_t>(e) * NPE] = 0.0;
    }
    __syncthreads();
    return was ? 1 : 0;
}

// Steady-state test after an accepted step (the allDerivPass test of DiffEqCallbacks'
// TerminateSteadyState): every |du_i| <= ss_abstol or <= ss_reltol*|u_i|.  Block-uniform.
template <int V, int D>
__device__ __forceinline__ bool at_steady_state(const SolveArgs& a, const InstDesc& I, const Lds& L,
                                                const double (&y)[V][D], const double (&du)[V][D]) {
    const int T = static_cast<int>(blockDim.x), t = static_cast<int>(threadIdx.x);
    double fail = 0.0;
#pragma unroll
    for (int q = 0; q < V; ++q)
        if (t + q * T < I.n)
#pragma unroll
            for (int j = 0; j < D; ++j) {
                const double d = fabs(du[q][j]);
                if (!(d <= a.ss_abstol || d <= a.ss_reltol * fabs(y[q][j]))) fail = 1.0;
            }
    return block_sum(fail, L.red) == 0.0;
}

// Line crossings (ND_LINES_CROSSING).  sign8 codes the sign of a condition value (2 for
// NaN, which never crosses); crossed(p, n) is the event test of DiffEqBase's
// VectorContinuousCallback when both directions have an affect: the sign p at the step
// start is nonzero and p*n <= 0.
__device__ __forceinline__ int8_t sign8(double g) {
    return static_cast<int8_t>(g > 0.0 ? 1 : g < 0.0 ? -1 : g == 0.0 ? 0 : 2);
}
__device__ __forceinline__ bool crossed(int8_t p, int8_t n) {
    return (p == -1 && (n == 1 || n == 0)) || (p == 1 && (n == -1 || n == 0));
}

// the condition |P_line| - limit of edge e (docs/examples/cascading_failure.jl:104-106)
// at the phases th
template <int EM>
__device__ __forceinline__ double line_cond(const Lds& L, const double* th, const double* lim, int e) {
    constexpr int NPE = EModel<EM>::np > 0 ? EModel<EM>::np : 1;
    return fabs(g_dst<EM>(th[L.ends[2 * e]], th[L.ends[2 * e + 1]], L.epar + e * NPE)) - lim[e];
}

// sg[e] = sign8 of every edge's condition at the state y (two block barriers)
template <int VM, int EM, int V>
__device__ __forceinline__ void line_signs(const InstDesc& I, const Lds& L, const double* lim,
                                           const double (&y)[V][VModel<VM>::dim], int8_t* sg) {
    const int T = static_cast<int>(blockDim.x), t = static_cast<int>(threadIdx.x);
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int v = t + k * T;
        if (v < I.n) L.theta[v] = y[k][0];
    }
    __syncthreads();
    for (int e = t; e < I.ne; e += T) sg[e] = sign8(line_cond<EM>(L, L.theta, lim, e));
    __syncthreads();
}

__device__ __forceinline__ double block_min(double v, double* red) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, kWave));
    const int t = static_cast<int>(threadIdx.x), nw = static_cast<int>(blockDim.x) / kWave;
    if ((t & (kWave - 1)) == 0) red[t / kWave] = v;
    __syncthreads();
    if (t == 0) {
        double s = red[0];
        for (int w = 1; w < nw; ++w) s = fmin(s, red[w]);
        red[16] = s;
    }
    __syncthreads();
    const double r = red[16];
    __syncthreads();
    return r;
}

// Tsitouras (2011) 5(4) coefficients: stage matrix a (row s = stage s+1), error weights
// btilde = b - bhat (b = the last row, first same as last) and nodes c = a * 1.  The
// kernel reads these names and nd_tsit5_tableau exports the same table, so the order
// conditions checked on the exported copy hold for the kernel's.
struct Tsit5Tableau {
    double a[7][6];
    double bt[7];
    double c[7];
};
constexpr Tsit5Tableau kTsit5 = {
    {{0, 0, 0, 0, 0, 0},
     {0.161, 0, 0, 0, 0, 0},
     {-0.008480655492356989, 0.335480655492357, 0, 0, 0, 0},
     {2.897153057105493, -6.359448489975075, 4.3622954328695815, 0, 0, 0},
     {5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525, 0, 0},
     {5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401, -0.028269050394068383, 0},
     {0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081, 2.324710524099774}},
    {-0.00178001105222577714, -0.0008164344596567469, 0.007880878010261995, -0.1447110071732629,
     0.5823571654525552, -0.45808210592918697, 0.015151515151515152},
    {0.0, 0.161, 0.327, 0.9, 0.9800255409045097, 1.0, 1.0}};
constexpr double c_a21 = kTsit5.a[1][0];
constexpr double c_a31 = kTsit5.a[2][0], c_a32 = kTsit5.a[2][1];
constexpr double c_a41 = kTsit5.a[3][0], c_a42 = kTsit5.a[3][1], c_a43 = kTsit5.a[3][2];
constexpr double c_a51 = kTsit5.a[4][0], c_a52 = kTsit5.a[4][1], c_a53 = kTsit5.a[4][2], c_a54 = kTsit5.a[4][3];
constexpr double c_a61 = kTsit5.a[5][0], c_a62 = kTsit5.a[5][1], c_a63 = kTsit5.a[5][2], c_a64 = kTsit5.a[5][3],
                 c_a65 = kTsit5.a[5][4];
constexpr double c_a71 = kTsit5.a[6][0], c_a72 = kTsit5.a[6][1], c_a73 = kTsit5.a[6][2], c_a74 = kTsit5.a[6][3],
                 c_a75 = kTsit5.a[6][4], c_a76 = kTsit5.a[6][5];
constexpr double c_bt1 = kTsit5.bt[0], c_bt2 = kTsit5.bt[1], c_bt3 = kTsit5.bt[2], c_bt4 = kTsit5.bt[3],
                 c_bt5 = kTsit5.bt[4], c_bt6 = kTsit5.bt[5], c_bt7 = kTsit5.bt[6];

// Tsitouras' continuous extension of the pair, the dense output OrdinaryDiffEqTsit5
// interpolates solutions and locates callback events on: u(t + s h) = u + h sum_i b_i(s) k_i,
// b_i(s) = sum_j r[i][j-1] s^j (only stage 1 has a linear term; b_i(1) is the last row of
// a).  nd_tsit5_dense exports this table, so its order conditions are checked on the
// kernel's copy.
struct Tsit5Dense {
    double r[7][4];
};
constexpr Tsit5Dense kTsit5Dense = {{{1.0, -2.763706197274826, 2.9132554618219126, -1.0530884977290216},
                                     {0.0, 0.13169999999999998, -0.2234, 0.1017},
                                     {0.0, 3.9302962368947516, -5.941033872131505, 2.490627285651253},
                                     {0.0, -12.411077166933676, 30.33818863028232, -16.548102889244902},
                                     {0.0, 37.50931341651104, -88.1789048947664, 47.37952196281928},
                                     {0.0, -27.896526289197286, 65.09189467479366, -34.8706578614966},
                                     {0.0, 1.5, -4.0, 2.5}}};
#define ND_R(i, j) kTsit5Dense.r[i - 1][j - 1]

// the dense output at s from the step's start value y0, first derivative k1 and the
// coefficients d_j = sum_i r_ij k_i of s^2, s^3, s^4
__device__ __forceinline__ double dense_at(double y0, double k1, double d2, double d3, double d4, double h, double s) {
    return y0 + h * (s * (k1 + s * (d2 + s * (d3 + s * d4))));
}

template <int VM, int EM, int METHOD, int V, int LB>
__global__ void __launch_bounds__(LB) k_ensemble(SolveArgs a) {
    constexpr int D = VModel<VM>::dim;
    constexpr int NPV = VPar<VM>::np > 0 ? VPar<VM>::np : 1;
    constexpr int NPE = EModel<EM>::np;
    const int T = static_cast<int>(blockDim.x), t = static_cast<int>(threadIdx.x);
    const int inst = a.order[a.first + static_cast<int>(blockIdx.x)];
    const InstDesc I = a.inst[inst];
    extern __shared__ double smem[];
    Lds L;
    L.theta = smem;
    L.epar = L.theta + I.n;
    L.fval = L.epar + I.ne * NPE;
    L.partial = L.fval + I.ne;
    L.red = L.partial + I.nchunk;
    L.ends = reinterpret_cast<uint16_t*>(L.red + 32);
    L.eidx = L.ends + I.nslot;
    L.cs = L.eidx + I.nslot;
    L.cv = L.cs + I.nchunk + 1;
    L.cptr = L.cv + I.nchunk;
    for (int i = t; i < I.nslot; i += T) {
        L.ends[i] = a.ends[I.soff + i];
        L.eidx[i] = a.slot_e[I.soff + i];
    }
    for (int i = t; i <= I.nchunk; i += T) L.cs[i] = a.cs[I.csoff + i];
    for (int i = t; i < I.nchunk; i += T) L.cv[i] = a.cv[I.cvoff + i];
    for (int i = t; i <= I.n; i += T) L.cptr[i] = a.cptr[I.cpoff + i];
    const double* pe_g = a.p + I.poff + static_cast<int64_t>(I.n) * VPar<VM>::np;
    for (int i = t; i < I.ne * NPE; i += T) L.epar[i] = pe_g[i];

    double y[V][D], pv[V][NPV];
    const double* pg = a.p + I.poff;                 // the instance's vertex parameters
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int v = t + k * T;
#pragma unroll
        for (int j = 0; j < D; ++j) y[k][j] = v < I.n ? a.u[I.uoff + static_cast<int64_t>(v) * D + j] : 0.0;
#pragma unroll
        for (int j = 0; j < NPV; ++j)
            pv[k][j] = (!kParamsInMemory<V> && v < I.n && VPar<VM>::np > 0)
                           ? a.p[I.poff + static_cast<int64_t>(v) * VPar<VM>::np + j] : 0.0;
    }
    __syncthreads();

    double tcur = a.t0;
    int64_t naccept = 0, nreject = 0, nf = 0;
    int retcode = 0, ntrip = 0;
    // preset trip: fires once at te (t0 < te <= t1), else never
    double te = __builtin_inf();
    int32_t ee = -1;
    if (a.event_t) {
        te = a.event_t[inst];
        ee = a.event_e[inst];
        if (!(te > a.t0 && te <= a.t1) || ee < 0 || ee >= I.ne) {
            te = __builtin_inf();
            ee = -1;
        }
    }
    double dt = a.dt0;
    const double inv_len = 1.0 / (static_cast<double>(I.n) * D);

    if constexpr (METHOD == 1) {
        // fixed-step classic RK4: u + dt/6*(((k1 + 2k2) + 2k3) + k4)
        double k[V][D], acc[V][D], ys[V][D];
        const int64_t nsteps = static_cast<int64_t>(llround((a.t1 - a.t0) / dt));
        const double half = 0.5 * dt, sixth = dt / 6.0;
        int64_t done = nsteps;
        for (int64_t step = 0; step < nsteps; ++step) {
            rhs<VM, EM, V>(I, L, y, pv, pg, k);
#pragma unroll
            for (int q = 0; q < V; ++q)
#pragma unroll
                for (int j = 0; j < D; ++j) { acc[q][j] = k[q][j]; ys[q][j] = y[q][j] + half * k[q][j]; }
            rhs<VM, EM, V>(I, L, ys, pv, pg, k);
#pragma unroll
            for (int q = 0; q < V; ++q)
#pragma unroll
                for (int j = 0; j < D; ++j) { acc[q][j] = acc[q][j] + 2.0 * k[q][j]; ys[q][j] = y[q][j] + half * k[q][j]; }
            rhs<VM, EM, V>(I, L, ys, pv, pg, k);
#pragma unroll
            for (int q = 0; q < V; ++q)
#pragma unroll
                for (int j = 0; j < D; ++j) { acc[q][j] = acc[q][j] + 2.0 * k[q][j]; ys[q][j] = y[q][j] + dt * k[q][j]; }
            rhs<VM, EM, V>(I, L, ys, pv, pg, k);
#pragma unroll
            for (int q = 0; q < V; ++q)
#pragma unroll
                for (int j = 0; j < D; ++j) y[q][j] = y[q][j] + sixth * (acc[q][j] + k[q][j]);
            nf += 4;
            ++naccept;
            if (a.limit) {
                edge_flows<VM, EM, V>(I, L, y);
                ntrip += trip_lines<VM, EM>(a, I, L);
            }
            if (ee >= 0 && step + 1 == static_cast<int64_t>(llround((te - a.t0) / dt))) ntrip += trip_edge<VM, EM>(a, I, L, ee);
            if (a.ss) {
                rhs<VM, EM, V>(I, L, y, pv, pg, k);
                ++nf;
                if (at_steady_state<V, D>(a, I, L, y, k)) {
                    done = step + 1;
                    retcode = 4;
                    break;
                }
            }
        }
        tcur = a.t0 + static_cast<double>(done) * dt;
    } else {
        double k1[V][D], k2[V][D], k3[V][D], k4[V][D], k5[V][D], k6[V][D], k7[V][D], ys[V][D];
        double d2[V][D], d3[V][D], d4[V][D];      // dense output (crossing location only)
        rhs<VM, EM, V>(I, L, y, pv, pg, k1);
        ++nf;
        auto wnorm2 = [&](const double (&x)[V][D], const double (&ref)[V][D]) {
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < V; ++q)
                if (t + q * T < I.n)
#pragma unroll
                    for (int j = 0; j < D; ++j) {
                        const double sc = a.abstol + fabs(ref[q][j]) * a.reltol;
                        const double r = x[q][j] / sc;
                        s += r * r;
                    }
            return block_sum(s, L.red);
        };
        if (!(dt > 0.0)) {
            // initial step (Hairer, Norsett & Wanner, Solving ODEs I, II.4)
            const double d0 = sqrt(wnorm2(y, y) * inv_len), d1 = sqrt(wnorm2(k1, y) * inv_len);
            double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
            h0 = fmin(h0, a.t1 - a.t0);
#pragma unroll
            for (int q = 0; q < V; ++q)
#pragma unroll
                for (int j = 0; j < D; ++j) ys[q][j] = y[q][j] + h0 * k1[q][j];
            rhs<VM, EM, V>(I, L, ys, pv, pg, k2);
            ++nf;
#pragma unroll
            for (int q = 0; q < V; ++q)
#pragma unroll
                for (int j = 0; j < D; ++j) k3[q][j] = k2[q][j] - k1[q][j];
            const double d2 = sqrt(wnorm2(k3, y) * inv_len) / h0;
            const double m = fmax(d1, d2);
            const double h1 = m <= 1e-15 ? fmax(1e-6, h0 * 1e-3) : pow(0.01 / m, 1.0 / 6.0);
            dt = fmin(100.0 * h0, h1);
        }
        dt = fmin(dt, a.dtmax);
        double qold = 1e-4;
        int64_t iters = 0;
        double tstop = ee >= 0 ? te : a.t1;       // the next time a step must end on
        // crossing location: a second phase buffer and, per edge, the condition's sign at
        // the step start (gsign) and end (gnext) and the first point of the step where it
        // crossed (gev), in LDS after the u16 tables
        constexpr bool kLocate = METHOD == 2;
        const double* lim = a.limit ? a.limit + I.soff / 2 : nullptr;
        double* theta2 = nullptr;
        int8_t *gsign = nullptr, *gnext = nullptr, *gev = nullptr;
        if constexpr (kLocate) {
            char* base = reinterpret_cast<char*>(smem);
            const size_t off = (static_cast<size_t>(reinterpret_cast<char*>(L.cptr + I.n + 1) - base) + 7) & ~size_t(7);
            theta2 = reinterpret_cast<double*>(base + off);
            gsign = reinterpret_cast<int8_t*>(theta2 + I.n);
            gnext = gsign + I.ne;
            gev = gnext + I.ne;
            if (lim) line_signs<VM, EM, V>(I, L, lim, y, gsign);
        }
        while (tcur < a.t1) {
            if (iters++ >= a.maxiters) { retcode = 1; break; }
            if (dt < a.dtmin) { retcode = 2; break; }
            const bool last = tcur + dt >= tstop;
            const double h = last ? tstop - tcur : dt;
#define ND_STAGE(expr)                                                        \
    _Pragma("unroll") for (int q = 0; q < V; ++q)                             \
        _Pragma("unroll") for (int j = 0; j < D; ++j) ys[q][j] = y[q][j] + h * (expr);
            ND_STAGE(c_a21 * k1[q][j]);
            rhs<VM, EM, V>(I, L, ys, pv, pg, k2);
            ND_STAGE(c_a31 * k1[q][j] + c_a32 * k2[q][j]);
            rhs<VM, EM, V>(I, L, ys, pv, pg, k3);
            ND_STAGE(c_a41 * k1[q][j] + c_a42 * k2[q][j] + c_a43 * k3[q][j]);
            rhs<VM, EM, V>(I, L, ys, pv, pg, k4);
            ND_STAGE(c_a51 * k1[q][j] + c_a52 * k2[q][j] + c_a53 * k3[q][j] + c_a54 * k4[q][j]);
            rhs<VM, EM, V>(I, L, ys, pv, pg, k5);
            ND_STAGE(c_a61 * k1[q][j] + c_a62 * k2[q][j] + c_a63 * k3[q][j] + c_a64 * k4[q][j] + c_a65 * k5[q][j]);
            rhs<VM, EM, V>(I, L, ys, pv, pg, k6);
            ND_STAGE(c_a71 * k1[q][j] + c_a72 * k2[q][j] + c_a73 * k3[q][j] + c_a74 * k4[q][j] + c_a75 * k5[q][j] +
                     c_a76 * k6[q][j]);
#undef ND_STAGE
            if constexpr (kLocate) {
                // dense-output coefficients from k1..k6 (k7 follows); they take the place
                // of k3..k6 in registers
#pragma unroll
                for (int q = 0; q < V; ++q)
#pragma unroll
                    for (int j = 0; j < D; ++j) {
                        d2[q][j] = ND_R(1, 2) * k1[q][j] + ND_R(2, 2) * k2[q][j] + ND_R(3, 2) * k3[q][j] +
                                   ND_R(4, 2) * k4[q][j] + ND_R(5, 2) * k5[q][j] + ND_R(6, 2) * k6[q][j];
                        d3[q][j] = ND_R(1, 3) * k1[q][j] + ND_R(2, 3) * k2[q][j] + ND_R(3, 3) * k3[q][j] +
                                   ND_R(4, 3) * k4[q][j] + ND_R(5, 3) * k5[q][j] + ND_R(6, 3) * k6[q][j];
                        d4[q][j] = ND_R(1, 4) * k1[q][j] + ND_R(2, 4) * k2[q][j] + ND_R(3, 4) * k3[q][j] +
                                   ND_R(4, 4) * k4[q][j] + ND_R(5, 4) * k5[q][j] + ND_R(6, 4) * k6[q][j];
                    }
            }
            // error-estimate combination of k1..k6 before the last evaluation, so only
            // y, ynew and this sum stay live across it
#pragma unroll
            for (int q = 0; q < V; ++q)
#pragma unroll
                for (int j = 0; j < D; ++j)
                    k2[q][j] = c_bt1 * k1[q][j] + c_bt2 * k2[q][j] + c_bt3 * k3[q][j] + c_bt4 * k4[q][j] +
                               c_bt5 * k5[q][j] + c_bt6 * k6[q][j];
            rhs<VM, EM, V>(I, L, ys, pv, pg, k7);
            nf += 6;
            if constexpr (kLocate) {
#pragma unroll
                for (int q = 0; q < V; ++q)
#pragma unroll
                    for (int j = 0; j < D; ++j) {
                        d2[q][j] = d2[q][j] + ND_R(7, 2) * k7[q][j];
                        d3[q][j] = d3[q][j] + ND_R(7, 3) * k7[q][j];
                        d4[q][j] = d4[q][j] + ND_R(7, 4) * k7[q][j];
                    }
            }
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < V; ++q)
                if (t + q * T < I.n)
#pragma unroll
                    for (int j = 0; j < D; ++j) {
                        const double e = h * (k2[q][j] + c_bt7 * k7[q][j]);
                        const double sc = a.abstol + fmax(fabs(y[q][j]), fabs(ys[q][j])) * a.reltol;
                        const double r = e / sc;
                        s += r * r;
                    }
            const double eest = sqrt(block_sum(s, L.red) * inv_len);
            const double q11 = pow(eest, 0.14);
            if (eest <= 1.0) {
                double qq = eest == 0.0 ? 0.1 : fmax(0.1, fmin(5.0, q11 / pow(qold, 0.08) / 0.9));
                // a step that ends within 100 ulp of the stop time ends on it (as
                // OrdinaryDiffEq's fixed_t_for_floatingpoint_error! does)
                const double tn = tcur + h;
                const bool at_stop = last || fabs(tn - tstop) < 100.0 * 2.220446049250313e-16 * fmax(fabs(tcur), fabs(tstop));
                qold = fmax(eest, 1e-4);
                dt = fmin(h / qq, a.dtmax);
                ++naccept;
                int changed = 0;
                bool located = false;
                if constexpr (kLocate) {
                    if (lim) {
                        // the condition at the points s = m/9 of the step (m = 9: the accepted
                        // state), against its sign at the step start
                        double first = 10.0;
#pragma unroll 1
                        for (int m = 1; m <= 9; ++m) {
                            double* th = (m & 1) ? theta2 : L.theta;   // alternate: one barrier a point
                            const double s = static_cast<double>(m) / 9.0;
#pragma unroll
                            for (int q = 0; q < V; ++q) {
                                const int v = t + q * T;
                                if (v < I.n)
                                    th[v] = m == 9 ? ys[q][0]
                                                   : dense_at(y[q][0], k1[q][0], d2[q][0], d3[q][0], d4[q][0], h, s);
                            }
                            __syncthreads();
                            for (int e = t; e < I.ne; e += T) {
                                const int8_t ns = sign8(line_cond<EM>(L, th, lim, e));
                                if (m == 1) gev[e] = 0;
                                if (gev[e] == 0 && crossed(gsign[e], ns)) {
                                    gev[e] = static_cast<int8_t>(m);
                                    first = fmin(first, static_cast<double>(m));
                                }
                                if (m == 9) gnext[e] = ns;
                            }
                        }
                        first = block_min(first, L.red);
                        if (first < 10.0) {
                            // bisect [point mf-1, point mf] on the dense output, over the edges
                            // that crossed at point mf, down to adjacent doubles
                            const int mf = static_cast<int>(first);
                            double lo = tcur + h * (static_cast<double>(mf - 1) / 9.0);
                            double hi = tcur + h * (static_cast<double>(mf) / 9.0);
                            bool hi_end = mf == 9;                  // hi: still the accepted state
                            for (int it = 0; it < 256; ++it) {
                                const double mid = lo + 0.5 * (hi - lo);
                                if (!(mid > lo && mid < hi)) break;
                                const double s = (mid - tcur) / h;
#pragma unroll
                                for (int q = 0; q < V; ++q) {
                                    const int v = t + q * T;
                                    if (v < I.n) L.theta[v] = dense_at(y[q][0], k1[q][0], d2[q][0], d3[q][0], d4[q][0], h, s);
                                }
                                __syncthreads();
                                double hit = 0.0;
                                for (int e = t; e < I.ne; e += T)
                                    if (gev[e] == mf && crossed(gsign[e], sign8(line_cond<EM>(L, L.theta, lim, e)))) hit = 1.0;
                                if (block_sum(hit, L.red) > 0.0) {
                                    hi = mid;
                                    hi_end = false;
                                } else {
                                    lo = mid;
                                }
                            }
                            // the edges that crossed by hi trip; the solution moves to lo, the
                            // left end of the bracket (DiffEqBase's LeftRootFind)
                            const double shi = (hi - tcur) / h;
#pragma unroll
                            for (int q = 0; q < V; ++q) {
                                const int v = t + q * T;
                                if (v < I.n)
                                    L.theta[v] = hi_end ? ys[q][0]
                                                        : dense_at(y[q][0], k1[q][0], d2[q][0], d3[q][0], d4[q][0], h, shi);
                            }
                            __syncthreads();
                            constexpr int NPE1 = NPE > 0 ? NPE : 1;
                            double* pw = a.p + I.poff + static_cast<int64_t>(I.n) * VPar<VM>::np;
                            double cnt = 0.0;
                            for (int e = t; e < I.ne; e += T)
                                if (gev[e] == mf && crossed(gsign[e], sign8(line_cond<EM>(L, L.theta, lim, e)))) {
                                    L.epar[e * NPE1] = 0.0;
                                    pw[static_cast<int64_t>(e) * NPE1] = 0.0;
                                    cnt += 1.0;
                                }
                            ntrip += static_cast<int>(block_sum(cnt, L.red));
                            const double slo = (lo - tcur) / h;
#pragma unroll
                            for (int q = 0; q < V; ++q)
#pragma unroll
                                for (int j = 0; j < D; ++j)
                                    y[q][j] = dense_at(y[q][j], k1[q][j], d2[q][j], d3[q][j], d4[q][j], h, slo);
                            tcur = lo;
                            located = true;
                            changed = 1;                            // the state moved: new derivative
                        }
                    }
                }
                if (!located) {
                    tcur = at_stop ? tstop : tn;
#pragma unroll
                    for (int q = 0; q < V; ++q)
#pragma unroll
                        for (int j = 0; j < D; ++j) { y[q][j] = ys[q][j]; k1[q][j] = k7[q][j]; }
                    if constexpr (kLocate) {
                        int8_t* sw = gsign;                         // the end signs start the next step
                        gsign = gnext;
                        gnext = sw;
                    } else if (a.limit) {
                        // fval holds the flows at the accepted state (the last evaluation was at ys)
                        const int tr = trip_lines<VM, EM>(a, I, L);
                        ntrip += tr;
                        changed += tr;
                    }
                    if (at_stop && ee >= 0 && tstop == te) {   // the preset trip's time is reached
                        const int tr = trip_edge<VM, EM>(a, I, L, ee);
                        ntrip += tr;
                        changed += tr;
                        ee = -1;
                        tstop = a.t1;
                    }
                }
                // after a trip or a move to a crossing the first-same-as-last derivative
                // (and the condition signs) are stale: re-evaluate them
                if (changed > 0) {
                    rhs<VM, EM, V>(I, L, y, pv, pg, k1);
                    ++nf;
                    if constexpr (kLocate)
                        if (lim) line_signs<VM, EM, V>(I, L, lim, y, gsign);
                }
                if (a.ss && at_steady_state<V, D>(a, I, L, y, k1)) {
                    retcode = 4;
                    break;
                }
            } else {
                dt = h / fmin(5.0, q11 / 0.9);
                ++nreject;
            }
            if (!(dt == dt)) { retcode = 3; break; }
        }
    }
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int v = t + k * T;
        if (v < I.n)
#pragma unroll
            for (int j = 0; j < D; ++j) a.u[I.uoff + static_cast<int64_t>(v) * D + j] = y[k][j];
    }
    if (t == 0) {
        nd_ensemble_stat st;
        st.t = tcur;
        st.dt = dt;
        st.naccept = naccept;
        st.nreject = nreject;
        st.nf = nf;
        st.retcode = retcode;
        st.ntrip = ntrip;
        a.stats[inst] = st;
    }
}

using EnsKernel = void (*)(SolveArgs);

// Workgroups of up to 512 threads take any vper; 1024-thread workgroups (16 waves, at
// most 128 VGPRs each) take one or two vertices per thread.
template <int VM, int EM, int METHOD>
EnsKernel pick_v(int vper, int threads) {
    if (threads > 512) return vper == 1 ? k_ensemble<VM, EM, METHOD, 1, 1024> : k_ensemble<VM, EM, METHOD, 2, 1024>;
    switch (vper) {
    case 1: return k_ensemble<VM, EM, METHOD, 1, 512>;
    case 3: return k_ensemble<VM, EM, METHOD, 3, 512>;
    case 4: return k_ensemble<VM, EM, METHOD, 4, 512>;
    default: return k_ensemble<VM, EM, METHOD, 2, 512>;
    }
}

// kernel method: 0 Tsit5, 1 RK4, 2 Tsit5 with line crossings located on the dense output
template <int VM, int EM>
EnsKernel pick_m(int kmethod, int vper, int threads) {
    switch (kmethod) {
    case 1: return pick_v<VM, EM, 1>(vper, threads);
    case 2: return pick_v<VM, EM, 2>(vper, threads);
    default: return pick_v<VM, EM, 0>(vper, threads);
    }
}

EnsKernel pick_ensemble(int vm, int em, int kmethod, int vper, int threads) {
    if (vm == ND_VM_SWING && em == ND_EM_KURAMOTO) return pick_m<ND_VM_SWING, ND_EM_KURAMOTO>(kmethod, vper, threads);
    if (vm == ND_VM_SWING && em == ND_EM_STATIC_LINE) return pick_m<ND_VM_SWING, ND_EM_STATIC_LINE>(kmethod, vper, threads);
    if (vm == ND_VM_KURAMOTO && em == ND_EM_KURAMOTO) return pick_m<ND_VM_KURAMOTO, ND_EM_KURAMOTO>(kmethod, vper, threads);
    if (vm == ND_VM_KURAMOTO && em == ND_EM_STATIC_LINE)
        return pick_m<ND_VM_KURAMOTO, ND_EM_STATIC_LINE>(kmethod, vper, threads);
    return nullptr;
}

// the kernel method of a solve: crossings are located for Tsit5 with line limits unless
// step-end checks were asked for
int kernel_method(const nd_ensemble_opts* o) {
    if (o->method == ND_METHOD_RK4) return 1;
    return (o->line_limit && o->line_events != ND_LINES_STEP_END) ? 2 : 0;
}

int vpar(int vm) {
    switch (vm) {
    case ND_VM_KURAMOTO: return 1;
    case ND_VM_SWING: return 3;
    default: return -1;
    }
}
int epar(int em) { return em == ND_EM_KURAMOTO ? 1 : em == ND_EM_STATIC_LINE ? 2 : -1; }

int hip_err(hipError_t e, const char* what) {
    if (e == hipSuccess) return ND_OK;
    return nd::set_error(ND_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

}  // namespace

struct nd_ensemble {
    int device = 0;
    int lds_max = 0;                                // dynamic LDS a workgroup may take
    int32_t ninst = 0;
    nd_kind vk{}, ek{};
    int32_t dim = 0;
    int64_t total_states = 0, total_params = 0;
    std::vector<InstDesc> inst;
    int32_t cls_threads[ND_ENSEMBLE_CLASSES] = {128, 256, 512, 512, 512, 512};
    int32_t cls_vper[ND_ENSEMBLE_CLASSES] = {2, 2, 2, 3, 4, 4};
    std::vector<void*> allocs;
    InstDesc* d_inst = nullptr;
    int32_t* d_order = nullptr;
    uint16_t *d_ends = nullptr, *d_e = nullptr, *d_cs = nullptr, *d_cv = nullptr, *d_cptr = nullptr;
    std::vector<int32_t> order;                     // instances grouped by class, largest first
    int32_t cls_first[ND_ENSEMBLE_CLASSES + 1] = {};
    int64_t generation = 0;                         // bumped whenever the class layout changes
    // one stream per class so the class launches overlap; forked from and joined back
    // to the caller's stream with events (default priority, non-blocking)
    hipStream_t cls_stream[ND_ENSEMBLE_CLASSES] = {};
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_join[ND_ENSEMBLE_CLASSES] = {};
    hipStream_t cap_stream = nullptr;               // origin stream of graph captures
    ~nd_ensemble() {
        for (void* p : allocs) (void)hipFree(p);
        for (auto& x : cls_stream) if (x) (void)hipStreamDestroy(x);
        for (auto& x : ev_join) if (x) (void)hipEventDestroy(x);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
    }
};

struct nd_ensemble_graph {
    nd_ensemble* en = nullptr;
    int64_t generation = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~nd_ensemble_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

namespace {

template <typename T>
int up(nd_ensemble* en, const std::vector<T>& h, T** d) {
    *d = nullptr;
    if (h.empty()) return ND_OK;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, h.size() * sizeof(T));
    if (e != hipSuccess) return nd::set_error(ND_ERR_ALLOC, "hipMalloc: %s", hipGetErrorString(e));
    en->allocs.push_back(p);
    e = hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_err(e, "hipMemcpy");
    *d = static_cast<T*>(p);
    return ND_OK;
}

int class_of(const nd_ensemble* en, int32_t n) {
    for (int c = 0; c < ND_ENSEMBLE_CLASSES; ++c)
        if (n <= en->cls_threads[c] * en->cls_vper[c]) return c;
    return -1;
}

int regroup(nd_ensemble* en) {
    ++en->generation;
    std::vector<std::vector<std::pair<int32_t, int32_t>>> byc(ND_ENSEMBLE_CLASSES);
    for (int32_t i = 0; i < en->ninst; ++i) {
        const int c = class_of(en, en->inst[i].n);
        if (c < 0) return nd::set_error(ND_ERR_UNSUPPORTED, "instance %d: %d vertices exceed the largest class", i, en->inst[i].n);
        en->inst[i].cls = c;
        byc[c].emplace_back(-en->inst[i].n, i);
    }
    en->order.clear();
    for (int c = 0; c < ND_ENSEMBLE_CLASSES; ++c) {
        en->cls_first[c] = static_cast<int32_t>(en->order.size());
        std::sort(byc[c].begin(), byc[c].end());
        for (auto& x : byc[c]) en->order.push_back(x.second);
    }
    en->cls_first[ND_ENSEMBLE_CLASSES] = static_cast<int32_t>(en->order.size());
    if (en->d_order) {
        auto it = std::find(en->allocs.begin(), en->allocs.end(), static_cast<void*>(en->d_order));
        if (it != en->allocs.end()) { (void)hipFree(*it); en->allocs.erase(it); }
        en->d_order = nullptr;
    }
    int rc = up(en, en->order, &en->d_order);
    if (rc) return rc;
    if (en->d_inst) {
        return hip_err(hipMemcpy(en->d_inst, en->inst.data(), en->inst.size() * sizeof(InstDesc), hipMemcpyHostToDevice),
                       "hipMemcpy");
    }
    return ND_OK;
}

int make_args(nd_ensemble* en, double* u, double* p, double t0, double t1, const nd_ensemble_opts* o,
              nd_ensemble_stat* stats, SolveArgs* a) {
    if (!en || !u || !p || !o || !stats) return nd::set_error(ND_ERR_ARGUMENT, "NULL argument");
    if (!(t1 >= t0)) return nd::set_error(ND_ERR_ARGUMENT, "t1 < t0");
    if (o->method != ND_METHOD_TSIT5 && o->method != ND_METHOD_RK4) return nd::set_error(ND_ERR_ARGUMENT, "unknown method");
    if (o->method == ND_METHOD_RK4 && !(o->dt0 > 0.0)) return nd::set_error(ND_ERR_ARGUMENT, "RK4 needs dt0 > 0");
    if (o->line_events != ND_LINES_DEFAULT && o->line_events != ND_LINES_STEP_END && o->line_events != ND_LINES_CROSSING)
        return nd::set_error(ND_ERR_ARGUMENT, "line_events must be ND_LINES_DEFAULT, ND_LINES_STEP_END or ND_LINES_CROSSING");
    if (o->method == ND_METHOD_RK4 && o->line_events == ND_LINES_CROSSING)
        return nd::set_error(ND_ERR_UNSUPPORTED, "line crossings are located on Tsit5's dense output; RK4 checks limits "
                             "at step ends (ND_LINES_STEP_END)");
    int dev = -1;
    (void)hipGetDevice(&dev);
    if (dev != en->device) return nd::set_error(ND_ERR_ARGUMENT, "ensemble lives on device %d", en->device);
    *a = SolveArgs{};
    a->inst = en->d_inst;
    a->order = en->d_order;
    a->ends = en->d_ends;
    a->slot_e = en->d_e;
    a->cs = en->d_cs;
    a->cv = en->d_cv;
    a->cptr = en->d_cptr;
    a->u = u;
    a->p = p;
    a->limit = o->line_limit;
    if ((o->event_t == nullptr) != (o->event_e == nullptr))
        return nd::set_error(ND_ERR_ARGUMENT, "event_t and event_e go together");
    a->event_t = o->event_t;
    a->event_e = o->event_e;
    a->stats = stats;
    a->t0 = t0;
    a->t1 = t1;
    a->dt0 = o->dt0;
    a->reltol = o->reltol > 0 ? o->reltol : 1e-3;
    a->abstol = o->abstol > 0 ? o->abstol : 1e-6;
    a->dtmax = o->dtmax > 0 ? o->dtmax : (t1 - t0);
    a->dtmin = o->dtmin > 0 ? o->dtmin : 1e-14 * std::max(1.0, std::fabs(t1));
    a->maxiters = o->maxiters > 0 ? o->maxiters : 1000000;
    if (!(o->ss_abstol >= 0.0) || !(o->ss_reltol >= 0.0))
        return nd::set_error(ND_ERR_ARGUMENT, "steady-state tolerances must be >= 0");
    a->ss_abstol = o->ss_abstol;
    a->ss_reltol = o->ss_reltol;
    a->ss = (o->ss_abstol > 0.0 || o->ss_reltol > 0.0) ? 1 : 0;
    return ND_OK;
}

// dynamic LDS of class c: the largest instance's tables, plus for crossing location a
// second phase buffer and three bytes per edge (after 8-byte alignment)
int32_t class_lds(const nd_ensemble* en, int c, int kmethod) {
    int32_t lds = 0;
    for (int32_t q = en->cls_first[c]; q < en->cls_first[c + 1]; ++q) {
        const InstDesc& I = en->inst[en->order[q]];
        lds = std::max(lds, I.lds_bytes + (kmethod == 2 ? 8 + 8 * I.n + 3 * I.ne : 0));
    }
    return lds;
}

// Streams, events and the dynamic-LDS limit of every class kernel; host-side only,
// so a capture that follows records nothing but the launches and the event edges.
int prepare(nd_ensemble* en, int kmethod) {
    int rc = ND_OK;
    if (!en->ev_fork) rc = hip_err(hipEventCreateWithFlags(&en->ev_fork, hipEventDisableTiming), "hipEventCreate");
    for (int c = 0; c < ND_ENSEMBLE_CLASSES && !rc; ++c) {
        if (!en->cls_stream[c])
            rc = hip_err(hipStreamCreateWithFlags(&en->cls_stream[c], hipStreamNonBlocking), "hipStreamCreateWithFlags");
        if (!rc && !en->ev_join[c])
            rc = hip_err(hipEventCreateWithFlags(&en->ev_join[c], hipEventDisableTiming), "hipEventCreate");
    }
    if (rc) return rc;
    EnsKernel ks[ND_ENSEMBLE_CLASSES] = {};
    int32_t lds[ND_ENSEMBLE_CLASSES] = {};
    for (int c = 0; c < ND_ENSEMBLE_CLASSES; ++c) {
        if (en->cls_first[c + 1] == en->cls_first[c]) continue;
        ks[c] = pick_ensemble(en->vk.model, en->ek.model, kmethod, en->cls_vper[c], en->cls_threads[c]);
        if (!ks[c]) return nd::set_error(ND_ERR_UNSUPPORTED, "no ensemble kernel for these models");
        lds[c] = class_lds(en, c, kmethod);
        if (lds[c] > en->lds_max)
            return nd::set_error(ND_ERR_UNSUPPORTED, "class %d needs %d bytes of LDS with line crossings located "
                                 "(device limit %d)", c, lds[c], en->lds_max);
    }
    // one template instance may serve several classes: give it the largest need
    for (int c = 0; c < ND_ENSEMBLE_CLASSES; ++c) {
        if (!ks[c]) continue;
        int32_t need = 0;
        for (int d = 0; d < ND_ENSEMBLE_CLASSES; ++d)
            if (ks[d] == ks[c]) need = std::max(need, lds[d]);
        rc = hip_err(hipFuncSetAttribute(reinterpret_cast<const void*>(ks[c]), hipFuncAttributeMaxDynamicSharedMemorySize,
                                         need),
                     "hipFuncSetAttribute(LDS)");
        if (rc) return rc;
    }
    return ND_OK;
}

// Fork: every non-empty class runs on its own stream after the caller's prior work;
// join: the caller's stream waits for all of them.  Largest class first.
int launch_classes(nd_ensemble* en, SolveArgs a, int kmethod, hipStream_t s) {
    int rc = hip_err(hipEventRecord(en->ev_fork, s), "hipEventRecord");
    for (int c = ND_ENSEMBLE_CLASSES - 1; c >= 0 && !rc; --c) {
        const int32_t cnt = en->cls_first[c + 1] - en->cls_first[c];
        if (cnt == 0) continue;
        EnsKernel k = pick_ensemble(en->vk.model, en->ek.model, kmethod, en->cls_vper[c], en->cls_threads[c]);
        hipStream_t cs = en->cls_stream[c];
        rc = hip_err(hipStreamWaitEvent(cs, en->ev_fork, 0), "hipStreamWaitEvent");
        if (rc) break;
        a.first = en->cls_first[c];
        a.count = cnt;
        hipLaunchKernelGGL(k, dim3(static_cast<unsigned>(cnt)), dim3(static_cast<unsigned>(en->cls_threads[c])),
                           static_cast<unsigned>(class_lds(en, c, kmethod)), cs, a);
        rc = hip_err(hipGetLastError(), "ensemble launch");
        rc = rc ? rc : hip_err(hipEventRecord(en->ev_join[c], cs), "hipEventRecord");
        rc = rc ? rc : hip_err(hipStreamWaitEvent(s, en->ev_join[c], 0), "hipStreamWaitEvent");
    }
    return rc;
}

}  // namespace

extern "C" {

int nd_ensemble_create(int32_t ninst, const int64_t* nv, const int64_t* ne, const int64_t* src, const int64_t* dst,
                       const nd_kind* vkind, const nd_kind* ekind, nd_ensemble** out) {
    nd::clear_error();
    if (!out || ninst < 1 || !nv || !ne || !vkind || !ekind) return nd::set_error(ND_ERR_ARGUMENT, "bad arguments");
    *out = nullptr;
    if (vkind->is_edge || !ekind->is_edge) return nd::set_error(ND_ERR_ARGUMENT, "kinds: need a vertex and an edge kind");
    if (vpar(vkind->model) < 0 || vkind->pdim != vpar(vkind->model) || vkind->ff != ND_FF_PURESTATEMAP ||
        vkind->outdim_dst != 1 || vkind->mask_dst[0] != 1)
        return nd::set_error(ND_ERR_UNSUPPORTED, "ensemble vertex model %d (needs Kuramoto or swing with output x[1])",
                             vkind->model);
    if (epar(ekind->model) < 0 || ekind->pdim != epar(ekind->model) || ekind->wrapper != ND_WRAP_ANTISYMMETRIC ||
        ekind->outdim_dst != 1 || ekind->dim != 0)
        return nd::set_error(ND_ERR_UNSUPPORTED, "ensemble edge model %d (needs an AntiSymmetric Kuramoto or static line)",
                             ekind->model);
    nd_ensemble* en = new (std::nothrow) nd_ensemble;
    if (!en) return nd::set_error(ND_ERR_ALLOC, "allocation failed");
    (void)hipGetDevice(&en->device);
    int lds_max = 0;
    if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, en->device) != hipSuccess || lds_max <= 0)
        lds_max = 64 * 1024;
    en->lds_max = lds_max;
    en->ninst = ninst;
    en->vk = *vkind;
    en->ek = *ekind;
    en->dim = vkind->dim;
    std::vector<uint16_t> ends, eix, cs, cv, cptr;
    int64_t eoff = 0, uoff = 0, poff = 0;
    for (int32_t i = 0; i < ninst; ++i) {
        const int64_t n = nv[i], m = ne[i];
        if (n < 1 || n > 32767 || m < 0 || 2 * m > 65535) {
            delete en;
            return nd::set_error(ND_ERR_UNSUPPORTED, "instance %d: sizes outside the ensemble kernel limits", i);
        }
        // destination lists in the reference aggregation order: increasing edge index,
        // src output before dst output (aggregators.jl:54-86, 140-151)
        std::vector<int32_t> deg(static_cast<size_t>(n) + 1, 0);
        for (int64_t e = 0; e < m; ++e) {
            const int64_t s = src[eoff + e] - 1, d = dst[eoff + e] - 1;
            if (s < 0 || s >= n || d < 0 || d >= n) {
                delete en;
                return nd::set_error(ND_ERR_ARGUMENT, "instance %d: edge endpoint out of range", i);
            }
            deg[s + 1]++;
            deg[d + 1]++;
        }
        for (int64_t v = 0; v < n; ++v) deg[v + 1] += deg[v];
        std::vector<int32_t> fill(deg.begin(), deg.end() - 1);
        std::vector<uint16_t> le(static_cast<size_t>(2 * m));
        for (int64_t e = 0; e < m; ++e) {
            const int64_t s = src[eoff + e] - 1, d = dst[eoff + e] - 1;
            le[fill[s]++] = static_cast<uint16_t>(e | 0x8000);
            le[fill[d]++] = static_cast<uint16_t>(e);
            ends.push_back(static_cast<uint16_t>(s));
            ends.push_back(static_cast<uint16_t>(d));
        }
        InstDesc I{};
        I.n = static_cast<int32_t>(n);
        I.ne = static_cast<int32_t>(m);
        I.nslot = static_cast<int32_t>(2 * m);
        I.uoff = uoff;
        I.poff = poff;
        I.soff = static_cast<int64_t>(eix.size());      // offset of both eix and ends (2 m entries each)
        I.csoff = static_cast<int64_t>(cs.size());
        I.cvoff = static_cast<int64_t>(cv.size());
        I.cpoff = static_cast<int64_t>(cptr.size());

        eix.insert(eix.end(), le.begin(), le.end());
        int32_t nch = 0;
        for (int64_t v = 0; v < n; ++v) {
            cptr.push_back(static_cast<uint16_t>(nch));
            for (int32_t b = deg[v]; b < deg[v + 1]; b += kChunk) {
                cs.push_back(static_cast<uint16_t>(b));
                cv.push_back(static_cast<uint16_t>(v));
                ++nch;
            }
        }
        cptr.push_back(static_cast<uint16_t>(nch));
        cs.push_back(static_cast<uint16_t>(2 * m));
        if (nch > 65535) {
            delete en;
            return nd::set_error(ND_ERR_UNSUPPORTED, "instance %d: too many chunks", i);
        }
        I.nchunk = nch;
        const int npe = epar(ekind->model);
        I.lds_bytes = static_cast<int32_t>(8 * (n + m * npe + m + nch + 32) + 2 * (2 * 2 * m + 2 * nch + 1 + n + 1));
        if (I.lds_bytes > lds_max) {
            delete en;
            return nd::set_error(ND_ERR_UNSUPPORTED, "instance %d needs %d bytes of LDS (device limit %d)", i, I.lds_bytes,
                                 lds_max);
        }
        en->inst.push_back(I);
        eoff += m;
        uoff += n * vkind->dim;
        poff += n * vkind->pdim + m * ekind->pdim;
    }
    en->total_states = uoff;
    en->total_params = poff;
    int rc = regroup(en);
    rc = rc ? rc : up(en, en->inst, &en->d_inst);
    rc = rc ? rc : up(en, ends, &en->d_ends);
    rc = rc ? rc : up(en, eix, &en->d_e);
    rc = rc ? rc : up(en, cs, &en->d_cs);
    rc = rc ? rc : up(en, cv, &en->d_cv);
    rc = rc ? rc : up(en, cptr, &en->d_cptr);
    if (rc) { delete en; return rc; }
    *out = en;
    return ND_OK;
}

void nd_ensemble_destroy(nd_ensemble* en) { delete en; }

int nd_ensemble_sizes(const nd_ensemble* en, int64_t* total_states, int64_t* total_params, int64_t* state_offsets,
                      int64_t* param_offsets) {
    if (!en) return nd::set_error(ND_ERR_ARGUMENT, "NULL ensemble");
    if (total_states) *total_states = en->total_states;
    if (total_params) *total_params = en->total_params;
    for (int32_t i = 0; i < en->ninst; ++i) {
        if (state_offsets) state_offsets[i] = en->inst[i].uoff;
        if (param_offsets) param_offsets[i] = en->inst[i].poff;
    }
    return ND_OK;
}

int nd_ensemble_set_class(nd_ensemble* en, int32_t cls, int32_t threads, int32_t vper) {
    if (!en || cls < 0 || cls >= ND_ENSEMBLE_CLASSES) return nd::set_error(ND_ERR_ARGUMENT, "bad class");
    if (threads < kWave || threads > 1024 || threads % kWave || vper < 1 || vper > 4 || (threads > 512 && vper > 2))
        return nd::set_error(ND_ERR_ARGUMENT, "threads must be a multiple of 64 in [64,1024], vper in {1,2,3,4} "
                             "(in {1,2} above 512 threads)");
    // regrouping replaces the instance-order array: queued solves finish with the old one
    int rc = hip_err(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (rc) return rc;
    en->cls_threads[cls] = threads;
    en->cls_vper[cls] = vper;
    return regroup(en);
}

int nd_tsit5_dense(double* r) {
    if (!r) return nd::set_error(ND_ERR_ARGUMENT, "NULL argument");
    for (int i = 0; i < 7; ++i)
        for (int j = 0; j < 4; ++j) r[i * 4 + j] = kTsit5Dense.r[i][j];
    return ND_OK;
}

int nd_tsit5_tableau(double* a, double* btilde, double* c) {
    if (!a || !btilde || !c) return nd::set_error(ND_ERR_ARGUMENT, "NULL argument");
    for (int i = 0; i < 7; ++i) {
        for (int j = 0; j < 6; ++j) a[i * 6 + j] = kTsit5.a[i][j];
        btilde[i] = kTsit5.bt[i];
        c[i] = kTsit5.c[i];
    }
    return ND_OK;
}

int nd_ensemble_solve(nd_ensemble* en, double* u, double* p, double t0, double t1, const nd_ensemble_opts* o,
                      nd_ensemble_stat* stats, void* stream) {
    SolveArgs a{};
    int rc = make_args(en, u, p, t0, t1, o, stats, &a);
    rc = rc ? rc : prepare(en, kernel_method(o));
    return rc ? rc : launch_classes(en, a, kernel_method(o), static_cast<hipStream_t>(stream));
}

int nd_ensemble_graph_create(nd_ensemble* en, double* u, double* p, double t0, double t1,
                             const nd_ensemble_opts* o, nd_ensemble_stat* stats, nd_ensemble_graph** out) {
    if (!out) return nd::set_error(ND_ERR_ARGUMENT, "NULL argument");
    *out = nullptr;
    SolveArgs a{};
    int rc = make_args(en, u, p, t0, t1, o, stats, &a);
    rc = rc ? rc : prepare(en, kernel_method(o));
    if (rc) return rc;
    if (!en->cap_stream) {
        rc = hip_err(hipStreamCreateWithFlags(&en->cap_stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
        if (rc) return rc;
    }
    nd_ensemble_graph* gr = new (std::nothrow) nd_ensemble_graph;
    if (!gr) return nd::set_error(ND_ERR_ALLOC, "allocation failed");
    gr->en = en;
    gr->generation = en->generation;
    rc = hip_err(hipStreamBeginCapture(en->cap_stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (rc) { delete gr; return rc; }
    const int lrc = launch_classes(en, a, kernel_method(o), en->cap_stream);
    hipGraph_t g = nullptr;
    rc = hip_err(hipStreamEndCapture(en->cap_stream, &g), "hipStreamEndCapture");
    gr->graph = g;
    if (lrc || rc) { delete gr; return lrc ? lrc : rc; }
    rc = hip_err(hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    if (rc) { delete gr; return rc; }
    *out = gr;
    return ND_OK;
}

int nd_ensemble_graph_launch(nd_ensemble_graph* gr, void* stream) {
    if (!gr || !gr->exec) return nd::set_error(ND_ERR_ARGUMENT, "NULL graph");
    if (gr->generation != gr->en->generation)
        return nd::set_error(ND_ERR_ARGUMENT, "the ensemble's size classes changed after this graph was captured");
    int dev = -1;
    (void)hipGetDevice(&dev);
    if (dev != gr->en->device) return nd::set_error(ND_ERR_ARGUMENT, "graph lives on device %d", gr->en->device);
    return hip_err(hipGraphLaunch(gr->exec, static_cast<hipStream_t>(stream)), "hipGraphLaunch");
}

void nd_ensemble_graph_destroy(nd_ensemble_graph* gr) { delete gr; }

}  // extern "C"
