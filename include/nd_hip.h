/*
 * nd_hip.h -- C ABI of libnd_hip.so, the MI355X (gfx950) implementation of
 * NetworkDynamics.jl's network right-hand side `nw(du, u, p, t)`.
 *
 * Plain C: pointers, sizes, status codes.  Device pointers are HIP device
 * addresses owned by the caller; `stream` is a hipStream_t passed as void*
 * (NULL = the null stream).  Every entry point returns ND_OK or an ND_ERR_*
 * code; nd_last_error() returns the message of the last failure on the
 * calling thread.
 *
 * Reference interfaces replaced (NetworkDynamics.jl v1.1.0):
 *   nd_network_create        Network(g, vertexm, edgem; ...)        src/construction.jl:31-236
 *                            IndexManager + register_vertices!/edges! src/network_structure.jl:1-55,224-289
 *   nd_network_export_*      AggregationMap / EagerGBufProvider map   src/aggregators.jl:54-86, src/gbufs.jl:11-22
 *                            _inv_aggregation_map                     src/aggregators.jl:203-235
 *   nd_plan_create           Adapt.adapt_structure(to, nw)            ext/NetworkDynamicsCUDAExt.jl:16-45
 *   nd_rhs                   (nw::Network)(du, u, p, t)               src/coreloop.jl:1-102
 *   nd_get_buffers           get_buffers(nw, u, p, t)                 src/coreloop.jl:103-109
 *   nd_rhs_perturbed_jvp     nw(du, u, p, t; perturb, perturb_maps)  src/coreloop.jl:1,49-87 (and its
 *   nd_plan_set_perturbation   ForwardDiff pass in linearize_network, src/linear_analysis.jl:86-122)
 *   nd_gather                gather!(::EagerGBufProvider, gbuf, o)    src/gbufs.jl:25
 *   nd_aggregate             aggregate!(::Aggregator, aggbuf, data)   src/aggregators.jl:140-151 (additive)
 *   nd_rk4, nd_tsit5         solve(ODEProblem(nw, u0, tspan, p), RK4() / Tsit5())  src/post_utils.jl:43-82
 *                            (the OrdinaryDiffEq solves of the network's ODEFunction)
 */
#ifndef ND_HIP_H
#define ND_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
enum {
    ND_OK = 0,
    ND_ERR_ARGUMENT = 1,      /* the reference's ArgumentError / DimensionMismatch cases  */
    ND_ERR_UNSUPPORTED = 2,   /* component model or option without a HIP implementation    */
    ND_ERR_DEVICE = 3,        /* HIP runtime failure                                        */
    ND_ERR_ALLOC = 4,         /* host or device allocation failure                          */
    ND_ERR_NOT_CONVERGED = 5  /* iterative solver stopped before its tolerance              */
};

/* ---- component description (src/component_functions.jl) ----------------- */

/* Output wrappers of edge models, component_functions.jl:81-203. */
enum {
    ND_WRAP_NONE = 0,          /* vertex models; or a two-sided edge g(osrc, odst, ...)     */
    ND_WRAP_ANTISYMMETRIC = 1, /* y_src = -y_dst                                            */
    ND_WRAP_SYMMETRIC = 2,     /* y_src =  y_dst                                            */
    ND_WRAP_DIRECTED = 3,      /* no src output                                             */
    ND_WRAP_FIDUCIAL = 4       /* independent src/dst outputs (StateMask pair)              */
};

/* Feed-forward traits, component_functions.jl:6-64. */
enum {
    ND_FF_PURESTATEMAP = 0,
    ND_FF_PUREFEEDFORWARD = 1,
    ND_FF_FEEDFORWARD = 2,
    ND_FF_NOFEEDFORWARD = 3
};

/* Vertex models (the equations a fused kernel implements).  Reference twins:
 *   DIFFUSION          test/ComponentLibrary.jl:42-46       dv = acc
 *   KURAMOTO           test/ComponentLibrary.jl:69-72       dtheta = omega + acc,        p = [omega]
 *   SWING              test/ComponentLibrary.jl:59-67       ddelta = w; dw = 1/M*(Pm - D*w + acc), p = [M, D, Pm]
 *   SWING_UNIT_DAMPING benchmark/benchmark_models.jl:37-42  dw = P - 1.0*w; dw += acc,   p = [P]
 *   LOAD               (no reference twin) first-order load  dtheta = 1/tau*(P + acc),    p = [tau, P]
 *   CAPACITOR_SINK     docs/src/external_inputs.md:76-82     dV = 1/C*(acc - (1 + 0.1*sin(t))), p = [C]
 *   PI_SOURCE          docs/src/external_inputs.md:86-99     d = Vref - ext[1]; dxi = d;
 *                      (external input: a remote voltage)    dV = 1/C*(acc + (Kp*d + Ki*xi)), p = [C, Vref, Ki, Kp]
 *   CAPACITOR          test/loopback_test.jl:53-63           dv = acc / C,                 p = [C]
 *   VOLTAGE_SOURCE     test/loopback_test.jl:19-34           no state; output v = V (NoFeedForward), p = [V]
 *   RESISTOR_INJECTOR  test/loopback_test.jl:66-78           no state; output i = in / R (PureFeedForward:
 *                                                            an injector fed the hub potential), p = [R]
 *   INDUCTOR_INJECTOR  test/loopback_test.jl:80-92           di = in / L, output i (StateMask), p = [L]
 *   CRL                test/loopback_test.jl:113-131         dv = (acc - v/R - iL) / C; diL = v / L,
 *                                                            output v, p = [C, R, L]
 *   FHN                docs/examples/directed_and_weighted_graphs.jl:89-95  FitzHugh-Nagumo:
 *                      du = u - u^3/3 - v + acc; dv = (u - a)*eps, output u, p = [a, eps]
 *   TRUSS_FREE         docs/examples/stress_on_truss.jl:55-62  a free truss node, states (vx, vy, x, y),
 *                      outputs (x, y), inputs (Fx, Fy): dv = (F - gamma v)/M, dvy -= g, dpos = v,
 *                      p = [M, gamma, g]
 *   TRUSS_FIXED        docs/examples/stress_on_truss.jl:45-48  a fixed node: no state, outputs (x, y) = p
 *                      (NoFeedForward), p = [xfix, yfix]
 *   NOOP               f does nothing (aggregator tests, test/aggregators_test.jl:15-18)  */
enum {
    ND_VM_NOOP = 0,
    ND_VM_DIFFUSION = 1,
    ND_VM_KURAMOTO = 2,
    ND_VM_SWING = 3,
    ND_VM_SWING_UNIT_DAMPING = 4,
    ND_VM_LOAD = 5,
    ND_VM_CAPACITOR_SINK = 6,
    ND_VM_PI_SOURCE = 7,
    ND_VM_CAPACITOR = 8,
    ND_VM_VOLTAGE_SOURCE = 9,
    ND_VM_RESISTOR_INJECTOR = 10,
    ND_VM_INDUCTOR_INJECTOR = 11,
    ND_VM_CRL = 12,
    ND_VM_FHN = 13,
    ND_VM_TRUSS_FREE = 14,
    ND_VM_TRUSS_FIXED = 15
};

/* Edge models.  Reference twins:
 *   DIFFUSION            test/diffusion_test.jl:57-60          y = v_s - v_d
 *   DIFFUSION_WEIGHTED   test/ComponentLibrary.jl:8-13         y = p*(v_s - v_d),  p = [w]
 *   KURAMOTO             test/ComponentLibrary.jl:51-57        y = K*sin(v_s - v_d), p = [K]
 *   STATIC_LINE          (no reference twin)                   y = B*sin(v_s - v_d - phi), p = [B, phi]
 *   DIFFUSION_FIDUCIAL   test/ComponentLibrary.jl:22-28        y_dst = p*(v_s - v_d); y_src = -y_dst
 *   ODE_SIN_DIFFUSION    test/ComponentLibrary.jl:30-40        de = 1/tau*(sin(+-(v_s - v_d)) - e), p = [tau]
 *   ODE_SIN_DIFFUSION_FAST benchmark/benchmark_models.jl:11-16 de = 100.0*(sin(+-(v_s - v_d)) - e)
 *   ODE_LINEAR           test/diffusion_test.jl:102-107        de = +-(v_s - v_d) - e
 *   RESISTOR             docs/src/external_inputs.md:66-72     y = (v_s - v_d)/R,   p = [R]
 *   LOOPBACK             src/post_utils.jl:105-108             Directed: y_dst = -1 * v_s, the
 *                        LoopbackConnection from an injector (src) to its hub (dst); the hub's
 *                        output is copied to the injector's input (post_utils.jl:221-228)
 *   ODE_LAG              one-state form of ComponentLibrary.jl:30-34: de = 1/tau*(sin(v_s - v_d) - e),
 *                        p = [tau]; output a wrapped StateMask (AntiSymmetric(1): y_dst = e,
 *                        y_src = -e; Symmetric(1); Directed(1)), component_functions.jl:117-175
 *   EXT_DRIVEN           test/external_inputs_test.jl:45-47: de = ext[1] (one external input),
 *                        any StateMask output
 *   ELECTRICAL           docs/examples/directed_and_weighted_graphs.jl:97-101: Directed gap
 *                        junction y = w*(v_s - v_d)*sigma, p = [w, sigma]
 *   BEAM                 docs/examples/stress_on_truss.jl:77-86: AntiSymmetric spring between two
 *                        positions, F = K*(L - d)*(dst - src)/d, d = |dst - src|, two outputs
 *                        (Fx, Fy), p = [K, L]
 *   NOOP                 g does nothing (test/aggregators_test.jl:20-30)
 * ODE edge models (state dimension 2 for ODE_SIN_DIFFUSION, ODE_SIN_DIFFUSION_FAST and
 * ODE_LINEAR, 1 for ODE_LAG and EXT_DRIVEN) take Fiducial StateMask outputs or a StateMask
 * inside AntiSymmetric / Symmetric / Directed.                                              */
enum {
    ND_EM_NOOP = 0,
    ND_EM_DIFFUSION = 1,
    ND_EM_DIFFUSION_WEIGHTED = 2,
    ND_EM_KURAMOTO = 3,
    ND_EM_STATIC_LINE = 4,
    ND_EM_DIFFUSION_FIDUCIAL = 5,
    ND_EM_ODE_SIN_DIFFUSION = 6,
    ND_EM_ODE_SIN_DIFFUSION_FAST = 7,
    ND_EM_ODE_LINEAR = 8,
    ND_EM_RESISTOR = 9,
    ND_EM_LOOPBACK = 10,
    ND_EM_ODE_LAG = 11,
    ND_EM_EXT_DRIVEN = 12,
    ND_EM_ELECTRICAL = 13,
    ND_EM_BEAM = 14
};

#define ND_MAX_OUT 4
#define ND_MAX_EXT 8

/* One component kind: what the reference's ComponentBatch type parameters
 * carry (network_structure.jl:176-198): dims, ff trait, wrapper, and the
 * StateMask indices of a PureStateMap output (1-based, like StateMask), and the
 * number of external inputs (extdim, the ComponentBatch's extbufstride).
 * Vertices use outdim_dst as their outdim and outdim_src = 0. */
typedef struct {
    int32_t is_edge;
    int32_t model;
    int32_t wrapper;
    int32_t ff;
    int32_t dim;
    int32_t pdim;
    int32_t outdim_src;
    int32_t outdim_dst;
    int32_t mask_src[ND_MAX_OUT];
    int32_t mask_dst[ND_MAX_OUT];
    int32_t extdim;
} nd_kind;

/* ---- host index layer ------------------------------------------------------ */

typedef struct nd_network nd_network;

/* Network(g, vertexm, edgem) -- construction.jl:31-236.
 * src/dst: the graph's edge list in Graphs.jl `edges(g)` order, 1-based (Int64),
 * as IndexManager.edgevec (network_structure.jl:37).
 * vertex_kind[nv], edge_kind[ne]: index into kinds[] per component; components
 * with equal kind index are batched together in first-appearance order
 * (_find_identical_components, construction.jl:238-256; utils.jl:197-217).
 * Errors (ND_ERR_ARGUMENT): unequal vertex outdims or edge dst outdims
 * (construction.jl:99-106), out-of-range vertex ids, bad kind indices. */
int nd_network_create(int64_t nv, int64_t ne, const int64_t* src, const int64_t* dst,
                      int32_t nkinds, const nd_kind* kinds,
                      const int32_t* vertex_kind, const int32_t* edge_kind,
                      nd_network** out);
void nd_network_destroy(nd_network* nw);

typedef struct {
    int64_t nv, ne;
    int64_t dim;            /* lastidx_dynamic */
    int64_t pdim;           /* lastidx_p       */
    int64_t lastidx_out;
    int64_t lastidx_aggr;
    int64_t lastidx_gbuf;
    int64_t agg_first;      /* AggregationMap.range (1-based, inclusive); empty: agg_last < agg_first */
    int64_t agg_last;
    int32_t vdepth, edepth;
    int32_t nvbatches, nebatches;
    int64_t lastidx_extbuf; /* external-input buffer length (0: no external inputs) */
} nd_network_info;

int nd_network_get_info(const nd_network* nw, nd_network_info* info);

/* ComponentBatch + its BatchStride firsts (network_structure.jl:176-222,
 * utils.jl:1-42).  `indices` points into the network (valid while it lives).
 * in_first is the v_aggr first for vertex batches and the e_gbufr first for
 * edge batches. */
typedef struct {
    int32_t kind;
    int32_t pad_;
    int64_t count;
    const int64_t* indices;
    int64_t state_first, param_first, out_first, in_first;
} nd_batch;

int nd_network_get_batch(const nd_network* nw, int32_t is_edge, int32_t b, nd_batch* out);

/* Per-component ranges of the IndexManager, 1-based (first, last) pairs:
 * v_ranges[nv][4][2]: v_data, v_out, v_para, v_aggr;
 * e_ranges[ne][6][2]: e_data, e_out.src, e_out.dst, e_para, e_gbufr.src, e_gbufr.dst. */
int nd_network_export_ranges(const nd_network* nw, int64_t* v_ranges, int64_t* e_ranges);

/* AggregationMap.map over AggregationMap.range (agg_last-agg_first+1 entries)
 * and the EagerGBufProvider map (lastidx_gbuf entries), bit-exact, 1-based. */
int nd_network_export_maps(const nd_network* nw, int64_t* aggmap, int64_t* gbufmap);

/* The destination-ordered reduction lists (CSR): for aggregation index a
 * (1-based), the output-buffer indices summed into aggbuf[a], in summation
 * order (increasing output index, as SequentialAggregator, aggregators.jl:140-151).
 * rowptr[lastidx_aggr+1] (0-based offsets), cols[rowptr[lastidx_aggr]] (1-based). */
int nd_network_csr_size(const nd_network* nw, int64_t* nnz);
int nd_network_export_csr(const nd_network* nw, int64_t* rowptr, int64_t* cols);

/* ---- external inputs (src/external_inputs.jl) -------------------------------- */

/* v_ext / e_ext ranges of the IndexManager (network_structure.jl:224-258), 1-based
 * (first, last) pairs: v_ext[nv][2], e_ext[ne][2] (either may be NULL). */
int nd_network_export_ext_ranges(const nd_network* nw, int64_t* v_ext, int64_t* e_ext);

/* ExtMap(im), external_inputs.jl:10-46: for each external-buffer index k (1-based),
 * map[k-1] > 0 is StateBufIdx(map[k-1]) (an index into u) and map[k-1] < 0 is
 * OutBufIdx(-map[k-1]) (an index into the output buffer o).  Required before
 * nd_plan_create when the network has external inputs.  ND_ERR_ARGUMENT: no
 * external inputs ("There are no external inputs."), len != lastidx_extbuf, an index
 * out of range, or an output of a feed-forward component
 * (external_inputs.jl:38-41). */
int nd_network_set_extmap(nd_network* nw, int64_t len, const int64_t* map);

/* One component's mass matrix, in the forms VertexModel/EdgeModel normalise a user's
 * value to (component_functions.jl:948-963): I, UniformScaling(m[0]), Diagonal(m[0:dim])
 * or a dense dim x dim matrix (column-major, m[0:dim*dim]). */
enum { ND_MASS_IDENTITY = 0, ND_MASS_UNIFORM = 1, ND_MASS_DIAGONAL = 2, ND_MASS_DENSE = 3 };
typedef struct {
    int32_t form;
    int32_t pad_;
    const double* m;
} nd_mass;

/* construct_mass_matrix(im), construction.jl:409-453: the network's mass matrix from
 * its components' (vertex_mass[nv], edge_mass[ne]; a NULL array means I for all).
 * Components whose matrix equals I are skipped, like the reference's filter; the others
 * are written into their v_data / e_data blocks.  *shape = ND_MASS_IDENTITY when every
 * component's matrix equals I (nothing written), ND_MASS_DIAGONAL when every other one
 * is diagonal (out: the dim diagonal entries), else ND_MASS_DENSE (out: dim*dim,
 * column-major).  out == NULL only reports the shape.  ND_ERR_ARGUMENT for an unknown
 * form or a missing m. */
int nd_network_mass_matrix(const nd_network* nw, const nd_mass* vertex_mass, const nd_mass* edge_mass,
                           int32_t* shape, double* out);

/* ---- device plan ------------------------------------------------------------ */

typedef struct nd_plan nd_plan;

/* Uploads the network's index structures to the current HIP device and selects
 * the fused kernels.  ND_ERR_UNSUPPORTED if a model has no HIP kernel (output
 * depth > 1 except for diffusion vertices and edges over consecutive states and, at
 * depth 2, the truss nodes and beams);
 * ND_ERR_ARGUMENT if the network has external inputs and no ExtMap
 * (nd_network_set_extmap) or a model's external-input count differs from its kind's.
 * External inputs (of vertices and of ODE edges) resolve to state indices: a vertex
 * output and an ODE edge's output are StateMasks of states, the src side of an
 * AntiSymmetric one the negated state. */
int nd_plan_create(const nd_network* nw, nd_plan** out);
void nd_plan_destroy(nd_plan* plan);

/* Tuning knobs of a plan (0 = keep current). */
typedef struct {
    int32_t hub_degree;      /* segments longer than this: one 64-lane wave per vertex    */
    int32_t block_size;      /* threads per block of the vertex kernel (64..256)         */
    int32_t big_hub_degree;  /* segments longer than this: one block per vertex          */
    int32_t rk4_path;        /* nd_rk4: 0 keep, 1 automatic (one workgroup for networks that
                                fit it), 2 always four launches per step                  */
    int32_t gather_mode;     /* neighbour outputs: 0 keep, 1 automatic (by size), 2 read
                                from u directly, 3 compact copy per call (with the slot-
                                ordered parameter copy); get_tuning reports 2 or 3        */
} nd_plan_tuning;
int nd_plan_tune(nd_plan* plan, const nd_plan_tuning* t);
int nd_plan_get_tuning(const nd_plan* plan, nd_plan_tuning* t);

/* nw(du, u, p, t): coreloop.jl:1-102.  du fully overwritten.  Lengths are
 * checked like coreloop.jl:2-7 (len_du == len_u == dim, len_p == pdim when
 * pdim > 0) -> ND_ERR_ARGUMENT.  Asynchronous on `stream`. */
int nd_rhs(nd_plan* plan, double* du, const double* u, const double* p, double t,
           int64_t len_du, int64_t len_u, int64_t len_p, void* stream);

/* Fixed-step classic Runge-Kutta 4, in place on u (the "plain method" stepper that
 * replaces an OrdinaryDiffEq RK4 solve of ODEProblem(nw, u0, tspan, p),
 * src/post_utils.jl:43-82).  Each stage is one fused RHS launch whose epilogue
 * forms the stage state and the weighted derivative sum, in the order
 * u + dt/6*(((k1 + 2k2) + 2k3) + k4).  work: 3*dim doubles of device scratch, or
 * NULL for a plan-owned buffer.  The library's models are autonomous, so t0 only
 * documents the time axis. */
int nd_rk4(nd_plan* plan, double* u, const double* p, double t0, double dt, int64_t nsteps, double* work,
           int64_t len_u, int64_t len_p, void* stream);
/* nd_rk4 without allocating: the same fixed-step RK4 by the stage launches (never the
 * one-workgroup path), with its three stage vectors and second compact-output buffer
 * taken from a library-owned device array of 2^23 doubles (64 MiB) per GPU.  No
 * counterpart in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The plan must exist already; a network needing more than the array (3 x dim plus the
 * compact outputs) -> ND_ERR_ARGUMENT.  Calls on one GPU share the array: calls that
 * overlap on a GPU (other streams, other threads) race.  Results equal nd_rk4's
 * stage-launch path (rk4_path 2) bit for bit. */
int nd_rk4_noalloc(nd_plan* plan, double* u, const double* p, double t0, double dt, int64_t nsteps, int64_t len_u,
                   int64_t len_p, void* stream);

/* get_buffers(nw, u, p, t; initbufs=true), coreloop.jl:103-109: fills the
 * output buffer o[lastidx_out] and the aggregation buffer aggbuf[lastidx_aggr]. */
int nd_get_buffers(nd_plan* plan, double* o, double* aggbuf, const double* u,
                   const double* p, double t, void* stream);

/* The perturbation maps of nw(du, u, p, t; perturb, perturb_maps), coreloop.jl:1,49-87
 * (built by _classify_perturbation_channels, linear_analysis.jl:629-680): each map is
 * n pairs (buffer index, channel index), both 1-based, flattened.  vo: output buffer
 * (vertex outputs), vi: aggregation buffer (vertex inputs), eo: output buffer (edge
 * outputs), ei: gather buffer (edge inputs).  ndelta: number of channels.  Replaces the
 * plan's previous maps (zero lengths clear them).  Parameter channels are the caller's
 * (added to p before the call, as linearize_network does). */
int nd_plan_set_perturbation(nd_plan* plan, int64_t ndelta, int64_t nvo, const int64_t* vo, int64_t nvi,
                             const int64_t* vi, int64_t neo, const int64_t* eo, int64_t nei, const int64_t* ei);

/* nw(du, u, p, t; perturb=delta, perturb_maps) on first-order dual numbers
 * u + eps*v, delta + eps*ddelta: the reference's buffered loop (coreloop.jl:1-102) with
 * delta added at its four points -- after the loopback copy (vi), after the external
 * inputs are collected (vo), after the gather (ei) and before the aggregation (eo).
 * du, jv: f and its directional derivative, also along the parameter tangent dp (the
 * models evaluate their parameters as dual numbers here).  v, dp, ddelta: NULL for
 * zero tangents; delta:
 * ndelta values (NULL allowed only without channels).  o, dout, aggbuf, dagg: NULL, or
 * the caller's buffers [lastidx_out] / [lastidx_aggr] receiving the perturbed buffers
 * and their tangents (the observables of linearize_network's output channels); jv may
 * be NULL.  Stream-ordered; the plan's work buffers are shared by its calls. */
int nd_rhs_perturbed_jvp(nd_plan* plan, double* du, double* jv, const double* u, const double* v, const double* p,
                         const double* dp, double t, const double* delta, const double* ddelta, double* o,
                         double* dout, double* aggbuf, double* dagg, void* stream);

/* nw(du, u + eps*v, p, t) on first-order dual numbers: du = f(u) and jv = J(u) v, the
 * Jacobian-vector product ForwardDiff forms when linearize_network
 * (src/linear_analysis.jl:74-85) or an implicit solver calls nw with Dual inputs through
 * the DiffCache buffers (src/network_structure.jl:152-155, src/gbufs.jl:20-24).  Every
 * aggregation segment is summed in the reference order.  v, jv: device vectors of
 * length dim. */
int nd_rhs_jvp(nd_plan* plan, double* du, double* jv, const double* u, const double* v, const double* p, double t,
               int64_t len_du, int64_t len_u, int64_t len_p, void* stream);

/* nw(du, u + sum_j eps_j v_j, p, t) on dual numbers with k partials (1 <= k <= 8): du =
 * f(u) and jv[:, j] = J(u) v[:, j] in one pass, the chunk of ForwardDiff duals the
 * reference sizes its DiffCache buffers for (src/construction.jl:205-208, pickchunksize;
 * src/network_structure.jl:152-155), so a colored Jacobian takes ceil(colors / k)
 * launches instead of one per color.  v, jv: k device columns of length dim, column j at
 * v + j*dim.  Column j equals nd_rhs_jvp along v[:, j] bit for bit.  k = 1 is nd_rhs_jvp.
 * Networks whose vertex batches take the generic or vector-output kernels, or with
 * feed-forward vertex outputs -> ND_ERR_UNSUPPORTED (nd_rhs_jvp per tangent serves them). */
int nd_rhs_jvp_multi(nd_plan* plan, double* du, double* jv, const double* u, const double* v, int32_t k,
                     const double* p, double t, int64_t len_du, int64_t len_u, int64_t len_p, void* stream);

/* collect_externals!(map, extbuf, u, o), external_inputs.jl:49-77:
 * extbuf[k] = u[i] for StateBufIdx(i), o[i] for OutBufIdx(i).  extbuf: lastidx_extbuf
 * doubles on the device; o: the output buffer (nd_get_buffers). */
int nd_collect_externals(nd_plan* plan, double* extbuf, const double* u, const double* o, void* stream);

/* gather!(EagerGBufProvider, gbuf, o): gbuf[k] = o[map[k]], gbufs.jl:25. */
int nd_gather(nd_plan* plan, double* gbuf, const double* o, void* stream);

/* aggregate!(aggregator, aggbuf, o): aggbuf[a] += sum over its CSR segment of o,
 * additive like every reference aggregator (aggregators_test.jl:69-79),
 * deterministic (no atomics). */
int nd_aggregate(nd_plan* plan, double* aggbuf, const double* o, void* stream);

/* ---- batched ensemble integration ------------------------------------------------
 * Many independent networks (instances) of one vertex kind and one edge kind each,
 * integrated in one launch per size class: one workgroup per instance, the whole
 * trajectory inside the kernel (register-resident Runge-Kutta state, LDS-resident
 * topology).  The reference has no ensemble code; the caller it replaces is a SciML
 * EnsembleProblem over ODEProblem(nw, u0, tspan, p) (src/post_utils.jl:43-82)
 * solved with OrdinaryDiffEqTsit5.  Instance i: nv[i] vertices, ne[i] edges whose
 * 1-based endpoints follow in src/dst (concatenated, Graphs.jl edge order), state
 * and parameters in the reference's flat layout for a one-batch network,
 * concatenated in instance order (offsets from nd_ensemble_sizes).  Vertex kinds:
 * KURAMOTO, SWING; edge kinds: AntiSymmetric KURAMOTO, STATIC_LINE; up to 2048
 * vertices per instance with the default classes (threads x vertices per thread:
 * 128x2, 256x2, 512x2, 512x3, 512x4, 512x4). */
#define ND_ENSEMBLE_CLASSES 6
enum { ND_METHOD_TSIT5 = 0, ND_METHOD_RK4 = 1 };
/* When a line-limit condition is checked (nd_ensemble_opts.line_events) */
enum { ND_LINES_DEFAULT = 0, ND_LINES_STEP_END = 1, ND_LINES_CROSSING = 2 };

typedef struct {
    int32_t method;       /* ND_METHOD_TSIT5 (adaptive) or ND_METHOD_RK4 (fixed dt0)          */
    int32_t line_events;  /* with line_limit: ND_LINES_CROSSING locates each crossing of
                             |edge output| - limit inside the step, as the reference's
                             ContinuousComponentCallback does (src/callbacks.jl:113-140,
                             batched into a VectorContinuousCallback, callbacks.jl:450-476):
                             after an accepted Tsit5 step the condition is compared at the
                             nine dense-output points t + h*m/9 (m = 1..9) with its sign at
                             the step start; an up-crossing (< 0 to >= 0) or down-crossing
                             (> 0 to <= 0) at the first such point is bisected on the dense
                             output to adjacent doubles, the solution moves to the left end
                             of the bracket, the crossing edges get K = 0 and the step
                             restarts there with the step size the controller proposed.
                             ND_LINES_STEP_END: |edge output| > limit is checked at the end
                             of each accepted step.  ND_LINES_DEFAULT: crossing for Tsit5,
                             step end for RK4 (which has no dense output here).            */
    double reltol, abstol;
    double dt0;           /* RK4: the step; Tsit5: initial step, 0 = automatic                  */
    double dtmax, dtmin;  /* 0 = t1-t0 / 1e-14*max(1,|t1|)                                       */
    int64_t maxiters;     /* steps per instance, 0 = 1e6                                         */
    const double* line_limit; /* per edge, concatenated in instance order, or NULL.  Line
                             tripping (the cascading-failure callback of
                             docs/examples/cascading_failure.jl:104-111, condition
                             |P| - limit, affect K = 0 on either crossing direction; when
                             it is checked: line_events): a tripped edge's first parameter
                             (K) is set to 0, in p as well                                  */
    double ss_abstol, ss_reltol; /* steady-state termination, both 0 = off: after an accepted
                             step, an instance whose every |du_i| <= ss_abstol or
                             <= ss_reltol*|u_i| stops there with retcode 4 (the test of the
                             TerminateSteadyState callback with which basin-stability runs
                             stop at sync).  RK4 pays one extra RHS per step for du.       */
    /* Per instance, or both NULL: the PresetTimeComponentCallback of
       docs/examples/cascading_failure.jl:125-131.  At t = event_t[i] (only
       t0 < event_t <= t1 fires) edge event_e[i] (0-based within the instance, -1 = none)
       gets its first parameter (K) set to 0, in p as well, and counts in ntrip if it was
       not 0.  Tsit5 steps to event_t exactly (a tstop) and restarts from the new
       derivative; RK4 fires at the end of the step that ends at event_t (the time must
       lie on the step grid t0 + k*dt0). */
    const double* event_t;
    const int32_t* event_e;
} nd_ensemble_opts;

typedef struct {
    double t, dt;         /* final time and next step                                           */
    int64_t naccept, nreject, nf;
    int32_t retcode;      /* 0 success, 1 maxiters, 2 dt below dtmin, 3 NaN step,
                             4 terminated at a steady state (ss_abstol / ss_reltol)          */
    int32_t ntrip;        /* lines tripped (line_limit set, or the preset event)               */
} nd_ensemble_stat;

typedef struct nd_ensemble nd_ensemble;
int nd_ensemble_create(int32_t ninst, const int64_t* nv, const int64_t* ne, const int64_t* src, const int64_t* dst,
                       const nd_kind* vertex_kind, const nd_kind* edge_kind, nd_ensemble** out);
void nd_ensemble_destroy(nd_ensemble* en);
int nd_ensemble_sizes(const nd_ensemble* en, int64_t* total_states, int64_t* total_params,
                      int64_t* state_offsets, int64_t* param_offsets);
/* size class c holds instances with nv <= threads*vper (first class that fits);
 * threads: a multiple of 64 up to 1024, vper 1-4 (1-2 above 512 threads) */
int nd_ensemble_set_class(nd_ensemble* en, int32_t cls, int32_t threads, int32_t vper);
/* integrates every instance from t0 to t1 in place on u (device); stats: device array
 * of ninst entries.  p is read only, except that with opts->line_limit the first
 * parameter of each tripped edge is set to 0 (the callback's affect on the parameters) */
int nd_ensemble_solve(nd_ensemble* en, double* u, double* p, double t0, double t1,
                      const nd_ensemble_opts* opts, nd_ensemble_stat* stats, void* stream);
/* Adaptive Tsit5 for one network (the Tsitouras 5(4) pair, OrdinaryDiffEq's PI
 * controller and initial-dt heuristic), in place on u from t0 to t1, every step in one
 * workgroup: the single-instance ensemble kernel, whose flat layout is the plan's.  For
 * the networks nd_rk4 runs in one workgroup (one vertex batch of Kuramoto or swing
 * vertices, one batch of AntiSymmetric Kuramoto or static-line edges, up to 2048
 * vertices); any other network -> ND_ERR_UNSUPPORTED, and the host steps it on nd_rhs.
 * dt0 = 0: the initial-dt heuristic; maxiters <= 0: 1e6.  stats: device pointer to one
 * nd_ensemble_stat (final t, accepted and rejected steps, retcode), or NULL. */
int nd_tsit5(nd_plan* plan, double* u, const double* p, double t0, double t1, double reltol, double abstol,
             double dt0, int64_t maxiters, nd_ensemble_stat* stats, int64_t len_u, int64_t len_p, void* stream);
/* The Tsit5 coefficients the ensemble kernel and nd_tsit5 step with (Tsitouras 2011;
 * the method of OrdinaryDiffEqTsit5's Tsit5(), docs/examples/cascading_failure.jl:145):
 * a[7*6] the stage matrix row-major (row s: the coefficients of stage s+1), btilde[7] the
 * error weights b - bhat, c[7] the nodes (stage s+1 is evaluated at t + c[s]*h).  Host
 * only; no device work. */
int nd_tsit5_tableau(double* a, double* btilde, double* c);
/* The dense output the ensemble kernel locates line crossings on (Tsitouras' continuous
 * extension of the same pair, the interpolant OrdinaryDiffEqTsit5 gives Tsit5
 * solutions and callbacks): u(t + s*h) = u + h * sum_i b_i(s) k_i with
 * b_i(s) = sum_{j=1..4} r[4*i + j-1] s^j, i = 0..6 over the seven stage derivatives.
 * Host only; no device work. */
int nd_tsit5_dense(double* r);
/* The same solve captured once into a hipGraph (class launches on parallel branches)
 * and replayed with nd_ensemble_graph_launch on any stream; u, p, stats, tspan and
 * opts are fixed at capture.  A graph refuses to launch after nd_ensemble_set_class
 * changed the classes; destroy graphs before their ensemble.  No reference
 * counterpart (a solve repeated over many perturbation rounds is its caller). */
typedef struct nd_ensemble_graph nd_ensemble_graph;
int nd_ensemble_graph_create(nd_ensemble* en, double* u, double* p, double t0, double t1,
                             const nd_ensemble_opts* opts, nd_ensemble_stat* stats, nd_ensemble_graph** out);
int nd_ensemble_graph_launch(nd_ensemble_graph* graph, void* stream);
void nd_ensemble_graph_destroy(nd_ensemble_graph* graph);

/* ---- batched fixpoint (Newton) ---------------------------------------------------
 * Synchronous fixpoints of many Kuramoto-type grids at once (one workgroup per grid),
 * replacing find_fixpoint's NonlinearSolve Newton (src/initialization.jl:73-144) for
 * networks of swing / first-order load vertices with AntiSymmetric K*sin lines: with
 * omega = 0 a state is a fixpoint iff F_i = P_i + sum_j K_ij sin(theta_j - theta_i) = 0.
 * Grid g: nv[g] vertices, ne[g] edges (1-based endpoints concatenated in src/dst).
 * theta (in/out), P and scale (residual divisor per vertex: M for swing, tau for load,
 * so F_i/scale_i is the vertex's du) are concatenated per vertex, K per edge.
 * Converged when max_i |F_i / scale_i| <= tol; the acceptance of a fixpoint is the
 * network RHS itself (isfixpoint, src/linear_analysis.jl:8-15).  Up to 5120 vertices
 * per grid.  stats: device array of ngrid entries. */
typedef struct nd_fixpoint nd_fixpoint;
typedef struct {
    double tol;            /* max |F_i/scale_i| to stop at, 0 = 1e-10                    */
    double cg_rtol;        /* relative residual of each linear solve, 0 = 1e-12           */
    int32_t max_newton;    /* 0 = 50                                                       */
    int32_t cg_maxiter;    /* 0 = 20000                                                    */
    int32_t linearized;    /* 1: solve the DC power flow P + sum K (theta_j - theta_i) = 0 */
    int32_t inexact;       /* 1: each Newton step's CG stops at the relative residual
                              max(cg_rtol, min(eta_max, current residual)) (inexact Newton
                              with a residual-based forcing term); 0: always cg_rtol       */
    double eta_max;        /* cap of that forcing term, 0 = 1e-2                           */
} nd_newton_opts;
typedef struct {
    double residual;       /* final max |F_i/scale_i|                                      */
    int32_t newton_iters;  /* linear solves (Newton steps) taken                           */
    int32_t cg_iters;      /* conjugate-gradient iterations over all steps                 */
    int32_t retcode;       /* 0 converged, 1 max_newton reached, 2 line search stalled,
                              3 non-finite residual                                      */
    int32_t pad_;
} nd_newton_stat;
int nd_fixpoint_create(int32_t ngrid, const int64_t* nv, const int64_t* ne, const int64_t* src, const int64_t* dst,
                       nd_fixpoint** out);
void nd_fixpoint_destroy(nd_fixpoint* f);
int nd_fixpoint_solve(nd_fixpoint* f, double* theta, const double* P, const double* scale, const double* K,
                      const nd_newton_opts* opts, nd_newton_stat* stats, void* stream);

/* ---- learned warm start: message-passing network (FP32) -------------------------
 * No counterpart in the reference (its find_fixpoint starts from the given state).
 * A 4-layer message-passing network, hidden width 64, over the stacked nodes of a
 * batch of grids: H0 = relu(X W_in + b_in); H_l = relu(A (H_{l-1} W_l) + H_{l-1} U_l + b_l);
 * out = H_4 w_out + b_out, with A the row-normalised K-weighted adjacency (CSR rowptr
 * / nbr / a_fwd; a_bwd = weights of its transpose on the same pattern).  Parameter
 * layout: W_in[nfeat*64], b_in[64], {[W_l|U_l][64*128], b_l[64]} x 4, w_out[64], b_out.
 * work: nd_mp_workspace_floats(n) floats, shared by a forward and its backward; params and
 * work 16-byte aligned. */
int64_t nd_mp_param_count(int32_t nfeat);
int64_t nd_mp_workspace_floats(int64_t nnodes);
int nd_mp_forward(int64_t n, const int32_t* rowptr, const int32_t* nbr, const float* a_fwd, const float* X,
                  int32_t nfeat, const float* params, float* work, float* out, void* stream);
/* gradient of mean((out - target)^2) into grads; loss_part: n/4+1 per-block sums of squared errors */
int nd_mp_backward(int64_t n, const int32_t* rowptr, const int32_t* nbr, const float* a_bwd, const float* X,
                   int32_t nfeat, const float* params, const float* target, float* work, float* grads,
                   float* loss_part, void* stream);
int nd_mp_adam(int64_t np, float* params, const float* grads, float* m, float* v, float lr, float b1, float b2,
               float eps, int64_t step, void* stream);
/* The network's per-node inputs for the stacked nodes of a batch of grids (CSR rowptr /
 * nbr of the symmetric adjacency, w = K of each entry, theta = a phase estimate such as
 * the DC solution): with D_s = theta[nbr_s] - theta[i] and kdeg_i = sum_s w_s,
 * s0_i = sum_s w_s (sin D_s - D_s), the nonlinear part of the line flows, and
 * X[i] = {P_i, is_gen_i, log kdeg_i, deg_i / 6, s0_i, sum_s w_s s0[nbr_s] / kdeg_i,
 * sum_s w_s D_s^2, max_s |D_s|} (n x 8, row-major).  Sums in CSR order; a node without
 * entries gets 0 for log kdeg_i and the neighbour mean. */
int nd_ws_features(int64_t n, const int32_t* rowptr, const int32_t* nbr, const double* w, const double* theta,
                   const double* P, const double* is_gen, double* s0, double* X, void* stream);
/* Per-node fixpoint residual of a phase vector on the same stacked CSR: r_i = |F_i| / scale_i
 * with F_i = P_i + sum_s w_s sin(theta[nbr_s] - theta_i), the quantity nd_fixpoint_solve
 * drives to its tolerance; its per-grid maximum is the residual of a start. */
int nd_ws_residual(int64_t n, const int32_t* rowptr, const int32_t* nbr, const double* w, const double* theta,
                   const double* P, const double* scale, double* r, void* stream);

/* ---- synthetic inputs ----------------------------------------------------------
 * Seeded graph generators for the workloads.  No counterpart in the reference's
 * RHS path (its tests use Graphs.jl generators, whose random streams need Julia);
 * edge lists come back simple and sorted as Graphs.jl `edges(g)` orders a
 * SimpleGraph: src < dst, lexicographic, 1-based. */
typedef struct nd_edgelist nd_edgelist;
int nd_graph_erdos_renyi(int64_t n, int64_t m, uint64_t seed, nd_edgelist** out);
int nd_graph_barabasi_albert(int64_t n, int64_t m, uint64_t seed, nd_edgelist** out);
int nd_graph_random_geometric(int64_t n, double radius, uint64_t seed, double* xy, nd_edgelist** out);
int nd_graph_watts_strogatz(int64_t n, int64_t k, double beta, uint64_t seed, nd_edgelist** out);
int64_t nd_edgelist_size(const nd_edgelist* el);
int nd_edgelist_copy(const nd_edgelist* el, int64_t* src, int64_t* dst);
void nd_edgelist_free(nd_edgelist* el);

/* ---- misc -------------------------------------------------------------------- */

const char* nd_last_error(void);
int nd_device_synchronize(void);

/* Device memory for a host shim that has no GPU array package of its own: the
 * reference's adapt(CuArray, nw) path (ext/NetworkDynamicsCUDAExt.jl:16-45) moves u, p
 * and the caches to the device; these are the plain-pointer equivalents.  Copies are
 * synchronous. */
int nd_malloc(size_t bytes, void** out);
int nd_free(void* ptr);
int nd_memcpy_h2d(void* dst, const void* src, size_t bytes);
int nd_memcpy_d2h(void* dst, const void* src, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* ND_HIP_H */
