// Device plan and kernels of the network right-hand side on gfx950.
//
// The reference's coreloop (coreloop.jl:1-102) runs ~9-11 phases per call:
// fills, vertex g, gather (NNlib), edge g, scatter-add aggregation, vertex f.
// Here one fused kernel per vertex batch does all of it for the component
// kinds of the library:
//
//   * vertex outputs are StateMasks, so the output buffer o is never formed:
//     neighbour outputs are read straight from u through precomputed u-indices;
//   * each vertex walks its destination list (the CSR of the aggregation map,
//     rows in vertex order, entries in increasing output index -- the order of
//     SequentialAggregator, aggregators.jl:140-151), evaluates the edge model
//     for its own side on the fly (AntiSymmetric src = -dst exactly) and sums
//     in that order; no float atomics, no aggbuf, no gbuf;
//   * the vertex model f is applied in the epilogue and du is written once.
//
// Segments longer than hub_degree ("hubs") are reduced by one 64-lane wave:
// lane-strided partial sums, then a fixed xor-butterfly, so results are
// reproducible run to run (a deterministic tree order instead of the
// sequential one).  Edges with their own state (ODE edges) get a separate
// per-edge kernel for du_e.
//
// Generic device primitives mirror the reference's internal dispatch points
// for any output depth: gather! (gbufs.jl:25), aggregate! (aggregators.jl:140)
// and get_buffers (coreloop.jl:103-109).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <new>
#include <vector>

#include "nd_internal.h"
#include "nd_models.h"

namespace {

using namespace nd;
constexpr int kWave = 64;

struct VDesc {            // one vertex batch as the kernels see it (0-based offsets)
    int32_t model, dim, pdim, depth;
    int32_t state0, param0, row0, count;
    int32_t out0;         // o index of the first vertex output (get_buffers)
    int32_t nbig, nmid;   // hub vertices: block-per-vertex / wave-per-vertex classes
    int32_t mask[ND_MAX_OUT];
    const int32_t* hubs;  // positions q of hub vertices: nbig big ones, then nmid mid ones
    const int32_t* lrp;   // hub-free destination lists: count+1 offsets into lslots
    const int2* lslots;
    const uint8_t* lseb;  // edge batch of each lslots entry (several edge batches only)
    const int32_t* lnbr;  // with cached parameters instead of lslots: neighbour output index | (is_src << 31)
    const double* kfull;  // edge parameters per destination-list entry (rowptr order), or null
    const double* klow;   // the same for the hub-free lists (lrp order), or null
    // one-thread-per-vertex path (models with external inputs or time dependence, edge
    // models outside the fused set): external inputs as u-indices, and the time
    const int32_t* xmap;  // u-index of every external-buffer entry
    int32_t generic;
    int32_t ext0, extdim; // first external-buffer entry of the batch (0-based), per vertex
    double t;
    // computed outputs: 0 StateMask, 1 NoFeedForward g(x, p), 2 an injector's
    // PureFeedForward g(in, p); lbin[lb0 + q]: the o-index of the hub output copied to
    // an injector's input (post_utils.jl:221-228), -1 for other vertices
    int32_t outkind, lb0;
    const int32_t* lbin;
};

// Slot-ordered edge parameters: a copy of the edge batch's parameters laid out like the
// destination lists, so the vertex kernels stream them instead of fetching each entry's
// parameters from p at a random position.  Each call compares p's parameters with the
// cached copy kc bit for bit (streamed) and rewrites the slots of the edges that differ,
// so the copy always equals p when the vertex kernels run.
struct KSync {
    const double* pe;     // p + first parameter of the edge batch (count * np, contiguous)
    double* kc;           // cached copy, same layout
    double* kfull;
    double* klow;
    const int4* kpos;     // per edge: full-list slots of its src / dst entries, then hub-free-list
                          // slots; -1 where the edge has no such entry
    int32_t count, np, vblocks;
};

struct EDesc {            // one edge batch
    int32_t model, wrap, ode, dim, pdim;
    int32_t state0, param0, count;
    int32_t out0;         // o index of the first edge output
    int32_t outdim_src, outdim_dst;
    int32_t mask_src[ND_MAX_OUT];
    int32_t mask_dst[ND_MAX_OUT];
    const int2* ends;     // per edge (batch order): u-index of the src / dst vertex output
    int32_t ext0, extdim; // ODE edges with external inputs: first external-buffer entry of
    const int32_t* xmap;  // the batch (0-based), per edge; entries resolved as ext_index does
};

// An ODE edge's output as one side of it sees it: a StateMask of its state, negated on
// the src side of an AntiSymmetric wrapper (component_functions.jl:117-127; Symmetric and
// Fiducial pass the state, Directed has no src side).  is_src: the entry's role bit.
template <int WRAP>
__device__ __forceinline__ double ode_side(double x, bool is_src) {
    return (WRAP == ND_WRAP_ANTISYMMETRIC && is_src) ? -x : x;
}
__device__ __forceinline__ double ode_side_rt(int wrap, double x, bool is_src) {
    return (wrap == ND_WRAP_ANTISYMMETRIC && is_src) ? -x : x;
}
__device__ __forceinline__ Dual ode_side_rt(int wrap, Dual x, bool is_src) {
    return (wrap == ND_WRAP_ANTISYMMETRIC && is_src) ? -x : x;
}

// External-input entry k from the resolved map: a u-index, or ~u-index where the source
// is the src-side output of an AntiSymmetric StateMask edge (the negated state).
__device__ __forceinline__ double ext_read(const int32_t* __restrict__ xmap, int64_t k, const double* __restrict__ u) {
    const int32_t xi = xmap[k];
    return xi >= 0 ? u[xi] : -u[~xi];
}
__device__ __forceinline__ Dual ext_read_dual(const int32_t* __restrict__ xmap, int64_t k, const double* __restrict__ u,
                                              const double* __restrict__ v) {
    const int32_t xi = xmap[k];
    return xi >= 0 ? Dual{u[xi], v[xi]} : Dual{-u[~xi], -v[~xi]};
}

// Value of one destination-list entry for the vertex whose own output is `own`.
// s.x: u-index of the other endpoint's output (feed-forward edges) or of the edge
//      state element this side outputs (ODE edges);
// s.y: (edge parameter offset << 1) | (1 if this vertex is the edge's src); ODE edges
//      carry the role bit only.
template <int EM, int WRAP>
__device__ __forceinline__ double slot_value(int2 s, double own, const double* __restrict__ u,
                                             const double* __restrict__ vo, const double* __restrict__ p) {
    if constexpr (EM == ND_EM_NOOP) {
        return __builtin_nan("");          // g writes nothing: o keeps its NaN fill
    } else if constexpr (is_ode_model(EM)) {
        return ode_side<WRAP>(u[s.x], s.y & 1);
    } else {
        const double nbr = vo[s.x];
        const double* pe = p + (s.y >> 1);
        if (s.y & 1) {
            const double y = g_dst<EM>(own, nbr, pe);
            return WRAP == ND_WRAP_SYMMETRIC ? y : -y;
        }
        return g_dst<EM>(nbr, own, pe);
    }
}

__device__ __forceinline__ double slot_value_rt(const EDesc& e, int2 s, double own, const double* __restrict__ u,
                                                const double* __restrict__ vo, const double* __restrict__ p) {
    if (e.model == ND_EM_NOOP) return __builtin_nan("");
    if (e.ode) return ode_side_rt(e.wrap, u[s.x], s.y & 1);
    const double nbr = vo[s.x];
    const double* pe = p + (s.y >> 1);
    if (s.y & 1) {
        const double y = g_dst_rt(e.model, own, nbr, pe);
        return e.wrap == ND_WRAP_SYMMETRIC ? y : -y;
    }
    return g_dst_rt(e.model, nbr, own, pe);
}

// Explicit Runge-Kutta stage epilogue fused into the RHS kernels.  kind 0: plain
// du = f(u).  For a stage derivative k at index i:
//   kind 1 (first stage):  acc[i] = k;             yout[i] = base[i] + h*k
//   kind 2 (middle stage): acc[i] = acc[i] + w*k;  yout[i] = base[i] + h*k
//   kind 3 (final stage):  yout[i] = base[i] + h*(acc[i] + k)
// With h, w chosen on the host this is the classic RK4 in the operation order
// u + dt/6*(((k1 + 2k2) + 2k3) + k4).
struct StageArgs {
    const double* base;
    double* acc;
    double* yout;
    double* vnext;        // compact vertex outputs of yout for the next stage, or null
    double h, w;
    int32_t kind;
};

// returns the value written to yout (stage) or du (plain)
template <bool STAGE>
__device__ __forceinline__ double emit1(double* __restrict__ du, const StageArgs& sa, int64_t i, double k) {
    if constexpr (!STAGE) {
        du[i] = k;
        return k;
    } else {
        double y;
        if (sa.kind == 3) {
            y = sa.base[i] + sa.h * (sa.acc[i] + k);
        } else {
            sa.acc[i] = sa.kind == 1 ? k : sa.acc[i] + sa.w * k;
            y = sa.base[i] + sa.h * k;
        }
        sa.yout[i] = y;
        return y;
    }
}

// Vertex f and its epilogue.  In a stage with sa.vnext, the vertex's output component
// (state mask0) of the new stage state also goes to vnext[oidx], the compact vertex
// outputs the next stage gathers from.
template <int VM, bool STAGE>
__device__ __forceinline__ void vertex_emit(double* __restrict__ du, const StageArgs& sa, int64_t x0,
                                            const double* __restrict__ x, double acc,
                                            const double* __restrict__ pv, int dim, int mask0, int64_t oidx) {
    constexpr int DIMC = VModel<VM>::dim;
    if constexpr (DIMC > 0) {
        double dx[DIMC];
        vertex_f<VM>(dx, x, acc, pv, DIMC);
#pragma unroll
        for (int j = 0; j < DIMC; ++j) {
            const double y = emit1<STAGE>(du, sa, x0 + j, dx[j]);
            if (STAGE && sa.vnext && j == mask0) sa.vnext[oidx] = y;
        }
    } else {
        for (int j = 0; j < dim; ++j) {
            const double y = emit1<STAGE>(du, sa, x0 + j, 0.0);
            if (STAGE && sa.vnext && j == mask0) sa.vnext[oidx] = y;
        }
    }
}

// Orders LDS accesses across the lanes of one wave (no block-wide barrier): earlier LDS
// writes of any lane are visible to later reads of every lane of the wave.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Entry value from operands already in registers (nbr: u[s.x]; k: the edge parameters).
template <int EM, int WRAP>
__device__ __forceinline__ double slot_eval(int2 s, double nbr, const double* k, double own) {
    if constexpr (EM == ND_EM_NOOP) {
        return __builtin_nan("");
    } else if constexpr (is_ode_model(EM)) {
        return ode_side<WRAP>(nbr, s.y & 1);
    } else {
        if (s.y & 1) {
            const double y = g_dst<EM>(own, nbr, k);
            return WRAP == ND_WRAP_SYMMETRIC ? y : -y;
        }
        return g_dst<EM>(nbr, own, k);
    }
}

constexpr int kUnroll = 4;   // entries whose loads are issued together per thread

// Partial sum of the entries beg+t, beg+t+G, beg+t+2G, ... of one segment, in that
// order.  Loads of kUnroll entries are issued before any of them is used, so a
// thread keeps several gathers in flight instead of one dependent chain.
template <int EM, int WRAP, bool MULTI, bool KC>
__device__ __forceinline__ double segment_sum(int32_t beg, int32_t end, int32_t t, int32_t G, double own,
                                              const int2* __restrict__ slots, const uint8_t* __restrict__ slot_eb,
                                              const EDesc* __restrict__ ed, const double* __restrict__ u,
                                              const double* __restrict__ vo, const double* __restrict__ p,
                                              const double* __restrict__ kslot) {
    double acc = 0.0;
    if constexpr (MULTI) {
        for (int32_t s = beg + t; s < end; s += G) acc += slot_value_rt(ed[slot_eb[s]], slots[s], own, u, vo, p);
    } else {
        constexpr int NP = EModel<EM>::np;
        // Entries past the segment's end load its last entry instead (a valid index): with
        // no branch around them the loads all go out before the first wait, where loads
        // guarded by a condition are each followed by a wait for their result.
        for (int32_t s0 = beg + t; s0 < end; s0 += kUnroll * G) {
            int2 sl[kUnroll];
            double nb[kUnroll];
            double kk[kUnroll][NP > 0 ? NP : 1];
#pragma unroll
            for (int j = 0; j < kUnroll; ++j) sl[j] = slots[min(s0 + j * G, end - 1)];
#pragma unroll
            for (int j = 0; j < kUnroll; ++j) {
                const int32_t sc = min(s0 + j * G, end - 1);
#pragma unroll
                for (int c = 0; c < NP; ++c)
                    kk[j][c] = KC ? kslot[static_cast<int64_t>(sc) * NP + c] : p[(sl[j].y >> 1) + c];
            }
#pragma unroll
            for (int j = 0; j < kUnroll; ++j) {
                if constexpr (is_ode_model(EM)) nb[j] = u[sl[j].x];
                else if constexpr (EM != ND_EM_NOOP) nb[j] = vo[sl[j].x];
            }
#pragma unroll
            for (int j = 0; j < kUnroll; ++j)
                if (s0 + j * G < end) acc += slot_eval<EM, WRAP>(sl[j], nb[j], kk[j], own);
        }
    }
    return acc;
}

// Streamed reads of the low path (entry lists, parameter copy, own states and
// parameters): nontemporal, so that they displace fewer neighbour outputs from L2
constexpr bool kStreamNT = true;
template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
    if constexpr (kStreamNT) return __builtin_nontemporal_load(p);
    else return *p;
}

// LDS of the fused vertex kernels (each kernel that uses them gets its own copy): the
// low path's per-wave entry windows (values, owners) and own outputs, the hub blocks'
// wave partials (blocks of at most 256 threads)
__shared__ double rhs_val[4 * 256];
__shared__ double rhs_own[256];
__shared__ uint8_t rhs_owner[4 * 256];
__shared__ double rhs_wpart[256 / kWave];

// Fused RHS for one vertex batch.  Block layout (hubs first, so the longest
// segments start earliest):
//   [0, nbig)                 one block per big hub (segment > big_hub_degree),
//   [nbig, nbig + nmidb)      one wave per mid hub (segment > hub_degree),
//   [nbig + nmidb, ...)       one vertex per thread, segment summed in reference order.
// Hub segments: per-thread partial sums in a fixed order, xor-butterfly per wave,
// wave partials combined in wave order -- deterministic run to run.
template <int VM, int EM, int WRAP, bool MULTI, bool STAGE, bool KC>
__device__ __forceinline__ void vertex_rhs_body(const VDesc& vd, int32_t b,
                                                const int32_t* __restrict__ rowptr, const int2* __restrict__ slots,
                                                const uint8_t* __restrict__ slot_eb, const EDesc* __restrict__ ed,
                                                const double* __restrict__ u, const double* __restrict__ vo,
                                                const double* __restrict__ p, double* __restrict__ du,
                                                const StageArgs& sa) {
    constexpr int DIMC = VModel<VM>::dim;
    const int dim = DIMC > 0 ? DIMC : vd.dim;
    const int32_t wpb = static_cast<int32_t>(blockDim.x / kWave);
    const int32_t nmidb = (vd.nmid + wpb - 1) / wpb;
    if constexpr (VM != ND_VM_NOOP) {
        if (b < vd.nbig) {
            const int32_t q = vd.hubs[b];
            const int64_t x0 = static_cast<int64_t>(vd.state0) + static_cast<int64_t>(q) * dim;
            const int32_t beg = rowptr[vd.row0 + q], end = rowptr[vd.row0 + q + 1];
            const double own = u[x0 + vd.mask[0]];
            const int32_t t = static_cast<int32_t>(threadIdx.x);
            const double part = wave_sum(segment_sum<EM, WRAP, MULTI, KC>(beg, end, t, static_cast<int32_t>(blockDim.x),
                                                                          own, slots, slot_eb, ed, u, vo, p, vd.kfull));
            if ((t & (kWave - 1)) == 0) rhs_wpart[t / kWave] = part;
            __syncthreads();
            if (t == 0) {
                double acc = rhs_wpart[0];
                for (int32_t w = 1; w < wpb; ++w) acc += rhs_wpart[w];
                vertex_emit<VM, STAGE>(du, sa, x0, u + x0, acc, p + vd.param0 + static_cast<int64_t>(q) * vd.pdim, dim,
                                       vd.mask[0], static_cast<int64_t>(vd.out0) + q);
            }
            return;
        }
        b -= vd.nbig;
        if (b < nmidb) {
            const int32_t w = b * wpb + static_cast<int32_t>(threadIdx.x / kWave);
            const int32_t lane = static_cast<int32_t>(threadIdx.x & (kWave - 1));
            if (w >= vd.nmid) return;
            const int32_t q = vd.hubs[vd.nbig + w];
            const int64_t x0 = static_cast<int64_t>(vd.state0) + static_cast<int64_t>(q) * dim;
            const int32_t beg = rowptr[vd.row0 + q], end = rowptr[vd.row0 + q + 1];
            const double own = u[x0 + vd.mask[0]];
            const double acc = wave_sum(segment_sum<EM, WRAP, MULTI, KC>(beg, end, lane, kWave, own, slots, slot_eb,
                                                                         ed, u, vo, p, vd.kfull));
            if (lane == 0)
                vertex_emit<VM, STAGE>(du, sa, x0, u + x0, acc, p + vd.param0 + static_cast<int64_t>(q) * vd.pdim, dim,
                                       vd.mask[0], static_cast<int64_t>(vd.out0) + q);
            return;
        }
        b -= nmidb;
    }
    // Low-degree vertices: wave w of block b takes the 64 consecutive vertices
    // q0..q0+63 and their hub-free destination lists lrp/lslots.  Entries are evaluated
    // lane-densely in per-wave LDS windows of 4*64 (loads issued together, no lane idles
    // on a neighbour's long segment); then each lane sums its own segment from LDS in
    // order, so the summation order is the reference's.  The waves of a block share no
    // data and synchronise only within themselves.
    const int32_t lane = static_cast<int32_t>(threadIdx.x & (kWave - 1));
    const int32_t wib = static_cast<int32_t>(threadIdx.x / kWave);
    const int32_t q0 = (b * static_cast<int32_t>(blockDim.x / kWave) + wib) * kWave;
    const int32_t q = q0 + lane;
    if (q0 >= vd.count) return;
    const bool valid = q < vd.count;
    const int32_t qv = valid ? q : q0;             // a valid vertex for the loads of idle lanes
    const int64_t x0 = static_cast<int64_t>(vd.state0) + static_cast<int64_t>(qv) * dim;
    if constexpr (VM == ND_VM_NOOP) {
        if (valid) vertex_emit<VM, STAGE>(du, sa, x0, u + x0, 0.0, p, dim, vd.mask[0], static_cast<int64_t>(vd.out0) + q);
    } else {
        const int32_t* __restrict__ lrp = vd.lrp;
        // lrp[q] = offset | (hub ? 1 << 31 : 0): hub vertices keep an empty segment here and
        // their du is written by the hub paths.  One offset load per lane: a segment ends
        // where the next lane's begins, the wave's last one at lrp[q0 + 64].
        constexpr int32_t kOff = 0x7fffffff;
        const int32_t raw = lrp[valid ? q : vd.count];
        const int32_t tail = lrp[min(q0 + kWave, vd.count)];
        const bool hub = raw < 0;
        const int32_t sb = raw & kOff;
        const int32_t nxt = __shfl_down(raw, 1, kWave);
        const int32_t se = valid ? ((lane == kWave - 1 ? tail : nxt) & kOff) : sb;
        const int32_t S0 = __shfl(sb, 0, kWave);
        const int32_t S1 = tail & kOff;
        // own state and parameters are loaded up front, in parallel with the offsets
        constexpr int DC = VModel<VM>::dim > 0 ? VModel<VM>::dim : 1;
        constexpr int NPV = VPar<VM>::np > 0 ? VPar<VM>::np : 1;
        double xs[DC], pvs[NPV];
#pragma unroll
        for (int j = 0; j < DC; ++j) xs[j] = ld_stream(u + x0 + j);
#pragma unroll
        for (int j = 0; j < NPV; ++j)
            pvs[j] = VPar<VM>::np > 0 ? ld_stream(p + vd.param0 + static_cast<int64_t>(qv) * vd.pdim + j) : 0.0;
        const double own = (DC == 1 || vd.mask[0] == 0) ? xs[0] : xs[DC - 1];
        double* __restrict__ wval = rhs_val + wib * 4 * kWave;
        uint8_t* __restrict__ wowner = rhs_owner + wib * 4 * kWave;
        const double* __restrict__ wown = rhs_own + wib * kWave;
        rhs_own[threadIdx.x] = own;
        double acc = 0.0;
        constexpr int32_t W = 4 * kWave;
        for (int32_t base = S0; base < S1; base += W) {
            const int32_t wend = min(base + W, S1);
            for (int32_t s2 = max(sb, base); s2 < min(se, wend); ++s2) wowner[s2 - base] = static_cast<uint8_t>(lane);
            wave_sync();
            constexpr int NP = EModel<EM>::np;
            int2 sl[4];
            double nb[4];
            double kk[4][NP > 0 ? NP : 1];
            // entries past the window's end load its last entry instead (a valid index), so
            // the loads carry no branch and all go out before the first wait
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int32_t s2c = min(base + j * kWave + lane, wend - 1);
                if constexpr (KC) {
                    const uint32_t w = static_cast<uint32_t>(ld_stream(vd.lnbr + s2c));
                    sl[j] = make_int2(static_cast<int32_t>(w & 0x7fffffffu), static_cast<int32_t>(w >> 31));
                } else {
                    sl[j] = vd.lslots[s2c];
                }
            }
            if constexpr (!MULTI) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int32_t s2c = min(base + j * kWave + lane, wend - 1);
#pragma unroll
                    for (int c = 0; c < NP; ++c)
                        kk[j][c] = KC ? ld_stream(vd.klow + static_cast<int64_t>(s2c) * NP + c) : p[(sl[j].y >> 1) + c];
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if constexpr (is_ode_model(EM)) nb[j] = u[sl[j].x];
                    else if constexpr (EM != ND_EM_NOOP) nb[j] = vo[sl[j].x];
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int32_t s2 = base + j * kWave + lane;
                if (s2 < wend) {
                    const double ow = wown[wowner[s2 - base]];
                    double v;
                    if constexpr (MULTI) v = slot_value_rt(ed[vd.lseb[s2]], sl[j], ow, u, vo, p);
                    else v = slot_eval<EM, WRAP>(sl[j], nb[j], kk[j], ow);
                    wval[s2 - base] = v;
                }
            }
            wave_sync();
            for (int32_t s2 = max(sb, base); s2 < min(se, wend); ++s2) acc += wval[s2 - base];
            wave_sync();
        }
        if (valid && !hub) vertex_emit<VM, STAGE>(du, sa, x0, xs, acc, pvs, dim, vd.mask[0], static_cast<int64_t>(vd.out0) + q);
    }
}

template <int VM, int EM, int WRAP, bool MULTI, bool STAGE, bool KC>
__global__ void __launch_bounds__(256) k_vertex_rhs(VDesc vd, const int32_t* __restrict__ rowptr,
                                                    const int2* __restrict__ slots,
                                                    const uint8_t* __restrict__ slot_eb,
                                                    const EDesc* __restrict__ ed,
                                                    const double* __restrict__ u, const double* __restrict__ vo,
                                                    const double* __restrict__ p,
                                                    double* __restrict__ du, StageArgs sa) {
    vertex_rhs_body<VM, EM, WRAP, MULTI, STAGE, KC>(vd, static_cast<int32_t>(blockIdx.x), rowptr, slots, slot_eb,
                                                    ed, u, vo, p, du, sa);
}

// Several vertex batches of the fused set in one launch: batch i owns the blocks
// [first[i], first[i+1]) and runs them exactly as its own k_vertex_rhs launch would.
// Small batches then overlap their load-latency chains instead of running one after
// the other.
constexpr int kGroupMax = 4;
struct VGroup {
    VDesc vd[kGroupMax];
    int32_t first[kGroupMax + 1];
    int32_t n;
};

__device__ __forceinline__ int group_of_block(const VGroup& g, int32_t bx) {
    int i = 0;
    while (i + 1 < g.n && bx >= g.first[i + 1]) ++i;
    return i;
}

template <int EM, int WRAP, bool MULTI, bool STAGE, bool KC>
__global__ void __launch_bounds__(256) k_vertex_rhs_group(VGroup g, const int32_t* __restrict__ rowptr,
                                                          const int2* __restrict__ slots,
                                                          const uint8_t* __restrict__ slot_eb,
                                                          const EDesc* __restrict__ ed,
                                                          const double* __restrict__ u, const double* __restrict__ vo,
                                                          const double* __restrict__ p,
                                                          double* __restrict__ du, StageArgs sa) {
    const int32_t bx = static_cast<int32_t>(blockIdx.x);
    const int i = group_of_block(g, bx);
    const VDesc& vd = g.vd[i];
    const int32_t b = bx - g.first[i];
    switch (vd.model) {
    case ND_VM_DIFFUSION:
        vertex_rhs_body<ND_VM_DIFFUSION, EM, WRAP, MULTI, STAGE, KC>(vd, b, rowptr, slots, slot_eb, ed, u, vo, p, du, sa);
        break;
    case ND_VM_KURAMOTO:
        vertex_rhs_body<ND_VM_KURAMOTO, EM, WRAP, MULTI, STAGE, KC>(vd, b, rowptr, slots, slot_eb, ed, u, vo, p, du, sa);
        break;
    case ND_VM_SWING:
        vertex_rhs_body<ND_VM_SWING, EM, WRAP, MULTI, STAGE, KC>(vd, b, rowptr, slots, slot_eb, ed, u, vo, p, du, sa);
        break;
    case ND_VM_SWING_UNIT_DAMPING:
        vertex_rhs_body<ND_VM_SWING_UNIT_DAMPING, EM, WRAP, MULTI, STAGE, KC>(vd, b, rowptr, slots, slot_eb, ed, u, vo,
                                                                             p, du, sa);
        break;
    case ND_VM_LOAD:
        vertex_rhs_body<ND_VM_LOAD, EM, WRAP, MULTI, STAGE, KC>(vd, b, rowptr, slots, slot_eb, ed, u, vo, p, du, sa);
        break;
    default:
        vertex_rhs_body<ND_VM_NOOP, ND_EM_NOOP, ND_WRAP_NONE, false, STAGE, false>(vd, b, rowptr, slots, slot_eb, ed,
                                                                                  u, vo, p, du, sa);
        break;
    }
}

// du of ODE edges (coreloop.jl:76, edge f without feed-forward); external inputs are
// read from the states they resolve to (collect_externals! fills them from u and from
// StateMask outputs, external_inputs.jl:49-77).  eo: where e.ends point -- u, or the
// compact vertex outputs when some vertex outputs are computed by a g.
template <int EM, bool STAGE>
__global__ void __launch_bounds__(256) k_edge_ode(EDesc e, const double* __restrict__ u,
                                                  const double* __restrict__ eo,
                                                  const double* __restrict__ p, double* __restrict__ du,
                                                  StageArgs sa) {
    constexpr int D = ode_dim(EM);
    constexpr int NX = edge_ext_inputs(EM) > 0 ? edge_ext_inputs(EM) : 1;
    const int32_t q = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= e.count) return;
    const int2 en = e.ends[q];
    const double vs = eo[en.x], vd = eo[en.y];
    const int64_t x0 = static_cast<int64_t>(e.state0) + static_cast<int64_t>(q) * D;
    double x[D], ext[NX], dx[D];
#pragma unroll
    for (int j = 0; j < D; ++j) x[j] = u[x0 + j];
#pragma unroll
    for (int j = 0; j < NX; ++j)
        ext[j] = j < edge_ext_inputs(EM) ? ext_read(e.xmap, e.ext0 + static_cast<int64_t>(q) * e.extdim + j, u) : 0.0;
    edge_ode_f<EM, double>(dx, x, vs, vd, ext, p + e.param0 + static_cast<int64_t>(q) * e.pdim);
#pragma unroll
    for (int j = 0; j < D; ++j) emit1<STAGE>(du, sa, x0 + j, dx[j]);
}

// ---- Jacobian-vector product: the RHS evaluated on dual numbers u + eps*v -------------
// du = f(u) and jv = J(u) v in one pass; the value part is the plain RHS with every
// segment summed in the reference order (no hub trees), the tangent part follows the
// forward-mode rules of each operation.  One thread per vertex.
template <int EM, int WRAP>
__device__ __forceinline__ Dual slot_value_dual(int2 s, Dual own, const double* __restrict__ u,
                                                const double* __restrict__ v, const double* __restrict__ vo,
                                                const double* __restrict__ dvo, const double* __restrict__ p) {
    if constexpr (EM == ND_EM_NOOP) {
        return dual_nan();
    } else if constexpr (is_ode_model(EM)) {
        return ode_side_rt(WRAP, Dual{u[s.x], v[s.x]}, s.y & 1);
    } else {
        const Dual nbr{vo[s.x], dvo[s.x]};
        const double* pe = p + (s.y >> 1);
        if (s.y & 1) {
            const Dual y = g_dst<EM, Dual>(own, nbr, pe);
            return WRAP == ND_WRAP_SYMMETRIC ? y : -y;
        }
        return g_dst<EM, Dual>(nbr, own, pe);
    }
}

__device__ __forceinline__ Dual slot_value_dual_rt(const EDesc& e, int2 s, Dual own, const double* __restrict__ u,
                                                   const double* __restrict__ v, const double* __restrict__ vo,
                                                   const double* __restrict__ dvo, const double* __restrict__ p) {
    if (e.model == ND_EM_NOOP) return dual_nan();
    if (e.ode) return ode_side_rt(e.wrap, Dual{u[s.x], v[s.x]}, s.y & 1);
    const Dual nbr{vo[s.x], dvo[s.x]};
    const double* pe = p + (s.y >> 1);
    if (s.y & 1) {
        const Dual y = g_dst_rt<Dual>(e.model, own, nbr, pe);
        return e.wrap == ND_WRAP_SYMMETRIC ? y : -y;
    }
    return g_dst_rt<Dual>(e.model, nbr, own, pe);
}

template <int VM, int EM, int WRAP, bool MULTI>
__global__ void __launch_bounds__(256) k_vertex_jvp(VDesc vd, const int32_t* __restrict__ rowptr,
                                                    const int2* __restrict__ slots, const uint8_t* __restrict__ slot_eb,
                                                    const EDesc* __restrict__ ed, const double* __restrict__ u,
                                                    const double* __restrict__ v, const double* __restrict__ vo,
                                                    const double* __restrict__ dvo, const double* __restrict__ p,
                                                    double* __restrict__ du, double* __restrict__ jv) {
    constexpr int DIMC = VModel<VM>::dim;
    const int32_t q = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= vd.count) return;
    const int dim = DIMC > 0 ? DIMC : vd.dim;
    const int64_t x0 = static_cast<int64_t>(vd.state0) + static_cast<int64_t>(q) * dim;
    if constexpr (VM == ND_VM_NOOP) {
        for (int j = 0; j < dim; ++j) { du[x0 + j] = 0.0; jv[x0 + j] = 0.0; }
    } else {
        Dual x[DIMC];
#pragma unroll
        for (int j = 0; j < DIMC; ++j) x[j] = Dual{u[x0 + j], v[x0 + j]};
        const Dual own = x[vd.mask[0]];
        Dual acc{0.0, 0.0};
        const int32_t beg = rowptr[vd.row0 + q], end = rowptr[vd.row0 + q + 1];
        for (int32_t s = beg; s < end; ++s) {
            if constexpr (MULTI) acc = acc + slot_value_dual_rt(ed[slot_eb[s]], slots[s], own, u, v, vo, dvo, p);
            else acc = acc + slot_value_dual<EM, WRAP>(slots[s], own, u, v, vo, dvo, p);
        }
        Dual dx[DIMC];
        vertex_f<VM, Dual>(dx, x, acc, p + vd.param0 + static_cast<int64_t>(q) * vd.pdim, DIMC);
#pragma unroll
        for (int j = 0; j < DIMC; ++j) { du[x0 + j] = dx[j].v; jv[x0 + j] = dx[j].d; }
    }
}

// du, jv of ODE edges (the dual form of k_edge_ode); eo, deo: the endpoint outputs and
// their tangents (u and v, or the compact vertex outputs)
template <int EM>
__global__ void __launch_bounds__(256) k_edge_ode_jvp(EDesc e, const double* __restrict__ u, const double* __restrict__ v,
                                                      const double* __restrict__ eo, const double* __restrict__ deo,
                                                      const double* __restrict__ p, double* __restrict__ du,
                                                      double* __restrict__ jv) {
    constexpr int D = ode_dim(EM);
    constexpr int NX = edge_ext_inputs(EM) > 0 ? edge_ext_inputs(EM) : 1;
    const int32_t q = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= e.count) return;
    const int2 en = e.ends[q];
    const Dual vs{eo[en.x], deo[en.x]}, vdn{eo[en.y], deo[en.y]};
    const int64_t x0 = static_cast<int64_t>(e.state0) + static_cast<int64_t>(q) * D;
    Dual x[D], ext[NX], dx[D];
#pragma unroll
    for (int j = 0; j < D; ++j) x[j] = Dual{u[x0 + j], v[x0 + j]};
#pragma unroll
    for (int j = 0; j < NX; ++j)
        ext[j] = j < edge_ext_inputs(EM) ? ext_read_dual(e.xmap, e.ext0 + static_cast<int64_t>(q) * e.extdim + j, u, v)
                                         : Dual{0.0, 0.0};
    edge_ode_f<EM, Dual>(dx, x, vs, vdn, ext, p + e.param0 + static_cast<int64_t>(q) * e.pdim);
#pragma unroll
    for (int j = 0; j < D; ++j) {
        du[x0 + j] = dx[j].v;
        jv[x0 + j] = dx[j].d;
    }
}

// ---- K tangents per launch (the chunked duals of ForwardDiff) -----------------------------
// du = f(u) and, for k <= K tangent columns, jv[:, j] = J(u) v[:, j]: column j of v and
// jv at + j*ldu (ldu = the state length), of the tangent outputs dvo at + j*ldo.  Columns
// k..K-1 are carried as zeros and not written.  Each column equals the single-tangent
// kernels' result along v[:, j] bit for bit (the same per-partial operations).
template <int K>
__device__ __forceinline__ DualK<K> dk_load(const double* __restrict__ x, const double* __restrict__ dx, int64_t ld,
                                            int k, int64_t i) {
    DualK<K> r;
    r.v = x[i];
#pragma unroll
    for (int j = 0; j < K; ++j) r.d[j] = j < k ? dx[j * ld + i] : 0.0;
    return r;
}
template <int K>
__device__ __forceinline__ void dk_store(double* __restrict__ y, double* __restrict__ dy, int64_t ld, int k, int64_t i,
                                         const DualK<K>& r) {
    y[i] = r.v;
#pragma unroll
    for (int j = 0; j < K; ++j)
        if (j < k) dy[j * ld + i] = r.d[j];
}
template <int K>
__device__ __forceinline__ DualK<K> ode_side_k(int wrap, DualK<K> x, bool is_src) {
    return (wrap == ND_WRAP_ANTISYMMETRIC && is_src) ? -x : x;
}
template <int K>
__device__ __forceinline__ DualK<K> ext_read_k(const int32_t* __restrict__ xmap, int64_t q, const double* __restrict__ u,
                                               const double* __restrict__ v, int64_t ldu, int k) {
    const int32_t xi = xmap[q];
    return xi >= 0 ? dk_load<K>(u, v, ldu, k, xi) : -dk_load<K>(u, v, ldu, k, ~xi);
}

struct JvpKArgs {
    const double* u;
    const double* v;              // k tangent columns of the state, stride ldu
    const double* vo;
    const double* dvo;            // k tangent columns of the vertex outputs, stride ldo
    const double* p;
    double* du;
    double* jv;                   // k columns, stride ldu
    int64_t ldu, ldo;
    int32_t k;
};

template <int EM, int WRAP, int K>
__device__ __forceinline__ DualK<K> slot_value_k(int2 s, DualK<K> own, const JvpKArgs& a) {
    if constexpr (EM == ND_EM_NOOP) {
        return nan_of<DualK<K>>();
    } else if constexpr (is_ode_model(EM)) {
        return ode_side_k<K>(WRAP, dk_load<K>(a.u, a.v, a.ldu, a.k, s.x), s.y & 1);
    } else {
        const DualK<K> nbr = dk_load<K>(a.vo, a.dvo, a.ldo, a.k, s.x);
        const double* pe = a.p + (s.y >> 1);
        if (s.y & 1) {
            const DualK<K> y = g_dst<EM, DualK<K>>(own, nbr, pe);
            return WRAP == ND_WRAP_SYMMETRIC ? y : -y;
        }
        return g_dst<EM, DualK<K>>(nbr, own, pe);
    }
}

template <int K>
__device__ __forceinline__ DualK<K> slot_value_k_rt(const EDesc& e, int2 s, DualK<K> own, const JvpKArgs& a) {
    if (e.model == ND_EM_NOOP) return nan_of<DualK<K>>();
    if (e.ode) return ode_side_k<K>(e.wrap, dk_load<K>(a.u, a.v, a.ldu, a.k, s.x), s.y & 1);
    const DualK<K> nbr = dk_load<K>(a.vo, a.dvo, a.ldo, a.k, s.x);
    const double* pe = a.p + (s.y >> 1);
    if (s.y & 1) {
        const DualK<K> y = g_dst_rt<DualK<K>>(e.model, own, nbr, pe);
        return e.wrap == ND_WRAP_SYMMETRIC ? y : -y;
    }
    return g_dst_rt<DualK<K>>(e.model, nbr, own, pe);
}

// the K-tangent form of k_vertex_jvp: one thread per vertex
template <int VM, int EM, int WRAP, bool MULTI, int K>
__global__ void __launch_bounds__(256) k_vertex_jvpk(VDesc vd, const int32_t* __restrict__ rowptr,
                                                     const int2* __restrict__ slots, const uint8_t* __restrict__ slot_eb,
                                                     const EDesc* __restrict__ ed, JvpKArgs a) {
    constexpr int DIMC = VModel<VM>::dim;
    const int32_t q = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= vd.count) return;
    const int dim = DIMC > 0 ? DIMC : vd.dim;
    const int64_t x0 = static_cast<int64_t>(vd.state0) + static_cast<int64_t>(q) * dim;
    if constexpr (VM == ND_VM_NOOP) {
        for (int j = 0; j < dim; ++j) dk_store<K>(a.du, a.jv, a.ldu, a.k, x0 + j, dualk_const<K>(0.0));
    } else {
        DualK<K> x[DIMC];
#pragma unroll
        for (int j = 0; j < DIMC; ++j) x[j] = dk_load<K>(a.u, a.v, a.ldu, a.k, x0 + j);
        DualK<K> own = x[0];          // x[mask] by selects: a runtime index would put x in scratch
#pragma unroll
        for (int j = 1; j < DIMC; ++j)
            if (vd.mask[0] == j) own = x[j];
        DualK<K> acc = dualk_const<K>(0.0);
        const int32_t beg = rowptr[vd.row0 + q], end = rowptr[vd.row0 + q + 1];
        for (int32_t s = beg; s < end; ++s) {
            if constexpr (MULTI) acc = acc + slot_value_k_rt<K>(ed[slot_eb[s]], slots[s], own, a);
            else acc = acc + slot_value_k<EM, WRAP, K>(slots[s], own, a);
        }
        DualK<K> dx[DIMC];
        vertex_f<VM, DualK<K>>(dx, x, acc, a.p + vd.param0 + static_cast<int64_t>(q) * vd.pdim, DIMC);
#pragma unroll
        for (int j = 0; j < DIMC; ++j) dk_store<K>(a.du, a.jv, a.ldu, a.k, x0 + j, dx[j]);
    }
}

// the K-tangent form of k_edge_ode_jvp; eo/deo: a.vo/a.dvo (compact outputs) or u/v
template <int EM, int K>
__global__ void __launch_bounds__(256) k_edge_ode_jvpk(EDesc e, JvpKArgs a, int compact) {
    constexpr int D = ode_dim(EM);
    constexpr int NX = edge_ext_inputs(EM) > 0 ? edge_ext_inputs(EM) : 1;
    const int32_t q = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= e.count) return;
    const int2 en = e.ends[q];
    const double* eo = compact ? a.vo : a.u;
    const double* deo = compact ? a.dvo : a.v;
    const int64_t ldo = compact ? a.ldo : a.ldu;
    const DualK<K> vs = dk_load<K>(eo, deo, ldo, a.k, en.x), vdn = dk_load<K>(eo, deo, ldo, a.k, en.y);
    const int64_t x0 = static_cast<int64_t>(e.state0) + static_cast<int64_t>(q) * D;
    DualK<K> x[D], ext[NX], dx[D];
#pragma unroll
    for (int j = 0; j < D; ++j) x[j] = dk_load<K>(a.u, a.v, a.ldu, a.k, x0 + j);
#pragma unroll
    for (int j = 0; j < NX; ++j)
        ext[j] = j < edge_ext_inputs(EM) ? ext_read_k<K>(e.xmap, e.ext0 + static_cast<int64_t>(q) * e.extdim + j, a.u, a.v,
                                                         a.ldu, a.k)
                                         : dualk_const<K>(0.0);
    edge_ode_f<EM, DualK<K>>(dx, x, vs, vdn, ext, a.p + e.param0 + static_cast<int64_t>(q) * e.pdim);
#pragma unroll
    for (int j = 0; j < D; ++j) dk_store<K>(a.du, a.jv, a.ldu, a.k, x0 + j, dx[j]);
}

__global__ void k_fill(double* __restrict__ x, int64_t n, double v) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        x[i] = v;
}

__device__ __forceinline__ bool same_bits(double a, double b) {
    return __double_as_longlong(a) == __double_as_longlong(b);
}

// Blocks [0, vblocks): o[v_out] = u[StateMask] for one vertex batch (coreloop.jl:39).
// Blocks [vblocks, ...): the parameter-cached edge batch; an edge whose parameters differ
// (bitwise) from the cached copy gets the copy and its slot entries rewritten.  Both
// parts loop grid-stride over two elements per thread (16-byte accesses where the
// layout allows), so a modest grid covers any size.
__device__ __forceinline__ void ksync_edge(const KSync& ks, int64_t e, const double* nv) {
    bool changed = false;
    for (int c = 0; c < ks.np; ++c) changed |= !same_bits(nv[c], ks.kc[e * ks.np + c]);
    if (!changed) return;
    const int4 ps = ks.kpos[e];
    for (int c = 0; c < ks.np; ++c) {
        ks.kc[e * ks.np + c] = nv[c];
        if (ps.x >= 0) ks.kfull[static_cast<int64_t>(ps.x) * ks.np + c] = nv[c];
        if (ps.y >= 0) ks.kfull[static_cast<int64_t>(ps.y) * ks.np + c] = nv[c];
        if (ps.z >= 0) ks.klow[static_cast<int64_t>(ps.z) * ks.np + c] = nv[c];
        if (ps.w >= 0) ks.klow[static_cast<int64_t>(ps.w) * ks.np + c] = nv[c];
    }
}

__device__ __forceinline__ void ksync_part(const KSync& ks, int64_t tid, int64_t nthr) {
    const bool vec = ks.np == 1 && (reinterpret_cast<uintptr_t>(ks.pe) & 15) == 0;
    for (int64_t e = 2 * tid; e < ks.count; e += 2 * nthr) {
        if (vec && e + 1 < ks.count) {
            const double2 k2 = *reinterpret_cast<const double2*>(ks.pe + e);
            const double a = k2.x, b = k2.y;
            ksync_edge(ks, e, &a);
            ksync_edge(ks, e + 1, &b);
        } else {
            for (int64_t f = e; f < e + 2 && f < ks.count; ++f) {
                double nv[2];
                for (int c = 0; c < ks.np; ++c) nv[c] = ks.pe[f * ks.np + c];
                ksync_edge(ks, f, nv);
            }
        }
    }
}

// the outputs of one vertex batch, grid-stride over nthr threads
__device__ __forceinline__ void vertex_out_part(const VDesc& vd, int64_t tid, int64_t nthr, const double* __restrict__ u,
                                                double* __restrict__ o, const double* __restrict__ p, int tangent) {
    if (vd.outkind == 2) return;                 // injector outputs: k_vertex_out_ff, after the hubs'
    if (vd.outkind == 1) {                       // NoFeedForward outputs from the parameters
        for (int64_t q = tid; q < vd.count; q += nthr)
            for (int j = 0; j < vd.depth; ++j)
                o[static_cast<int64_t>(vd.out0) + q * vd.depth + j] =
                    tangent ? 0.0 : vertex_out_nff<double>(vd.model, p + vd.param0 + q * vd.pdim, j);
        return;
    }
    // two-state vertices with output state 1 (swing): two vertices = one 32-byte read
    const bool vec2 = vd.dim == 2 && vd.depth == 1 && vd.mask[0] == 0 && (vd.state0 & 1) == 0 &&
                      (reinterpret_cast<uintptr_t>(u) & 15) == 0 && ((vd.out0 & 1) == 0) &&
                      (reinterpret_cast<uintptr_t>(o) & 15) == 0;
    for (int64_t q = 2 * tid; q < vd.count; q += 2 * nthr) {
        if (vec2 && q + 1 < vd.count) {
            const double2* up = reinterpret_cast<const double2*>(u + vd.state0 + 2 * q);
            const double2 s0 = up[0], s1 = up[1];
            *reinterpret_cast<double2*>(o + vd.out0 + q) = make_double2(s0.x, s1.x);
        } else {
            for (int64_t r = q; r < q + 2 && r < vd.count; ++r) {
                const int64_t x0 = static_cast<int64_t>(vd.state0) + r * vd.dim;
                const int64_t o0 = static_cast<int64_t>(vd.out0) + r * vd.depth;
                for (int j = 0; j < vd.depth; ++j) o[o0 + j] = u[x0 + vd.mask[j]];
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_vertex_out(VDesc vd, const double* __restrict__ u, double* __restrict__ o,
                                                    KSync ks, const double* __restrict__ p, int tangent) {
    if (static_cast<int32_t>(blockIdx.x) >= ks.vblocks) {
        ksync_part(ks, static_cast<int64_t>(blockIdx.x - ks.vblocks) * blockDim.x + threadIdx.x,
                   static_cast<int64_t>(gridDim.x - ks.vblocks) * blockDim.x);
        return;
    }
    vertex_out_part(vd, static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x,
                    static_cast<int64_t>(ks.vblocks) * blockDim.x, u, o, p, tangent);
}

// The same for up to kGroupMax vertex batches in one launch: batch i owns the blocks
// [g.first[i], g.first[i+1]) and ks.vblocks = g.first[g.n].
__global__ void __launch_bounds__(256) k_vertex_out_group(VGroup g, const double* __restrict__ u,
                                                          double* __restrict__ o, KSync ks,
                                                          const double* __restrict__ p, int tangent) {
    const int32_t bx = static_cast<int32_t>(blockIdx.x);
    if (bx >= ks.vblocks) {
        ksync_part(ks, static_cast<int64_t>(bx - ks.vblocks) * blockDim.x + threadIdx.x,
                   static_cast<int64_t>(gridDim.x - ks.vblocks) * blockDim.x);
        return;
    }
    const int i = group_of_block(g, bx);
    vertex_out_part(g.vd[i], static_cast<int64_t>(bx - g.first[i]) * blockDim.x + threadIdx.x,
                    static_cast<int64_t>(g.first[i + 1] - g.first[i]) * blockDim.x, u, o, p, tangent);
}

// Injector outputs (coreloop.jl:44-55): the loopback copy of the hub's output is the
// injector's input, and its PureFeedForward g turns it into the output.  vo: the
// output buffer holding the hubs' outputs (written by k_vertex_out before).
__global__ void __launch_bounds__(256) k_vertex_out_ff(VDesc vd, double* __restrict__ vo,
                                                       const double* __restrict__ p) {
    const int32_t q = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= vd.count) return;
    const double in = vo[vd.lbin[vd.lb0 + q]];
    vo[static_cast<int64_t>(vd.out0) + q] = vertex_out_ff<double>(vd.model, in, p + vd.param0 + static_cast<int64_t>(q) * vd.pdim);
}
// and their tangents, from the hubs' tangent outputs
__global__ void __launch_bounds__(256) k_vertex_out_ff_jvp(VDesc vd, double* __restrict__ vo, double* __restrict__ dvo,
                                                           const double* __restrict__ p) {
    const int32_t q = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= vd.count) return;
    const int32_t h = vd.lbin[vd.lb0 + q];
    const Dual y = vertex_out_ff<Dual>(vd.model, Dual{vo[h], dvo[h]}, p + vd.param0 + static_cast<int64_t>(q) * vd.pdim);
    vo[static_cast<int64_t>(vd.out0) + q] = y.v;
    dvo[static_cast<int64_t>(vd.out0) + q] = y.d;
}

// apply_loopback! into the aggregation buffer (post_utils.jl:221-228): an injector's
// input row holds its hub's output before the edge outputs are added
__global__ void __launch_bounds__(256) k_loopback_copy(VDesc vd, const double* __restrict__ o,
                                                       double* __restrict__ aggbuf) {
    const int32_t q = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= vd.count) return;
    const int32_t h = vd.lbin[vd.lb0 + q];
    if (h >= 0) aggbuf[vd.row0 + q] = o[h];
}

// both outputs of one edge batch (coreloop.jl:41,78 with the wrappers of
// component_functions.jl:81-203).  ODE edges copy their states from u; feed-forward
// edges read the endpoint outputs from eo (u, or o where vertex outputs are computed).
__global__ void k_edge_out(EDesc e, const double* __restrict__ u, const double* __restrict__ eo,
                           const double* __restrict__ p, double* __restrict__ o) {
    const int32_t q = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= e.count) return;
    if (e.model == ND_EM_NOOP) return;
    const int64_t o0 = static_cast<int64_t>(e.out0) + static_cast<int64_t>(q) * (e.outdim_src + e.outdim_dst);
    if (e.ode) {
        const int64_t x0 = static_cast<int64_t>(e.state0) + static_cast<int64_t>(q) * e.dim;
        for (int j = 0; j < e.outdim_src; ++j) o[o0 + j] = ode_side_rt(e.wrap, u[x0 + e.mask_src[j]], true);
        for (int j = 0; j < e.outdim_dst; ++j) o[o0 + e.outdim_src + j] = u[x0 + e.mask_dst[j]];
        return;
    }
    const int2 en = e.ends[q];
    const double* pe = p + e.param0 + static_cast<int64_t>(q) * e.pdim;
    // component j of each side (models at output depth > 1: the endpoint outputs are
    // consecutive)
    for (int j = 0; j < e.outdim_dst; ++j) {
        const double y = e.model == ND_EM_BEAM ? beam_dst<double>(eo + en.x, eo + en.y, pe, j)
                                               : g_dst_rt(e.model, eo[en.x + j], eo[en.y + j], pe);
        o[o0 + e.outdim_src + j] = y;
        if (j < e.outdim_src) {
            if (e.wrap == ND_WRAP_SYMMETRIC) o[o0 + j] = y;
            else if (e.wrap == ND_WRAP_DIRECTED) o[o0 + j] = 0.0;
            else o[o0 + j] = -y;
        }
    }
}

__global__ void k_gather(const int32_t* __restrict__ gmap, const double* __restrict__ o, double* __restrict__ gbuf,
                         int64_t n) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        gbuf[i] = o[gmap[i]];
}

// aggbuf[a] += sum of o over row a.  Rows up to hub_degree: one thread, reference
// order; longer rows: one wave, lane-strided + butterfly.
__global__ void __launch_bounds__(256) k_aggregate(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
                                                   const double* __restrict__ o, double* __restrict__ aggbuf,
                                                   int32_t nrows, const int32_t* __restrict__ hub_rows, int32_t nhub,
                                                   int32_t hub_degree, int32_t nlow) {
    if (static_cast<int32_t>(blockIdx.x) < nlow) {
        const int32_t r = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
        if (r >= nrows) return;
        const int32_t beg = rowptr[r], end = rowptr[r + 1];
        if (end - beg > hub_degree) return;
        double acc = aggbuf[r];
        for (int32_t s = beg; s < end; ++s) acc = acc + o[cols[s]];
        aggbuf[r] = acc;
    } else {
        const int32_t w = static_cast<int32_t>((blockIdx.x - nlow) * (blockDim.x / kWave) + threadIdx.x / kWave);
        const int lane = static_cast<int>(threadIdx.x & (kWave - 1));
        if (w >= nhub) return;
        const int32_t r = hub_rows[w];
        const int32_t beg = rowptr[r], end = rowptr[r + 1];
        double part = 0.0;
        for (int32_t s = beg + lane; s < end; s += kWave) part += o[cols[s]];
        const double sum = wave_sum(part);
        if (lane == 0) aggbuf[r] = aggbuf[r] + sum;
    }
}

// Vertex batches outside the fused kernels' model set -- models that read external
// inputs (external_inputs.jl: collect_externals! fills them after the :g passes, from
// states and StateMask outputs, so here they are read from u directly) or the time,
// computed outputs, and every batch of a network whose edge model the fused kernels do
// not carry.  Same block layout as k_vertex_rhs: a block per big hub, a wave per mid
// hub (fixed-order partial sums), then one thread per remaining vertex summing its
// destination list in the reference order.
__device__ __forceinline__ double gen_segment(int32_t beg, int32_t end, int32_t t, int32_t G, double own,
                                              const int2* __restrict__ slots, const uint8_t* __restrict__ slot_eb,
                                              const EDesc* __restrict__ ed, const double* __restrict__ u,
                                              const double* __restrict__ vo, const double* __restrict__ p) {
    double acc = 0.0;
    for (int32_t k = beg + t; k < end; k += G) acc += slot_value_rt(ed[slot_eb ? slot_eb[k] : 0], slots[k], own, u, vo, p);
    return acc;
}

template <int VM>
__device__ __forceinline__ double gen_own(const VDesc& vd, int32_t q, const double* __restrict__ u,
                                          const double* __restrict__ vo) {
    constexpr int DIMC = VModel<VM>::dim;
    return vd.outkind == 0 ? u[static_cast<int64_t>(vd.state0) + static_cast<int64_t>(q) * DIMC + vd.mask[0]]
                           : vo[static_cast<int64_t>(vd.out0) + q];
}

// the injector input, the external inputs, f and the stage epilogue of vertex q
template <int VM, bool STAGE>
__device__ __forceinline__ void gen_finish(const VDesc& vd, int32_t q, double acc, const double* __restrict__ u,
                                           const double* __restrict__ vo, const double* __restrict__ p,
                                           double* __restrict__ du, const StageArgs& sa) {
    constexpr int DIMC = VModel<VM>::dim;
    constexpr int NX = VExt<VM>::n > 0 ? VExt<VM>::n : 1;
    const int64_t x0 = static_cast<int64_t>(vd.state0) + static_cast<int64_t>(q) * DIMC;
    double x[DIMC];
#pragma unroll
    for (int j = 0; j < DIMC; ++j) x[j] = u[x0 + j];
    if (vd.lbin) {                               // an injector: its input is the hub's output
        const int32_t h = vd.lbin[vd.lb0 + q];
        if (h >= 0) acc = acc + vo[h];
    }
    double ext[NX];
#pragma unroll
    for (int j = 0; j < NX; ++j)
        ext[j] = j < VExt<VM>::n ? ext_read(vd.xmap, vd.ext0 + static_cast<int64_t>(q) * vd.extdim + j, u) : 0.0;
    double dx[DIMC];
    vertex_f_ext<VM>(dx, x, acc, p + vd.param0 + static_cast<int64_t>(q) * vd.pdim, ext, vd.t, DIMC);
#pragma unroll
    for (int j = 0; j < DIMC; ++j) {
        const double y = emit1<STAGE>(du, sa, x0 + j, dx[j]);
        if (STAGE && sa.vnext && j == vd.mask[0]) sa.vnext[static_cast<int64_t>(vd.out0) + q] = y;
    }
}

template <int VM, bool STAGE>
__global__ void __launch_bounds__(256) k_vertex_gen(VDesc vd, const int32_t* __restrict__ rowptr,
                                                    const int2* __restrict__ slots, const uint8_t* __restrict__ slot_eb,
                                                    const EDesc* __restrict__ ed, const double* __restrict__ u,
                                                    const double* __restrict__ vo, const double* __restrict__ p,
                                                    double* __restrict__ du, StageArgs sa) {
    constexpr int DIMC = VModel<VM>::dim;
    if constexpr (DIMC <= 0) {
        return;                                  // no states: only an output (k_vertex_out*)
    } else {
    const int32_t wpb = static_cast<int32_t>(blockDim.x / kWave);
    const int32_t nmidb = (vd.nmid + wpb - 1) / wpb;
    int32_t b = static_cast<int32_t>(blockIdx.x);
    const int32_t t = static_cast<int32_t>(threadIdx.x);
    if (b < vd.nbig) {
        __shared__ double wpart[256 / kWave];
        const int32_t q = vd.hubs[b];
        const int32_t beg = rowptr[vd.row0 + q], end = rowptr[vd.row0 + q + 1];
        const double part = wave_sum(gen_segment(beg, end, t, static_cast<int32_t>(blockDim.x), gen_own<VM>(vd, q, u, vo),
                                                 slots, slot_eb, ed, u, vo, p));
        if ((t & (kWave - 1)) == 0) wpart[t / kWave] = part;
        __syncthreads();
        if (t == 0) {
            double acc = wpart[0];
            for (int32_t w = 1; w < wpb; ++w) acc += wpart[w];
            gen_finish<VM, STAGE>(vd, q, acc, u, vo, p, du, sa);
        }
        return;
    }
    b -= vd.nbig;
    if (b < nmidb) {
        const int32_t w = b * wpb + t / kWave;
        const int32_t lane = t & (kWave - 1);
        if (w >= vd.nmid) return;
        const int32_t q = vd.hubs[vd.nbig + w];
        const int32_t beg = rowptr[vd.row0 + q], end = rowptr[vd.row0 + q + 1];
        const double acc = wave_sum(gen_segment(beg, end, lane, kWave, gen_own<VM>(vd, q, u, vo), slots, slot_eb, ed, u,
                                                vo, p));
        if (lane == 0) gen_finish<VM, STAGE>(vd, q, acc, u, vo, p, du, sa);
        return;
    }
    b -= nmidb;
    const int32_t q = b * static_cast<int32_t>(blockDim.x) + t;
    if (q >= vd.count) return;
    if (vd.lrp && (static_cast<uint32_t>(vd.lrp[q]) & 0x80000000u)) return;   // a hub, summed above
    const double own = gen_own<VM>(vd, q, u, vo);
    double acc = 0.0;
    const int32_t beg = rowptr[vd.row0 + q], end = rowptr[vd.row0 + q + 1];
    for (int32_t k = beg; k < end; ++k) acc = acc + slot_value_rt(ed[slot_eb ? slot_eb[k] : 0], slots[k], own, u, vo, p);
    gen_finish<VM, STAGE>(vd, q, acc, u, vo, p, du, sa);
    }
}

// the dual-number form of k_vertex_gen (du and J(u) v)
template <int VM>
__global__ void __launch_bounds__(256) k_vertex_gen_jvp(VDesc vd, const int32_t* __restrict__ rowptr,
                                                        const int2* __restrict__ slots,
                                                        const uint8_t* __restrict__ slot_eb,
                                                        const EDesc* __restrict__ ed, const double* __restrict__ u,
                                                        const double* __restrict__ v, const double* __restrict__ vo,
                                                        const double* __restrict__ dvo, const double* __restrict__ p,
                                                        double* __restrict__ du, double* __restrict__ jv) {
    constexpr int DIMC = VModel<VM>::dim;
    constexpr int NX = VExt<VM>::n > 0 ? VExt<VM>::n : 1;
    const int32_t q = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if constexpr (DIMC <= 0) {
        return;
    } else {
    if (q >= vd.count) return;
    const int64_t x0 = static_cast<int64_t>(vd.state0) + static_cast<int64_t>(q) * DIMC;
    Dual x[DIMC];
#pragma unroll
    for (int j = 0; j < DIMC; ++j) x[j] = Dual{u[x0 + j], v[x0 + j]};
    const Dual own = vd.outkind == 0 ? x[vd.mask[0]]
                                      : Dual{vo[static_cast<int64_t>(vd.out0) + q], dvo[static_cast<int64_t>(vd.out0) + q]};
    Dual acc{0.0, 0.0};
    const int32_t beg = rowptr[vd.row0 + q], end = rowptr[vd.row0 + q + 1];
    for (int32_t k = beg; k < end; ++k)
        acc = acc + slot_value_dual_rt(ed[slot_eb ? slot_eb[k] : 0], slots[k], own, u, v, vo, dvo, p);
    if (vd.lbin) {
        const int32_t h = vd.lbin[vd.lb0 + q];
        if (h >= 0) acc = acc + Dual{vo[h], dvo[h]};
    }
    Dual ext[NX];
#pragma unroll
    for (int j = 0; j < NX; ++j) {
        ext[j] = Dual{0.0, 0.0};
        if (j < VExt<VM>::n) ext[j] = ext_read_dual(vd.xmap, vd.ext0 + static_cast<int64_t>(q) * vd.extdim + j, u, v);
    }
    Dual dx[DIMC];
    vertex_f_ext<VM, Dual>(dx, x, acc, p + vd.param0 + static_cast<int64_t>(q) * vd.pdim, ext, vd.t, DIMC);
#pragma unroll
    for (int j = 0; j < DIMC; ++j) { du[x0 + j] = dx[j].v; jv[x0 + j] = dx[j].d; }
    }
}

// One entry of a destination row at output depth d: component j of the edge output that
// this vertex receives.  own: this vertex's output vector; vo + (s.x - j): the other
// endpoint's (its d components are consecutive).  Componentwise models read component j
// of both; the beam reads both vectors.
template <typename S>
__device__ __forceinline__ S vec_slot_value(const EDesc& e, int2 s, int j, const S* own, const S* nb,
                                            const double* __restrict__ p) {
    if (e.model == ND_EM_NOOP) return nan_of<S>();
    const double* pe = p + (s.y >> 1);
    const bool is_src = s.y & 1;
    S y;
    if (e.model == ND_EM_BEAM) y = is_src ? beam_dst<S>(own, nb, pe, j) : beam_dst<S>(nb, own, pe, j);
    else y = is_src ? g_dst_rt<S>(e.model, own[j], nb[j], pe) : g_dst_rt<S>(e.model, nb[j], own[j], pe);
    if (is_src) return e.wrap == ND_WRAP_SYMMETRIC ? y : -y;
    return y;
}

// f at output depth d: diffusion dv .= esum, the free truss node
template <typename S>
__device__ __forceinline__ void vec_vertex_f(const VDesc& vd, S* dx, const S* x, const S* acc,
                                             const double* __restrict__ pv) {
    if (vd.model == ND_VM_TRUSS_FREE) truss_free_f<S>(dx, x, acc, pv);
    else for (int j = 0; j < vd.depth; ++j) dx[j] = acc[j];
}

// Vertex batches with d > 1 output components (vdepth = edepth = d): componentwise
// diffusion (getting_started_with_network_dynamics.jl:218-221) and the truss of
// stress_on_truss.jl.  One thread per vertex; component j's destination row
// (row0 + q*d + j) is summed in the reference order.
template <bool STAGE>
__global__ void __launch_bounds__(256) k_vertex_vec(VDesc vd, const int32_t* __restrict__ rowptr,
                                                    const int2* __restrict__ slots, const uint8_t* __restrict__ slot_eb,
                                                    const EDesc* __restrict__ ed, const double* __restrict__ u,
                                                    const double* __restrict__ vo, const double* __restrict__ p,
                                                    double* __restrict__ du, StageArgs sa) {
    const int32_t q = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= vd.count) return;
    const int d = vd.depth;
    const int64_t x0 = static_cast<int64_t>(vd.state0) + static_cast<int64_t>(q) * vd.dim;
    if (vd.model == ND_VM_NOOP || vd.dim == 0) { // f does nothing (or no states): du keeps zero
        for (int j = 0; j < vd.dim; ++j) emit1<STAGE>(du, sa, x0 + j, 0.0);
        return;
    }
    double own[ND_MAX_OUT], acc[ND_MAX_OUT], x[4], dx[4];
    for (int k = 0; k < d; ++k)
        own[k] = vd.outkind == 0 ? u[x0 + vd.mask[k]] : vo[static_cast<int64_t>(vd.out0) + static_cast<int64_t>(q) * d + k];
    for (int j = 0; j < d; ++j) {
        const int32_t row = vd.row0 + q * d + j;
        double a = 0.0;
        for (int32_t k = rowptr[row]; k < rowptr[row + 1]; ++k) {
            const int2 sl = slots[k];
            const EDesc& e = ed[slot_eb ? slot_eb[k] : 0];
            // NoOp entries read no endpoint (their slot holds none)
            a = a + (e.model == ND_EM_NOOP ? __builtin_nan("") : vec_slot_value<double>(e, sl, j, own, vo + (sl.x - j), p));
        }
        acc[j] = a;
    }
    for (int j = 0; j < vd.dim && j < 4; ++j) x[j] = u[x0 + j];
    vec_vertex_f<double>(vd, dx, x, acc, p + vd.param0 + static_cast<int64_t>(q) * vd.pdim);
    for (int j = 0; j < vd.dim && j < 4; ++j) {
        const double y = emit1<STAGE>(du, sa, x0 + j, dx[j]);
        if (STAGE && sa.vnext)
            for (int k = 0; k < d; ++k)
                if (vd.mask[k] == j) sa.vnext[static_cast<int64_t>(vd.out0) + static_cast<int64_t>(q) * d + k] = y;
    }
}

// its dual-number form (du and J(u) v)
__global__ void __launch_bounds__(256) k_vertex_vec_jvp(VDesc vd, const int32_t* __restrict__ rowptr,
                                                        const int2* __restrict__ slots,
                                                        const uint8_t* __restrict__ slot_eb,
                                                        const EDesc* __restrict__ ed, const double* __restrict__ u,
                                                        const double* __restrict__ v, const double* __restrict__ vo,
                                                        const double* __restrict__ dvo, const double* __restrict__ p,
                                                        double* __restrict__ du, double* __restrict__ jv) {
    const int32_t q = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= vd.count) return;
    const int d = vd.depth;
    const int64_t x0 = static_cast<int64_t>(vd.state0) + static_cast<int64_t>(q) * vd.dim;
    if (vd.model == ND_VM_NOOP || vd.dim == 0) {
        for (int j = 0; j < vd.dim; ++j) { du[x0 + j] = 0.0; jv[x0 + j] = 0.0; }
        return;
    }
    Dual own[ND_MAX_OUT], acc[ND_MAX_OUT], x[4], dx[4];
    for (int k = 0; k < d; ++k) {
        const int64_t oi = static_cast<int64_t>(vd.out0) + static_cast<int64_t>(q) * d + k;
        own[k] = vd.outkind == 0 ? Dual{u[x0 + vd.mask[k]], v[x0 + vd.mask[k]]} : Dual{vo[oi], dvo[oi]};
    }
    for (int j = 0; j < d; ++j) {
        const int32_t row = vd.row0 + q * d + j;
        Dual a{0.0, 0.0};
        for (int32_t k = rowptr[row]; k < rowptr[row + 1]; ++k) {
            const int2 sl = slots[k];
            const EDesc& e = ed[slot_eb ? slot_eb[k] : 0];
            if (e.model == ND_EM_NOOP) {         // g writes nothing (its slot holds no endpoint)
                a = a + dual_nan();
                continue;
            }
            Dual nb[ND_MAX_OUT];
            for (int c = 0; c < d; ++c) nb[c] = Dual{vo[sl.x - j + c], dvo[sl.x - j + c]};
            a = a + vec_slot_value<Dual>(e, sl, j, own, nb, p);
        }
        acc[j] = a;
    }
    for (int j = 0; j < vd.dim && j < 4; ++j) x[j] = Dual{u[x0 + j], v[x0 + j]};
    vec_vertex_f<Dual>(vd, dx, x, acc, p + vd.param0 + static_cast<int64_t>(q) * vd.pdim);
    for (int j = 0; j < vd.dim && j < 4; ++j) { du[x0 + j] = dx[j].v; jv[x0 + j] = dx[j].d; }
}

// collect_externals!(map, extbuf, u, o), external_inputs.jl:49-77: map > 0 reads u,
// map < 0 reads o (1-based magnitudes)
__global__ void k_collect_externals(const int32_t* __restrict__ map, int64_t n, double* __restrict__ extbuf,
                                    const double* __restrict__ u, const double* __restrict__ o) {
    for (int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; k < n;
         k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t m = map[k];
        extbuf[k] = m > 0 ? u[m - 1] : o[-m - 1];
    }
}

// ---- perturbed right-hand side (coreloop.jl:1-102 with perturb / perturb_maps) ----------
// The buffered form of the RHS on dual numbers: every phase writes the reference's
// buffers (o, aggbuf, gbuf, extbuf) and their tangents, so the four additive
// perturbations of coreloop.jl:49-87 enter where the reference adds them.  Linear phases
// (gather, aggregation, loopback copy, external inputs) run the plain kernels on the
// value and on the tangent buffers; the kernels below evaluate the models on Duals.

// apply_perturb! (coreloop.jl:258-262): buf[m.x] += delta[m.y] in map order.  One thread:
// a map holds a handful of channels, and the order stays the reference's.
__global__ void k_perturb(const int2* __restrict__ map, int32_t n, const double* __restrict__ delta,
                          const double* __restrict__ ddelta, double* __restrict__ buf, double* __restrict__ dbuf) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int32_t k = 0; k < n; ++k) {
        const int2 m = map[k];
        buf[m.x] += delta[m.y];
        dbuf[m.x] += ddelta[m.y];
    }
}

// One component's parameters p[first .. first+n) with their tangents dp as Duals (the
// models read at most kMaxPar of them)
constexpr int kMaxPar = 8;
__device__ __forceinline__ void dual_params(Dual* pd, const double* __restrict__ p, const double* __restrict__ dp,
                                            int64_t first, int n) {
#pragma unroll
    for (int c = 0; c < kMaxPar; ++c) pd[c] = c < n ? Dual{p[first + c], dp[first + c]} : Dual{0.0, 0.0};
}

// feed-forward vertex outputs g(in, p) from the (perturbed) input buffer (coreloop.jl:55)
__global__ void __launch_bounds__(256) k_vertex_out_ff_buf(VDesc vd, const double* __restrict__ agg,
                                                           const double* __restrict__ dagg,
                                                           const double* __restrict__ p, const double* __restrict__ dp,
                                                           double* __restrict__ o, double* __restrict__ dout) {
    const int32_t q = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= vd.count) return;
    Dual pd[kMaxPar];
    dual_params(pd, p, dp, vd.param0 + static_cast<int64_t>(q) * vd.pdim, vd.pdim);
    const Dual y = vertex_out_ff<Dual, Dual>(vd.model, Dual{agg[vd.row0 + q], dagg[vd.row0 + q]}, pd);
    o[static_cast<int64_t>(vd.out0) + q] = y.v;
    dout[static_cast<int64_t>(vd.out0) + q] = y.d;
}

// NoFeedForward vertex outputs g(p) (voltage sources): their tangents in the parameters
__global__ void __launch_bounds__(256) k_vertex_out_nff_tangent(VDesc vd, const double* __restrict__ p,
                                                                const double* __restrict__ dp,
                                                                double* __restrict__ dout) {
    const int32_t q = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= vd.count) return;
    Dual pd[kMaxPar];
    dual_params(pd, p, dp, vd.param0 + static_cast<int64_t>(q) * vd.pdim, vd.pdim);
    for (int j = 0; j < vd.depth; ++j)
        dout[static_cast<int64_t>(vd.out0) + static_cast<int64_t>(q) * vd.depth + j] =
            vertex_out_nff<Dual, Dual>(vd.model, pd, j).d;
}

// edges from the gathered buffer (coreloop.jl:76-78): ODE edges' f into du, feed-forward
// edges' g into o.  egb: first gbuf index of each edge of the batch (src output, then dst).
__global__ void __launch_bounds__(256) k_edge_buf(EDesc e, const int32_t* __restrict__ egb,
                                                  const double* __restrict__ gbuf, const double* __restrict__ dgbuf,
                                                  const double* __restrict__ u, const double* __restrict__ v,
                                                  const double* __restrict__ ext, const double* __restrict__ dext,
                                                  const double* __restrict__ p, const double* __restrict__ dp,
                                                  double* __restrict__ o, double* __restrict__ dout,
                                                  double* __restrict__ du, double* __restrict__ jv) {
    const int32_t q = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= e.count || e.model == ND_EM_NOOP) return;
    const int32_t g0 = egb[q];
    const int32_t gd = e.ode ? 1 : e.outdim_dst;  // gbuf: the src output's components, then the dst's
    const Dual vs{gbuf[g0], dgbuf[g0]}, vdn{gbuf[g0 + gd], dgbuf[g0 + gd]};
    Dual pe[kMaxPar];
    dual_params(pe, p, dp, e.param0 + static_cast<int64_t>(q) * e.pdim, e.pdim);
    if (e.ode) {                                 // outputs: StateMasks, written before the gather
        const int64_t x0 = static_cast<int64_t>(e.state0) + static_cast<int64_t>(q) * e.dim;
        Dual x[2] = {Dual{0.0, 0.0}, Dual{0.0, 0.0}}, dx[2];
        for (int j = 0; j < e.dim && j < 2; ++j) x[j] = Dual{u[x0 + j], v[x0 + j]};
        Dual xe[1] = {Dual{0.0, 0.0}};
        if (e.extdim > 0) {
            const int64_t k = static_cast<int64_t>(e.ext0) + static_cast<int64_t>(q) * e.extdim;
            xe[0] = Dual{ext[k], dext[k]};
        }
        edge_ode_f_rt<Dual, Dual>(e.model, dx, x, vs, vdn, xe, pe);
        for (int j = 0; j < e.dim && j < 2; ++j) {
            du[x0 + j] = dx[j].v;
            jv[x0 + j] = dx[j].d;
        }
        return;
    }
    const int64_t o0 = static_cast<int64_t>(e.out0) + static_cast<int64_t>(q) * (e.outdim_src + e.outdim_dst);
    Dual vsv[ND_MAX_OUT], vdv[ND_MAX_OUT];
    for (int j = 0; j < e.outdim_dst && j < ND_MAX_OUT; ++j) {
        vsv[j] = Dual{gbuf[g0 + j], dgbuf[g0 + j]};
        vdv[j] = Dual{gbuf[g0 + gd + j], dgbuf[g0 + gd + j]};
    }
    for (int j = 0; j < e.outdim_dst; ++j) {
        const Dual y = e.model == ND_EM_BEAM ? beam_dst<Dual, Dual>(vsv, vdv, pe, j)
                                             : g_dst_rt<Dual, Dual>(e.model, vsv[j], vdv[j], pe);
        o[o0 + e.outdim_src + j] = y.v;
        dout[o0 + e.outdim_src + j] = y.d;
        if (j < e.outdim_src) {
            const Dual ys = e.wrap == ND_WRAP_SYMMETRIC ? y : (e.wrap == ND_WRAP_DIRECTED ? Dual{0.0, 0.0} : -y);
            o[o0 + j] = ys.v;
            dout[o0 + j] = ys.d;
        }
    }
}

// vertex f from the aggregated (perturbed) inputs and the collected external inputs
// (coreloop.jl:97)
template <int VM>
__global__ void __launch_bounds__(256) k_vertex_f_buf(VDesc vd, const double* __restrict__ agg,
                                                      const double* __restrict__ dagg, const double* __restrict__ u,
                                                      const double* __restrict__ v, const double* __restrict__ ext,
                                                      const double* __restrict__ dext, const double* __restrict__ p,
                                                      const double* __restrict__ dp, double* __restrict__ du,
                                                      double* __restrict__ jv) {
    constexpr int DIMC = VModel<VM>::dim;
    const int32_t q = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= vd.count) return;
    if (VM == ND_VM_TRUSS_FREE) {                // the free truss node, inputs (Fx, Fy)
        const int64_t x0 = static_cast<int64_t>(vd.state0) + static_cast<int64_t>(q) * 4;
        const int64_t r = static_cast<int64_t>(vd.row0) + static_cast<int64_t>(q) * 2;
        Dual x[4], acc[2], dx[4], pd[kMaxPar];
        for (int j = 0; j < 4; ++j) x[j] = Dual{u[x0 + j], v[x0 + j]};
        for (int j = 0; j < 2; ++j) acc[j] = agg ? Dual{agg[r + j], dagg[r + j]} : Dual{0.0, 0.0};
        dual_params(pd, p, dp, vd.param0 + static_cast<int64_t>(q) * vd.pdim, vd.pdim);
        truss_free_f<Dual, Dual>(dx, x, acc, pd);
        for (int j = 0; j < 4; ++j) { du[x0 + j] = dx[j].v; jv[x0 + j] = dx[j].d; }
        return;
    }
    if (VM == ND_VM_DIFFUSION && vd.depth > 1) { // d components: dv .= esum
        const int64_t x0 = static_cast<int64_t>(vd.state0) + static_cast<int64_t>(q) * vd.dim;
        for (int j = 0; j < vd.depth; ++j) {
            const int64_t r = static_cast<int64_t>(vd.row0) + static_cast<int64_t>(q) * vd.depth + j;
            du[x0 + j] = agg ? agg[r] : 0.0;
            jv[x0 + j] = agg ? dagg[r] : 0.0;
        }
        return;
    }
    if constexpr (DIMC <= 0) {                   // no f (NoOp): du stays zero
        const int64_t x0 = static_cast<int64_t>(vd.state0) + static_cast<int64_t>(q) * vd.dim;
        for (int j = 0; j < vd.dim; ++j) {
            du[x0 + j] = 0.0;
            jv[x0 + j] = 0.0;
        }
    } else {
        constexpr int NX = VExt<VM>::n > 0 ? VExt<VM>::n : 1;
        const int64_t x0 = static_cast<int64_t>(vd.state0) + static_cast<int64_t>(q) * DIMC;
        Dual x[DIMC], xe[NX], dx[DIMC];
#pragma unroll
        for (int j = 0; j < DIMC; ++j) x[j] = Dual{u[x0 + j], v[x0 + j]};
        const Dual acc = agg ? Dual{agg[vd.row0 + q], dagg[vd.row0 + q]} : Dual{0.0, 0.0};
#pragma unroll
        for (int j = 0; j < NX; ++j) {
            xe[j] = Dual{0.0, 0.0};
            if (j < VExt<VM>::n) {
                const int64_t k = static_cast<int64_t>(vd.ext0) + static_cast<int64_t>(q) * vd.extdim + j;
                xe[j] = Dual{ext[k], dext[k]};
            }
        }
        Dual pd[kMaxPar];
        dual_params(pd, p, dp, vd.param0 + static_cast<int64_t>(q) * vd.pdim, vd.pdim);
        vertex_f_ext<VM, Dual, Dual>(dx, x, acc, pd, xe, vd.t, DIMC);
#pragma unroll
        for (int j = 0; j < DIMC; ++j) {
            du[x0 + j] = dx[j].v;
            jv[x0 + j] = dx[j].d;
        }
    }
}

// ---- kernel selection ----------------------------------------------------------

using VertexKernel = void (*)(VDesc, const int32_t*, const int2*, const uint8_t*, const EDesc*, const double*,
                              const double*, const double*, double*, StageArgs);

template <int VM, int EM, int WRAP, bool STAGE>
VertexKernel pick3(bool kc) {
    if constexpr (EModel<EM>::np > 0 && !is_ode_model(EM))
        if (kc) return k_vertex_rhs<VM, EM, WRAP, false, STAGE, true>;
    return k_vertex_rhs<VM, EM, WRAP, false, STAGE, false>;
}

template <int VM, bool STAGE>
VertexKernel pick_edge(int em, int wrap, bool multi, bool kc) {
    if (multi) return k_vertex_rhs<VM, ND_EM_NOOP, ND_WRAP_NONE, true, STAGE, false>;
    const bool sym = wrap == ND_WRAP_SYMMETRIC;
    switch (em) {
    case ND_EM_KURAMOTO:
        return sym ? pick3<VM, ND_EM_KURAMOTO, ND_WRAP_SYMMETRIC, STAGE>(kc)
                   : pick3<VM, ND_EM_KURAMOTO, ND_WRAP_ANTISYMMETRIC, STAGE>(kc);
    case ND_EM_STATIC_LINE:
        return sym ? pick3<VM, ND_EM_STATIC_LINE, ND_WRAP_SYMMETRIC, STAGE>(kc)
                   : pick3<VM, ND_EM_STATIC_LINE, ND_WRAP_ANTISYMMETRIC, STAGE>(kc);
    case ND_EM_DIFFUSION:
        return sym ? pick3<VM, ND_EM_DIFFUSION, ND_WRAP_SYMMETRIC, STAGE>(false)
                   : pick3<VM, ND_EM_DIFFUSION, ND_WRAP_ANTISYMMETRIC, STAGE>(false);
    case ND_EM_DIFFUSION_WEIGHTED:
        return sym ? pick3<VM, ND_EM_DIFFUSION_WEIGHTED, ND_WRAP_SYMMETRIC, STAGE>(kc)
                   : pick3<VM, ND_EM_DIFFUSION_WEIGHTED, ND_WRAP_ANTISYMMETRIC, STAGE>(kc);
    case ND_EM_DIFFUSION_FIDUCIAL:
        return pick3<VM, ND_EM_DIFFUSION_FIDUCIAL, ND_WRAP_ANTISYMMETRIC, STAGE>(kc);
    case ND_EM_ODE_SIN_DIFFUSION:
    case ND_EM_ODE_SIN_DIFFUSION_FAST:
    case ND_EM_ODE_LINEAR:
    case ND_EM_ODE_LAG:
    case ND_EM_EXT_DRIVEN:   // ODE edges: entries read the edge state, negated on an AntiSymmetric src side
        return wrap == ND_WRAP_ANTISYMMETRIC ? pick3<VM, ND_EM_ODE_LINEAR, ND_WRAP_ANTISYMMETRIC, STAGE>(false)
                                             : pick3<VM, ND_EM_ODE_LINEAR, ND_WRAP_FIDUCIAL, STAGE>(false);
    default:
        return pick3<VM, ND_EM_NOOP, ND_WRAP_NONE, STAGE>(false);
    }
}

template <bool STAGE>
VertexKernel pick_vertex_kernel(int vm, int em, int wrap, bool multi, bool kc) {
    switch (vm) {
    case ND_VM_DIFFUSION: return pick_edge<ND_VM_DIFFUSION, STAGE>(em, wrap, multi, kc);
    case ND_VM_KURAMOTO: return pick_edge<ND_VM_KURAMOTO, STAGE>(em, wrap, multi, kc);
    case ND_VM_SWING: return pick_edge<ND_VM_SWING, STAGE>(em, wrap, multi, kc);
    case ND_VM_SWING_UNIT_DAMPING: return pick_edge<ND_VM_SWING_UNIT_DAMPING, STAGE>(em, wrap, multi, kc);
    case ND_VM_LOAD: return pick_edge<ND_VM_LOAD, STAGE>(em, wrap, multi, kc);
    default: return k_vertex_rhs<ND_VM_NOOP, ND_EM_NOOP, ND_WRAP_NONE, false, STAGE, false>;
    }
}

using GroupKernel = void (*)(VGroup, const int32_t*, const int2*, const uint8_t*, const EDesc*, const double*,
                             const double*, const double*, double*, StageArgs);

template <int EM, int WRAP, bool STAGE>
GroupKernel gpick3(bool kc) {
    if constexpr (EModel<EM>::np > 0 && !is_ode_model(EM))
        if (kc) return k_vertex_rhs_group<EM, WRAP, false, STAGE, true>;
    return k_vertex_rhs_group<EM, WRAP, false, STAGE, false>;
}

// the group kernel with the template arguments pick_edge gives each member batch
template <bool STAGE>
GroupKernel pick_group_kernel(int em, int wrap, bool multi, bool kc) {
    if (multi) return k_vertex_rhs_group<ND_EM_NOOP, ND_WRAP_NONE, true, STAGE, false>;
    const bool sym = wrap == ND_WRAP_SYMMETRIC;
    switch (em) {
    case ND_EM_KURAMOTO:
        return sym ? gpick3<ND_EM_KURAMOTO, ND_WRAP_SYMMETRIC, STAGE>(kc)
                   : gpick3<ND_EM_KURAMOTO, ND_WRAP_ANTISYMMETRIC, STAGE>(kc);
    case ND_EM_STATIC_LINE:
        return sym ? gpick3<ND_EM_STATIC_LINE, ND_WRAP_SYMMETRIC, STAGE>(kc)
                   : gpick3<ND_EM_STATIC_LINE, ND_WRAP_ANTISYMMETRIC, STAGE>(kc);
    case ND_EM_DIFFUSION:
        return sym ? gpick3<ND_EM_DIFFUSION, ND_WRAP_SYMMETRIC, STAGE>(false)
                   : gpick3<ND_EM_DIFFUSION, ND_WRAP_ANTISYMMETRIC, STAGE>(false);
    case ND_EM_DIFFUSION_WEIGHTED:
        return sym ? gpick3<ND_EM_DIFFUSION_WEIGHTED, ND_WRAP_SYMMETRIC, STAGE>(kc)
                   : gpick3<ND_EM_DIFFUSION_WEIGHTED, ND_WRAP_ANTISYMMETRIC, STAGE>(kc);
    case ND_EM_DIFFUSION_FIDUCIAL:
        return gpick3<ND_EM_DIFFUSION_FIDUCIAL, ND_WRAP_ANTISYMMETRIC, STAGE>(kc);
    case ND_EM_ODE_SIN_DIFFUSION:
    case ND_EM_ODE_SIN_DIFFUSION_FAST:
    case ND_EM_ODE_LINEAR:
    case ND_EM_ODE_LAG:
    case ND_EM_EXT_DRIVEN:
        return wrap == ND_WRAP_ANTISYMMETRIC ? gpick3<ND_EM_ODE_LINEAR, ND_WRAP_ANTISYMMETRIC, STAGE>(false)
                                             : gpick3<ND_EM_ODE_LINEAR, ND_WRAP_FIDUCIAL, STAGE>(false);
    default:
        return gpick3<ND_EM_NOOP, ND_WRAP_NONE, STAGE>(false);
    }
}

template <bool STAGE>
VertexKernel pick_gen_kernel(int vm) {
    switch (vm) {
    case ND_VM_DIFFUSION: return k_vertex_gen<ND_VM_DIFFUSION, STAGE>;
    case ND_VM_KURAMOTO: return k_vertex_gen<ND_VM_KURAMOTO, STAGE>;
    case ND_VM_SWING: return k_vertex_gen<ND_VM_SWING, STAGE>;
    case ND_VM_SWING_UNIT_DAMPING: return k_vertex_gen<ND_VM_SWING_UNIT_DAMPING, STAGE>;
    case ND_VM_LOAD: return k_vertex_gen<ND_VM_LOAD, STAGE>;
    case ND_VM_CAPACITOR_SINK: return k_vertex_gen<ND_VM_CAPACITOR_SINK, STAGE>;
    case ND_VM_PI_SOURCE: return k_vertex_gen<ND_VM_PI_SOURCE, STAGE>;
    case ND_VM_CAPACITOR: return k_vertex_gen<ND_VM_CAPACITOR, STAGE>;
    case ND_VM_VOLTAGE_SOURCE: return k_vertex_gen<ND_VM_VOLTAGE_SOURCE, STAGE>;
    case ND_VM_RESISTOR_INJECTOR: return k_vertex_gen<ND_VM_RESISTOR_INJECTOR, STAGE>;
    case ND_VM_INDUCTOR_INJECTOR: return k_vertex_gen<ND_VM_INDUCTOR_INJECTOR, STAGE>;
    case ND_VM_CRL: return k_vertex_gen<ND_VM_CRL, STAGE>;
    case ND_VM_FHN: return k_vertex_gen<ND_VM_FHN, STAGE>;
    default: return k_vertex_rhs<ND_VM_NOOP, ND_EM_NOOP, ND_WRAP_NONE, false, STAGE, false>;
    }
}

using VertexBufKernel = void (*)(VDesc, const double*, const double*, const double*, const double*, const double*,
                                 const double*, const double*, const double*, double*, double*);
VertexBufKernel pick_vertex_f_buf(int vm) {
    switch (vm) {
    case ND_VM_DIFFUSION: return k_vertex_f_buf<ND_VM_DIFFUSION>;
    case ND_VM_KURAMOTO: return k_vertex_f_buf<ND_VM_KURAMOTO>;
    case ND_VM_SWING: return k_vertex_f_buf<ND_VM_SWING>;
    case ND_VM_SWING_UNIT_DAMPING: return k_vertex_f_buf<ND_VM_SWING_UNIT_DAMPING>;
    case ND_VM_LOAD: return k_vertex_f_buf<ND_VM_LOAD>;
    case ND_VM_CAPACITOR_SINK: return k_vertex_f_buf<ND_VM_CAPACITOR_SINK>;
    case ND_VM_PI_SOURCE: return k_vertex_f_buf<ND_VM_PI_SOURCE>;
    case ND_VM_CAPACITOR: return k_vertex_f_buf<ND_VM_CAPACITOR>;
    case ND_VM_VOLTAGE_SOURCE: return k_vertex_f_buf<ND_VM_VOLTAGE_SOURCE>;
    case ND_VM_RESISTOR_INJECTOR: return k_vertex_f_buf<ND_VM_RESISTOR_INJECTOR>;
    case ND_VM_INDUCTOR_INJECTOR: return k_vertex_f_buf<ND_VM_INDUCTOR_INJECTOR>;
    case ND_VM_CRL: return k_vertex_f_buf<ND_VM_CRL>;
    case ND_VM_FHN: return k_vertex_f_buf<ND_VM_FHN>;
    case ND_VM_TRUSS_FREE: return k_vertex_f_buf<ND_VM_TRUSS_FREE>;
    default: return k_vertex_f_buf<ND_VM_NOOP>;
    }
}

using JvpKernel = void (*)(VDesc, const int32_t*, const int2*, const uint8_t*, const EDesc*, const double*,
                           const double*, const double*, const double*, const double*, double*, double*);

template <int VM>
JvpKernel pick_jvp_edge(int em, int wrap, bool multi) {
    if (multi) return k_vertex_jvp<VM, ND_EM_NOOP, ND_WRAP_NONE, true>;
    const bool sym = wrap == ND_WRAP_SYMMETRIC;
    switch (em) {
    case ND_EM_KURAMOTO:
        return sym ? k_vertex_jvp<VM, ND_EM_KURAMOTO, ND_WRAP_SYMMETRIC, false>
                   : k_vertex_jvp<VM, ND_EM_KURAMOTO, ND_WRAP_ANTISYMMETRIC, false>;
    case ND_EM_STATIC_LINE:
        return sym ? k_vertex_jvp<VM, ND_EM_STATIC_LINE, ND_WRAP_SYMMETRIC, false>
                   : k_vertex_jvp<VM, ND_EM_STATIC_LINE, ND_WRAP_ANTISYMMETRIC, false>;
    case ND_EM_DIFFUSION:
        return sym ? k_vertex_jvp<VM, ND_EM_DIFFUSION, ND_WRAP_SYMMETRIC, false>
                   : k_vertex_jvp<VM, ND_EM_DIFFUSION, ND_WRAP_ANTISYMMETRIC, false>;
    case ND_EM_DIFFUSION_WEIGHTED:
        return sym ? k_vertex_jvp<VM, ND_EM_DIFFUSION_WEIGHTED, ND_WRAP_SYMMETRIC, false>
                   : k_vertex_jvp<VM, ND_EM_DIFFUSION_WEIGHTED, ND_WRAP_ANTISYMMETRIC, false>;
    case ND_EM_DIFFUSION_FIDUCIAL:
        return k_vertex_jvp<VM, ND_EM_DIFFUSION_FIDUCIAL, ND_WRAP_ANTISYMMETRIC, false>;
    case ND_EM_ODE_SIN_DIFFUSION:
    case ND_EM_ODE_SIN_DIFFUSION_FAST:
    case ND_EM_ODE_LINEAR:
    case ND_EM_ODE_LAG:
    case ND_EM_EXT_DRIVEN:
        return wrap == ND_WRAP_ANTISYMMETRIC ? k_vertex_jvp<VM, ND_EM_ODE_LINEAR, ND_WRAP_ANTISYMMETRIC, false>
                                             : k_vertex_jvp<VM, ND_EM_ODE_LINEAR, ND_WRAP_FIDUCIAL, false>;
    default:
        return k_vertex_jvp<VM, ND_EM_NOOP, ND_WRAP_NONE, false>;
    }
}

JvpKernel pick_jvp_kernel(int vm, int em, int wrap, bool multi) {
    switch (vm) {
    case ND_VM_DIFFUSION: return pick_jvp_edge<ND_VM_DIFFUSION>(em, wrap, multi);
    case ND_VM_KURAMOTO: return pick_jvp_edge<ND_VM_KURAMOTO>(em, wrap, multi);
    case ND_VM_SWING: return pick_jvp_edge<ND_VM_SWING>(em, wrap, multi);
    case ND_VM_SWING_UNIT_DAMPING: return pick_jvp_edge<ND_VM_SWING_UNIT_DAMPING>(em, wrap, multi);
    case ND_VM_LOAD: return pick_jvp_edge<ND_VM_LOAD>(em, wrap, multi);
    default: return k_vertex_jvp<ND_VM_NOOP, ND_EM_NOOP, ND_WRAP_NONE, false>;
    }
}

JvpKernel pick_gen_jvp(int vm) {
    switch (vm) {
    case ND_VM_DIFFUSION: return k_vertex_gen_jvp<ND_VM_DIFFUSION>;
    case ND_VM_KURAMOTO: return k_vertex_gen_jvp<ND_VM_KURAMOTO>;
    case ND_VM_SWING: return k_vertex_gen_jvp<ND_VM_SWING>;
    case ND_VM_SWING_UNIT_DAMPING: return k_vertex_gen_jvp<ND_VM_SWING_UNIT_DAMPING>;
    case ND_VM_LOAD: return k_vertex_gen_jvp<ND_VM_LOAD>;
    case ND_VM_CAPACITOR_SINK: return k_vertex_gen_jvp<ND_VM_CAPACITOR_SINK>;
    case ND_VM_PI_SOURCE: return k_vertex_gen_jvp<ND_VM_PI_SOURCE>;
    case ND_VM_CAPACITOR: return k_vertex_gen_jvp<ND_VM_CAPACITOR>;
    case ND_VM_VOLTAGE_SOURCE: return k_vertex_gen_jvp<ND_VM_VOLTAGE_SOURCE>;
    case ND_VM_RESISTOR_INJECTOR: return k_vertex_gen_jvp<ND_VM_RESISTOR_INJECTOR>;
    case ND_VM_INDUCTOR_INJECTOR: return k_vertex_gen_jvp<ND_VM_INDUCTOR_INJECTOR>;
    case ND_VM_CRL: return k_vertex_gen_jvp<ND_VM_CRL>;
    case ND_VM_FHN: return k_vertex_gen_jvp<ND_VM_FHN>;
    default: return k_vertex_jvp<ND_VM_NOOP, ND_EM_NOOP, ND_WRAP_NONE, false>;
    }
}

using EdgeOdeJvpKernel = void (*)(EDesc, const double*, const double*, const double*, const double*, const double*,
                                  double*, double*);
EdgeOdeJvpKernel pick_edge_ode_jvp(int em) {
    switch (em) {
    case ND_EM_ODE_SIN_DIFFUSION: return k_edge_ode_jvp<ND_EM_ODE_SIN_DIFFUSION>;
    case ND_EM_ODE_SIN_DIFFUSION_FAST: return k_edge_ode_jvp<ND_EM_ODE_SIN_DIFFUSION_FAST>;
    case ND_EM_ODE_LAG: return k_edge_ode_jvp<ND_EM_ODE_LAG>;
    case ND_EM_EXT_DRIVEN: return k_edge_ode_jvp<ND_EM_EXT_DRIVEN>;
    default: return k_edge_ode_jvp<ND_EM_ODE_LINEAR>;
    }
}

using JvpKKernel = void (*)(VDesc, const int32_t*, const int2*, const uint8_t*, const EDesc*, JvpKArgs);

template <int VM, int K>
JvpKKernel pick_jvpk_edge(int em, int wrap, bool multi) {
    if (multi) return k_vertex_jvpk<VM, ND_EM_NOOP, ND_WRAP_NONE, true, K>;
    const bool sym = wrap == ND_WRAP_SYMMETRIC;
    switch (em) {
    case ND_EM_KURAMOTO:
        return sym ? k_vertex_jvpk<VM, ND_EM_KURAMOTO, ND_WRAP_SYMMETRIC, false, K>
                   : k_vertex_jvpk<VM, ND_EM_KURAMOTO, ND_WRAP_ANTISYMMETRIC, false, K>;
    case ND_EM_STATIC_LINE:
        return sym ? k_vertex_jvpk<VM, ND_EM_STATIC_LINE, ND_WRAP_SYMMETRIC, false, K>
                   : k_vertex_jvpk<VM, ND_EM_STATIC_LINE, ND_WRAP_ANTISYMMETRIC, false, K>;
    case ND_EM_DIFFUSION:
        return sym ? k_vertex_jvpk<VM, ND_EM_DIFFUSION, ND_WRAP_SYMMETRIC, false, K>
                   : k_vertex_jvpk<VM, ND_EM_DIFFUSION, ND_WRAP_ANTISYMMETRIC, false, K>;
    case ND_EM_DIFFUSION_WEIGHTED:
        return sym ? k_vertex_jvpk<VM, ND_EM_DIFFUSION_WEIGHTED, ND_WRAP_SYMMETRIC, false, K>
                   : k_vertex_jvpk<VM, ND_EM_DIFFUSION_WEIGHTED, ND_WRAP_ANTISYMMETRIC, false, K>;
    case ND_EM_DIFFUSION_FIDUCIAL:
        return k_vertex_jvpk<VM, ND_EM_DIFFUSION_FIDUCIAL, ND_WRAP_ANTISYMMETRIC, false, K>;
    case ND_EM_ODE_SIN_DIFFUSION:
    case ND_EM_ODE_SIN_DIFFUSION_FAST:
    case ND_EM_ODE_LINEAR:
    case ND_EM_ODE_LAG:
    case ND_EM_EXT_DRIVEN:
        return wrap == ND_WRAP_ANTISYMMETRIC ? k_vertex_jvpk<VM, ND_EM_ODE_LINEAR, ND_WRAP_ANTISYMMETRIC, false, K>
                                             : k_vertex_jvpk<VM, ND_EM_ODE_LINEAR, ND_WRAP_FIDUCIAL, false, K>;
    default:
        return k_vertex_jvpk<VM, ND_EM_NOOP, ND_WRAP_NONE, false, K>;
    }
}

// the K-tangent vertex kernel of a batch, or null where the batch takes the generic or
// vector-output kernels (those have single-tangent forms only)
template <int K>
JvpKKernel pick_jvpk_kernel(int vm, int em, int wrap, bool multi) {
    switch (vm) {
    case ND_VM_DIFFUSION: return pick_jvpk_edge<ND_VM_DIFFUSION, K>(em, wrap, multi);
    case ND_VM_KURAMOTO: return pick_jvpk_edge<ND_VM_KURAMOTO, K>(em, wrap, multi);
    case ND_VM_SWING: return pick_jvpk_edge<ND_VM_SWING, K>(em, wrap, multi);
    case ND_VM_SWING_UNIT_DAMPING: return pick_jvpk_edge<ND_VM_SWING_UNIT_DAMPING, K>(em, wrap, multi);
    case ND_VM_LOAD: return pick_jvpk_edge<ND_VM_LOAD, K>(em, wrap, multi);
    case ND_VM_NOOP: return k_vertex_jvpk<ND_VM_NOOP, ND_EM_NOOP, ND_WRAP_NONE, false, K>;
    default: return nullptr;
    }
}

using EdgeOdeJvpKKernel = void (*)(EDesc, JvpKArgs, int);
template <int K>
EdgeOdeJvpKKernel pick_edge_ode_jvpk(int em) {
    switch (em) {
    case ND_EM_ODE_SIN_DIFFUSION: return k_edge_ode_jvpk<ND_EM_ODE_SIN_DIFFUSION, K>;
    case ND_EM_ODE_SIN_DIFFUSION_FAST: return k_edge_ode_jvpk<ND_EM_ODE_SIN_DIFFUSION_FAST, K>;
    case ND_EM_ODE_LAG: return k_edge_ode_jvpk<ND_EM_ODE_LAG, K>;
    case ND_EM_EXT_DRIVEN: return k_edge_ode_jvpk<ND_EM_EXT_DRIVEN, K>;
    default: return k_edge_ode_jvpk<ND_EM_ODE_LINEAR, K>;
    }
}

using EdgeOdeKernel = void (*)(EDesc, const double*, const double*, const double*, double*, StageArgs);
template <bool STAGE>
EdgeOdeKernel pick_edge_ode(int em) {
    switch (em) {
    case ND_EM_ODE_SIN_DIFFUSION: return k_edge_ode<ND_EM_ODE_SIN_DIFFUSION, STAGE>;
    case ND_EM_ODE_SIN_DIFFUSION_FAST: return k_edge_ode<ND_EM_ODE_SIN_DIFFUSION_FAST, STAGE>;
    case ND_EM_ODE_LAG: return k_edge_ode<ND_EM_ODE_LAG, STAGE>;
    case ND_EM_EXT_DRIVEN: return k_edge_ode<ND_EM_EXT_DRIVEN, STAGE>;
    default: return k_edge_ode<ND_EM_ODE_LINEAR, STAGE>;
    }
}

bool vertex_model_known(int vm) { return vm >= ND_VM_NOOP && vm <= ND_VM_TRUSS_FIXED; }
int VModelDim(int vm) {
    switch (vm) {
    case ND_VM_SWING:
    case ND_VM_SWING_UNIT_DAMPING:
    case ND_VM_PI_SOURCE:
    case ND_VM_CRL:
    case ND_VM_FHN: return 2;
    case ND_VM_DIFFUSION:
    case ND_VM_KURAMOTO:
    case ND_VM_LOAD:
    case ND_VM_CAPACITOR_SINK:
    case ND_VM_CAPACITOR:
    case ND_VM_INDUCTOR_INJECTOR: return 1;
    case ND_VM_TRUSS_FREE: return 4;
    case ND_VM_VOLTAGE_SOURCE:
    case ND_VM_RESISTOR_INJECTOR:
    case ND_VM_TRUSS_FIXED: return 0;
    default: return -1;
    }
}
// the output kind each vertex model has: StateMask, NoFeedForward or PureFeedForward
int vertex_model_ff(int vm) {
    if (vm == ND_VM_VOLTAGE_SOURCE || vm == ND_VM_TRUSS_FIXED) return ND_FF_NOFEEDFORWARD;
    if (vm == ND_VM_RESISTOR_INJECTOR) return ND_FF_PUREFEEDFORWARD;
    return ND_FF_PURESTATEMAP;
}
int vertex_ext_inputs(int vm) { return vm == ND_VM_PI_SOURCE ? VExt<ND_VM_PI_SOURCE>::n : 0; }
// vertex models only the one-thread-per-vertex kernels carry
bool vertex_model_generic(int vm) { return vm >= ND_VM_CAPACITOR_SINK && vm <= ND_VM_TRUSS_FIXED; }
bool edge_model_known(int em) { return em >= ND_EM_NOOP && em <= ND_EM_BEAM; }
bool edge_model_generic(int em) {
    return em == ND_EM_RESISTOR || em == ND_EM_LOOPBACK || em == ND_EM_ELECTRICAL || em == ND_EM_BEAM;
}
int edge_params(int em) {
    switch (em) {
    case ND_EM_KURAMOTO: return EModel<ND_EM_KURAMOTO>::np;
    case ND_EM_STATIC_LINE: return EModel<ND_EM_STATIC_LINE>::np;
    case ND_EM_DIFFUSION_WEIGHTED: return EModel<ND_EM_DIFFUSION_WEIGHTED>::np;
    case ND_EM_DIFFUSION_FIDUCIAL: return EModel<ND_EM_DIFFUSION_FIDUCIAL>::np;
    case ND_EM_RESISTOR: return EModel<ND_EM_RESISTOR>::np;
    default: return 0;
    }
}

template <typename T>
struct DevBuf {
    T* ptr = nullptr;
    size_t n = 0;
};

}  // namespace

struct nd_plan {
    int device = 0;
    int64_t dim = 0, pdim = 0, lastidx_out = 0, lastidx_aggr = 0, lastidx_gbuf = 0;
    int32_t vdepth = 0, edepth = 0;
    bool multi = false;
    int32_t hub_degree = 48;        // longer segments: one wave per vertex
    int32_t big_hub_degree = 128;   // longer segments: one block per vertex
    int32_t block = 256;
    std::vector<VDesc> vdesc;
    std::vector<EDesc> edesc;
    std::vector<VertexKernel> vkern, vkern_stage;
    GroupKernel gkern = nullptr, gkern_stage = nullptr;   // several fused-set batches per launch
    std::vector<JvpKernel> vkern_jvp;
    std::vector<JvpKKernel> vkern_jvpk4, vkern_jvpk8;   // null: no K-tangent form for the batch
    double* d_dvoutk = nullptr;                          // 8 tangent columns of the vertex outputs
    double* d_dvout = nullptr;     // compact tangent outputs (Jacobian-vector products)
    double* rk_work = nullptr;
    double* d_vout = nullptr;      // compact vertex outputs (vdepth 1 networks with functional edges)
    size_t vout_len = 0;
    double* d_vout2 = nullptr;     // second compact-output buffer (RK stages ping-pong)
    // networks small enough for one workgroup: RK4 runs in the single-instance ensemble
    // kernel (one launch for all steps) instead of four launches per step
    bool rk4_single = false;       // the network fits the single-workgroup kernel
    int32_t rk4_path = 1;          // 1 automatic, 2 per-stage launches
    std::vector<int64_t> small_src, small_dst;
    nd_kind small_vk{}, small_ek{};
    nd_ensemble* small_en = nullptr;      // RK4: 1024-thread workgroup
    nd_ensemble* small_en_ts = nullptr;   // Tsit5: the default size classes
    nd_ensemble_stat* d_small_stat = nullptr;
    std::vector<int32_t> rowptr_host;
    std::vector<int2> slots_host;
    std::vector<uint8_t> slot_eb_host;
    std::vector<void*> allocations;
    int32_t* d_rowptr = nullptr;
    int2* d_slots = nullptr;
    uint8_t* d_slot_eb = nullptr;
    EDesc* d_edesc = nullptr;
    int32_t* d_cols = nullptr;
    int32_t* d_gmap = nullptr;
    int32_t* d_hubs_all = nullptr;
    int32_t* d_agg_hub_rows = nullptr;
    int32_t* d_lrp = nullptr;
    int2* d_lslots = nullptr;
    int32_t* d_lnbr = nullptr;
    uint8_t* d_lseb = nullptr;
    int32_t n_agg_hub = 0;
    std::vector<size_t> hub_offset;
    // slot-ordered edge parameters (one feed-forward edge batch with parameters)
    bool kcache = false;           // in use: kcache_ok and the compact gather
    bool kcache_ok = false;        // the edge batch qualifies for the parameter copy
    // neighbour outputs: gathered from the compact array (rebuilt per call) or straight
    // from u; slot .x holds an index into one or the other
    bool use_vout = false;
    int32_t gather_mode = 1;       // 1 automatic, 2 direct from u, 3 compact
    std::vector<int32_t> slot_oidx, slot_uidx;
    int em0 = ND_EM_NOOP, wrap0 = ND_WRAP_NONE;
    int32_t knp = 0, kcount = 0, kparam0 = 0;
    std::vector<int32_t> slot_eq_host;   // per full-list entry: (edge position << 1) | is_src
    double* d_kc = nullptr;
    double* d_kfull = nullptr;
    double* d_klow = nullptr;
    int4* d_kpos = nullptr;
    // external inputs: u-index per external-buffer entry (vertex kernels), and the ExtMap
    // as +u / -o 1-based indices (collect_externals!)
    int64_t lastidx_extbuf = 0;
    int32_t* d_xmap = nullptr;
    int32_t* d_extmap = nullptr;
    // vertex outputs computed by a g (voltage sources, injectors): every call writes the
    // compact outputs, the injectors' after their hubs'
    bool computed_out = false, ff_out = false;
    int32_t* d_lbin = nullptr;
    // perturbed RHS: first gbuf index of every edge (edge batches in order, batch order
    // within), the four perturbation maps (buffer index, channel; 0-based) and work buffers
    std::vector<int32_t> egb_host;
    std::vector<int64_t> egb_off;
    int32_t* d_egb = nullptr;
    int2* d_pmap[4] = {nullptr, nullptr, nullptr, nullptr};   // vo, vi, eo, ei
    int32_t n_pmap[4] = {0, 0, 0, 0};
    int64_t ndelta = 0;
    double* d_pwork = nullptr;
    size_t pwork_len = 0;

    ~nd_plan() {
        if (small_en) nd_ensemble_destroy(small_en);
        if (small_en_ts) nd_ensemble_destroy(small_en_ts);
        for (void* p : allocations) (void)hipFree(p);
    }
};

namespace {

int hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return ND_OK;
    return nd::set_error(ND_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

template <typename T>
int upload(nd_plan* pl, const std::vector<T>& host, T** dev) {
    *dev = nullptr;
    if (host.empty()) return ND_OK;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, host.size() * sizeof(T));
    if (e != hipSuccess) return nd::set_error(ND_ERR_ALLOC, "hipMalloc(%zu): %s", host.size() * sizeof(T), hipGetErrorString(e));
    pl->allocations.push_back(p);
    e = hipMemcpy(p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_check(e, "hipMemcpy");
    *dev = static_cast<T*>(p);
    return ND_OK;
}

void release(nd_plan* pl, void* ptr) {
    if (!ptr) return;
    auto it = std::find(pl->allocations.begin(), pl->allocations.end(), ptr);
    if (it != pl->allocations.end()) { (void)hipFree(*it); pl->allocations.erase(it); }
}

// Hub lists and the hub-free destination lists follow hub_degree / big_hub_degree;
// rebuilt when they change.  Per vertex batch: big hubs then mid hubs, each by
// decreasing segment length; every other vertex keeps its segment in lrp/lslots.
int rebuild_hubs(nd_plan* pl) {
    const int32_t H = pl->hub_degree, HB = pl->big_hub_degree;
    std::vector<int32_t> all, lrp;
    std::vector<int2> lslots;
    std::vector<int32_t> lnbr;       // compact form of lslots when the parameters are cached
    size_t nlow = 0;
    std::vector<uint8_t> lseb;
    std::vector<size_t> lrp_offset(pl->vdesc.size(), 0);
    pl->hub_offset.assign(pl->vdesc.size() + 1, 0);
    std::vector<int4> kpos(pl->kcache ? pl->kcount : 0, make_int4(-1, -1, -1, -1));
    if (pl->kcache)
        for (size_t k = 0; k < pl->slot_eq_host.size(); ++k) {
            const int32_t eq = pl->slot_eq_host[k];
            (eq & 1 ? kpos[eq >> 1].x : kpos[eq >> 1].y) = static_cast<int32_t>(k);
        }
    for (size_t b = 0; b < pl->vdesc.size(); ++b) {
        VDesc& vd = pl->vdesc[b];
        pl->hub_offset[b] = all.size();
        lrp_offset[b] = lrp.size();
        std::vector<std::pair<int32_t, int32_t>> big, mid;   // (-length, q)
        const bool functional = vd.model != ND_VM_NOOP && vd.depth == 1;
        for (int32_t q = 0; q < vd.count; ++q) {
            lrp.push_back(static_cast<int32_t>(nlow));
            if (!functional) continue;
            const int32_t r = vd.row0 + q;
            const int32_t beg = pl->rowptr_host[r], len = pl->rowptr_host[r + 1] - beg;
            if (len > H) lrp.back() = static_cast<int32_t>(static_cast<uint32_t>(lrp.back()) | 0x80000000u);
            if (len > HB) big.emplace_back(-len, q);
            else if (len > H) mid.emplace_back(-len, q);
            else
                for (int32_t k = beg; k < beg + len; ++k) {
                    if (pl->kcache) {
                        const int32_t eq = pl->slot_eq_host[k];
                        (eq & 1 ? kpos[eq >> 1].z : kpos[eq >> 1].w) = static_cast<int32_t>(nlow);
                    }
                    if (pl->kcache) {
                        const int2 sl = pl->slots_host[k];
                        lnbr.push_back(static_cast<int32_t>(static_cast<uint32_t>(sl.x) | (static_cast<uint32_t>(sl.y & 1) << 31)));
                    } else {
                        lslots.push_back(pl->slots_host[k]);
                    }
                    ++nlow;
                    if (pl->multi) lseb.push_back(pl->slot_eb_host[k]);
                }
        }
        lrp.push_back(static_cast<int32_t>(nlow));
        std::sort(big.begin(), big.end());
        std::sort(mid.begin(), mid.end());
        for (auto& x : big) all.push_back(x.second);
        for (auto& x : mid) all.push_back(x.second);
        vd.nbig = static_cast<int32_t>(big.size());
        vd.nmid = static_cast<int32_t>(mid.size());
    }
    pl->hub_offset[pl->vdesc.size()] = all.size();
    std::vector<int32_t> rows;
    for (int64_t r = 0; r < pl->lastidx_aggr; ++r)
        if (pl->rowptr_host[r + 1] - pl->rowptr_host[r] > H) rows.push_back(static_cast<int32_t>(r));
    pl->n_agg_hub = static_cast<int32_t>(rows.size());
    release(pl, pl->d_hubs_all);
    release(pl, pl->d_agg_hub_rows);
    release(pl, pl->d_lrp);
    release(pl, pl->d_lslots);
    release(pl, pl->d_lnbr);
    pl->d_lnbr = nullptr;
    release(pl, pl->d_lseb);
    pl->d_hubs_all = nullptr;
    pl->d_agg_hub_rows = nullptr;
    pl->d_lrp = nullptr;
    pl->d_lslots = nullptr;
    pl->d_lseb = nullptr;
    int rc = upload(pl, all, &pl->d_hubs_all);
    rc = rc ? rc : upload(pl, rows, &pl->d_agg_hub_rows);
    rc = rc ? rc : upload(pl, lrp, &pl->d_lrp);
    rc = rc ? rc : upload(pl, lslots, &pl->d_lslots);
    rc = rc ? rc : upload(pl, lnbr, &pl->d_lnbr);
    rc = rc ? rc : upload(pl, lseb, &pl->d_lseb);
    if (rc) return rc;
    if (pl->kcache) {
        // the cached copy and both slot layouts restart from one NaN pattern, so they agree
        // entry for entry and the next call rewrites every edge's slots
        for (void* x : {static_cast<void*>(pl->d_kc), static_cast<void*>(pl->d_kfull), static_cast<void*>(pl->d_klow),
                        static_cast<void*>(pl->d_kpos)})
            release(pl, x);
        pl->d_kc = pl->d_kfull = pl->d_klow = nullptr;
        pl->d_kpos = nullptr;
        const double nan = std::numeric_limits<double>::quiet_NaN();
        const size_t np = static_cast<size_t>(pl->knp);
        rc = upload(pl, std::vector<double>(static_cast<size_t>(pl->kcount) * np, nan), &pl->d_kc);
        rc = rc ? rc : upload(pl, std::vector<double>(pl->slot_eq_host.size() * np, nan), &pl->d_kfull);
        rc = rc ? rc : upload(pl, std::vector<double>(std::max<size_t>(nlow, 1) * np, nan), &pl->d_klow);
        rc = rc ? rc : upload(pl, kpos, &pl->d_kpos);
        if (rc) return rc;
    }
    for (size_t b = 0; b < pl->vdesc.size(); ++b) {
        pl->vdesc[b].hubs = pl->d_hubs_all ? pl->d_hubs_all + pl->hub_offset[b] : nullptr;
        pl->vdesc[b].lrp = pl->d_lrp + lrp_offset[b];
        pl->vdesc[b].lslots = pl->d_lslots;
        pl->vdesc[b].lnbr = pl->d_lnbr;
        pl->vdesc[b].lseb = pl->d_lseb;
        pl->vdesc[b].kfull = pl->d_kfull;
        pl->vdesc[b].klow = pl->d_klow;
    }
    return ND_OK;
}

// Chooses how the edge phase reads neighbour outputs and rebuilds what depends on it:
// the slot indices, the hub-free lists, the parameter copy and the kernel variants.
// Automatic: the compact array (one extra launch per call, 8 B per neighbour read)
// once the network is large enough that its states do not stay in L2; below that the
// vertex kernel reads u directly and one RHS is one launch per vertex batch.
constexpr int64_t kCompactMinVertices = 1 << 18;
int apply_gather_mode(nd_plan* pl) {
    int64_t nv = 0;
    for (const VDesc& vd : pl->vdesc) nv += vd.count;
    const bool possible = pl->d_vout != nullptr;
    pl->use_vout = possible && (pl->gather_mode == 3 || (pl->gather_mode == 1 && nv >= kCompactMinVertices) ||
                                pl->computed_out);   // computed outputs exist only in the compact array
    pl->kcache = pl->kcache_ok && pl->use_vout;
    const std::vector<int32_t>& x = pl->use_vout ? pl->slot_oidx : pl->slot_uidx;
    for (size_t k = 0; k < pl->slots_host.size(); ++k) pl->slots_host[k].x = x[k];
    release(pl, pl->d_slots);
    pl->d_slots = nullptr;
    int rc = upload(pl, pl->slots_host, &pl->d_slots);
    if (rc) return rc;
    pl->vkern.clear();
    pl->vkern_stage.clear();
    const int em = pl->edesc.empty() ? ND_EM_NOOP : pl->em0;
    pl->gkern = pick_group_kernel<false>(em, pl->wrap0, pl->multi, pl->kcache);
    pl->gkern_stage = pick_group_kernel<true>(em, pl->wrap0, pl->multi, pl->kcache);
    for (const VDesc& vd : pl->vdesc) {
        pl->vkern.push_back(vd.depth > 1 ? k_vertex_vec<false>
                            : vd.generic ? pick_gen_kernel<false>(vd.model)
                                         : pick_vertex_kernel<false>(vd.model, em, pl->wrap0, pl->multi, pl->kcache));
        pl->vkern_stage.push_back(vd.depth > 1 ? k_vertex_vec<true>
                                  : vd.generic ? pick_gen_kernel<true>(vd.model)
                                               : pick_vertex_kernel<true>(vd.model, em, pl->wrap0, pl->multi, pl->kcache));
    }
    return rebuild_hubs(pl);
}

int check_device(const nd_plan* pl) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_check(e, "hipGetDevice");
    if (dev != pl->device) return nd::set_error(ND_ERR_ARGUMENT, "plan lives on device %d, current device is %d", pl->device, dev);
    return ND_OK;
}

inline unsigned blocks_for(int64_t n, int block) { return static_cast<unsigned>((n + block - 1) / block); }

// grid of the grid-stride parts of k_vertex_out: two elements per thread, at most 2048
// blocks of 256 per part
unsigned vout_blocks(int64_t count) {
    const int64_t b = (count + 511) / 512;
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(b, 2048)));
}

KSync no_sync(unsigned vblocks) {
    KSync k{};
    k.vblocks = static_cast<int32_t>(vblocks);
    return k;
}

// One right-hand side (stage = false: du = f(u)) or one fused RK stage (stage = true).
// sync: compare p's edge parameters with the slot-ordered copy first (skipped only
// where p is known to be unchanged since the previous launch of this plan).
// vin: the compact vertex outputs of u (default: the plan's buffer, built here first);
// build = false when a previous stage epilogue already wrote them.
void launch_rhs(nd_plan* pl, bool stage, double* du, const double* u, const double* p, const StageArgs& sa,
                hipStream_t s, double t, bool sync = true, double* vin = nullptr, bool build = true) {
    const int B = pl->block;
    if (!pl->use_vout) {
        vin = const_cast<double*>(u);    // direct gather: slot indices address u itself
        build = false;
    } else if (!vin) {
        vin = pl->d_vout;
    }
    // vertex outputs o_v = u[StateMask] into a compact array (coreloop.jl:39): the edge
    // evaluations gather neighbour outputs from it instead of from the interleaved state
    if (pl->use_vout && build) {
        // every batch's outputs, up to kGroupMax batches per launch; the first launch also
        // brings the slot-ordered parameter copy up to date
        bool pending = sync && pl->kcache;
        VGroup g{};
        auto flush = [&]() {
            if (!g.n) return;
            KSync ks = no_sync(static_cast<unsigned>(g.first[g.n]));
            unsigned nb = static_cast<unsigned>(g.first[g.n]);
            if (pending) {
                ks.pe = p + pl->kparam0;
                ks.kc = pl->d_kc;
                ks.kfull = pl->d_kfull;
                ks.klow = pl->d_klow;
                ks.kpos = pl->d_kpos;
                ks.count = pl->kcount;
                ks.np = pl->knp;
                nb += vout_blocks(pl->kcount);
                pending = false;
            }
            if (g.n == 1) hipLaunchKernelGGL(k_vertex_out, dim3(nb), dim3(256), 0, s, g.vd[0], u, vin, ks, p, 0);
            else hipLaunchKernelGGL(k_vertex_out_group, dim3(nb), dim3(256), 0, s, g, u, vin, ks, p, 0);
            g.n = 0;
        };
        for (const VDesc& vd : pl->vdesc) {
            if (!vd.count) continue;
            if (g.n == 0) g.first[0] = 0;
            g.vd[g.n] = vd;
            g.first[g.n + 1] = g.first[g.n] + static_cast<int32_t>(vout_blocks(vd.count));
            if (++g.n == kGroupMax) flush();
        }
        flush();
        if (pl->ff_out)
            for (const VDesc& vd : pl->vdesc)
                if (vd.outkind == 2 && vd.count)
                    hipLaunchKernelGGL(k_vertex_out_ff, dim3(blocks_for(vd.count, 256)), dim3(256), 0, s, vd, vin, p);
    }
    // fused-set batches go out together, up to kGroupMax per launch; a lone one (and every
    // per-vertex-kernel batch) takes its own kernel
    VGroup g{};
    size_t gb[kGroupMax] = {};
    auto single = [&](size_t b, const VDesc& vd, int64_t nblocks) {
        hipLaunchKernelGGL(stage ? pl->vkern_stage[b] : pl->vkern[b], dim3(static_cast<unsigned>(nblocks)),
                           dim3(static_cast<unsigned>(B)), 0, s, vd, pl->d_rowptr, pl->d_slots, pl->d_slot_eb,
                           pl->d_edesc, u, vin, p, du, sa);
    };
    auto flush = [&]() {
        if (g.n == 1) single(gb[0], g.vd[0], g.first[1]);
        else if (g.n > 1)
            hipLaunchKernelGGL(stage ? pl->gkern_stage : pl->gkern, dim3(static_cast<unsigned>(g.first[g.n])),
                               dim3(static_cast<unsigned>(B)), 0, s, g, pl->d_rowptr, pl->d_slots, pl->d_slot_eb,
                               pl->d_edesc, u, vin, p, du, sa);
        g.n = 0;
    };
    for (size_t b = 0; b < pl->vdesc.size(); ++b) {
        VDesc vd = pl->vdesc[b];
        if (vd.count == 0) continue;
        vd.t = t;
        const int64_t nblocks = vd.nbig + blocks_for(static_cast<int64_t>(vd.nmid) * kWave, B) + blocks_for(vd.count, B);
        if (vd.generic || !pl->gkern) {
            single(b, vd, nblocks);
            continue;
        }
        if (g.n == 0) g.first[0] = 0;
        g.vd[g.n] = vd;
        gb[g.n] = b;
        g.first[g.n + 1] = g.first[g.n] + static_cast<int32_t>(nblocks);
        if (++g.n == kGroupMax) flush();
    }
    flush();
    for (const EDesc& e : pl->edesc) {
        if (!e.ode || e.count == 0) continue;
        hipLaunchKernelGGL(stage ? pick_edge_ode<true>(e.model) : pick_edge_ode<false>(e.model),
                           dim3(blocks_for(e.count, B)), dim3(static_cast<unsigned>(B)), 0, s, e, u,
                           pl->computed_out ? vin : u, p, du, sa);
    }
}

}  // namespace

// Scratch of nd_rk4_noalloc, the allocation-free RK4: the three stage vectors and the
// second compact-output buffer of one call, in one file-scope device array.  Every GPU the
// library runs on holds its own copy until the process ends; calls on one GPU share it
// (and race if they overlap).  Host code never passes the array's name to a kernel: its
// device address comes from hipGetSymbolAddress, looked up once per device with that
// device current, and kept in noalloc_ptr.
constexpr int64_t kNoAllocScratch = int64_t(1) << 23;   // doubles (64 MiB)
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) double scratch[kNoAllocScratch];
static std::mutex noalloc_mu;
static double* noalloc_ptr[kNoAllocDevices] = {};

static int noalloc_scratch(int dev, double** out) {
    if (dev < 0 || dev >= kNoAllocDevices)
        return nd::set_error(ND_ERR_UNSUPPORTED, "nd_rk4_noalloc: device %d beyond the %d it keeps scratch for", dev,
                             kNoAllocDevices);
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (!noalloc_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return nd::set_error(ND_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        noalloc_ptr[dev] = static_cast<double*>(ptr);
    }
    *out = noalloc_ptr[dev];
    return ND_OK;
}

extern "C" {

int nd_plan_create(const nd_network* h, nd_plan** out) {
    nd::clear_error();
    if (!h || !out) return nd::set_error(ND_ERR_ARGUMENT, "NULL argument");
    *out = nullptr;
    const nd::Network& n = h->net;
    const int64_t lim = std::numeric_limits<int32_t>::max() / 2;
    if (n.dim > lim || n.pdim > lim || n.lastidx_out > lim || static_cast<int64_t>(n.csr_cols.size()) > lim ||
        n.lastidx_gbuf > lim)
        return nd::set_error(ND_ERR_UNSUPPORTED, "network too large for 32-bit device indices");

    bool functional_vertex = false, functional_edge = false;
    for (const nd::Batch& b : n.vb) {
        const nd_kind& k = n.kinds[b.kind];
        if (!vertex_model_known(k.model)) return nd::set_error(ND_ERR_UNSUPPORTED, "vertex model %d has no HIP kernel", k.model);
        if (k.ff != vertex_model_ff(k.model))
            return nd::set_error(ND_ERR_UNSUPPORTED, "vertex model %d: this output form has no HIP kernel", k.model);
        if (k.model != ND_VM_NOOP) {
            functional_vertex = true;
            // with d > 1 output components, diffusion vertices carry d states (dv .= esum)
            if (k.dim != VModelDim(k.model) && !(n.vdepth > 1 && k.model == ND_VM_DIFFUSION && k.dim == n.vdepth))
                return nd::set_error(ND_ERR_ARGUMENT, "vertex model %d needs dim %d", k.model, VModelDim(k.model));
            if (k.extdim != vertex_ext_inputs(k.model))
                return nd::set_error(ND_ERR_ARGUMENT, "vertex model %d reads %d external inputs, its kind declares %d",
                                     k.model, vertex_ext_inputs(k.model), k.extdim);
        }
    }
    bool generic_edges = false;
    for (const nd::Batch& b : n.eb) {
        const nd_kind& k = n.kinds[b.kind];
        if (!edge_model_known(k.model)) return nd::set_error(ND_ERR_UNSUPPORTED, "edge model %d has no HIP kernel", k.model);
        if (k.model != ND_EM_NOOP) functional_edge = true;
        if (edge_model_generic(k.model)) generic_edges = true;
        if (k.extdim != edge_ext_inputs(k.model))
            return nd::set_error(ND_ERR_ARGUMENT, "edge model %d reads %d external inputs, its kind declares %d", k.model,
                                 edge_ext_inputs(k.model), k.extdim);
        if (k.model == ND_EM_LOOPBACK && (k.wrapper != ND_WRAP_DIRECTED || k.ff != ND_FF_PUREFEEDFORWARD))
            return nd::set_error(ND_ERR_ARGUMENT, "a LoopbackConnection is a Directed feed-forward edge");
        if (k.ff == ND_FF_PURESTATEMAP && k.wrapper != ND_WRAP_FIDUCIAL) {
            // a wrapped StateMask: one mask, both sides (AntiSymmetric, Symmetric) or dst only
            const int ns = k.wrapper == ND_WRAP_DIRECTED ? 0 : k.outdim_dst;
            bool same = k.outdim_src == ns;
            for (int j = 0; same && j < ns; ++j) same = k.mask_src[j] == k.mask_dst[j];
            if (!same || (k.wrapper != ND_WRAP_ANTISYMMETRIC && k.wrapper != ND_WRAP_SYMMETRIC &&
                          k.wrapper != ND_WRAP_DIRECTED))
                return nd::set_error(ND_ERR_ARGUMENT, "a wrapped StateMask edge output has one mask for both sides");
        }
        if (k.model == ND_EM_NOOP && k.dim > 0)
            return nd::set_error(ND_ERR_UNSUPPORTED, "edge kinds with states but no state function have no HIP kernel");
        if (is_ode_model(k.model)) {
            if (k.ff != ND_FF_PURESTATEMAP || k.dim != ode_dim(k.model))
                return nd::set_error(ND_ERR_UNSUPPORTED, "ODE edge model %d needs dim %d and a StateMask output", k.model,
                                     ode_dim(k.model));
        } else if (k.model != ND_EM_NOOP) {
            if (k.ff != ND_FF_PUREFEEDFORWARD)
                return nd::set_error(ND_ERR_UNSUPPORTED, "edge model %d needs a pure feed-forward output", k.model);
            if (k.dim != 0) return nd::set_error(ND_ERR_UNSUPPORTED, "edge model %d has no state", k.model);
        }
    }
    if ((functional_vertex || functional_edge) && (n.vdepth != 1 || (n.ne > 0 && n.edepth != 1))) {
        // d output components per vertex and per edge side: componentwise diffusion
        // (vertices dv .= esum, edges e .= v_s .- v_d) and, at d = 2, the truss's free and
        // fixed nodes and beams; every vertex output a StateMask of d consecutive states or,
        // for fixed nodes, the parameters
        const int d = n.vdepth;
        bool ok = d >= 1 && d <= ND_MAX_OUT && (n.ne == 0 || n.edepth == d);
        for (const nd::Batch& b : n.vb) {
            const nd_kind& k = n.kinds[b.kind];
            const bool fixed = k.model == ND_VM_TRUSS_FIXED && d == 2;
            ok = ok && (k.model == ND_VM_NOOP || k.model == ND_VM_DIFFUSION || (k.model == ND_VM_TRUSS_FREE && d == 2) ||
                        fixed) && k.outdim_dst == d && k.extdim == 0 && (fixed || k.ff == ND_FF_PURESTATEMAP);
            for (int j = 1; ok && !fixed && j < d; ++j) ok = k.mask_dst[j] == k.mask_dst[0] + j;
        }
        for (const nd::Batch& b : n.eb) {
            const nd_kind& k = n.kinds[b.kind];
            ok = ok && (k.model == ND_EM_NOOP || k.model == ND_EM_DIFFUSION || (k.model == ND_EM_BEAM && d == 2)) &&
                 k.outdim_dst == d && k.dim == 0;
        }
        if (!ok)
            return nd::set_error(ND_ERR_UNSUPPORTED, "output depth %d (vdepth=%d, edepth=%d): only componentwise models "
                                 "(diffusion vertices and edges over consecutive states) have a HIP kernel",
                                 d, n.vdepth, n.edepth);
    }
    if (n.lastidx_extbuf > 0 && static_cast<int64_t>(n.extmap.size()) != n.lastidx_extbuf)
        return nd::set_error(ND_ERR_ARGUMENT, "the network has external inputs: set its ExtMap (nd_network_set_extmap) first");
    bool computed_out = false;
    for (const nd::Batch& b : n.vb) computed_out = computed_out || n.kinds[b.kind].ff != ND_FF_PURESTATEMAP;

    nd_plan* pl = new (std::nothrow) nd_plan;
    if (!pl) return nd::set_error(ND_ERR_ALLOC, "host allocation failed");
    hipError_t he = hipGetDevice(&pl->device);
    if (he != hipSuccess) { delete pl; return hip_check(he, "hipGetDevice"); }
    pl->dim = n.dim; pl->pdim = n.pdim; pl->lastidx_out = n.lastidx_out;
    pl->lastidx_aggr = n.lastidx_aggr; pl->lastidx_gbuf = n.lastidx_gbuf;
    pl->vdepth = n.vdepth; pl->edepth = n.edepth;
    pl->multi = n.eb.size() > 1;

    // edge batch descriptors + per-edge endpoint output u-indices (batch order)
    std::vector<std::vector<int2>> ends(n.eb.size());
    for (size_t b = 0; b < n.eb.size(); ++b) {
        const nd::Batch& B = n.eb[b];
        const nd_kind& k = n.kinds[B.kind];
        EDesc e{};
        e.model = k.model;
        e.wrap = k.model == ND_EM_DIFFUSION_FIDUCIAL ? ND_WRAP_ANTISYMMETRIC : k.wrapper;
        e.ode = is_ode_model(k.model) ? 1 : 0;
        e.dim = k.dim; e.pdim = k.pdim;
        e.state0 = static_cast<int32_t>(B.state_first - 1);
        e.param0 = static_cast<int32_t>(B.param_first - 1);
        e.count = static_cast<int32_t>(B.indices.size());
        e.out0 = static_cast<int32_t>(B.out_first - 1);
        e.outdim_src = k.outdim_src; e.outdim_dst = k.outdim_dst;
        for (int j = 0; j < ND_MAX_OUT; ++j) { e.mask_src[j] = k.mask_src[j] - 1; e.mask_dst[j] = k.mask_dst[j] - 1; }
        e.ext0 = static_cast<int32_t>(B.ext_first - 1);
        e.extdim = k.extdim;
        if (n.vdepth >= 1) {
            // endpoint outputs (component 0; further components follow it) as u-indices
            // (StateMask outputs), or as o-indices where outputs are computed (get_buffers
            // then reads them from o)
            ends[b].resize(B.indices.size());
            for (size_t q = 0; q < B.indices.size(); ++q) {
                const int64_t ei = B.indices[q] - 1;
                const int64_t s0 = n.src[ei] - 1, d0 = n.dst[ei] - 1;
                ends[b][q] = computed_out
                    ? make_int2(static_cast<int32_t>(n.v_out_first(s0) - 1), static_cast<int32_t>(n.v_out_first(d0) - 1))
                    : make_int2(static_cast<int32_t>(nd::vertex_out_u_index(n, s0, 0)),
                                static_cast<int32_t>(nd::vertex_out_u_index(n, d0, 0)));
            }
        }
        pl->edesc.push_back(e);
    }

    // CSR rows (aggregation indices) and the per-entry slots of the fused kernel
    const size_t nnz = n.csr_cols.size();
    pl->rowptr_host.resize(static_cast<size_t>(n.lastidx_aggr) + 1);
    for (size_t r = 0; r < pl->rowptr_host.size(); ++r) pl->rowptr_host[r] = static_cast<int32_t>(n.csr_rowptr[r]);
    // Without edges (edepth 0) there are no aggregation rows; vertex kernels then read
    // empty segments from a zero row table spanning the largest vertex batch.
    bool no_rows = n.edepth == 0;
    if (no_rows) {
        size_t mx = 0;
        for (const nd::Batch& B : n.vb) mx = std::max(mx, B.indices.size());
        pl->rowptr_host.assign(mx + 1, 0);
    }
    std::vector<int32_t> cols(nnz);
    std::vector<int2> slots(nnz);
    std::vector<uint8_t> slot_eb(pl->multi ? nnz : 0);
    std::vector<int32_t> slot_eq(nnz);
    std::vector<int32_t> uidx(nnz, -1);   // u-index of the neighbour output (direct gather)
    if (pl->multi && n.eb.size() > 255) { delete pl; return nd::set_error(ND_ERR_UNSUPPORTED, "more than 255 edge batches"); }
    {
        // output index -> (edge batch, position, side): batches hold ascending output blocks
        std::vector<int64_t> bfirst(n.eb.size());
        for (size_t b = 0; b < n.eb.size(); ++b) bfirst[b] = n.eb[b].out_first;
        for (int64_t r = 0; r < n.lastidx_aggr; ++r) {
            for (int64_t s = n.csr_rowptr[r]; s < n.csr_rowptr[r + 1]; ++s) {
                const int64_t k = n.csr_cols[s];
                cols[s] = static_cast<int32_t>(k - 1);
                const size_t b = static_cast<size_t>(std::upper_bound(bfirst.begin(), bfirst.end(), k) - bfirst.begin()) - 1;
                const nd::Batch& B = n.eb[b];
                const nd_kind& kk = n.kinds[B.kind];
                const int64_t stride = kk.outdim_src + kk.outdim_dst;
                const int64_t q = (k - B.out_first) / stride, rem = (k - B.out_first) % stride;
                const bool is_src = rem < kk.outdim_src;
                const int j = static_cast<int>(is_src ? rem : rem - kk.outdim_src);
                const int64_t e = B.indices[q] - 1;
                int2 sl = make_int2(0, 0);
                if (is_ode_model(kk.model)) {
                    sl.x = static_cast<int32_t>(n.e_state_first(e) - 1 + (is_src ? kk.mask_src[j] : kk.mask_dst[j]) - 1);
                    sl.y = is_src ? 1 : 0;
                } else if (kk.model != ND_EM_NOOP) {
                    const int64_t other = is_src ? n.dst[e] - 1 : n.src[e] - 1;
                    // component j of the other endpoint's output (j = 0 at output depth 1)
                    sl.x = static_cast<int32_t>(n.v_out_first(other) - 1 + j);
                    sl.y = static_cast<int32_t>(((n.e_param_first(e) - 1) << 1) | (is_src ? 1 : 0));
                    uidx[s] = static_cast<int32_t>(nd::vertex_out_u_index(n, other, j));
                }
                if (uidx[s] < 0) uidx[s] = sl.x;
                slots[s] = sl;
                if (pl->multi) slot_eb[s] = static_cast<uint8_t>(b);
                slot_eq[s] = static_cast<int32_t>((q << 1) | (is_src ? 1 : 0));
            }
        }
    }
    for (const nd::Batch& B : n.eb) {
        pl->egb_off.push_back(static_cast<int64_t>(pl->egb_host.size()));
        for (int64_t ei : B.indices) pl->egb_host.push_back(static_cast<int32_t>(n.e_gbuf_first(ei - 1) - 1));
    }
    std::vector<int32_t> gmap(static_cast<size_t>(n.lastidx_gbuf));
    for (int64_t e = 0; e < n.ne; ++e) {
        const int64_t g = n.e_gbuf_first(e) - 1;
        const int64_t os = n.v_out_first(n.src[e] - 1) - 1, od = n.v_out_first(n.dst[e] - 1) - 1;
        for (int j = 0; j < n.vdepth; ++j) {
            gmap[static_cast<size_t>(g + j)] = static_cast<int32_t>(os + j);
            gmap[static_cast<size_t>(g + n.vdepth + j)] = static_cast<int32_t>(od + j);
        }
    }

    if (n.vb.size() == 1 && n.eb.size() == 1 && n.vdepth == 1 && n.nv <= 2048 && 2 * n.ne <= 65535 && n.ne > 0) {
        const nd_kind& vk = n.kinds[n.vb[0].kind];
        const nd_kind& ek = n.kinds[n.eb[0].kind];
        const bool vok = (vk.model == ND_VM_KURAMOTO || vk.model == ND_VM_SWING) && vk.ff == ND_FF_PURESTATEMAP &&
                         vk.outdim_dst == 1 && vk.mask_dst[0] == 1;
        const bool eok = (ek.model == ND_EM_KURAMOTO || ek.model == ND_EM_STATIC_LINE) &&
                         ek.wrapper == ND_WRAP_ANTISYMMETRIC && ek.dim == 0 && ek.outdim_dst == 1;
        if (vok && eok) {
            pl->rk4_single = true;
            pl->small_src.assign(n.src.begin(), n.src.end());
            pl->small_dst.assign(n.dst.begin(), n.dst.end());
            pl->small_vk = vk;
            pl->small_ek = ek;
        }
    }
    int rc = ND_OK;
    rc = rc ? rc : upload(pl, pl->rowptr_host, &pl->d_rowptr);
    rc = rc ? rc : upload(pl, slot_eb, &pl->d_slot_eb);
    pl->slot_oidx.resize(nnz);
    for (size_t k = 0; k < nnz; ++k) pl->slot_oidx[k] = slots[k].x;
    pl->slot_uidx = std::move(uidx);
    pl->slots_host = std::move(slots);
    pl->slot_eb_host = std::move(slot_eb);
    if (n.eb.size() == 1 && n.vdepth == 1 && n.ne > 0) {
        const nd_kind& k = n.kinds[n.eb[0].kind];
        const int np = edge_params(k.model);
        if (np > 0 && !is_ode_model(k.model) && k.pdim == np && !no_rows) {
            pl->kcache_ok = true;
            pl->knp = np;
            pl->kcount = pl->edesc[0].count;
            pl->kparam0 = pl->edesc[0].param0;
            pl->slot_eq_host = std::move(slot_eq);
        }
    }
    rc = rc ? rc : upload(pl, cols, &pl->d_cols);
    rc = rc ? rc : upload(pl, gmap, &pl->d_gmap);
    for (size_t b = 0; b < n.eb.size() && !rc; ++b) {
        int2* d = nullptr;
        rc = upload(pl, ends[b], &d);
        pl->edesc[b].ends = d;
    }
    rc = rc ? rc : upload(pl, pl->edesc, &pl->d_edesc);
    if (rc) { delete pl; return rc; }

    // vertex batch descriptors + kernels
    const int em0 = n.eb.empty() ? ND_EM_NOOP : n.kinds[n.eb[0].kind].model;
    const int wrap0 = n.eb.empty() ? ND_WRAP_NONE : pl->edesc[0].wrap;
    pl->em0 = em0;
    pl->wrap0 = wrap0;
    for (const nd::Batch& B : n.vb) {
        const nd_kind& k = n.kinds[B.kind];
        VDesc vd{};
        vd.model = k.model; vd.dim = k.dim; vd.pdim = k.pdim; vd.depth = k.outdim_dst;
        vd.state0 = static_cast<int32_t>(B.state_first - 1);
        vd.param0 = static_cast<int32_t>(B.param_first - 1);
        vd.row0 = no_rows ? 0 : static_cast<int32_t>(B.in_first - 1);
        vd.count = static_cast<int32_t>(B.indices.size());
        vd.out0 = static_cast<int32_t>(B.out_first - 1);
        for (int j = 0; j < ND_MAX_OUT; ++j) vd.mask[j] = k.mask_dst[j] - 1;
        vd.generic = (k.model != ND_VM_NOOP && (vertex_model_generic(k.model) || generic_edges)) || n.vdepth > 1 ? 1 : 0;
        vd.ext0 = static_cast<int32_t>(B.ext_first - 1);
        vd.extdim = k.extdim;
        vd.outkind = k.ff == ND_FF_NOFEEDFORWARD ? 1 : k.ff == ND_FF_PUREFEEDFORWARD ? 2 : 0;
        pl->vdesc.push_back(vd);
        pl->vkern_jvp.push_back(vd.depth > 1 ? k_vertex_vec_jvp : vd.generic ? pick_gen_jvp(k.model)
                                           : pick_jvp_kernel(k.model, n.eb.empty() ? ND_EM_NOOP : em0, wrap0, pl->multi));
        const bool kform = vd.depth <= 1 && !vd.generic;
        pl->vkern_jvpk4.push_back(kform ? pick_jvpk_kernel<4>(k.model, n.eb.empty() ? ND_EM_NOOP : em0, wrap0, pl->multi)
                                        : nullptr);
        pl->vkern_jvpk8.push_back(kform ? pick_jvpk_kernel<8>(k.model, n.eb.empty() ? ND_EM_NOOP : em0, wrap0, pl->multi)
                                        : nullptr);
    }
    // loopback: each injector's input is its hub's output (gen_loopback_map, post_utils.jl:205-217)
    pl->computed_out = computed_out;
    {
        std::vector<int32_t> lbin;
        bool any = false;
        for (size_t b = 0; b < n.vb.size(); ++b) {
            pl->vdesc[b].lb0 = static_cast<int32_t>(lbin.size());
            lbin.insert(lbin.end(), n.vb[b].indices.size(), -1);
            if (pl->vdesc[b].outkind == 2) pl->ff_out = true;
        }
        for (int64_t e = 0; e < n.ne; ++e) {
            if (n.kinds[n.ekind[e]].model != ND_EM_LOOPBACK) continue;
            const int64_t inj = n.src[e] - 1, hub = n.dst[e] - 1;
            lbin[static_cast<size_t>(pl->vdesc[n.v_batch[inj]].lb0 + n.v_pos[inj])] = static_cast<int32_t>(n.v_out_first(hub) - 1);
            any = true;
        }
        if (any) {
            rc = upload(pl, lbin, &pl->d_lbin);
            if (rc) { delete pl; return rc; }
            for (VDesc& vd : pl->vdesc) vd.lbin = pl->d_lbin;
        }
        for (const VDesc& vd : pl->vdesc)
            if (vd.outkind == 2 && !any) { delete pl; return nd::set_error(ND_ERR_ARGUMENT, "injector outputs need a LoopbackConnection"); }
    }
    // external inputs: every entry resolves to a state (StateBufIdx, or a StateMask output)
    pl->lastidx_extbuf = n.lastidx_extbuf;
    if (n.lastidx_extbuf > 0) {
        std::vector<int32_t> xmap(static_cast<size_t>(n.lastidx_extbuf)), emap(static_cast<size_t>(n.lastidx_extbuf));
        for (int64_t k = 0; k < n.lastidx_extbuf; ++k) {
            const int64_t m = n.extmap[k];
            bool neg = false;
            const int64_t xi = m > 0 ? m - 1 : nd::output_state_index(n, -m, &neg);
            xmap[k] = neg ? ~static_cast<int32_t>(xi) : static_cast<int32_t>(xi);
            emap[k] = static_cast<int32_t>(m);
        }
        rc = upload(pl, xmap, &pl->d_xmap);
        rc = rc ? rc : upload(pl, emap, &pl->d_extmap);
        if (rc) { delete pl; return rc; }
        for (VDesc& vd : pl->vdesc) vd.xmap = pl->d_xmap;
        bool edge_ext = false;
        for (EDesc& e : pl->edesc) {
            e.xmap = pl->d_xmap;
            edge_ext = edge_ext || e.extdim > 0;
        }
        if (edge_ext) {                          // the device copy of the edge descriptors too
            release(pl, pl->d_edesc);
            pl->d_edesc = nullptr;
            rc = upload(pl, pl->edesc, &pl->d_edesc);
            if (rc) { delete pl; return rc; }
        }
    }
    if (n.ne > 0) {                              // compact vertex outputs: nv * vdepth
        void* w = nullptr;
        hipError_t e = hipMalloc(&w, static_cast<size_t>(n.nv) * n.vdepth * sizeof(double));
        if (e != hipSuccess) { delete pl; return nd::set_error(ND_ERR_ALLOC, "hipMalloc(vertex outputs): %s", hipGetErrorString(e)); }
        pl->allocations.push_back(w);
        pl->d_vout = static_cast<double*>(w);
        pl->vout_len = static_cast<size_t>(n.nv) * n.vdepth;
    }
    rc = apply_gather_mode(pl);
    if (rc) { delete pl; return rc; }
    *out = pl;
    return ND_OK;
}

void nd_plan_destroy(nd_plan* plan) { delete plan; }

int nd_plan_tune(nd_plan* pl, const nd_plan_tuning* t) {
    if (!pl || !t) return nd::set_error(ND_ERR_ARGUMENT, "NULL argument");
    // re-tuning frees and re-uploads device arrays: let work already queued on any
    // stream finish with the old ones first
    int src = check_device(pl);
    if (src) return src;
    src = hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (src) return src;
    if (t->block_size) {
        if (t->block_size < kWave || t->block_size > 256 || t->block_size % kWave)
            return nd::set_error(ND_ERR_ARGUMENT, "block_size must be a multiple of 64 in [64, 256]");
        pl->block = t->block_size;
    }
    if (t->hub_degree < 0 || t->big_hub_degree < 0)
        return nd::set_error(ND_ERR_ARGUMENT, "hub degrees must be positive");
    if (t->rk4_path < 0 || t->rk4_path > 2) return nd::set_error(ND_ERR_ARGUMENT, "rk4_path must be 0, 1 or 2");
    if (t->rk4_path) pl->rk4_path = t->rk4_path;
    if (t->gather_mode < 0 || t->gather_mode > 3) return nd::set_error(ND_ERR_ARGUMENT, "gather_mode must be 0..3");
    if (t->gather_mode && t->gather_mode != pl->gather_mode) {
        pl->gather_mode = t->gather_mode;
        int rc = apply_gather_mode(pl);
        if (rc) return rc;
    }
    if (t->hub_degree || t->big_hub_degree) {
        if (t->hub_degree) pl->hub_degree = t->hub_degree;
        if (t->big_hub_degree) pl->big_hub_degree = t->big_hub_degree;
        if (pl->big_hub_degree < pl->hub_degree) pl->big_hub_degree = pl->hub_degree;
        return rebuild_hubs(pl);
    }
    return ND_OK;
}

int nd_plan_get_tuning(const nd_plan* pl, nd_plan_tuning* t) {
    if (!pl || !t) return nd::set_error(ND_ERR_ARGUMENT, "NULL argument");
    t->hub_degree = pl->hub_degree;
    t->block_size = pl->block;
    t->big_hub_degree = pl->big_hub_degree;
    t->rk4_path = pl->rk4_path;
    t->gather_mode = pl->use_vout ? 3 : 2;
    return ND_OK;
}

int nd_rhs(nd_plan* pl, double* du, const double* u, const double* p, double t, int64_t len_du, int64_t len_u,
           int64_t len_p, void* stream) {
    if (!pl) return nd::set_error(ND_ERR_ARGUMENT, "plan is NULL");
    if (!(len_du == len_u && len_u == pl->dim))   // coreloop.jl:2-4
        return nd::set_error(ND_ERR_ARGUMENT, "du or u does not have expected size %lld", (long long)pl->dim);
    if (pl->pdim > 0 && len_p != pl->pdim)        // coreloop.jl:5-7
        return nd::set_error(ND_ERR_ARGUMENT, "p does not has expecte size %lld", (long long)pl->pdim);
    if (pl->dim > 0 && (!du || !u)) return nd::set_error(ND_ERR_ARGUMENT, "NULL state pointer");
    if (pl->pdim > 0 && !p) return nd::set_error(ND_ERR_ARGUMENT, "NULL parameter pointer");
    int rc = check_device(pl);
    if (rc) return rc;
    launch_rhs(pl, false, du, u, p, StageArgs{}, static_cast<hipStream_t>(stream), t);
    return hip_check(hipGetLastError(), "nd_rhs launch");
}

int nd_rhs_jvp(nd_plan* pl, double* du, double* jv, const double* u, const double* v, const double* p, double t,
               int64_t len_du, int64_t len_u, int64_t len_p, void* stream) {
    if (!pl) return nd::set_error(ND_ERR_ARGUMENT, "plan is NULL");
    if (!(len_du == len_u && len_u == pl->dim))   // coreloop.jl:2-4
        return nd::set_error(ND_ERR_ARGUMENT, "du or u does not have expected size %lld", (long long)pl->dim);
    if (pl->pdim > 0 && len_p != pl->pdim)        // coreloop.jl:5-7
        return nd::set_error(ND_ERR_ARGUMENT, "p does not has expecte size %lld", (long long)pl->pdim);
    if (pl->dim > 0 && (!du || !u || !jv || !v)) return nd::set_error(ND_ERR_ARGUMENT, "NULL state pointer");
    if (pl->pdim > 0 && !p) return nd::set_error(ND_ERR_ARGUMENT, "NULL parameter pointer");
    int rc = check_device(pl);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const double* vo = u;
    const double* dvo = v;
    if (pl->use_vout && !pl->d_dvout) {
        void* w = nullptr;
        hipError_t e = hipMalloc(&w, pl->vout_len * sizeof(double));
        if (e != hipSuccess) return nd::set_error(ND_ERR_ALLOC, "hipMalloc(tangent outputs): %s", hipGetErrorString(e));
        pl->allocations.push_back(w);
        pl->d_dvout = static_cast<double*>(w);
    }
    if (pl->use_vout) {
        for (const VDesc& vd : pl->vdesc) {
            if (!vd.count) continue;
            const unsigned vb = blocks_for(vd.count, 256);
            hipLaunchKernelGGL(k_vertex_out, dim3(vb), dim3(256), 0, s, vd, u, pl->d_vout, no_sync(vb), p, 0);
            hipLaunchKernelGGL(k_vertex_out, dim3(vb), dim3(256), 0, s, vd, v, pl->d_dvout, no_sync(vb), p, 1);
        }
        if (pl->ff_out)
            for (const VDesc& vd : pl->vdesc)
                if (vd.outkind == 2 && vd.count)
                    hipLaunchKernelGGL(k_vertex_out_ff_jvp, dim3(blocks_for(vd.count, 256)), dim3(256), 0, s, vd,
                                       pl->d_vout, pl->d_dvout, p);
        vo = pl->d_vout;
        dvo = pl->d_dvout;
    }
    for (size_t b = 0; b < pl->vdesc.size(); ++b) {
        VDesc vd = pl->vdesc[b];
        if (vd.count == 0) continue;
        vd.t = t;
        hipLaunchKernelGGL(pl->vkern_jvp[b], dim3(blocks_for(vd.count, 256)), dim3(256), 0, s, vd, pl->d_rowptr,
                           pl->d_slots, pl->d_slot_eb, pl->d_edesc, u, v, vo, dvo, p, du, jv);
    }
    for (const EDesc& e : pl->edesc) {
        if (!e.ode || e.count == 0) continue;
        hipLaunchKernelGGL(pick_edge_ode_jvp(e.model), dim3(blocks_for(e.count, 256)), dim3(256), 0, s, e, u, v,
                           pl->computed_out ? vo : u, pl->computed_out ? dvo : v, p, du, jv);
    }
    return hip_check(hipGetLastError(), "nd_rhs_jvp launch");
}

int nd_rhs_jvp_multi(nd_plan* pl, double* du, double* jv, const double* u, const double* v, int32_t k,
                     const double* p, double t, int64_t len_du, int64_t len_u, int64_t len_p, void* stream) {
    if (!pl) return nd::set_error(ND_ERR_ARGUMENT, "plan is NULL");
    if (k < 1 || k > 8) return nd::set_error(ND_ERR_ARGUMENT, "k = %d tangents: need 1 <= k <= 8", k);
    if (k == 1) return nd_rhs_jvp(pl, du, jv, u, v, p, t, len_du, len_u, len_p, stream);
    if (!(len_du == len_u && len_u == pl->dim))   // coreloop.jl:2-4
        return nd::set_error(ND_ERR_ARGUMENT, "du or u does not have expected size %lld", (long long)pl->dim);
    if (pl->pdim > 0 && len_p != pl->pdim)        // coreloop.jl:5-7
        return nd::set_error(ND_ERR_ARGUMENT, "p does not has expecte size %lld", (long long)pl->pdim);
    if (pl->dim > 0 && (!du || !u || !jv || !v)) return nd::set_error(ND_ERR_ARGUMENT, "NULL state pointer");
    if (pl->pdim > 0 && !p) return nd::set_error(ND_ERR_ARGUMENT, "NULL parameter pointer");
    if (pl->ff_out) return nd::set_error(ND_ERR_UNSUPPORTED, "feed-forward vertex outputs: one tangent per nd_rhs_jvp");
    for (size_t b = 0; b < pl->vdesc.size(); ++b)
        if (pl->vdesc[b].count && !pl->vkern_jvpk8[b])
            return nd::set_error(ND_ERR_UNSUPPORTED, "vertex batch %zu takes the generic kernels: one tangent per "
                                 "nd_rhs_jvp", b);
    int rc = check_device(pl);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    JvpKArgs a{};
    a.u = u;
    a.v = v;
    a.vo = u;
    a.dvo = v;
    a.p = p;
    a.du = du;
    a.jv = jv;
    a.ldu = pl->dim;
    a.ldo = pl->dim;
    a.k = k;
    if (pl->use_vout) {
        if (!pl->d_dvoutk) {
            void* w = nullptr;
            hipError_t e = hipMalloc(&w, 8 * pl->vout_len * sizeof(double));
            if (e != hipSuccess) return nd::set_error(ND_ERR_ALLOC, "hipMalloc(tangent outputs): %s", hipGetErrorString(e));
            pl->allocations.push_back(w);
            pl->d_dvoutk = static_cast<double*>(w);
        }
        for (const VDesc& vd : pl->vdesc) {
            if (!vd.count) continue;
            const unsigned vb = blocks_for(vd.count, 256);
            hipLaunchKernelGGL(k_vertex_out, dim3(vb), dim3(256), 0, s, vd, u, pl->d_vout, no_sync(vb), p, 0);
            for (int j = 0; j < k; ++j)
                hipLaunchKernelGGL(k_vertex_out, dim3(vb), dim3(256), 0, s, vd, v + static_cast<int64_t>(j) * pl->dim,
                                   pl->d_dvoutk + static_cast<int64_t>(j) * pl->vout_len, no_sync(vb), p, 1);
        }
        a.vo = pl->d_vout;
        a.dvo = pl->d_dvoutk;
        a.ldo = pl->vout_len;
    }
    for (size_t b = 0; b < pl->vdesc.size(); ++b) {
        VDesc vd = pl->vdesc[b];
        if (vd.count == 0) continue;
        vd.t = t;
        JvpKKernel kern = k <= 4 ? pl->vkern_jvpk4[b] : pl->vkern_jvpk8[b];
        hipLaunchKernelGGL(kern, dim3(blocks_for(vd.count, 256)), dim3(256), 0, s, vd, pl->d_rowptr, pl->d_slots,
                           pl->d_slot_eb, pl->d_edesc, a);
    }
    for (const EDesc& e : pl->edesc) {
        if (!e.ode || e.count == 0) continue;
        EdgeOdeJvpKKernel kern = k <= 4 ? pick_edge_ode_jvpk<4>(e.model) : pick_edge_ode_jvpk<8>(e.model);
        hipLaunchKernelGGL(kern, dim3(blocks_for(e.count, 256)), dim3(256), 0, s, e, a, pl->computed_out ? 1 : 0);
    }
    return hip_check(hipGetLastError(), "nd_rhs_jvp_multi launch");
}

// The single-instance ensemble of a network that fits one workgroup (created on first
// use): one vertex batch in vertex order and one edge batch in edge order, so the plan's
// flat layout is the ensemble's.  RK4 runs a 1024-thread workgroup with one or two
// vertices per thread (16 waves hide the per-stage barriers better than 512 x 2 or
// 512 x 4); Tsit5 keeps the default classes, since its 1024-thread kernels spill.
static int one_instance_ensemble(nd_plan* pl, bool rk4) {
    nd_ensemble*& en = rk4 ? pl->small_en : pl->small_en_ts;
    if (en) return ND_OK;
    const int64_t nv = static_cast<int64_t>(pl->vdesc[0].count), ne = static_cast<int64_t>(pl->small_src.size());
    nd_ensemble* made = nullptr;
    int rc = nd_ensemble_create(1, &nv, &ne, pl->small_src.data(), pl->small_dst.data(), &pl->small_vk,
                                &pl->small_ek, &made);
    if (!rc && rk4) rc = nd_ensemble_set_class(made, 0, 1024, nv <= 1024 ? 1 : 2);
    if (!rc && !pl->d_small_stat) {
        void* w = nullptr;
        hipError_t e = hipMalloc(&w, sizeof(nd_ensemble_stat));
        if (e != hipSuccess) rc = nd::set_error(ND_ERR_ALLOC, "hipMalloc(stats): %s", hipGetErrorString(e));
        else {
            pl->allocations.push_back(w);
            pl->d_small_stat = static_cast<nd_ensemble_stat*>(w);
        }
    }
    if (rc) {
        if (made) nd_ensemble_destroy(made);
        return rc;
    }
    en = made;
    return ND_OK;
}

int nd_tsit5(nd_plan* pl, double* u, const double* p, double t0, double t1, double reltol, double abstol,
             double dt0, int64_t maxiters, nd_ensemble_stat* stats, int64_t len_u, int64_t len_p, void* stream) {
    if (!pl) return nd::set_error(ND_ERR_ARGUMENT, "plan is NULL");
    if (len_u != pl->dim) return nd::set_error(ND_ERR_ARGUMENT, "u does not have expected size %lld", (long long)pl->dim);
    if (pl->pdim > 0 && len_p != pl->pdim)
        return nd::set_error(ND_ERR_ARGUMENT, "p does not has expecte size %lld", (long long)pl->pdim);
    if (!(t1 >= t0) || !(reltol > 0.0) || !(abstol > 0.0) || !(dt0 >= 0.0))
        return nd::set_error(ND_ERR_ARGUMENT, "need t1 >= t0, positive tolerances and dt0 >= 0");
    if (pl->dim > 0 && !u) return nd::set_error(ND_ERR_ARGUMENT, "NULL state pointer");
    if (pl->pdim > 0 && !p) return nd::set_error(ND_ERR_ARGUMENT, "NULL parameter pointer");
    if (!pl->rk4_single)
        return nd::set_error(ND_ERR_UNSUPPORTED, "nd_tsit5 runs networks of one vertex batch and one edge batch "
                             "that fit one workgroup (Kuramoto or swing vertices, AntiSymmetric Kuramoto or static "
                             "line edges, up to 2048 vertices); step others with nd_rhs");
    int rc = check_device(pl);
    if (rc) return rc;
    rc = one_instance_ensemble(pl, false);
    if (rc == ND_ERR_UNSUPPORTED) pl->rk4_single = false;   // too dense for one workgroup's LDS
    if (rc) return rc;
    nd_ensemble_opts o{};
    o.method = ND_METHOD_TSIT5;
    o.reltol = reltol;
    o.abstol = abstol;
    o.dt0 = dt0;
    o.maxiters = maxiters;
    // without line limits the ensemble kernel never writes p
    return nd_ensemble_solve(pl->small_en_ts, u, const_cast<double*>(p), t0, t1, &o,
                             stats ? stats : pl->d_small_stat, stream);
}

// Classic RK4 as four fused stage launches per step.  work: 3*dim doubles (the stage
// sum, two stage states); vout2: a second compact-output buffer of vout_len doubles when
// the plan gathers from compact outputs (unused otherwise).  Allocates nothing.
static int rk4_stage_launches(nd_plan* pl, double* u, const double* p, double t0, double dt, int64_t nsteps,
                              double* work, double* vout2, hipStream_t s) {
    double* acc = work;
    double* ya = work + pl->dim;
    double* yb = work + 2 * pl->dim;
    const double half = 0.5 * dt, sixth = dt / 6.0;
    // Compact vertex outputs ping-pong between two buffers: every stage epilogue writes the
    // outputs of the state it produces into the buffer the next stage reads, so only the
    // first stage of the call builds them.  p is const for the whole call: its parameters
    // are compared with the cached copy once.
    if (pl->computed_out) {
        // computed outputs (voltage sources, injectors) are rebuilt for every stage
        for (int64_t k = 0; k < nsteps; ++k) {
            const double tk = t0 + static_cast<double>(k) * dt;
            launch_rhs(pl, true, nullptr, u, p, StageArgs{u, acc, ya, nullptr, half, 0.0, 1}, s, tk, k == 0);
            launch_rhs(pl, true, nullptr, ya, p, StageArgs{u, acc, yb, nullptr, half, 2.0, 2}, s, tk + half, false);
            launch_rhs(pl, true, nullptr, yb, p, StageArgs{u, acc, ya, nullptr, dt, 2.0, 2}, s, tk + half, false);
            launch_rhs(pl, true, nullptr, ya, p, StageArgs{u, acc, u, nullptr, sixth, 0.0, 3}, s, tk + dt, false);
        }
        return hip_check(hipGetLastError(), "nd_rk4 launch");
    }
    double* va = pl->use_vout ? pl->d_vout : nullptr;
    double* vb = pl->use_vout ? vout2 : nullptr;
    for (int64_t k = 0; k < nsteps; ++k) {
        // stage times t_k, t_k + dt/2, t_k + dt/2, t_k + dt (time-dependent models)
        const double tk = t0 + static_cast<double>(k) * dt;
        launch_rhs(pl, true, nullptr, u, p, StageArgs{u, acc, ya, vb, half, 0.0, 1}, s, tk, k == 0, va, k == 0);
        launch_rhs(pl, true, nullptr, ya, p, StageArgs{u, acc, yb, va, half, 2.0, 2}, s, tk + half, false, vb, false);
        launch_rhs(pl, true, nullptr, yb, p, StageArgs{u, acc, ya, vb, dt, 2.0, 2}, s, tk + half, false, va, false);
        launch_rhs(pl, true, nullptr, ya, p, StageArgs{u, acc, u, va, sixth, 0.0, 3}, s, tk + dt, false, vb, false);
    }
    return hip_check(hipGetLastError(), "nd_rk4 launch");
}

int nd_rk4(nd_plan* pl, double* u, const double* p, double t0, double dt, int64_t nsteps, double* work,
           int64_t len_u, int64_t len_p, void* stream) {
    if (!pl) return nd::set_error(ND_ERR_ARGUMENT, "plan is NULL");
    if (len_u != pl->dim) return nd::set_error(ND_ERR_ARGUMENT, "u does not have expected size %lld", (long long)pl->dim);
    if (pl->pdim > 0 && len_p != pl->pdim)
        return nd::set_error(ND_ERR_ARGUMENT, "p does not has expecte size %lld", (long long)pl->pdim);
    if (nsteps < 0 || !(dt == dt)) return nd::set_error(ND_ERR_ARGUMENT, "bad step count or dt");
    if (pl->dim > 0 && !u) return nd::set_error(ND_ERR_ARGUMENT, "NULL state pointer");
    if (pl->pdim > 0 && !p) return nd::set_error(ND_ERR_ARGUMENT, "NULL parameter pointer");
    int rc = check_device(pl);
    if (rc) return rc;
    if (pl->rk4_single && pl->rk4_path == 1 && nsteps > 0 && dt > 0.0) {
        rc = one_instance_ensemble(pl, true);
        if (rc == ND_ERR_UNSUPPORTED) {
            // the network's topology needs more LDS than one workgroup has (dense
            // graphs): it steps through the stage launches below, and nd_tsit5 refuses it
            pl->rk4_single = false;
            rc = ND_OK;
        }
        if (rc) return rc;
    }
    if (pl->rk4_single && pl->rk4_path == 1 && nsteps > 0 && dt > 0.0) {
        nd_ensemble_opts o{};
        o.method = ND_METHOD_RK4;
        o.dt0 = dt;
        o.maxiters = nsteps + 1;
        // without line limits the ensemble kernel never writes p
        return nd_ensemble_solve(pl->small_en, u, const_cast<double*>(p), 0.0, static_cast<double>(nsteps) * dt, &o,
                                 pl->d_small_stat, stream);
    }
    if (!work) {
        if (!pl->rk_work) {
            void* w = nullptr;
            hipError_t e = hipMalloc(&w, static_cast<size_t>(3 * std::max<int64_t>(pl->dim, 1)) * sizeof(double));
            if (e != hipSuccess) return nd::set_error(ND_ERR_ALLOC, "hipMalloc(rk work): %s", hipGetErrorString(e));
            pl->allocations.push_back(w);
            pl->rk_work = static_cast<double*>(w);
        }
        work = pl->rk_work;
    }
    if (pl->use_vout && !pl->computed_out && !pl->d_vout2) {
        void* w = nullptr;
        hipError_t e = hipMalloc(&w, pl->vout_len * sizeof(double));
        if (e != hipSuccess) return nd::set_error(ND_ERR_ALLOC, "hipMalloc(vertex outputs): %s", hipGetErrorString(e));
        pl->allocations.push_back(w);
        pl->d_vout2 = static_cast<double*>(w);
    }
    return rk4_stage_launches(pl, u, p, t0, dt, nsteps, work, pl->d_vout2, static_cast<hipStream_t>(stream));
}

int nd_rk4_noalloc(nd_plan* pl, double* u, const double* p, double t0, double dt, int64_t nsteps, int64_t len_u,
                   int64_t len_p, void* stream) {
    if (!pl) return nd::set_error(ND_ERR_ARGUMENT, "plan is NULL");
    if (len_u != pl->dim) return nd::set_error(ND_ERR_ARGUMENT, "u does not have expected size %lld", (long long)pl->dim);
    if (pl->pdim > 0 && len_p != pl->pdim)
        return nd::set_error(ND_ERR_ARGUMENT, "p does not has expecte size %lld", (long long)pl->pdim);
    if (nsteps < 0 || !(dt == dt)) return nd::set_error(ND_ERR_ARGUMENT, "bad step count or dt");
    if (pl->dim > 0 && !u) return nd::set_error(ND_ERR_ARGUMENT, "NULL state pointer");
    if (pl->pdim > 0 && !p) return nd::set_error(ND_ERR_ARGUMENT, "NULL parameter pointer");
    const bool two_outputs = pl->use_vout && !pl->computed_out;
    const int64_t work = 3 * std::max<int64_t>(pl->dim, 1);
    const int64_t need = work + (two_outputs ? static_cast<int64_t>(pl->vout_len) : 0);
    if (need > kNoAllocScratch)
        return nd::set_error(ND_ERR_ARGUMENT,
                             "nd_rk4_noalloc: the network needs %lld doubles of scratch, above its maximum of %lld "
                             "(3 x the state dimension plus the compact outputs); use nd_rk4",
                             (long long)need, (long long)kNoAllocScratch);
    int rc = check_device(pl);
    if (rc) return rc;
    double* sp = nullptr;
    rc = noalloc_scratch(pl->device, &sp);
    if (rc) return rc;
    return rk4_stage_launches(pl, u, p, t0, dt, nsteps, sp, two_outputs ? sp + work : nullptr,
                              static_cast<hipStream_t>(stream));
}

int nd_collect_externals(nd_plan* pl, double* extbuf, const double* u, const double* o, void* stream) {
    if (!pl) return nd::set_error(ND_ERR_ARGUMENT, "plan is NULL");
    if (pl->lastidx_extbuf == 0) return ND_OK;
    if (!extbuf || !u || !o) return nd::set_error(ND_ERR_ARGUMENT, "NULL buffer");
    int rc = check_device(pl);
    if (rc) return rc;
    const unsigned nb = static_cast<unsigned>(std::min<int64_t>((pl->lastidx_extbuf + 255) / 256, 4096));
    hipLaunchKernelGGL(k_collect_externals, dim3(nb), dim3(256), 0, static_cast<hipStream_t>(stream), pl->d_extmap,
                       pl->lastidx_extbuf, extbuf, u, o);
    return hip_check(hipGetLastError(), "nd_collect_externals launch");
}

int nd_gather(nd_plan* pl, double* gbuf, const double* o, void* stream) {
    if (!pl) return nd::set_error(ND_ERR_ARGUMENT, "plan is NULL");
    if (pl->lastidx_gbuf == 0) return ND_OK;
    if (!gbuf || !o) return nd::set_error(ND_ERR_ARGUMENT, "NULL buffer");
    int rc = check_device(pl);
    if (rc) return rc;
    const unsigned nb = std::min<unsigned>(blocks_for(pl->lastidx_gbuf, 256), 65535u);
    hipLaunchKernelGGL(k_gather, dim3(nb), dim3(256), 0, static_cast<hipStream_t>(stream), pl->d_gmap, o, gbuf,
                       pl->lastidx_gbuf);
    return hip_check(hipGetLastError(), "nd_gather launch");
}

int nd_aggregate(nd_plan* pl, double* aggbuf, const double* o, void* stream) {
    if (!pl) return nd::set_error(ND_ERR_ARGUMENT, "plan is NULL");
    if (pl->lastidx_aggr == 0) return ND_OK;
    if (!aggbuf || !o) return nd::set_error(ND_ERR_ARGUMENT, "NULL buffer");
    int rc = check_device(pl);
    if (rc) return rc;
    const int32_t nrows = static_cast<int32_t>(pl->lastidx_aggr);
    const int32_t nlow = static_cast<int32_t>(blocks_for(nrows, 256));
    const int32_t nhub = static_cast<int32_t>(blocks_for(static_cast<int64_t>(pl->n_agg_hub) * kWave, 256));
    hipLaunchKernelGGL(k_aggregate, dim3(static_cast<unsigned>(nlow + nhub)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), pl->d_rowptr, pl->d_cols, o, aggbuf, nrows, pl->d_agg_hub_rows,
                       pl->n_agg_hub, pl->hub_degree, nlow);
    return hip_check(hipGetLastError(), "nd_aggregate launch");
}

int nd_get_buffers(nd_plan* pl, double* o, double* aggbuf, const double* u, const double* p, double t,
                   void* stream) {
    (void)t;
    if (!pl) return nd::set_error(ND_ERR_ARGUMENT, "plan is NULL");
    if (!o || (pl->lastidx_aggr > 0 && !aggbuf) || (pl->dim > 0 && !u))
        return nd::set_error(ND_ERR_ARGUMENT, "NULL buffer");
    int rc = check_device(pl);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const unsigned nbo = std::max(1u, std::min<unsigned>(blocks_for(pl->lastidx_out, 256), 65535u));
    hipLaunchKernelGGL(k_fill, dim3(nbo), dim3(256), 0, s, o, pl->lastidx_out, std::numeric_limits<double>::quiet_NaN());
    if (pl->lastidx_aggr > 0) {
        const unsigned nba = std::min<unsigned>(blocks_for(pl->lastidx_aggr, 256), 65535u);
        hipLaunchKernelGGL(k_fill, dim3(nba), dim3(256), 0, s, aggbuf, pl->lastidx_aggr, 0.0);
    }
    for (const VDesc& vd : pl->vdesc)
        if (vd.count)
            hipLaunchKernelGGL(k_vertex_out, dim3(blocks_for(vd.count, 256)), dim3(256), 0, s, vd, u, o,
                               no_sync(blocks_for(vd.count, 256)), p, 0);
    if (pl->ff_out)
        for (const VDesc& vd : pl->vdesc)
            if (vd.outkind == 2 && vd.count)
                hipLaunchKernelGGL(k_vertex_out_ff, dim3(blocks_for(vd.count, 256)), dim3(256), 0, s, vd, o, p);
    // with computed outputs the edges' endpoint indices address o
    for (const EDesc& e : pl->edesc)
        if (e.count)
            hipLaunchKernelGGL(k_edge_out, dim3(blocks_for(e.count, 256)), dim3(256), 0, s, e, u,
                               pl->computed_out ? static_cast<const double*>(o) : u, p, o);
    if (pl->d_lbin && pl->lastidx_aggr > 0)
        for (const VDesc& vd : pl->vdesc)
            if (vd.count)
                hipLaunchKernelGGL(k_loopback_copy, dim3(blocks_for(vd.count, 256)), dim3(256), 0, s, vd, o, aggbuf);
    rc = hip_check(hipGetLastError(), "nd_get_buffers launch");
    if (rc) return rc;
    return nd_aggregate(pl, aggbuf, o, stream);
}

int nd_plan_set_perturbation(nd_plan* pl, int64_t ndelta, int64_t nvo, const int64_t* vo, int64_t nvi,
                              const int64_t* vi, int64_t neo, const int64_t* eo, int64_t nei, const int64_t* ei) {
    nd::clear_error();
    if (!pl) return nd::set_error(ND_ERR_ARGUMENT, "plan is NULL");
    if (ndelta < 0 || nvo < 0 || nvi < 0 || neo < 0 || nei < 0) return nd::set_error(ND_ERR_ARGUMENT, "negative length");
    int rc = check_device(pl);
    if (rc) return rc;
    const int64_t* maps[4] = {vo, vi, eo, ei};
    const int64_t lens[4] = {nvo, nvi, neo, nei};
    const int64_t bufs[4] = {pl->lastidx_out, pl->lastidx_aggr, pl->lastidx_out, pl->lastidx_gbuf};
    const char* names[4] = {"vertex output", "vertex input", "edge output", "edge input"};
    std::vector<int2> host[4];
    for (int m = 0; m < 4; ++m) {
        if (lens[m] > 0 && !maps[m]) return nd::set_error(ND_ERR_ARGUMENT, "%s map is NULL", names[m]);
        for (int64_t k = 0; k < lens[m]; ++k) {
            const int64_t b = maps[m][2 * k], d = maps[m][2 * k + 1];
            if (b < 1 || b > bufs[m]) return nd::set_error(ND_ERR_ARGUMENT, "%s perturbation: buffer index %lld out of range", names[m], (long long)b);
            if (d < 1 || d > ndelta) return nd::set_error(ND_ERR_ARGUMENT, "%s perturbation: channel %lld out of range", names[m], (long long)d);
            host[m].push_back(make_int2(static_cast<int32_t>(b - 1), static_cast<int32_t>(d - 1)));
        }
    }
    rc = hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");   // earlier calls may still read the old maps
    if (rc) return rc;
    for (int m = 0; m < 4; ++m) {
        release(pl, pl->d_pmap[m]);
        pl->d_pmap[m] = nullptr;
        rc = upload(pl, host[m], &pl->d_pmap[m]);
        if (rc) return rc;
        pl->n_pmap[m] = static_cast<int32_t>(host[m].size());
    }
    pl->ndelta = ndelta;
    if (!pl->d_egb && !pl->egb_host.empty()) {
        rc = upload(pl, pl->egb_host, &pl->d_egb);
        if (rc) return rc;
    }
    return ND_OK;
}

int nd_rhs_perturbed_jvp(nd_plan* pl, double* du, double* jv, const double* u, const double* v, const double* p,
                         const double* dp, double t, const double* delta, const double* ddelta, double* o, double* dout,
                         double* aggbuf, double* dagg, void* stream) {
    if (!pl) return nd::set_error(ND_ERR_ARGUMENT, "plan is NULL");
    if ((pl->dim > 0 && (!du || !u)) || (pl->ndelta > 0 && !delta))
        return nd::set_error(ND_ERR_ARGUMENT, "NULL buffer");
    int rc = check_device(pl);
    if (rc) return rc;
    if (!pl->d_egb && !pl->egb_host.empty()) {
        rc = upload(pl, pl->egb_host, &pl->d_egb);
        if (rc) return rc;
    }
    // work buffers: o, do, agg, dagg, gbuf, dgbuf, ext, dext, zero tangents, jv
    const size_t no = static_cast<size_t>(pl->lastidx_out), na = static_cast<size_t>(pl->lastidx_aggr);
    const size_t ng = static_cast<size_t>(pl->lastidx_gbuf), nx = static_cast<size_t>(pl->lastidx_extbuf);
    const size_t nz = std::max<size_t>({static_cast<size_t>(pl->dim), static_cast<size_t>(pl->ndelta),
                                        static_cast<size_t>(pl->pdim), size_t{1}});
    const size_t need = 2 * (no + na + ng + nx) + nz + static_cast<size_t>(pl->dim) + 8;
    if (pl->pwork_len < need) {
        rc = hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        if (rc) return rc;
        release(pl, pl->d_pwork);
        pl->d_pwork = nullptr;
        void* w = nullptr;
        hipError_t e = hipMalloc(&w, need * sizeof(double));
        if (e != hipSuccess) return nd::set_error(ND_ERR_ALLOC, "hipMalloc(perturbed RHS buffers): %s", hipGetErrorString(e));
        pl->allocations.push_back(w);
        pl->d_pwork = static_cast<double*>(w);
        pl->pwork_len = need;
    }
    double* w = pl->d_pwork;
    double* wo = w;            w += no;
    double* wdo = w;           w += no;
    double* wa = w;            w += na;
    double* wda = w;           w += na;
    double* gb = w;            w += ng;
    double* dgb = w;           w += ng;
    double* ex = w;            w += nx;
    double* dex = w;           w += nx;
    double* zero = w;          w += nz;
    double* wjv = w;
    if (!o) o = wo;
    if (!dout) dout = wdo;
    if (!aggbuf) aggbuf = wa;
    if (!dagg) dagg = wda;
    if (!jv) jv = wjv;
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto fill = [&](double* x, size_t n, double val) {
        if (n) hipLaunchKernelGGL(k_fill, dim3(std::max(1u, std::min<unsigned>(blocks_for(static_cast<int64_t>(n), 256), 65535u))),
                                  dim3(256), 0, s, x, static_cast<int64_t>(n), val);
    };
    fill(zero, nz, 0.0);
    if (!v) v = zero;
    if (!dp) dp = zero;
    if (!p) p = zero;
    if (!ddelta) ddelta = zero;
    if (!delta) delta = zero;
    fill(du, static_cast<size_t>(pl->dim), 0.0);                              // coreloop.jl:24
    fill(jv, static_cast<size_t>(pl->dim), 0.0);
    fill(o, no, std::numeric_limits<double>::quiet_NaN());                     // network_structure.jl:153
    fill(dout, no, std::numeric_limits<double>::quiet_NaN());
    fill(aggbuf, na, 0.0);                                                     // coreloop.jl:30
    fill(dagg, na, 0.0);
    auto perturb = [&](int m, double* buf, double* dbuf) {
        if (pl->n_pmap[m]) hipLaunchKernelGGL(k_perturb, dim3(1), dim3(64), 0, s, pl->d_pmap[m], pl->n_pmap[m], delta, ddelta, buf, dbuf);
    };
    // :g of vertices and of ODE edges without feed-forward (coreloop.jl:39-41)
    for (const VDesc& vd : pl->vdesc)
        if (vd.count) {
            const unsigned nb = blocks_for(vd.count, 256);
            hipLaunchKernelGGL(k_vertex_out, dim3(nb), dim3(256), 0, s, vd, u, o, no_sync(nb), p, 0);
            hipLaunchKernelGGL(k_vertex_out, dim3(nb), dim3(256), 0, s, vd, v, dout, no_sync(nb), p, 1);
            if (vd.outkind == 1) hipLaunchKernelGGL(k_vertex_out_nff_tangent, dim3(nb), dim3(256), 0, s, vd, p, dp, dout);
        }
    for (const EDesc& e : pl->edesc)
        if (e.count && e.ode) {
            hipLaunchKernelGGL(k_edge_out, dim3(blocks_for(e.count, 256)), dim3(256), 0, s, e, u, u, p, o);
            hipLaunchKernelGGL(k_edge_out, dim3(blocks_for(e.count, 256)), dim3(256), 0, s, e, v, v, p, dout);
        }
    // loopback copies, vertex-input perturbation, feed-forward vertex g (coreloop.jl:47-55)
    if (pl->d_lbin && na > 0)
        for (const VDesc& vd : pl->vdesc)
            if (vd.count) {
                hipLaunchKernelGGL(k_loopback_copy, dim3(blocks_for(vd.count, 256)), dim3(256), 0, s, vd, o, aggbuf);
                hipLaunchKernelGGL(k_loopback_copy, dim3(blocks_for(vd.count, 256)), dim3(256), 0, s, vd, dout, dagg);
            }
    perturb(1, aggbuf, dagg);
    if (pl->ff_out && na > 0)
        for (const VDesc& vd : pl->vdesc)
            if (vd.outkind == 2 && vd.count)
                hipLaunchKernelGGL(k_vertex_out_ff_buf, dim3(blocks_for(vd.count, 256)), dim3(256), 0, s, vd, aggbuf, dagg,
                                   p, dp, o, dout);
    // external inputs, vertex-output perturbation, gather, edge-input perturbation (:56-72)
    if (nx > 0) {
        const unsigned nb = std::max(1u, std::min<unsigned>(blocks_for(pl->lastidx_extbuf, 256), 65535u));
        hipLaunchKernelGGL(k_collect_externals, dim3(nb), dim3(256), 0, s, pl->d_extmap, pl->lastidx_extbuf, ex, u, o);
        hipLaunchKernelGGL(k_collect_externals, dim3(nb), dim3(256), 0, s, pl->d_extmap, pl->lastidx_extbuf, dex, v, dout);
    }
    perturb(0, o, dout);
    if (ng > 0) {
        const unsigned nb = std::min<unsigned>(blocks_for(pl->lastidx_gbuf, 256), 65535u);
        hipLaunchKernelGGL(k_gather, dim3(nb), dim3(256), 0, s, pl->d_gmap, o, gb, pl->lastidx_gbuf);
        hipLaunchKernelGGL(k_gather, dim3(nb), dim3(256), 0, s, pl->d_gmap, dout, dgb, pl->lastidx_gbuf);
    }
    perturb(3, gb, dgb);
    // edges: f of ODE edges, f and g of feed-forward edges (:74-78); edge-output perturbation
    for (size_t b = 0; b < pl->edesc.size(); ++b) {
        const EDesc& e = pl->edesc[b];
        if (e.count)
            hipLaunchKernelGGL(k_edge_buf, dim3(blocks_for(e.count, 256)), dim3(256), 0, s, e, pl->d_egb + pl->egb_off[b],
                               gb, dgb, u, v, ex, dex, p, dp, o, dout, du, jv);
    }
    perturb(2, o, dout);
    // aggregation (:90) and vertex f (:97)
    rc = hip_check(hipGetLastError(), "perturbed RHS launch");
    if (rc) return rc;
    if (na > 0) {
        rc = nd_aggregate(pl, aggbuf, o, stream);
        rc = rc ? rc : nd_aggregate(pl, dagg, dout, stream);
        if (rc) return rc;
    }
    for (const VDesc& vd0 : pl->vdesc) {
        if (!vd0.count) continue;
        VDesc vd = vd0;
        vd.t = t;
        hipLaunchKernelGGL(pick_vertex_f_buf(vd.model), dim3(blocks_for(vd.count, 256)), dim3(256), 0, s, vd,
                           na > 0 ? aggbuf : nullptr, na > 0 ? dagg : nullptr, u, v, ex, dex, p, dp, du, jv);
    }
    return hip_check(hipGetLastError(), "perturbed RHS launch");
}

int nd_device_synchronize(void) { return hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

int nd_malloc(size_t bytes, void** out) {
    if (!out) return nd::set_error(ND_ERR_ARGUMENT, "NULL output pointer");
    *out = nullptr;
    if (bytes == 0) return ND_OK;
    hipError_t e = hipMalloc(out, bytes);
    if (e != hipSuccess) {
        *out = nullptr;
        return nd::set_error(ND_ERR_ALLOC, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    return ND_OK;
}

int nd_free(void* ptr) { return ptr ? hip_check(hipFree(ptr), "hipFree") : ND_OK; }

int nd_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    if (bytes && (!dst || !src)) return nd::set_error(ND_ERR_ARGUMENT, "NULL pointer");
    return bytes ? hip_check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy") : ND_OK;
}

int nd_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    if (bytes && (!dst || !src)) return nd::set_error(ND_ERR_ARGUMENT, "NULL pointer");
    return bytes ? hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy") : ND_OK;
}

}  // extern "C"
