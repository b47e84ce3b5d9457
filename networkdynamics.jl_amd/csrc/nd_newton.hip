// Batched FP64 Newton solver for the synchronous fixpoints of Kuramoto-type power
// grids (swing / first-order load vertices, AntiSymmetric K*sin lines): one workgroup
// per grid, all grids in one launch.
//
// A state with omega = 0 is a fixpoint of the swing and load models iff
//     F_i(theta) = P_i + sum_j K_ij sin(theta_j - theta_i) = 0   for every vertex,
// the condition find_fixpoint solves through NonlinearSolve (initialization.jl:73-144).
// Jacobian: dF/dtheta = -L_w, the Laplacian with weights w_ij = K_ij cos(theta_j - theta_i),
// singular along the global phase.  Each Newton step solves L_w d = F (so theta + d
// is the Newton update) with Jacobi-preconditioned conjugate gradients on the
// mean-zero subspace, then backtracks (halving) until the max-norm of the scaled
// residual F_i / scale_i decreases.  Converged when max_i |F_i / scale_i| <= tol.
//
// On chip: theta and the CG search vector in LDS (neighbour gathers), per-vertex
// CG vectors in registers (V vertices per thread), destination lists and weights
// in global memory (L2 / MALL resident per grid).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "nd_internal.h"

namespace {

constexpr int kWave = 64;
constexpr int kThreads = 512;

struct GridDesc {
    int32_t n, nslot;
    int64_t voff;        // vertex offset (theta, P, scale)
    int64_t roff;        // rowptr offset (n + 1 entries)
    int64_t soff;        // slot offset
};

struct NewtonArgs {
    const GridDesc* grid;
    const int32_t* rowptr;   // local slot offsets
    const int32_t* nbr;      // local neighbour
    const int32_t* slot_edge;// global edge index of each slot
    double* w;               // per-slot weight scratch
    double* theta;
    const double* P;
    const double* scale;
    const double* K;         // per edge
    nd_newton_stat* stats;   // original grid order
    const int32_t* order;    // launch position -> grid index
    double tol, cg_rtol, eta_max;
    int32_t max_newton, cg_maxiter;
    int32_t linearized, inexact;
    int32_t first;
};

// max that propagates NaN (fmax drops it): a residual with a NaN entry must read as NaN,
// never as converged
__device__ __forceinline__ double max_nan(double a, double b) { return (b > a || b != b) ? b : a; }

__device__ __forceinline__ double block_reduce(double v, double* red, bool is_max) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, kWave);
        v = is_max ? max_nan(v, o) : v + o;
    }
    const int t = static_cast<int>(threadIdx.x), nw = static_cast<int>(blockDim.x) / kWave;
    if ((t & (kWave - 1)) == 0) red[t / kWave] = v;
    __syncthreads();
    if (t == 0) {
        double s = red[0];
        for (int w = 1; w < nw; ++w) s = is_max ? max_nan(s, red[w]) : s + red[w];
        red[16] = s;
    }
    __syncthreads();
    const double r = red[16];
    __syncthreads();
    return r;
}

template <int V>
__global__ void __launch_bounds__(kThreads) k_newton(NewtonArgs a) {
    const int T = static_cast<int>(blockDim.x), t = static_cast<int>(threadIdx.x);
    const int g = a.first + static_cast<int>(blockIdx.x);
    const GridDesc G = a.grid[g];
    extern __shared__ double sm[];
    double* th = sm;            // theta (n)
    double* pv = th + G.n;      // CG search vector (n)
    double* red = pv + G.n;     // 32
    const int32_t* rp = a.rowptr + G.roff;
    const int32_t* nb = a.nbr + G.soff;
    const int32_t* se = a.slot_edge + G.soff;
    double* w = a.w + G.soff;

    double x[V], F[V], r[V], d[V], dg[V];
    const double* Pg = a.P + G.voff;
    const double* scg = a.scale + G.voff;
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int v = t + k * T;
        const bool ok = v < G.n;
        x[k] = ok ? a.theta[G.voff + v] : 0.0;
        if (ok) th[v] = x[k];
    }
    __syncthreads();

    // F(theta) for the thread's vertices from th[]; returns max |F/scale| over the grid.
    // The library sin/cos stay here: with V residual slots unrolled, the inlined
    // reduction of nd_math.h and its out-of-line call made the V = 10 instance spill.
    auto residual = [&](double (&out)[V]) {
        double m = 0.0;
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const int v = t + k * T;
            out[k] = 0.0;
            if (v < G.n) {
                const double own = th[v];
                double acc = 0.0;
                if (a.linearized)
                    for (int s = rp[v]; s < rp[v + 1]; ++s) acc += a.K[se[s]] * (th[nb[s]] - own);
                else
                    for (int s = rp[v]; s < rp[v + 1]; ++s) acc += a.K[se[s]] * sin(th[nb[s]] - own);
                out[k] = Pg[v] + acc;
                m = max_nan(m, fabs(out[k] / scg[v]));
            }
        }
        return block_reduce(m, red, true);
    };

    double res = residual(F);
    int it = 0, cg_total = 0, retcode = 0;
    while (res > a.tol) {
        if (it >= a.max_newton) { retcode = 1; break; }
        ++it;
        // weights and Jacobi diagonal at the current theta
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const int v = t + k * T;
            dg[k] = 1.0;
            if (v < G.n) {
                const double own = th[v];
                double dsum = 0.0;
                for (int s = rp[v]; s < rp[v + 1]; ++s) {
                    const double ws = a.linearized ? a.K[se[s]] : a.K[se[s]] * cos(th[nb[s]] - own);
                    w[s] = ws;
                    dsum += ws;
                }
                dg[k] = dsum > 0.0 ? dsum : 1.0;
            }
        }
        // CG for L_w d = F, d mean-zero (F sums to zero when P is balanced)
        double fmean = 0.0;
#pragma unroll
        for (int k = 0; k < V; ++k) if (t + k * T < G.n) fmean += F[k];
        fmean = block_reduce(fmean, red, false) / G.n;
        double rz = 0.0, rr0 = 0.0;
#pragma unroll
        for (int k = 0; k < V; ++k) {
            d[k] = 0.0;
            r[k] = (t + k * T < G.n) ? F[k] - fmean : 0.0;
            rz += r[k] * (r[k] / dg[k]);
            rr0 += r[k] * r[k];
        }
        rz = block_reduce(rz, red, false);
        rr0 = block_reduce(rr0, red, false);
        const double eta = a.inexact ? fmax(a.cg_rtol, fmin(a.eta_max, res)) : a.cg_rtol;
        const double stop2 = eta * eta * rr0;
        double rr = rr0;
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const int v = t + k * T;
            if (v < G.n) pv[v] = r[k] / dg[k];
        }
        __syncthreads();
        int cg = 0;
        while (rr > stop2 && cg < a.cg_maxiter) {
            ++cg;
            double q[V], pq = 0.0;
#pragma unroll
            for (int k = 0; k < V; ++k) {
                const int v = t + k * T;
                q[k] = 0.0;
                if (v < G.n) {
                    const double pi = pv[v];
                    double acc = 0.0;
                    for (int s = rp[v]; s < rp[v + 1]; ++s) acc += w[s] * (pi - pv[nb[s]]);
                    q[k] = acc;
                    pq += pi * acc;
                }
            }
            pq = block_reduce(pq, red, false);
            const double alpha = rz / pq;
            double rzn = 0.0, rrn = 0.0;
#pragma unroll
            for (int k = 0; k < V; ++k) {
                const int v = t + k * T;
                if (v < G.n) {
                    d[k] += alpha * pv[v];
                    r[k] -= alpha * q[k];
                    rzn += r[k] * (r[k] / dg[k]);
                    rrn += r[k] * r[k];
                }
            }
            rzn = block_reduce(rzn, red, false);
            rr = block_reduce(rrn, red, false);
            const double beta = rzn / rz;
            rz = rzn;
#pragma unroll
            for (int k = 0; k < V; ++k) {
                const int v = t + k * T;
                if (v < G.n) pv[v] = r[k] / dg[k] + beta * pv[v];
            }
            __syncthreads();
        }
        cg_total += cg;
        // backtracking on the residual max-norm
        double step = 1.0, nres = res;
        for (int ls = 0; ls < 30; ++ls) {
#pragma unroll
            for (int k = 0; k < V; ++k) {
                const int v = t + k * T;
                if (v < G.n) th[v] = x[k] + step * d[k];
            }
            __syncthreads();
            nres = residual(F);
            if (nres < res || ls == 29) break;
            step *= 0.5;
        }
#pragma unroll
        for (int k = 0; k < V; ++k) {
            x[k] = (t + k * T < G.n) ? x[k] + step * d[k] : 0.0;
        }
        if (!(nres < res)) { res = nres; retcode = 2; break; }
        res = nres;
    }
    if (!(res == res)) retcode = 3;         // a NaN residual never counts as converged
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int v = t + k * T;
        if (v < G.n) a.theta[G.voff + v] = x[k];
    }
    if (t == 0) {
        nd_newton_stat s;
        s.residual = res;
        s.newton_iters = it;
        s.cg_iters = cg_total;
        s.retcode = retcode;
        s.pad_ = 0;
        a.stats[a.order[g]] = s;
    }
}

int herr(hipError_t e, const char* w) {
    return e == hipSuccess ? ND_OK : nd::set_error(ND_ERR_DEVICE, "%s: %s", w, hipGetErrorString(e));
}

}  // namespace

struct nd_fixpoint {
    int device = 0;
    int32_t ngrid = 0;
    std::vector<GridDesc> grids;
    std::vector<int32_t> small, large;      // grid indices by class (V = 4 / V = 10)
    std::vector<void*> allocs;
    GridDesc* d_grid = nullptr;
    int32_t *d_rp = nullptr, *d_nb = nullptr, *d_se = nullptr, *d_order = nullptr;
    double* d_w = nullptr;
    int64_t total_vertices = 0, total_edges = 0;
    // the two size classes run on their own streams, forked from and joined back to the
    // caller's stream with events, so the large grids start while small ones still run
    hipStream_t cls_stream[2] = {};
    hipEvent_t ev_fork = nullptr, ev_join[2] = {};
    ~nd_fixpoint() {
        for (void* p : allocs) (void)hipFree(p);
        for (auto& x : cls_stream) if (x) (void)hipStreamDestroy(x);
        for (auto& x : ev_join) if (x) (void)hipEventDestroy(x);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
    }
};

namespace {
template <typename T>
int upv(nd_fixpoint* f, const std::vector<T>& h, T** d) {
    *d = nullptr;
    if (h.empty()) return ND_OK;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, h.size() * sizeof(T));
    if (e != hipSuccess) return nd::set_error(ND_ERR_ALLOC, "hipMalloc: %s", hipGetErrorString(e));
    f->allocs.push_back(p);
    e = hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return herr(e, "hipMemcpy");
    *d = static_cast<T*>(p);
    return ND_OK;
}
}  // namespace

extern "C" {

int nd_fixpoint_create(int32_t ngrid, const int64_t* nv, const int64_t* ne, const int64_t* src, const int64_t* dst,
                       nd_fixpoint** out) {
    nd::clear_error();
    if (!out || ngrid < 1 || !nv || !ne) return nd::set_error(ND_ERR_ARGUMENT, "bad arguments");
    *out = nullptr;
    nd_fixpoint* f = new (std::nothrow) nd_fixpoint;
    if (!f) return nd::set_error(ND_ERR_ALLOC, "allocation failed");
    (void)hipGetDevice(&f->device);
    f->ngrid = ngrid;
    std::vector<int32_t> rp, nb, sedge;
    int64_t voff = 0, eoff = 0;
    for (int32_t gi = 0; gi < ngrid; ++gi) {
        const int64_t n = nv[gi], m = ne[gi];
        if (n < 1 || n > 10 * kThreads || m < 0) {
            delete f;
            return nd::set_error(ND_ERR_UNSUPPORTED, "grid %d: %lld vertices (supported 1..%d)", gi, (long long)n, 10 * kThreads);
        }
        GridDesc G{};
        G.n = static_cast<int32_t>(n);
        G.nslot = static_cast<int32_t>(2 * m);
        G.voff = voff;
        G.roff = static_cast<int64_t>(rp.size());
        G.soff = static_cast<int64_t>(nb.size());
        std::vector<int32_t> deg(static_cast<size_t>(n) + 1, 0);
        for (int64_t e = 0; e < m; ++e) {
            const int64_t s = src[eoff + e] - 1, d = dst[eoff + e] - 1;
            if (s < 0 || s >= n || d < 0 || d >= n) { delete f; return nd::set_error(ND_ERR_ARGUMENT, "grid %d: bad edge", gi); }
            deg[s + 1]++;
            deg[d + 1]++;
        }
        for (int64_t v = 0; v < n; ++v) deg[v + 1] += deg[v];
        std::vector<int32_t> fill(deg.begin(), deg.end() - 1), lnb(static_cast<size_t>(2 * m)), lse(static_cast<size_t>(2 * m));
        for (int64_t e = 0; e < m; ++e) {
            const int64_t s = src[eoff + e] - 1, d = dst[eoff + e] - 1;
            lnb[fill[s]] = static_cast<int32_t>(d);
            lse[fill[s]++] = static_cast<int32_t>(eoff + e);
            lnb[fill[d]] = static_cast<int32_t>(s);
            lse[fill[d]++] = static_cast<int32_t>(eoff + e);
        }
        rp.insert(rp.end(), deg.begin(), deg.end());
        nb.insert(nb.end(), lnb.begin(), lnb.end());
        sedge.insert(sedge.end(), lse.begin(), lse.end());
        f->grids.push_back(G);
        (n <= 4 * kThreads ? f->small : f->large).push_back(gi);
        voff += n;
        eoff += m;
    }
    f->total_vertices = voff;
    f->total_edges = eoff;
    // one launch per class: order = small grids then large grids, each by decreasing size
    auto bysize = [&](std::vector<int32_t>& v) {
        std::sort(v.begin(), v.end(), [&](int32_t x, int32_t y) { return f->grids[x].n > f->grids[y].n; });
    };
    bysize(f->small);
    bysize(f->large);
    std::vector<GridDesc> ordered;
    for (int32_t x : f->small) ordered.push_back(f->grids[x]);
    for (int32_t x : f->large) ordered.push_back(f->grids[x]);
    std::vector<int32_t> order(f->small);
    order.insert(order.end(), f->large.begin(), f->large.end());
    int rc = upv(f, ordered, &f->d_grid);
    rc = rc ? rc : upv(f, order, &f->d_order);
    rc = rc ? rc : upv(f, rp, &f->d_rp);
    rc = rc ? rc : upv(f, nb, &f->d_nb);
    rc = rc ? rc : upv(f, sedge, &f->d_se);
    if (!rc && !nb.empty()) {
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, nb.size() * sizeof(double));
        if (e != hipSuccess) rc = nd::set_error(ND_ERR_ALLOC, "hipMalloc(w): %s", hipGetErrorString(e));
        else { f->allocs.push_back(p); f->d_w = static_cast<double*>(p); }
    }
    if (rc) { delete f; return rc; }
    *out = f;
    return ND_OK;
}

void nd_fixpoint_destroy(nd_fixpoint* f) { delete f; }

int nd_fixpoint_solve(nd_fixpoint* f, double* theta, const double* P, const double* scale, const double* K,
                      const nd_newton_opts* o, nd_newton_stat* stats, void* stream) {
    if (!f || !theta || !P || !scale || !o || !stats) return nd::set_error(ND_ERR_ARGUMENT, "NULL argument");
    if (f->total_edges > 0 && !K) return nd::set_error(ND_ERR_ARGUMENT, "NULL K");
    int dev = -1;
    (void)hipGetDevice(&dev);
    if (dev != f->device) return nd::set_error(ND_ERR_ARGUMENT, "solver lives on device %d", f->device);
    NewtonArgs a{};
    a.grid = f->d_grid;
    a.rowptr = f->d_rp;
    a.nbr = f->d_nb;
    a.slot_edge = f->d_se;
    a.w = f->d_w;
    a.theta = theta;
    a.P = P;
    a.scale = scale;
    a.K = K;
    a.stats = stats;
    a.order = f->d_order;
    a.tol = o->tol > 0 ? o->tol : 1e-10;
    a.cg_rtol = o->cg_rtol > 0 ? o->cg_rtol : 1e-12;
    a.eta_max = o->eta_max > 0 ? o->eta_max : 1e-2;
    a.max_newton = o->max_newton > 0 ? o->max_newton : 50;
    a.cg_maxiter = o->cg_maxiter > 0 ? o->cg_maxiter : 20000;
    a.linearized = o->linearized ? 1 : 0;
    a.inexact = o->inexact ? 1 : 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t ns = f->small.size(), nl = f->large.size();
    int rc = ND_OK;
    if (!f->ev_fork) rc = herr(hipEventCreateWithFlags(&f->ev_fork, hipEventDisableTiming), "hipEventCreate");
    for (int c = 0; c < 2 && !rc; ++c) {
        if (!f->cls_stream[c]) rc = herr(hipStreamCreateWithFlags(&f->cls_stream[c], hipStreamNonBlocking), "hipStreamCreate");
        if (!rc && !f->ev_join[c]) rc = herr(hipEventCreateWithFlags(&f->ev_join[c], hipEventDisableTiming), "hipEventCreate");
    }
    if (rc) return rc;
    size_t lds_s = 0, lds_l = 0;
    if (ns) {
        lds_s = (2 * static_cast<size_t>(f->grids[f->small[0]].n) + 32) * sizeof(double);
        rc = herr(hipFuncSetAttribute(reinterpret_cast<const void*>(k_newton<4>), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      static_cast<int>(lds_s)), "hipFuncSetAttribute");
    }
    if (!rc && nl) {
        lds_l = (2 * static_cast<size_t>(f->grids[f->large[0]].n) + 32) * sizeof(double);
        rc = herr(hipFuncSetAttribute(reinterpret_cast<const void*>(k_newton<10>), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      static_cast<int>(lds_l)), "hipFuncSetAttribute");
    }
    rc = rc ? rc : herr(hipEventRecord(f->ev_fork, s), "hipEventRecord");
    if (rc) return rc;
    // large grids first: they take longest
    if (nl) {
        hipStream_t cs = f->cls_stream[1];
        rc = herr(hipStreamWaitEvent(cs, f->ev_fork, 0), "hipStreamWaitEvent");
        if (rc) return rc;
        a.first = static_cast<int32_t>(ns);
        hipLaunchKernelGGL(k_newton<10>, dim3(static_cast<unsigned>(nl)), dim3(kThreads), lds_l, cs, a);
        rc = herr(hipGetLastError(), "newton launch");
        rc = rc ? rc : herr(hipEventRecord(f->ev_join[1], cs), "hipEventRecord");
        rc = rc ? rc : herr(hipStreamWaitEvent(s, f->ev_join[1], 0), "hipStreamWaitEvent");
        if (rc) return rc;
    }
    if (ns) {
        hipStream_t cs = f->cls_stream[0];
        rc = herr(hipStreamWaitEvent(cs, f->ev_fork, 0), "hipStreamWaitEvent");
        if (rc) return rc;
        a.first = 0;
        hipLaunchKernelGGL(k_newton<4>, dim3(static_cast<unsigned>(ns)), dim3(kThreads), lds_s, cs, a);
        rc = herr(hipGetLastError(), "newton launch");
        rc = rc ? rc : herr(hipEventRecord(f->ev_join[0], cs), "hipEventRecord");
        rc = rc ? rc : herr(hipStreamWaitEvent(s, f->ev_join[0], 0), "hipStreamWaitEvent");
    }
    return rc;
}

}  // extern "C"
