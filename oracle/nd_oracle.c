/*
 * nd_oracle.c -- CPU ORACLE.  Test infrastructure only: a plain-C restatement
 * of NetworkDynamics.jl v1.1.0's network right-hand side, used by tests/ to
 * check the HIP library and by bench.py's cpu_baseline leg.  Nothing in the
 * product links or calls this file.
 *
 * Built with -O2 -ffp-contract=off (no FMA contraction) so each model's
 * arithmetic happens in the order the reference writes it.
 *
 * What each part follows (paths relative to the reference repository):
 *   construction / batching    src/construction.jl:31-256, 367-401; src/utils.jl:197-217
 *   IndexManager, registration src/network_structure.jl:1-55, 224-289
 *   BatchStride                src/utils.jl:1-42
 *   AggregationMap             src/aggregators.jl:54-86
 *   inverse aggregation map    src/aggregators.jl:203-235
 *   Eager gather map           src/gbufs.jl:11-31; Lazy provider src/gbufs.jl:44-66
 *   coreloop (phase order)     src/coreloop.jl:1-102, 176-248
 *   SequentialAggregator       src/aggregators.jl:140-151
 *   Threaded{Execution,Aggregator} src/coreloop.jl:121-129, src/aggregators.jl:192-201
 *   output wrappers            src/component_functions.jl:81-203
 *   model equations            test/ComponentLibrary.jl:8-72, benchmark/benchmark_models.jl:5-42,
 *                              test/diffusion_test.jl:57-107, docs/src/external_inputs.md:62-99
 *   external inputs            src/external_inputs.jl:10-77 (ExtMap, collect_externals!),
 *                              collected after the :g passes as coreloop.jl:56-57 does
 *   loopback / injectors       src/post_utils.jl:105-228 (LOOPBACK_G, apply_loopback!),
 *                              coreloop.jl:44-55 (copy, then the feed-forward vertex g);
 *                              models test/loopback_test.jl:19-131
 *
 * Pinning (tests/test_oracle.py): BatchStride known answers (test/utils_test.jl:147-178);
 * the complete_graph(4) maps derived by hand from the source (fixture
 * tests/golden/complete4_layout.json, cross-checked with
 * test/symbolicindexing_test.jl:84-113); Eager==Lazy and Sequential==Threaded==Naive
 * agreement (test/testutils.jl:10-64, test/aggregators_test.jl:58-79);
 * closed forms du = -L x (test/diffusion_test.jl:80-90) and the Kuramoto
 * incidence form dtheta = omega - K B sin(B^T theta) (oldexamples/paper.jl:10-11);
 * dense O(N^2) brute force; conservation invariants.
 * Parity unpinned: the last ulp of sin (Julia's own sin vs glibc's).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/nd_hip.h"

typedef struct { int64_t first, last; } rng_t;   /* Julia UnitRange, 1-based, empty: last = first-1 */

static int64_t rlen(rng_t r) { return r.last - r.first + 1; }

/* BatchStride (utils.jl:1-42): element i (1-based) range = first + (i-1)*sum(strides) + offset(j). */
typedef struct { int64_t first; int32_t n; int32_t s[2]; } bstride_t;
static int64_t bs_full(const bstride_t* b) { int64_t t = 0; for (int j = 0; j < b->n; ++j) t += b->s[j]; return t; }
static rng_t bs_range(const bstride_t* b, int64_t i) {
    int64_t st = b->first + (i - 1) * bs_full(b);
    rng_t r = {st, st + bs_full(b) - 1};
    return r;
}
static rng_t bs_subrange(const bstride_t* b, int64_t i, int j /* 1-based */) {
    int64_t off = 0;
    for (int q = 0; q < j - 1; ++q) off += b->s[q];
    int64_t st = b->first + (i - 1) * bs_full(b) + off;
    rng_t r = {st, st + b->s[j - 1] - 1};
    return r;
}

/* exported for the BatchStride known-answer test (test/utils_test.jl:147-178) */
void ndo_batchstride_range(int64_t first, int32_t n, const int32_t* strides, int64_t i, int32_t j,
                           int64_t* out_first, int64_t* out_last) {
    bstride_t b; b.first = first; b.n = n; b.s[0] = strides[0]; b.s[1] = n > 1 ? strides[1] : 0;
    rng_t r = (j == 0) ? bs_range(&b, i) : bs_subrange(&b, i, j);
    *out_first = r.first; *out_last = r.last;
}
int64_t ndo_batchstride_fullstride(int32_t n, const int32_t* strides) {
    bstride_t b; b.first = 1; b.n = n; b.s[0] = strides[0]; b.s[1] = n > 1 ? strides[1] : 0;
    return bs_full(&b);
}

typedef struct {
    int32_t kind;
    int64_t n;
    int64_t* idx;                         /* 1-based component ids */
    bstride_t state, p, in, out;
} batch_t;

typedef struct {
    int64_t nv, ne;
    int64_t *src, *dst;
    int32_t nkinds;
    nd_kind* kinds;
    int32_t *vkind, *ekind;
    /* IndexManager, network_structure.jl:1-31 */
    rng_t *v_data, *v_out, *v_para, *v_aggr;
    rng_t *e_data, *e_out_src, *e_out_dst, *e_para, *e_gb_src, *e_gb_dst;
    rng_t *v_ext, *e_ext;
    int32_t vdepth, edepth;
    int64_t lastidx_dynamic, lastidx_out, lastidx_p, lastidx_aggr, lastidx_gbuf, lastidx_extbuf;
    /* ExtMap (external_inputs.jl:10-31): > 0 StateBufIdx, < 0 OutBufIdx; and the extbuf cache */
    int64_t* extmap;
    double* extbuf;
    int32_t nvb, neb;
    batch_t *vb, *eb;
    /* AggregationMap, aggregators.jl:54-86 (tightened) */
    int64_t agg_first, agg_last;
    int64_t* aggmap;
    /* EagerGBufProvider map, gbufs.jl:11-22 */
    int64_t* gmap;
    /* _inv_aggregation_map, aggregators.jl:203-235: one (dst, srcs) entry per aggr index, vertex order */
    int64_t n_inv;
    int64_t *inv_dst, *inv_ptr, *inv_src;
    /* caches (construction.jl:205-208, gbufs.jl:20) */
    double *o, *aggbuf, *gbuf;
    char err[256];
} ndo_net;

static void* xcalloc(size_t n, size_t s) { return calloc(n ? n : 1, s); }

void ndo_net_destroy(ndo_net* nw) {
    if (!nw) return;
    free(nw->src); free(nw->dst); free(nw->kinds); free(nw->vkind); free(nw->ekind);
    free(nw->v_data); free(nw->v_out); free(nw->v_para); free(nw->v_aggr);
    free(nw->e_data); free(nw->e_out_src); free(nw->e_out_dst); free(nw->e_para);
    free(nw->e_gb_src); free(nw->e_gb_dst);
    free(nw->v_ext); free(nw->e_ext); free(nw->extmap); free(nw->extbuf);
    for (int b = 0; b < nw->nvb; ++b) free(nw->vb[b].idx);
    for (int b = 0; b < nw->neb; ++b) free(nw->eb[b].idx);
    free(nw->vb); free(nw->eb);
    free(nw->aggmap); free(nw->gmap);
    free(nw->inv_dst); free(nw->inv_ptr); free(nw->inv_src);
    free(nw->o); free(nw->aggbuf); free(nw->gbuf);
    free(nw);
}

/* find_identical (utils.jl:197-217) over the kind ids: groups in first-appearance order */
static int find_identical(const int32_t* key, int64_t n, batch_t** out) {
    int32_t ng = 0;
    int32_t* gk = (int32_t*)xcalloc((size_t)n, sizeof(int32_t));
    int64_t* cnt = (int64_t*)xcalloc((size_t)n, sizeof(int64_t));
    int32_t* gid = (int32_t*)xcalloc((size_t)n, sizeof(int32_t));
    for (int64_t i = 0; i < n; ++i) {
        int found = -1;
        for (int32_t j = 0; j < ng; ++j) if (gk[j] == key[i]) { found = j; break; }
        if (found < 0) { found = ng; gk[ng++] = key[i]; }
        gid[i] = found; cnt[found]++;
    }
    batch_t* bs = (batch_t*)xcalloc((size_t)ng, sizeof(batch_t));
    for (int32_t j = 0; j < ng; ++j) { bs[j].kind = gk[j]; bs[j].idx = (int64_t*)xcalloc((size_t)cnt[j], sizeof(int64_t)); }
    for (int64_t i = 0; i < n; ++i) { batch_t* b = &bs[gid[i]]; b->idx[b->n++] = i + 1; }
    free(gk); free(cnt); free(gid);
    *out = bs;
    return ng;
}

/* _nextrange (network_structure.jl:289) */
static rng_t nextrange(int64_t* last, int64_t N) { rng_t r = {*last + 1, *last + N}; *last += N; return r; }

ndo_net* ndo_net_create(int64_t nv, int64_t ne, const int64_t* src, const int64_t* dst,
                        int32_t nkinds, const nd_kind* kinds, const int32_t* vkind, const int32_t* ekind,
                        char* errbuf, int64_t errlen) {
    ndo_net* nw = (ndo_net*)xcalloc(1, sizeof(ndo_net));
#define FAIL(msg) do { if (errbuf && errlen > 0) { strncpy(errbuf, msg, (size_t)errlen - 1); errbuf[errlen - 1] = 0; } ndo_net_destroy(nw); return NULL; } while (0)
    nw->nv = nv; nw->ne = ne; nw->nkinds = nkinds;
    nw->src = (int64_t*)xcalloc((size_t)ne, 8); nw->dst = (int64_t*)xcalloc((size_t)ne, 8);
    memcpy(nw->src, src, (size_t)ne * 8); memcpy(nw->dst, dst, (size_t)ne * 8);
    nw->kinds = (nd_kind*)xcalloc((size_t)nkinds, sizeof(nd_kind)); memcpy(nw->kinds, kinds, (size_t)nkinds * sizeof(nd_kind));
    nw->vkind = (int32_t*)xcalloc((size_t)nv, 4); nw->ekind = (int32_t*)xcalloc((size_t)ne, 4);
    memcpy(nw->vkind, vkind, (size_t)nv * 4); memcpy(nw->ekind, ekind, (size_t)ne * 4);
    for (int64_t i = 0; i < nv; ++i) if (vkind[i] < 0 || vkind[i] >= nkinds || kinds[vkind[i]].is_edge) FAIL("bad vertex kind");
    for (int64_t i = 0; i < ne; ++i) if (ekind[i] < 0 || ekind[i] >= nkinds || !kinds[ekind[i]].is_edge) FAIL("bad edge kind");
    for (int64_t i = 0; i < ne; ++i) if (src[i] < 1 || src[i] > nv || dst[i] < 1 || dst[i] > nv) FAIL("edge endpoint out of range");
    if (nv < 1) FAIL("empty graph");

    /* construction.jl:99-106: equal vertex outdims / edge dst outdims */
    nw->vdepth = kinds[vkind[0]].outdim_dst;
    for (int64_t i = 0; i < nv; ++i) if (kinds[vkind[i]].outdim_dst != nw->vdepth) FAIL("All vertex models must have the same output dimension!");
    nw->edepth = ne ? kinds[ekind[0]].outdim_dst : 0;
    for (int64_t i = 0; i < ne; ++i) if (kinds[ekind[i]].outdim_dst != nw->edepth) FAIL("All edge models must have the same output dimension!");

    nw->v_data = (rng_t*)xcalloc((size_t)nv, sizeof(rng_t)); nw->v_out = (rng_t*)xcalloc((size_t)nv, sizeof(rng_t));
    nw->v_para = (rng_t*)xcalloc((size_t)nv, sizeof(rng_t)); nw->v_aggr = (rng_t*)xcalloc((size_t)nv, sizeof(rng_t));
    nw->e_data = (rng_t*)xcalloc((size_t)ne, sizeof(rng_t)); nw->e_out_src = (rng_t*)xcalloc((size_t)ne, sizeof(rng_t));
    nw->e_out_dst = (rng_t*)xcalloc((size_t)ne, sizeof(rng_t)); nw->e_para = (rng_t*)xcalloc((size_t)ne, sizeof(rng_t));
    nw->e_gb_src = (rng_t*)xcalloc((size_t)ne, sizeof(rng_t)); nw->e_gb_dst = (rng_t*)xcalloc((size_t)ne, sizeof(rng_t));
    nw->v_ext = (rng_t*)xcalloc((size_t)nv, sizeof(rng_t)); nw->e_ext = (rng_t*)xcalloc((size_t)ne, sizeof(rng_t));

    nw->nvb = find_identical(vkind, nv, &nw->vb);
    nw->neb = find_identical(ekind, ne, &nw->eb);

    /* VertexBatch -> register_vertices! (construction.jl:367-383, network_structure.jl:224-239) */
    for (int b = 0; b < nw->nvb; ++b) {
        batch_t* B = &nw->vb[b];
        const nd_kind* k = &kinds[B->kind];
        for (int64_t q = 0; q < B->n; ++q) {
            int64_t i = B->idx[q] - 1;
            nw->v_data[i] = nextrange(&nw->lastidx_dynamic, k->dim);
            nw->v_out[i] = nextrange(&nw->lastidx_out, k->outdim_dst);
            nw->v_para[i] = nextrange(&nw->lastidx_p, k->pdim);
            nw->v_aggr[i] = nextrange(&nw->lastidx_aggr, nw->edepth);
            nw->v_ext[i] = nextrange(&nw->lastidx_extbuf, k->extdim);
        }
        int64_t f = B->idx[0] - 1;
        B->state.first = nw->v_data[f].first; B->state.n = 1; B->state.s[0] = k->dim;
        B->p.first = nw->v_para[f].first; B->p.n = 1; B->p.s[0] = k->pdim;
        B->in.first = nw->v_aggr[f].first; B->in.n = 1; B->in.s[0] = nw->edepth;
        B->out.first = nw->v_out[f].first; B->out.n = 1; B->out.s[0] = k->outdim_dst;
    }
    /* EdgeBatch -> register_edges! (construction.jl:385-401, network_structure.jl:240-258) */
    for (int b = 0; b < nw->neb; ++b) {
        batch_t* B = &nw->eb[b];
        const nd_kind* k = &kinds[B->kind];
        for (int64_t q = 0; q < B->n; ++q) {
            int64_t i = B->idx[q] - 1;
            nw->e_data[i] = nextrange(&nw->lastidx_dynamic, k->dim);
            nw->e_out_src[i] = nextrange(&nw->lastidx_out, k->outdim_src);
            nw->e_out_dst[i] = nextrange(&nw->lastidx_out, k->outdim_dst);
            nw->e_para[i] = nextrange(&nw->lastidx_p, k->pdim);
            nw->e_gb_src[i] = nextrange(&nw->lastidx_gbuf, nw->vdepth);
            nw->e_gb_dst[i] = nextrange(&nw->lastidx_gbuf, nw->vdepth);
            nw->e_ext[i] = nextrange(&nw->lastidx_extbuf, k->extdim);
        }
        int64_t f = B->idx[0] - 1;
        B->state.first = nw->e_data[f].first; B->state.n = 1; B->state.s[0] = k->dim;
        B->p.first = nw->e_para[f].first; B->p.n = 1; B->p.s[0] = k->pdim;
        B->in.first = nw->e_gb_src[f].first; B->in.n = 2; B->in.s[0] = nw->vdepth; B->in.s[1] = nw->vdepth;
        B->out.first = nw->e_out_src[f].first; B->out.n = 2; B->out.s[0] = k->outdim_src; B->out.s[1] = k->outdim_dst;
    }

    /* AggregationMap (aggregators.jl:58-79) */
    int64_t* m = (int64_t*)xcalloc((size_t)nw->lastidx_out, 8);
    for (int b = 0; b < nw->neb; ++b) {
        batch_t* B = &nw->eb[b];
        for (int64_t q = 0; q < B->n; ++q) {
            int64_t e = B->idx[q] - 1;
            rng_t s = nw->e_out_dst[e], t = nw->v_aggr[nw->dst[e] - 1];
            if (rlen(s) != rlen(t)) FAIL("DimensionMismatch in aggregation map");
            for (int64_t j = 0; j < rlen(s); ++j) m[s.first - 1 + j] = t.first + j;
            s = nw->e_out_src[e];
            if (rlen(s) > 0) {
                t = nw->v_aggr[nw->src[e] - 1];
                if (rlen(s) != rlen(t)) FAIL("DimensionMismatch in aggregation map");
                for (int64_t j = 0; j < rlen(s); ++j) m[s.first - 1 + j] = t.first + j;
            }
        }
    }
    /* _tighten_idxrange (aggregators.jl:80-86) */
    int64_t fst = -1, lst = -1;
    for (int64_t k = 0; k < nw->lastidx_out; ++k) if (m[k] != 0) { fst = k; break; }
    for (int64_t k = nw->lastidx_out - 1; k >= 0; --k) if (m[k] != 0) { lst = k; break; }
    if (fst < 0) { nw->agg_first = 1; nw->agg_last = 0; nw->aggmap = (int64_t*)xcalloc(1, 8); }
    else {
        nw->agg_first = fst + 1; nw->agg_last = lst + 1;
        nw->aggmap = (int64_t*)xcalloc((size_t)(lst - fst + 1), 8);
        memcpy(nw->aggmap, m + fst, (size_t)(lst - fst + 1) * 8);
    }
    free(m);

    /* EagerGBufProvider map (gbufs.jl:11-22): iterates pairs(edgevec) */
    nw->gmap = (int64_t*)xcalloc((size_t)nw->lastidx_gbuf, 8);
    for (int64_t e = 0; e < ne; ++e) {
        rng_t s = nw->e_gb_src[e], t = nw->v_out[nw->src[e] - 1];
        for (int64_t j = 0; j < rlen(s); ++j) nw->gmap[s.first - 1 + j] = t.first + j;
        s = nw->e_gb_dst[e]; t = nw->v_out[nw->dst[e] - 1];
        for (int64_t j = 0; j < rlen(s); ++j) nw->gmap[s.first - 1 + j] = t.first + j;
    }

    /* _inv_aggregation_map (aggregators.jl:203-235): per vertex, per aggr element a list;
       dst outputs pushed before src outputs for each edge, batches and components in order. */
    {
        int64_t na = nw->lastidx_aggr;
        int64_t* cnt = (int64_t*)xcalloc((size_t)na + 1, 8);
        for (int b = 0; b < nw->neb; ++b) {
            batch_t* B = &nw->eb[b];
            for (int64_t q = 0; q < B->n; ++q) {
                int64_t e = B->idx[q] - 1;
                rng_t t = nw->v_aggr[nw->dst[e] - 1];
                for (int64_t j = 0; j < rlen(nw->e_out_dst[e]); ++j) cnt[t.first - 1 + j]++;
                if (rlen(nw->e_out_src[e]) > 0) {
                    t = nw->v_aggr[nw->src[e] - 1];
                    for (int64_t j = 0; j < rlen(nw->e_out_src[e]); ++j) cnt[t.first - 1 + j]++;
                }
            }
        }
        int64_t* ptr_by_aggr = (int64_t*)xcalloc((size_t)na + 1, 8);
        for (int64_t a = 0; a < na; ++a) ptr_by_aggr[a + 1] = ptr_by_aggr[a] + cnt[a];
        int64_t* lists = (int64_t*)xcalloc((size_t)ptr_by_aggr[na], 8);
        int64_t* fill = (int64_t*)xcalloc((size_t)na + 1, 8);
        for (int b = 0; b < nw->neb; ++b) {
            batch_t* B = &nw->eb[b];
            for (int64_t q = 0; q < B->n; ++q) {
                int64_t e = B->idx[q] - 1;
                rng_t t = nw->v_aggr[nw->dst[e] - 1], s = nw->e_out_dst[e];
                for (int64_t j = 0; j < rlen(s); ++j) { int64_t a = t.first - 1 + j; lists[ptr_by_aggr[a] + fill[a]++] = s.first + j; }
                s = nw->e_out_src[e];
                if (rlen(s) > 0) {
                    t = nw->v_aggr[nw->src[e] - 1];
                    for (int64_t j = 0; j < rlen(s); ++j) { int64_t a = t.first - 1 + j; lists[ptr_by_aggr[a] + fill[a]++] = s.first + j; }
                }
            }
        }
        /* ret in vertex order: for (dstr, srcs) in zip(v_aggr, srcidxs) */
        nw->n_inv = na;
        nw->inv_dst = (int64_t*)xcalloc((size_t)na, 8);
        nw->inv_ptr = (int64_t*)xcalloc((size_t)na + 1, 8);
        nw->inv_src = (int64_t*)xcalloc((size_t)ptr_by_aggr[na], 8);
        int64_t w = 0, pos = 0;
        for (int64_t v = 0; v < nv; ++v) {
            rng_t r = nw->v_aggr[v];
            for (int64_t a = r.first - 1; a < r.last; ++a) {
                nw->inv_dst[w] = a + 1;
                nw->inv_ptr[w] = pos;
                for (int64_t q = ptr_by_aggr[a]; q < ptr_by_aggr[a + 1]; ++q) nw->inv_src[pos++] = lists[q];
                ++w;
            }
        }
        nw->inv_ptr[w] = pos;
        free(cnt); free(ptr_by_aggr); free(lists); free(fill);
    }

    nw->o = (double*)xcalloc((size_t)nw->lastidx_out, 8);
    nw->aggbuf = (double*)xcalloc((size_t)nw->lastidx_aggr, 8);
    nw->gbuf = (double*)xcalloc((size_t)nw->lastidx_gbuf, 8);
    nw->extbuf = (double*)xcalloc((size_t)nw->lastidx_extbuf, 8);
    return nw;
#undef FAIL
}

/* ---- accessors for the tests ---------------------------------------------- */
void ndo_net_info(const ndo_net* nw, int64_t* out /* 11 */) {
    out[0] = nw->lastidx_dynamic; out[1] = nw->lastidx_p; out[2] = nw->lastidx_out;
    out[3] = nw->lastidx_aggr; out[4] = nw->lastidx_gbuf; out[5] = nw->agg_first; out[6] = nw->agg_last;
    out[7] = nw->vdepth; out[8] = nw->edepth; out[9] = nw->nvb; out[10] = nw->neb;
}
void ndo_net_ranges(const ndo_net* nw, int64_t* v /* nv*4*2 */, int64_t* e /* ne*6*2 */) {
    for (int64_t i = 0; i < nw->nv; ++i) {
        rng_t rs[4] = {nw->v_data[i], nw->v_out[i], nw->v_para[i], nw->v_aggr[i]};
        for (int j = 0; j < 4; ++j) { v[i * 8 + 2 * j] = rs[j].first; v[i * 8 + 2 * j + 1] = rs[j].last; }
    }
    for (int64_t i = 0; i < nw->ne; ++i) {
        rng_t rs[6] = {nw->e_data[i], nw->e_out_src[i], nw->e_out_dst[i], nw->e_para[i], nw->e_gb_src[i], nw->e_gb_dst[i]};
        for (int j = 0; j < 6; ++j) { e[i * 12 + 2 * j] = rs[j].first; e[i * 12 + 2 * j + 1] = rs[j].last; }
    }
}
void ndo_net_maps(const ndo_net* nw, int64_t* aggmap, int64_t* gmap) {
    int64_t n = nw->agg_last - nw->agg_first + 1;
    if (n > 0) memcpy(aggmap, nw->aggmap, (size_t)n * 8);
    if (nw->lastidx_gbuf > 0) memcpy(gmap, nw->gmap, (size_t)nw->lastidx_gbuf * 8);
}
int64_t ndo_net_extbuf_len(const ndo_net* nw) { return nw->lastidx_extbuf; }
void ndo_net_ext_ranges(const ndo_net* nw, int64_t* v /* nv*2 */, int64_t* e /* ne*2 */) {
    for (int64_t i = 0; i < nw->nv; ++i) { v[2 * i] = nw->v_ext[i].first; v[2 * i + 1] = nw->v_ext[i].last; }
    for (int64_t i = 0; i < nw->ne; ++i) { e[2 * i] = nw->e_ext[i].first; e[2 * i + 1] = nw->e_ext[i].last; }
}
/* ExtMap entries as the caller resolved them (StateBufIdx > 0, OutBufIdx < 0) */
int ndo_net_set_extmap(ndo_net* nw, int64_t len, const int64_t* map) {
    if (len != nw->lastidx_extbuf) return 1;
    free(nw->extmap);
    nw->extmap = (int64_t*)xcalloc((size_t)len, 8);
    memcpy(nw->extmap, map, (size_t)len * 8);
    return 0;
}
/* collect_externals!(map, extbuf, u, o), CPU version (external_inputs.jl:49-61) */
void ndo_collect_externals(const ndo_net* nw, double* extbuf, const double* u, const double* o) {
    for (int64_t k = 0; k < nw->lastidx_extbuf; ++k) {
        int64_t m = nw->extmap[k];
        extbuf[k] = m > 0 ? u[m - 1] : o[-m - 1];
    }
}
int64_t ndo_net_inv_nnz(const ndo_net* nw) { return nw->inv_ptr[nw->n_inv]; }
void ndo_net_inv(const ndo_net* nw, int64_t* dst, int64_t* ptr, int64_t* srcs) {
    memcpy(dst, nw->inv_dst, (size_t)nw->n_inv * 8);
    memcpy(ptr, nw->inv_ptr, (size_t)(nw->n_inv + 1) * 8);
    memcpy(srcs, nw->inv_src, (size_t)nw->inv_ptr[nw->n_inv] * 8);
}
void ndo_net_batch(const ndo_net* nw, int32_t is_edge, int32_t b, int64_t* out /* 6 */, int64_t* idx) {
    const batch_t* B = is_edge ? &nw->eb[b] : &nw->vb[b];
    out[0] = B->kind; out[1] = B->n; out[2] = B->state.first; out[3] = B->p.first; out[4] = B->out.first; out[5] = B->in.first;
    if (idx) memcpy(idx, B->idx, (size_t)B->n * 8);
}

/* ---- models ---------------------------------------------------------------- */

static int hasff(const nd_kind* k) { return k->ff == ND_FF_PUREFEEDFORWARD || k->ff == ND_FF_FEEDFORWARD; }

/* vertex f (apply_compf, coreloop.jl:213-216); ext: the component's external inputs */
static void vertex_f(const nd_kind* k, double* du, const double* u, const double* acc, const double* ext,
                     const double* p, double t) {
    switch (k->model) {
    case ND_VM_DIFFUSION:          /* ComponentLibrary.jl:42-45; with d outputs dv .= esum
                                      (getting_started_with_network_dynamics.jl:129-133) */
        for (int j = 0; j < k->dim; ++j) du[j] = acc[j];
        break;
    case ND_VM_TRUSS_FREE: {       /* stress_on_truss.jl:55-61: free_f */
        du[0] = (acc[0] - p[1] * u[0]) / p[0];
        du[1] = (acc[1] - p[1] * u[1]) / p[0];
        du[1] = du[1] - p[2];
        du[2] = u[0];
        du[3] = u[1];
        break;
    }
    case ND_VM_FHN:                /* directed_and_weighted_graphs.jl:89-93; v^3 = v*v*v */
        du[0] = u[0] - u[0] * u[0] * u[0] / 3.0 - u[1] + acc[0];
        du[1] = (u[0] - p[0]) * p[1];
        break;
    case ND_VM_KURAMOTO:           /* ComponentLibrary.jl:69-71 */
        du[0] = p[0] + acc[0];
        break;
    case ND_VM_SWING: {            /* ComponentLibrary.jl:59-63 */
        double M = p[0], D = p[1], Pm = p[2];
        du[0] = u[1];
        du[1] = 1.0 / M * (Pm - D * u[1] + acc[0]);
        break;
    }
    case ND_VM_SWING_UNIT_DAMPING: /* benchmark_models.jl:37-41 */
        du[0] = u[1];
        du[1] = p[0] - 1.0 * u[1];
        du[1] += acc[0];
        break;
    case ND_VM_LOAD: {             /* build-defined first-order load */
        double tau = p[0], P = p[1];
        du[0] = 1.0 / tau * (P + acc[0]);
        break;
    }
    case ND_VM_CAPACITOR_SINK:     /* docs/src/external_inputs.md:76-79: f_load */
        du[0] = 1.0 / p[0] * (acc[0] - (1.0 + 0.1 * sin(t)));
        break;
    case ND_VM_CAPACITOR:          /* loopback_test.jl:61: D(p.v) ~ p.i / C */
        du[0] = acc[0] / p[0];
        break;
    case ND_VM_INDUCTOR_INJECTOR:  /* loopback_test.jl:89: D(p.i) ~ p.v / L */
        du[0] = acc[0] / p[0];
        break;
    case ND_VM_CRL: {              /* loopback_test.jl:113-131: CRLModel */
        double C = p[0], R = p[1], L = p[2];
        double v = u[0], iL = u[1];
        du[0] = (acc[0] - v / R - iL) / C;
        du[1] = v / L;
        break;
    }
    case ND_VM_PI_SOURCE: {        /* docs/src/external_inputs.md:87-92: f_source */
        double C = p[0], Vref = p[1], Ki = p[2], Kp = p[3];
        double d = Vref - ext[0];
        du[1] = d;
        double i_inj = Kp * d + Ki * u[1];
        du[0] = 1.0 / C * (acc[0] + i_inj);
        break;
    }
    default:
        break;
    }
}

/* single-sided (dst) output of a pure feed-forward edge model */
static void edge_gdst(const nd_kind* k, double* y, const double* vs, const double* vd, const double* p, int32_t depth) {
    switch (k->model) {
    case ND_EM_DIFFUSION:          /* diffusion_test.jl:57-60: e .= v_s .- v_d */
        for (int j = 0; j < depth; ++j) y[j] = vs[j] - vd[j];
        break;
    case ND_EM_DIFFUSION_WEIGHTED: /* ComponentLibrary.jl:8-10 */
        y[0] = p[0] * (vs[0] - vd[0]);
        break;
    case ND_EM_BEAM: {             /* stress_on_truss.jl:77-84: edge_g!; x^2 = x*x */
        const double dx = vd[0] - vs[0], dy = vd[1] - vs[1];
        const double d = sqrt(dx * dx + dy * dy);
        const double fabs_ = p[0] * (p[1] - d);
        y[0] = fabs_ * dx / d;
        y[1] = fabs_ * dy / d;
        break;
    }
    case ND_EM_ELECTRICAL:         /* directed_and_weighted_graphs.jl:97-100 */
        y[0] = p[0] * (vs[0] - vd[0]) * p[1];
        break;
    case ND_EM_KURAMOTO:           /* ComponentLibrary.jl:51-53 */
        y[0] = p[0] * sin(vs[0] - vd[0]);
        break;
    case ND_EM_STATIC_LINE:        /* build-defined static line with phase shift */
        y[0] = p[0] * sin(vs[0] - vd[0] - p[1]);
        break;
    case ND_EM_RESISTOR:           /* docs/src/external_inputs.md:66-68: resistor_g */
        y[0] = (vs[0] - vd[0]) / p[0];
        break;
    case ND_EM_LOOPBACK:           /* post_utils.jl:105-108: outdst .= -1 .* insrc */
        for (int j = 0; j < depth; ++j) y[j] = -1.0 * vs[j];
        break;
    default:
        break;
    }
}

/* edge g: the wrappers of component_functions.jl:81-203 around the model */
static void edge_g(const nd_kind* k, double* osrc, double* odst, const double* ue, const double* vs,
                   const double* vd, const double* p) {
    int32_t ns = k->outdim_src, nd = k->outdim_dst;
    if (k->model == ND_EM_NOOP) return;
    if (k->ff == ND_FF_PURESTATEMAP) {
        /* Fiducial(StateMask src, StateMask dst): each side its own mask; a StateMask
           inside AntiSymmetric / Symmetric / Directed: y_dst = u[mask], then
           y_src = -y_dst / y_dst / nothing (component_functions.jl:117-175) */
        for (int j = 0; j < nd; ++j) odst[j] = ue[k->mask_dst[j] - 1];
        if (k->wrapper == ND_WRAP_ANTISYMMETRIC)
            for (int j = 0; j < ns; ++j) osrc[j] = -odst[j];
        else
            for (int j = 0; j < ns; ++j) osrc[j] = ue[k->mask_src[j] - 1];
        return;
    }
    if (k->model == ND_EM_DIFFUSION_FIDUCIAL) {   /* ComponentLibrary.jl:22-25 */
        odst[0] = p[0] * (vs[0] - vd[0]);
        osrc[0] = -odst[0];
        return;
    }
    edge_gdst(k, odst, vs, vd, p, nd);
    switch (k->wrapper) {
    case ND_WRAP_ANTISYMMETRIC: for (int j = 0; j < ns; ++j) osrc[j] = -odst[j]; break;
    case ND_WRAP_SYMMETRIC:     for (int j = 0; j < ns; ++j) osrc[j] = odst[j]; break;
    case ND_WRAP_DIRECTED:      for (int j = 0; j < ns; ++j) osrc[j] = 0.0; break;
    default: break;
    }
}

/* edge f for ODE edges; ext: the edge's external inputs (external_inputs.jl) */
static void edge_f(const nd_kind* k, double* de, const double* e, const double* vs, const double* vd, const double* ext,
                   const double* p) {
    switch (k->model) {
    case ND_EM_ODE_SIN_DIFFUSION:      /* ComponentLibrary.jl:30-34 */
        de[0] = 1.0 / p[0] * (sin(vs[0] - vd[0]) - e[0]);
        de[1] = 1.0 / p[0] * (sin(vd[0] - vs[0]) - e[1]);
        break;
    case ND_EM_ODE_SIN_DIFFUSION_FAST: /* benchmark_models.jl:11-15 */
        de[0] = 100.0 * (sin(vs[0] - vd[0]) - e[0]);
        de[1] = 100.0 * (sin(vd[0] - vs[0]) - e[1]);
        break;
    case ND_EM_ODE_LINEAR:             /* diffusion_test.jl:102-106 */
        de[0] = vs[0] - vd[0] - e[0];
        de[1] = vd[0] - vs[0] - e[1];
        break;
    case ND_EM_ODE_LAG:                /* ComponentLibrary.jl:30-32 with the state of the dst side only */
        de[0] = 1.0 / p[0] * (sin(vs[0] - vd[0]) - e[0]);
        break;
    case ND_EM_EXT_DRIVEN:             /* external_inputs_test.jl:45-47: de[1] = ext[1] */
        de[0] = ext[0];
        break;
    default:
        break;
    }
}

static int edge_has_f(const nd_kind* k) {
    return k->model == ND_EM_ODE_SIN_DIFFUSION || k->model == ND_EM_ODE_SIN_DIFFUSION_FAST ||
           k->model == ND_EM_ODE_LINEAR || k->model == ND_EM_ODE_LAG || k->model == ND_EM_EXT_DRIVEN;
}
static int vertex_has_f(const nd_kind* k) { (void)k; return 1; }

/* ---- coreloop ----------------------------------------------------------------- */

/* apply_comp! for one vertex (coreloop.jl:176-192), BatchStride-indexed */
static void apply_vertex(const ndo_net* nw, const batch_t* B, int64_t i, int pass_g, double* du,
                         const double* u, double* o, const double* aggbuf, const double* p, double t) {
    const nd_kind* k = &nw->kinds[B->kind];
    rng_t sr = bs_range(&B->state, i), pr = bs_range(&B->p, i);
    if (pass_g) {
        rng_t orr = bs_range(&B->out, i);
        double* out = o + orr.first - 1;
        if (k->ff == ND_FF_PURESTATEMAP) {          /* StateMask: g(o, u) */
            for (int j = 0; j < k->outdim_dst; ++j) out[j] = u[sr.first - 1 + k->mask_dst[j] - 1];
        } else if (k->model == ND_VM_TRUSS_FIXED) {    /* stress_on_truss.jl:45-47: pos .= p */
            for (int j = 0; j < k->outdim_dst; ++j) out[j] = p[pr.first - 1 + j];
        } else if (k->model == ND_VM_VOLTAGE_SOURCE) { /* loopback_test.jl:31: p.v ~ V (NoFeedForward) */
            out[0] = p[pr.first - 1];
        } else if (k->model == ND_VM_RESISTOR_INJECTOR) { /* loopback_test.jl:75: p.i ~ p.v / R, input = hub */
            rng_t ir = bs_range(&B->in, i);
            out[0] = aggbuf[ir.first - 1] / p[pr.first - 1];
        }
    } else {
        rng_t ir = bs_range(&B->in, i);
        const double* ext = nw->extbuf + nw->v_ext[B->idx[i - 1] - 1].first - 1;
        vertex_f(k, du + sr.first - 1, u + sr.first - 1, aggbuf + ir.first - 1, ext, p + pr.first - 1, t);
    }
}

/* apply_comp! for one edge (coreloop.jl:194-211) */
static void apply_edge(const ndo_net* nw, const batch_t* B, int64_t i, int evalf, int evalg, int buffered,
                       double* du, const double* u, double* o, const double* gbuf, const double* p, double t) {
    (void)t;
    const nd_kind* k = &nw->kinds[B->kind];
    rng_t sr = bs_range(&B->state, i), pr = bs_range(&B->p, i);
    const double *vs, *vd;
    if (buffered) {   /* get_src_dst on the gathered buffer (gbufs.jl:27-31) */
        vs = gbuf + bs_subrange(&B->in, i, 1).first - 1;
        vd = gbuf + bs_subrange(&B->in, i, 2).first - 1;
    } else {          /* LazyGBuf: views into o through v_out (gbufs.jl:58-66) */
        int64_t e = B->idx[i - 1] - 1;
        vs = o + nw->v_out[nw->src[e] - 1].first - 1;
        vd = o + nw->v_out[nw->dst[e] - 1].first - 1;
    }
    if (evalf && edge_has_f(k))
        edge_f(k, du + sr.first - 1, u + sr.first - 1, vs, vd, nw->extbuf + nw->e_ext[B->idx[i - 1] - 1].first - 1,
               p + pr.first - 1);
    if (evalg) {
        rng_t os = bs_subrange(&B->out, i, 1), od = bs_subrange(&B->out, i, 2);
        edge_g(k, o + os.first - 1, o + od.first - 1, u + sr.first - 1, vs, vd, p + pr.first - 1);
    }
}

static void fill_d(double* x, int64_t n, double v) { for (int64_t i = 0; i < n; ++i) x[i] = v; }

/* SequentialAggregator aggregate! (aggregators.jl:140-151); additive */
void ndo_aggregate_seq(const ndo_net* nw, double* aggbuf, const double* data) {
    int64_t n = nw->agg_last - nw->agg_first + 1;
    for (int64_t k = 0; k < n; ++k) {
        int64_t d = nw->aggmap[k];
        if (d != 0) aggbuf[d - 1] = aggbuf[d - 1] + data[nw->agg_first - 1 + k];
    }
}
/* ThreadedAggregator aggregate! over the inverse map (aggregators.jl:192-201); additive */
void ndo_aggregate_inv(const ndo_net* nw, double* aggbuf, const double* data, int32_t nthreads) {
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int64_t w = 0; w < nw->n_inv; ++w) {
        int64_t d = nw->inv_dst[w] - 1;
        double acc = aggbuf[d];
        for (int64_t q = nw->inv_ptr[w]; q < nw->inv_ptr[w + 1]; ++q) acc = acc + data[nw->inv_src[q] - 1];
        aggbuf[d] = acc;
    }
    (void)nthreads;
}
/* NaiveAggregator (aggregators.jl:29-51); additive */
void ndo_aggregate_naive(const ndo_net* nw, double* aggbuf, const double* data) {
    for (int b = 0; b < nw->neb; ++b) {
        const batch_t* B = &nw->eb[b];
        for (int64_t q = 0; q < B->n; ++q) {
            int64_t e = B->idx[q] - 1;
            rng_t s = nw->e_out_dst[e], t = nw->v_aggr[nw->dst[e] - 1];
            for (int64_t j = 0; j < rlen(s); ++j) aggbuf[t.first - 1 + j] = aggbuf[t.first - 1 + j] + data[s.first - 1 + j];
            s = nw->e_out_src[e];
            if (rlen(s) > 0) {
                t = nw->v_aggr[nw->src[e] - 1];
                for (int64_t j = 0; j < rlen(s); ++j) aggbuf[t.first - 1 + j] = aggbuf[t.first - 1 + j] + data[s.first - 1 + j];
            }
        }
    }
}
/* gather!(EagerGBufProvider): NNlib.gather!(gbuf, o, map) (gbufs.jl:25) */
void ndo_gather(const ndo_net* nw, double* gbuf, const double* o) {
    for (int64_t k = 0; k < nw->lastidx_gbuf; ++k) gbuf[k] = o[nw->gmap[k] - 1];
}

/* apply_loopback! (post_utils.jl:221-228): each injector's input is its hub's output */
static void apply_loopback(const ndo_net* nw, double* aggbuf, const double* o) {
    for (int64_t e = 0; e < nw->ne; ++e) {
        if (nw->kinds[nw->ekind[e]].model != ND_EM_LOOPBACK) continue;
        rng_t a = nw->v_aggr[nw->src[e] - 1], b = nw->v_out[nw->dst[e] - 1];
        for (int64_t j = 0; j < rlen(a); ++j) aggbuf[a.first - 1 + j] = o[b.first - 1 + j];
    }
}

/* Full RHS, SequentialExecution{buffered} + SequentialAggregator(+). Returns 0, or 1 on size mismatch
   (the ArgumentError of coreloop.jl:2-7).  If o_out/agg_out are non-NULL they receive the buffers
   (get_buffers with initbufs=true, coreloop.jl:103-109; then du may be NULL). */
/* apply_perturb!(buf, perturb, map), coreloop.jl:258-262: map holds n pairs
   (buffer index, channel), both 1-based */
static void apply_perturb(double* buf, const double* delta, const int64_t* map, int64_t n) {
    for (int64_t k = 0; k < n; ++k) buf[map[2 * k] - 1] += delta[map[2 * k + 1] - 1];
}

/* The perturbation maps of nw(du, u, p, t; perturb, perturb_maps): vo (output buffer,
   vertex outputs), vi (aggregation buffer), eo (output buffer, edge outputs), ei (gather
   buffer); each n pairs (buffer index, channel), 1-based, flattened. */
typedef struct { const int64_t* map[4]; int64_t n[4]; const double* delta; } perturb_t;

static int rhs_impl(ndo_net* nw, double* du, const double* u, const double* p, double t, int32_t buffered,
                    double* o_out, double* agg_out, const perturb_t* pt);

int ndo_rhs(ndo_net* nw, double* du, const double* u, const double* p, double t, int32_t buffered,
            double* o_out, double* agg_out) {
    return rhs_impl(nw, du, u, p, t, buffered, o_out, agg_out, NULL);
}

/* nw(du, u, p, t; perturb=delta, perturb_maps) in the buffered loop (coreloop.jl:1-102):
   delta enters after the loopback copy (vi), after collect_externals! (vo), after the
   gather (ei) and before the aggregation (eo).  o_out/agg_out: the perturbed buffers. */
int ndo_rhs_perturbed(ndo_net* nw, double* du, const double* u, const double* p, double t, const double* delta,
                      const int64_t* vo, int64_t nvo, const int64_t* vi, int64_t nvi, const int64_t* eo, int64_t neo,
                      const int64_t* ei, int64_t nei, double* o_out, double* agg_out) {
    perturb_t pt = {{vo, vi, eo, ei}, {nvo, nvi, neo, nei}, delta};
    return rhs_impl(nw, du, u, p, t, 1, o_out, agg_out, &pt);
}

static int rhs_impl(ndo_net* nw, double* du, const double* u, const double* p, double t, int32_t buffered,
                    double* o_out, double* agg_out, const perturb_t* pt) {
    double* o = nw->o;
    double* aggbuf = nw->aggbuf;
    double* gbuf = nw->gbuf;
    if (du) fill_d(du, nw->lastidx_dynamic, 0.0);                                /* coreloop.jl:24 */
    fill_d(o, nw->lastidx_out, NAN);                                             /* network_structure.jl:153 */
    fill_d(aggbuf, nw->lastidx_aggr, 0.0);                                       /* coreloop.jl:30 */
    for (int b = 0; b < nw->nvb; ++b) {                                          /* coreloop.jl:39 */
        const batch_t* B = &nw->vb[b];
        if (hasff(&nw->kinds[B->kind])) continue;
        for (int64_t i = 1; i <= B->n; ++i) apply_vertex(nw, B, i, 1, du, u, o, aggbuf, p, t);
    }
    for (int b = 0; b < nw->neb; ++b) {                                          /* coreloop.jl:41 */
        const batch_t* B = &nw->eb[b];
        if (hasff(&nw->kinds[B->kind])) continue;
        for (int64_t i = 1; i <= B->n; ++i) apply_edge(nw, B, i, 0, 1, buffered, du, u, o, gbuf, p, t);
    }
    apply_loopback(nw, aggbuf, o);                                               /* coreloop.jl:47 */
    if (pt) apply_perturb(aggbuf, pt->delta, pt->map[1], pt->n[1]);               /* coreloop.jl:49-52 */
    for (int b = 0; b < nw->nvb; ++b) {                                          /* coreloop.jl:55 */
        const batch_t* B = &nw->vb[b];
        if (!hasff(&nw->kinds[B->kind])) continue;
        for (int64_t i = 1; i <= B->n; ++i) apply_vertex(nw, B, i, 1, du, u, o, aggbuf, p, t);
    }
    if (nw->lastidx_extbuf > 0) ndo_collect_externals(nw, nw->extbuf, u, o);     /* coreloop.jl:56-57 */
    if (pt) apply_perturb(o, pt->delta, pt->map[0], pt->n[0]);                    /* coreloop.jl:63-65 */
    if (buffered) ndo_gather(nw, gbuf, o);                                       /* coreloop.jl:67 */
    if (pt) apply_perturb(gbuf, pt->delta, pt->map[3], pt->n[3]);                 /* coreloop.jl:69-72 */
    for (int b = 0; b < nw->neb; ++b) {                                          /* coreloop.jl:76 */
        const batch_t* B = &nw->eb[b];
        if (hasff(&nw->kinds[B->kind])) continue;
        if (!du) continue;
        for (int64_t i = 1; i <= B->n; ++i) apply_edge(nw, B, i, 1, 0, buffered, du, u, o, gbuf, p, t);
    }
    for (int b = 0; b < nw->neb; ++b) {                                          /* coreloop.jl:78 */
        const batch_t* B = &nw->eb[b];
        if (!hasff(&nw->kinds[B->kind])) continue;
        for (int64_t i = 1; i <= B->n; ++i) apply_edge(nw, B, i, du != NULL, 1, buffered, du, u, o, gbuf, p, t);
    }
    if (pt) apply_perturb(o, pt->delta, pt->map[2], pt->n[2]);                    /* coreloop.jl:86-88 */
    ndo_aggregate_seq(nw, aggbuf, o);                                            /* coreloop.jl:90 */
    if (o_out) memcpy(o_out, o, (size_t)nw->lastidx_out * 8);
    if (agg_out) memcpy(agg_out, aggbuf, (size_t)nw->lastidx_aggr * 8);
    if (!du) return 0;
    for (int b = 0; b < nw->nvb; ++b) {                                          /* coreloop.jl:97 */
        const batch_t* B = &nw->vb[b];
        if (hasff(&nw->kinds[B->kind]) || !vertex_has_f(&nw->kinds[B->kind])) continue;
        for (int64_t i = 1; i <= B->n; ++i) apply_vertex(nw, B, i, 0, du, u, o, aggbuf, p, t);
    }
    return 0;
}

/* coreloop.jl:2-7 size checks, for the error-path tests */
int ndo_check_sizes(const ndo_net* nw, int64_t len_du, int64_t len_u, int64_t len_p) {
    if (!(len_du == len_u && len_u == nw->lastidx_dynamic)) return 1;
    if (nw->lastidx_p > 0 && len_p != nw->lastidx_p) return 2;
    return 0;
}

/* ThreadedExecution{true} + ThreadedAggregator(+): coreloop.jl:121-129, aggregators.jl:192-201.
   Per batch and phase, components are split over threads; fills and the NNlib gather stay serial
   as in the reference's CPU path. */
int ndo_rhs_threaded(ndo_net* nw, double* du, const double* u, const double* p, double t, int32_t nthreads) {
    double* o = nw->o;
    double* aggbuf = nw->aggbuf;
    double* gbuf = nw->gbuf;
    if (nthreads < 1) nthreads = 1;
    fill_d(du, nw->lastidx_dynamic, 0.0);
    fill_d(o, nw->lastidx_out, NAN);
    fill_d(aggbuf, nw->lastidx_aggr, 0.0);
    for (int b = 0; b < nw->nvb; ++b) {
        const batch_t* B = &nw->vb[b];
        if (hasff(&nw->kinds[B->kind])) continue;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
        for (int64_t i = 1; i <= B->n; ++i) apply_vertex(nw, B, i, 1, du, u, o, aggbuf, p, t);
    }
    for (int b = 0; b < nw->neb; ++b) {
        const batch_t* B = &nw->eb[b];
        if (hasff(&nw->kinds[B->kind])) continue;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
        for (int64_t i = 1; i <= B->n; ++i) apply_edge(nw, B, i, 0, 1, 1, du, u, o, gbuf, p, t);
    }
    apply_loopback(nw, aggbuf, o);
    for (int b = 0; b < nw->nvb; ++b) {
        const batch_t* B = &nw->vb[b];
        if (!hasff(&nw->kinds[B->kind])) continue;
        for (int64_t i = 1; i <= B->n; ++i) apply_vertex(nw, B, i, 1, du, u, o, aggbuf, p, t);
    }
    if (nw->lastidx_extbuf > 0) ndo_collect_externals(nw, nw->extbuf, u, o);
    ndo_gather(nw, gbuf, o);
    for (int b = 0; b < nw->neb; ++b) {
        const batch_t* B = &nw->eb[b];
        if (hasff(&nw->kinds[B->kind])) continue;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
        for (int64_t i = 1; i <= B->n; ++i) apply_edge(nw, B, i, 1, 0, 1, du, u, o, gbuf, p, t);
    }
    for (int b = 0; b < nw->neb; ++b) {
        const batch_t* B = &nw->eb[b];
        if (!hasff(&nw->kinds[B->kind])) continue;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
        for (int64_t i = 1; i <= B->n; ++i) apply_edge(nw, B, i, 1, 1, 1, du, u, o, gbuf, p, t);
    }
    ndo_aggregate_inv(nw, aggbuf, o, nthreads);
    for (int b = 0; b < nw->nvb; ++b) {
        const batch_t* B = &nw->vb[b];
        if (hasff(&nw->kinds[B->kind])) continue;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
        for (int64_t i = 1; i <= B->n; ++i) apply_vertex(nw, B, i, 0, du, u, o, aggbuf, p, t);
    }
    return 0;
}

/* Per-element magnitude scale for the parity tolerance: the same RHS with every
   term replaced by its absolute value (|terms| summed), so |du_gpu - du| <= rtol * S
   is meaningful where du itself cancels to ~0. */
void ndo_rhs_scale(ndo_net* nw, double* S, const double* u, const double* p, double t) {
    int64_t nout = nw->lastidx_out;
    double* o = nw->o;
    double* agg = nw->aggbuf;
    fill_d(S, nw->lastidx_dynamic, 0.0);
    ndo_rhs(nw, NULL, u, p, t, 1, NULL, NULL);    /* fills o */
    for (int64_t k = 0; k < nout; ++k) o[k] = fabs(o[k]);
    fill_d(agg, nw->lastidx_aggr, 0.0);
    apply_loopback(nw, agg, o);                   /* |hub output| as the injectors' input */
    ndo_aggregate_seq(nw, agg, o);
    for (int b = 0; b < nw->nvb; ++b) {
        const batch_t* B = &nw->vb[b];
        const nd_kind* k = &nw->kinds[B->kind];
        for (int64_t i = 1; i <= B->n; ++i) {
            rng_t sr = bs_range(&B->state, i), pr = bs_range(&B->p, i), ir = bs_range(&B->in, i);
            double* s = S + sr.first - 1;
            const double* x = u + sr.first - 1;
            const double* q = p + pr.first - 1;
            double a = agg[ir.first - 1];
            switch (k->model) {
            case ND_VM_DIFFUSION: for (int j = 0; j < k->dim; ++j) s[j] = agg[ir.first - 1 + j]; break;
            case ND_VM_KURAMOTO: s[0] = fabs(q[0]) + a; break;
            case ND_VM_TRUSS_FREE: {
                const double a0 = agg[ir.first - 1], a1 = agg[ir.first];
                s[0] = (a0 + fabs(q[1] * x[0])) / fabs(q[0]);
                s[1] = (a1 + fabs(q[1] * x[1])) / fabs(q[0]) + fabs(q[2]);
                s[2] = fabs(x[0]);
                s[3] = fabs(x[1]);
                break;
            }
            case ND_VM_FHN:
                s[0] = fabs(x[0]) + fabs(x[0] * x[0] * x[0] / 3.0) + fabs(x[1]) + a;
                s[1] = fabs(x[0] * q[1]) + fabs(q[0] * q[1]);
                break;
            case ND_VM_SWING: s[0] = fabs(x[1]); s[1] = fabs(1.0 / q[0]) * (fabs(q[2]) + fabs(q[1] * x[1]) + a); break;
            case ND_VM_SWING_UNIT_DAMPING: s[0] = fabs(x[1]); s[1] = fabs(q[0]) + fabs(x[1]) + a; break;
            case ND_VM_LOAD: s[0] = fabs(1.0 / q[0]) * (fabs(q[1]) + a); break;
            case ND_VM_CAPACITOR_SINK: s[0] = fabs(1.0 / q[0]) * (a + 1.0 + fabs(0.1 * sin(t))); break;
            case ND_VM_CAPACITOR: case ND_VM_INDUCTOR_INJECTOR: s[0] = a / fabs(q[0]); break;
            case ND_VM_CRL: s[0] = (a + fabs(x[0] / q[1]) + fabs(x[1])) / fabs(q[0]); s[1] = fabs(x[0] / q[2]); break;
            case ND_VM_PI_SOURCE: {
                const double ex = fabs(nw->extbuf[nw->v_ext[B->idx[i - 1] - 1].first - 1]);
                s[1] = fabs(q[1]) + ex;
                s[0] = fabs(1.0 / q[0]) * (a + fabs(q[3]) * s[1] + fabs(q[2] * x[1]));
                break;
            }
            default: break;
            }
        }
    }
    for (int b = 0; b < nw->neb; ++b) {
        const batch_t* B = &nw->eb[b];
        const nd_kind* k = &nw->kinds[B->kind];
        if (!edge_has_f(k)) continue;
        for (int64_t i = 1; i <= B->n; ++i) {
            int64_t e = B->idx[i - 1] - 1;
            rng_t sr = bs_range(&B->state, i), pr = bs_range(&B->p, i);
            double vs = u[nw->v_data[nw->src[e] - 1].first - 1 + nw->kinds[nw->vkind[nw->src[e] - 1]].mask_dst[0] - 1];
            double vd = u[nw->v_data[nw->dst[e] - 1].first - 1 + nw->kinds[nw->vkind[nw->dst[e] - 1]].mask_dst[0] - 1];
            double* s = S + sr.first - 1;
            const double* x = u + sr.first - 1;
            if (k->model == ND_EM_EXT_DRIVEN) {
                s[0] = fabs(nw->extbuf[nw->e_ext[e].first - 1]);
                continue;
            }
            double c = (k->model == ND_EM_ODE_SIN_DIFFUSION || k->model == ND_EM_ODE_LAG)
                           ? fabs(1.0 / p[pr.first - 1])
                           : (k->model == ND_EM_ODE_SIN_DIFFUSION_FAST ? 100.0 : 1.0);
            double dv = k->model == ND_EM_ODE_LINEAR ? fabs(vs) + fabs(vd) : fabs(sin(vs - vd));
            for (int j = 0; j < k->dim; ++j) s[j] = c * (dv + fabs(x[j]));
        }
    }
}

/* Classic fourth-order Runge-Kutta with fixed dt, in place on u (textbook RK4). */
int ndo_rk4(ndo_net* nw, double* u, const double* p, double t0, double dt, int64_t nsteps) {
    int64_t n = nw->lastidx_dynamic;
    double* k1 = (double*)xcalloc((size_t)n, 8);
    double* k2 = (double*)xcalloc((size_t)n, 8);
    double* k3 = (double*)xcalloc((size_t)n, 8);
    double* k4 = (double*)xcalloc((size_t)n, 8);
    double* y = (double*)xcalloc((size_t)n, 8);
    double t = t0;
    for (int64_t s = 0; s < nsteps; ++s) {
        ndo_rhs(nw, k1, u, p, t, 1, NULL, NULL);
        for (int64_t i = 0; i < n; ++i) y[i] = u[i] + 0.5 * dt * k1[i];
        ndo_rhs(nw, k2, y, p, t + 0.5 * dt, 1, NULL, NULL);
        for (int64_t i = 0; i < n; ++i) y[i] = u[i] + 0.5 * dt * k2[i];
        ndo_rhs(nw, k3, y, p, t + 0.5 * dt, 1, NULL, NULL);
        for (int64_t i = 0; i < n; ++i) y[i] = u[i] + dt * k3[i];
        ndo_rhs(nw, k4, y, p, t + dt, 1, NULL, NULL);
        for (int64_t i = 0; i < n; ++i) u[i] = u[i] + dt / 6.0 * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
        t = t0 + (double)(s + 1) * dt;
    }
    free(k1); free(k2); free(k3); free(k4); free(y);
    return 0;
}

/* ---- Tsit5 -------------------------------------------------------------------------
 * Adaptive Tsit5: the 5(4) pair of Ch. Tsitouras, "Runge-Kutta pairs of order 5(4)
 * satisfying only the first column simplifying assumption", Comput. Math. Appl. 62
 * (2011) 770-775, with the step control of OrdinaryDiffEqTsit5 (compat "2",
 * test/Project.toml:36,96; the solver of docs/examples/cascading_failure.jl:24,145):
 *   error norm   sqrt(mean((h*sum_i btilde_i k_i / (abstol + reltol*max(|y|,|ynew|)))^2))
 *   PI control   q11 = E^beta1, q = clamp(q11 / qold^beta2 / gamma, 1/qmax, 1/qmin),
 *                beta1 = 7/50, beta2 = 2/25, gamma = 9/10, qmin = 1/5, qmax = 10,
 *                qold starts at 1e-4 and becomes max(E, 1e-4) on accept; E == 0 gives
 *                q = 1/qmax; a reject takes h / min(1/qmin, q11/gamma)
 *   initial dt   Hairer, Norsett & Wanner, Solving ODEs I, II.4 (d0, d1, d2; order+1 = 6)
 * Stages are evaluated at t + c_i h.  The last step is shortened to end on t1 (or on a
 * preset trip time, a tstop); a step that ends within 100 ulp of the stop time ends on it
 * (fixed_t_for_floatingpoint_error! in OrdinaryDiffEq).  Callbacks after an accepted step, in this order: line
 * trips, the preset trip, then a re-evaluated first-same-as-last derivative if a
 * parameter or the state changed, then the steady-state test (every |du_i| <= ss_abstol
 * or <= ss_reltol*|u_i|).  Line trips set an edge's first parameter to 0:
 *   line_events = 1 (step end): where |edge dst output| > limit at the accepted state;
 *   line_events = 2 (crossing): the reference's ContinuousComponentCallback
 *     (src/callbacks.jl:113-140; the cascading-failure condition |P| - limit with the
 *     same affect on up- and down-crossings, docs/examples/cascading_failure.jl:104-111)
 *     run as DiffEqBase's VectorContinuousCallback (DiffEqBase 7, Project.toml:74; not in
 *     the reference tree): the condition at the step start (uprev) against the
 *     condition at t + h*m/9, m = 1..9 on the dense output (m = 9: the accepted state);
 *     an edge crosses where its start value g0 != 0 and g0*g <= 0 (sign-wise); at the
 *     first m with a crossing each crossing edge's root is bisected on the dense output
 *     between the points m-1 and m down to adjacent doubles, keeping the left end
 *     (LeftRootFind); the earliest root wins, every edge whose root equals it trips, the
 *     solution moves to the dense output there and the step size proposed by the
 *     controller is kept.  A crossing inside the step comes before the preset trip.
 * retcodes: 0 reached t1, 1 maxiters, 2 dt < dtmin, 3 NaN step size, 4 steady state.
 * stats[6] = t, dt, naccept, nreject, nf, ntrip.
 * -------------------------------------------------------------------------------------*/
static const double T5_C[7] = {0.0, 0.161, 0.327, 0.9, 0.9800255409045097, 1.0, 1.0};
static const double T5_A[7][6] = {
    {0, 0, 0, 0, 0, 0},
    {0.161, 0, 0, 0, 0, 0},
    {-0.008480655492356989, 0.335480655492357, 0, 0, 0, 0},
    {2.897153057105493, -6.359448489975075, 4.3622954328695815, 0, 0, 0},
    {5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525, 0, 0},
    {5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401, -0.028269050394068383, 0},
    {0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081, 2.324710524099774},
};
/* btilde = b - bhat (b = last row of A, first same as last) */
static const double T5_BT[7] = {-0.00178001105222577714, -0.0008164344596567469, 0.007880878010261995,
                                -0.1447110071732629, 0.5823571654525552, -0.45808210592918697,
                                0.015151515151515152};

/* the coefficients above, for the order-condition test: A[7*6] row-major, btilde[7], c[7] */
void ndo_tsit5_tableau(double* A, double* bt, double* c) {
    for (int i = 0; i < 7; ++i) {
        for (int j = 0; j < 6; ++j) A[i * 6 + j] = T5_A[i][j];
        bt[i] = T5_BT[i];
        c[i] = T5_C[i];
    }
}

/* Tsitouras' continuous extension (the dense output OrdinaryDiffEqTsit5 interpolates
 * with): b_i(s) = sum_j T5_R[i][j-1] s^j, i over the seven stage derivatives */
static const double T5_R[7][4] = {
    {1.0, -2.763706197274826, 2.9132554618219126, -1.0530884977290216},
    {0.0, 0.13169999999999998, -0.2234, 0.1017},
    {0.0, 3.9302962368947516, -5.941033872131505, 2.490627285651253},
    {0.0, -12.411077166933676, 30.33818863028232, -16.548102889244902},
    {0.0, 37.50931341651104, -88.1789048947664, 47.37952196281928},
    {0.0, -27.896526289197286, 65.09189467479366, -34.8706578614966},
    {0.0, 1.5, -4.0, 2.5},
};

void ndo_tsit5_dense(double* r) {
    for (int i = 0; i < 7; ++i)
        for (int j = 0; j < 4; ++j) r[i * 4 + j] = T5_R[i][j];
}

/* out = u + h * s * (k1 + s * (d2 + s * (d3 + s * d4))), d_j = sum_i r_ij k_i */
static void t5_dense_at(double* out, const double* u, double* const* k, const double* d, int64_t n, double h,
                        double s) {
    const double* d2 = d;
    const double* d3 = d + n;
    const double* d4 = d + 2 * n;
    for (int64_t i = 0; i < n; ++i) out[i] = u[i] + h * (s * (k[0][i] + s * (d2[i] + s * (d3[i] + s * d4[i]))));
}

/* the sign test of a crossing: start value g0 nonzero, g0*g <= 0 (NaN never crosses) */
static int crossed(double g0, double g) {
    return (g0 < 0.0 && g >= 0.0) || (g0 > 0.0 && g <= 0.0);
}

static double wnorm2(const double* x, const double* r1, const double* r2, int64_t n, double abstol, double reltol) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double m = r2 ? fmax(fabs(r1[i]), fabs(r2[i])) : fabs(r1[i]);
        const double q = x[i] / (abstol + m * reltol);
        s += q * q;
    }
    return s;
}

/* edge dst outputs at (u, p, t) into flow[ne] (the trip condition's |P_line|) */
static void edge_flows(ndo_net* nw, double* flow, const double* u, const double* p, double t, double* scratch_du) {
    rhs_impl(nw, scratch_du, u, p, t, 1, NULL, NULL, NULL);
    for (int64_t e = 0; e < nw->ne; ++e) flow[e] = nw->o[nw->e_out_dst[e].first - 1];
}

int ndo_tsit5(ndo_net* nw, double* u, double* p, double t0, double t1, double reltol, double abstol, double dt0,
              double dtmax, double dtmin, int64_t maxiters, const double* line_limit, int line_events,
              double ss_abstol, double ss_reltol, double event_t, int64_t event_e, double* stats) {
    const int64_t n = nw->lastidx_dynamic;
    const int64_t ne1 = nw->ne ? nw->ne : 1;
    double* k[7];
    for (int s = 0; s < 7; ++s) k[s] = (double*)xcalloc((size_t)n, 8);
    double* ys = (double*)xcalloc((size_t)n, 8);
    double* err = (double*)xcalloc((size_t)n, 8);
    double* flow = (double*)xcalloc((size_t)ne1, 8);
    const int locate = line_limit && line_events != 1;
    double* dc = (double*)xcalloc((size_t)(3 * n), 8);      /* dense coefficients d2, d3, d4 */
    double* ui = (double*)xcalloc((size_t)n, 8);            /* dense-output state */
    double* g0 = (double*)xcalloc((size_t)ne1, 8);          /* condition at the step start */
    double* root = (double*)xcalloc((size_t)ne1, 8);
    int* cand = (int*)xcalloc((size_t)ne1, sizeof(int));
    if (!(dtmax > 0)) dtmax = t1 - t0;
    if (!(dtmin > 0)) dtmin = 1e-14 * fmax(1.0, fabs(t1));
    if (maxiters <= 0) maxiters = 1000000;
    const int ss = ss_abstol > 0 || ss_reltol > 0;
    double te = INFINITY;
    int64_t ee = -1;
    if (event_e >= 0 && event_e < nw->ne && event_t > t0 && event_t <= t1) { te = event_t; ee = event_e; }
    int64_t naccept = 0, nreject = 0, nf = 0, ntrip = 0, iters = 0;
    int retcode = 0;
    double t = t0;
    rhs_impl(nw, k[0], u, p, t, 1, NULL, NULL, NULL);
    ++nf;
    double dt = dt0;
    const double inv_len = 1.0 / (double)n;
    if (!(dt > 0) && t1 > t0) {
        const double d0 = sqrt(wnorm2(u, u, NULL, n, abstol, reltol) * inv_len);
        const double d1 = sqrt(wnorm2(k[0], u, NULL, n, abstol, reltol) * inv_len);
        double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
        h0 = fmin(h0, t1 - t0);
        for (int64_t i = 0; i < n; ++i) ys[i] = u[i] + h0 * k[0][i];
        rhs_impl(nw, k[1], ys, p, t + h0, 1, NULL, NULL, NULL);
        ++nf;
        for (int64_t i = 0; i < n; ++i) err[i] = k[1][i] - k[0][i];
        const double d2 = sqrt(wnorm2(err, u, NULL, n, abstol, reltol) * inv_len) / h0;
        const double m = fmax(d1, d2);
        const double h1 = m <= 1e-15 ? fmax(1e-6, h0 * 1e-3) : pow(0.01 / m, 1.0 / 6.0);
        dt = fmin(100.0 * h0, h1);
    }
    dt = fmin(dt, dtmax);
    double qold = 1e-4;
    double tstop = ee >= 0 ? te : t1;
    while (t < t1) {
        if (iters++ >= maxiters) { retcode = 1; break; }
        if (dt < dtmin) { retcode = 2; break; }
        const int last = t + dt >= tstop;
        const double h = last ? tstop - t : dt;
        for (int s = 1; s < 7; ++s) {
            for (int64_t i = 0; i < n; ++i) {
                double acc = T5_A[s][0] * k[0][i];
                for (int j = 1; j < s; ++j) acc = acc + T5_A[s][j] * k[j][i];
                ys[i] = u[i] + h * acc;
            }
            rhs_impl(nw, k[s], ys, p, t + T5_C[s] * h, 1, NULL, NULL, NULL);
        }
        nf += 6;
        if (locate)
            for (int j = 1; j < 4; ++j)
                for (int64_t i = 0; i < n; ++i) {
                    double acc = T5_R[0][j] * k[0][i];
                    for (int s = 1; s < 7; ++s) acc = acc + T5_R[s][j] * k[s][i];
                    dc[(j - 1) * n + i] = acc;
                }
        for (int64_t i = 0; i < n; ++i) {
            double e = T5_BT[0] * k[0][i];
            for (int j = 1; j < 7; ++j) e = e + T5_BT[j] * k[j][i];
            err[i] = h * e;
        }
        const double eest = sqrt(wnorm2(err, u, ys, n, abstol, reltol) * inv_len);
        const double q11 = pow(eest, 0.14);
        if (eest <= 1.0) {
            const double qq = eest == 0.0 ? 0.1 : fmax(0.1, fmin(5.0, q11 / pow(qold, 0.08) / 0.9));
            const double tn = t + h;
            const int at_stop = last || fabs(tn - tstop) < 100.0 * 2.220446049250313e-16 * fmax(fabs(t), fabs(tstop));
            qold = fmax(eest, 1e-4);
            dt = fmin(h / qq, dtmax);
            ++naccept;
            int changed = 0, located = 0;
            if (locate) {
                /* condition at the step start (uprev, with the current parameters) */
                edge_flows(nw, flow, u, p, t, err);
                for (int64_t e = 0; e < nw->ne; ++e) g0[e] = fabs(flow[e]) - line_limit[e];
                int mf = 0;
                for (int m = 1; m <= 9 && !mf; ++m) {
                    if (m == 9) memcpy(ui, ys, (size_t)n * 8);
                    else t5_dense_at(ui, u, k, dc, n, h, (double)m / 9.0);
                    edge_flows(nw, flow, ui, p, t, err);
                    for (int64_t e = 0; e < nw->ne; ++e) {
                        cand[e] = crossed(g0[e], fabs(flow[e]) - line_limit[e]);
                        if (cand[e]) mf = m;
                    }
                }
                if (mf) {
                    double tmin = INFINITY;
                    for (int64_t e = 0; e < nw->ne; ++e) {
                        if (!cand[e]) continue;
                        double lo = t + h * ((double)(mf - 1) / 9.0), hi = t + h * ((double)mf / 9.0);
                        for (int it = 0; it < 256; ++it) {
                            const double mid = lo + 0.5 * (hi - lo);
                            if (!(mid > lo && mid < hi)) break;
                            t5_dense_at(ui, u, k, dc, n, h, (mid - t) / h);
                            edge_flows(nw, flow, ui, p, mid, err);
                            if (crossed(g0[e], fabs(flow[e]) - line_limit[e])) hi = mid;
                            else lo = mid;
                        }
                        root[e] = lo;
                        if (lo < tmin) tmin = lo;
                    }
                    for (int64_t e = 0; e < nw->ne; ++e)
                        if (cand[e] && root[e] == tmin) {
                            double* K = &p[nw->e_para[e].first - 1];
                            *K = 0.0;
                            ++ntrip;
                        }
                    t5_dense_at(ui, u, k, dc, n, h, (tmin - t) / h);
                    memcpy(u, ui, (size_t)n * 8);
                    t = tmin;
                    located = 1;
                    changed = 1;
                }
            }
            if (!located) {
                t = at_stop ? tstop : tn;
                memcpy(u, ys, (size_t)n * 8);
                memcpy(k[0], k[6], (size_t)n * 8);
                if (line_limit && !locate) {
                    edge_flows(nw, flow, u, p, t, err);
                    for (int64_t e = 0; e < nw->ne; ++e) {
                        double* K = &p[nw->e_para[e].first - 1];
                        if (*K != 0.0 && fabs(flow[e]) > line_limit[e]) { *K = 0.0; ++ntrip; ++changed; }
                    }
                }
                if (at_stop && ee >= 0 && tstop == te) {
                    double* K = &p[nw->e_para[ee].first - 1];
                    if (*K != 0.0) { ++ntrip; ++changed; }
                    *K = 0.0;
                    ee = -1;
                    tstop = t1;
                }
            }
            if (changed) {
                rhs_impl(nw, k[0], u, p, t, 1, NULL, NULL, NULL);
                ++nf;
            }
            if (ss) {
                int pass = 1;
                for (int64_t i = 0; i < n && pass; ++i) {
                    const double d = fabs(k[0][i]);
                    if (!(d <= ss_abstol || d <= ss_reltol * fabs(u[i]))) pass = 0;
                }
                if (pass) { retcode = 4; break; }
            }
        } else {
            dt = h / fmin(5.0, q11 / 0.9);
            ++nreject;
        }
        if (!(dt == dt)) { retcode = 3; break; }
    }
    stats[0] = t; stats[1] = dt; stats[2] = (double)naccept; stats[3] = (double)nreject;
    stats[4] = (double)nf; stats[5] = (double)ntrip;
    for (int s = 0; s < 7; ++s) free(k[s]);
    free(ys); free(err); free(flow);
    free(dc); free(ui); free(g0); free(root); free(cand);
    return retcode;
}
