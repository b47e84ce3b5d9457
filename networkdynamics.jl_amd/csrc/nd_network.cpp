// Host index layer of libnd_hip.so: batching, flat-layout registration and the
// integer maps of NetworkDynamics.jl's Network constructor.
//
//   batching              construction.jl:154-168, 238-256; utils.jl:197-217
//   registration          network_structure.jl:224-289 (independent running counters
//                         per range kind; all vertex batches before all edge batches)
//   BatchStride           utils.jl:1-42 (per-component ranges = first + pos*stride)
//   AggregationMap        aggregators.jl:54-86
//   destination lists     aggregators.jl:203-235 (here: summation order of
//                         SequentialAggregator, aggregators.jl:140-151)
//   Eager gather map      gbufs.jl:11-22
//
// Each counter advances by a whole batch at a time: a batch's components occupy
// one contiguous block per range kind, so registration is O(#batches) and the
// per-component ranges follow from (batch, position) by BatchStride arithmetic.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>

#include "nd_internal.h"

namespace nd {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}
void clear_error() { g_last_error.clear(); }

// Groups component ids by kind index in order of first appearance.
static void group_by_kind(const std::vector<int32_t>& kind_of, int32_t nkinds, std::vector<Batch>& out,
                          std::vector<int32_t>& batch_of, std::vector<int64_t>& pos_of) {
    std::vector<int32_t> batch_of_kind(static_cast<size_t>(nkinds), -1);
    out.clear();
    batch_of.assign(kind_of.size(), 0);
    pos_of.assign(kind_of.size(), 0);
    for (size_t i = 0; i < kind_of.size(); ++i) {
        int32_t k = kind_of[i];
        if (batch_of_kind[k] < 0) {
            batch_of_kind[k] = static_cast<int32_t>(out.size());
            out.emplace_back();
            out.back().kind = k;
        }
        Batch& b = out[batch_of_kind[k]];
        batch_of[i] = batch_of_kind[k];
        pos_of[i] = static_cast<int64_t>(b.indices.size());
        b.indices.push_back(static_cast<int64_t>(i) + 1);
    }
}

static int validate_kind(const nd_kind& k, int32_t idx) {
    if (k.dim < 0 || k.pdim < 0 || k.outdim_src < 0 || k.outdim_dst < 0)
        return set_error(ND_ERR_ARGUMENT, "kind %d: negative dimension", idx);
    if (k.outdim_src > ND_MAX_OUT || k.outdim_dst > ND_MAX_OUT)
        return set_error(ND_ERR_UNSUPPORTED, "kind %d: output dimension above %d", idx, ND_MAX_OUT);
    if (!k.is_edge && k.outdim_src != 0)
        return set_error(ND_ERR_ARGUMENT, "kind %d: vertex kinds have outdim_src = 0", idx);
    if (k.extdim < 0 || k.extdim > ND_MAX_EXT)
        return set_error(ND_ERR_UNSUPPORTED, "kind %d: %d external inputs (at most %d)", idx, k.extdim, ND_MAX_EXT);
    if (k.ff == ND_FF_PURESTATEMAP) {
        for (int j = 0; j < k.outdim_dst; ++j)
            if (k.mask_dst[j] < 1 || k.mask_dst[j] > k.dim)
                return set_error(ND_ERR_ARGUMENT, "kind %d: StateMask index %d outside 1:%d", idx, k.mask_dst[j], k.dim);
        if (k.is_edge)
            for (int j = 0; j < k.outdim_src; ++j)
                if (k.mask_src[j] < 1 || k.mask_src[j] > k.dim)
                    return set_error(ND_ERR_ARGUMENT, "kind %d: StateMask index %d outside 1:%d", idx, k.mask_src[j], k.dim);
    }
    return ND_OK;
}

static int build(Network& n) {
    const int64_t nv = n.nv, ne = n.ne;
    // construction.jl:99-106
    n.vdepth = n.kinds[n.vkind[0]].outdim_dst;
    for (int64_t v = 0; v < nv; ++v)
        if (n.kinds[n.vkind[v]].outdim_dst != n.vdepth)
            return set_error(ND_ERR_ARGUMENT, "All vertex models must have the same output dimension!");
    n.edepth = ne > 0 ? n.kinds[n.ekind[0]].outdim_dst : 0;
    for (int64_t e = 0; e < ne; ++e) {
        const nd_kind& k = n.kinds[n.ekind[e]];
        if (k.outdim_dst != n.edepth)
            return set_error(ND_ERR_ARGUMENT, "All edge models must have the same output dimension!");
        if (k.outdim_src != 0 && k.outdim_src != n.edepth)
            return set_error(ND_ERR_ARGUMENT, "DimensionMismatch: edge src output dimension %d vs aggregation depth %d",
                             k.outdim_src, n.edepth);
    }

    // feed-forward vertices and LoopbackConnection edges (construction.jl:55-83)
    {
        std::vector<int64_t> deg(static_cast<size_t>(nv), 0), only(static_cast<size_t>(nv), -1);
        for (int64_t e = 0; e < ne; ++e) {
            ++deg[n.src[e] - 1]; ++deg[n.dst[e] - 1];
            only[n.src[e] - 1] = e; only[n.dst[e] - 1] = e;
        }
        auto loopback = [&](int64_t e) { return n.kinds[n.ekind[e]].model == ND_EM_LOOPBACK; };
        for (int64_t v = 0; v < nv; ++v) {
            const int32_t ff = n.kinds[n.vkind[v]].ff;
            if (ff != ND_FF_PUREFEEDFORWARD && ff != ND_FF_FEEDFORWARD) continue;
            if (deg[v] != 1 || !loopback(only[v]) || n.src[only[v]] - 1 != v)
                return set_error(ND_ERR_ARGUMENT, "Feed forward vertex models are only allowed as leave nodes (1 "
                                 "neighbor) with single `LoopbackConnection` from satelite to cluster node! (vertex %lld)",
                                 (long long)(v + 1));
        }
        for (int64_t e = 0; e < ne; ++e)
            if (loopback(e) && deg[n.src[e] - 1] != 1)
                return set_error(ND_ERR_ARGUMENT, "All LoopbackConnection edges must originate from leaf nodes! "
                                 "Found at least one going from hub to leaf (edge %lld)", (long long)(e + 1));
    }

    group_by_kind(n.vkind, static_cast<int32_t>(n.kinds.size()), n.vb, n.v_batch, n.v_pos);
    group_by_kind(n.ekind, static_cast<int32_t>(n.kinds.size()), n.eb, n.e_batch, n.e_pos);

    // registration counters (network_structure.jl:259-289)
    int64_t c_dyn = 0, c_out = 0, c_par = 0, c_aggr = 0, c_gbuf = 0, c_ext = 0;
    for (Batch& b : n.vb) {
        const nd_kind& k = n.kinds[b.kind];
        const int64_t cnt = static_cast<int64_t>(b.indices.size());
        b.state_first = c_dyn + 1; c_dyn += cnt * k.dim;
        b.out_first = c_out + 1;   c_out += cnt * k.outdim_dst;
        b.param_first = c_par + 1; c_par += cnt * k.pdim;
        b.in_first = c_aggr + 1;   c_aggr += cnt * n.edepth;
        b.ext_first = c_ext + 1;   c_ext += cnt * k.extdim;
    }
    for (Batch& b : n.eb) {
        const nd_kind& k = n.kinds[b.kind];
        const int64_t cnt = static_cast<int64_t>(b.indices.size());
        b.state_first = c_dyn + 1; c_dyn += cnt * k.dim;
        b.out_first = c_out + 1;   c_out += cnt * (k.outdim_src + k.outdim_dst);
        b.param_first = c_par + 1; c_par += cnt * k.pdim;
        b.in_first = c_gbuf + 1;   c_gbuf += cnt * 2 * n.vdepth;
        b.ext_first = c_ext + 1;   c_ext += cnt * k.extdim;
    }
    n.dim = c_dyn; n.pdim = c_par; n.lastidx_out = c_out; n.lastidx_aggr = c_aggr; n.lastidx_gbuf = c_gbuf;
    n.lastidx_extbuf = c_ext;

    // Destination lists: emit (row, output index) in increasing output index, then a
    // stable counting sort by row keeps each row's summation order.
    std::vector<int64_t> cnt(static_cast<size_t>(n.lastidx_aggr) + 1, 0);
    int64_t lo = -1, hi = -1;   // tightened AggregationMap range (aggregators.jl:80-86)
    auto visit = [&](auto&& emit) {
        for (const Batch& b : n.eb) {
            const nd_kind& k = n.kinds[b.kind];
            const int64_t stride = k.outdim_src + k.outdim_dst;
            for (size_t q = 0; q < b.indices.size(); ++q) {
                const int64_t e = b.indices[q] - 1;
                const int64_t o = b.out_first + static_cast<int64_t>(q) * stride;
                const int64_t as = n.v_aggr_first(n.src[e] - 1), ad = n.v_aggr_first(n.dst[e] - 1);
                for (int j = 0; j < k.outdim_src; ++j) emit(as - 1 + j, o + j);
                for (int j = 0; j < k.outdim_dst; ++j) emit(ad - 1 + j, o + k.outdim_src + j);
            }
        }
    };
    visit([&](int64_t row, int64_t col) {
        cnt[row + 1]++;
        if (lo < 0) lo = col;
        hi = col;
    });
    for (int64_t a = 0; a < n.lastidx_aggr; ++a) cnt[a + 1] += cnt[a];
    n.csr_rowptr = cnt;
    n.csr_cols.assign(static_cast<size_t>(cnt[n.lastidx_aggr]), 0);
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    visit([&](int64_t row, int64_t col) { n.csr_cols[fill[row]++] = col; });
    if (lo < 0) { n.agg_first = 1; n.agg_last = 0; }
    else { n.agg_first = lo; n.agg_last = hi; }
    return ND_OK;
}

int64_t output_state_index(const Network& n, int64_t k, bool* negated) {
    if (negated) *negated = false;
    // output blocks: vertex batches, then edge batches, ascending (registration order)
    for (const Batch& b : n.vb) {
        const nd_kind& kd = n.kinds[b.kind];
        const int64_t len = static_cast<int64_t>(b.indices.size()) * kd.outdim_dst;
        if (k >= b.out_first && k < b.out_first + len) {
            if (kd.ff != ND_FF_PURESTATEMAP) return -1;
            const int64_t q = (k - b.out_first) / kd.outdim_dst, j = (k - b.out_first) % kd.outdim_dst;
            return b.state_first - 1 + q * kd.dim + kd.mask_dst[j] - 1;
        }
    }
    for (const Batch& b : n.eb) {
        const nd_kind& kd = n.kinds[b.kind];
        const int64_t stride = kd.outdim_src + kd.outdim_dst;
        const int64_t len = static_cast<int64_t>(b.indices.size()) * stride;
        if (k >= b.out_first && k < b.out_first + len) {
            if (kd.ff != ND_FF_PURESTATEMAP) return -1;
            const int64_t q = (k - b.out_first) / stride, j = (k - b.out_first) % stride;
            const int32_t m = j < kd.outdim_src ? kd.mask_src[j] : kd.mask_dst[j - kd.outdim_src];
            if (negated) *negated = j < kd.outdim_src && kd.wrapper == ND_WRAP_ANTISYMMETRIC;
            return b.state_first - 1 + q * kd.dim + m - 1;
        }
    }
    return -1;
}

}  // namespace nd

using namespace nd;

extern "C" {

const char* nd_last_error(void) { return g_last_error.c_str(); }

int nd_network_create(int64_t nv, int64_t ne, const int64_t* src, const int64_t* dst, int32_t nkinds,
                      const nd_kind* kinds, const int32_t* vertex_kind, const int32_t* edge_kind,
                      nd_network** out) {
    clear_error();
    if (!out) return set_error(ND_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (nv < 1) return set_error(ND_ERR_ARGUMENT, "a network needs at least one vertex");
    if (ne < 0) return set_error(ND_ERR_ARGUMENT, "negative edge count");
    if (nkinds < 1 || !kinds || !vertex_kind || (ne > 0 && (!src || !dst || !edge_kind)))
        return set_error(ND_ERR_ARGUMENT, "missing kinds / component kinds / edge list");
    for (int32_t i = 0; i < nkinds; ++i) {
        int rc = validate_kind(kinds[i], i);
        if (rc) return rc;
    }
    for (int64_t v = 0; v < nv; ++v)
        if (vertex_kind[v] < 0 || vertex_kind[v] >= nkinds || kinds[vertex_kind[v]].is_edge)
            return set_error(ND_ERR_ARGUMENT, "vertex %lld: kind index %d is not a vertex kind", (long long)(v + 1), vertex_kind[v]);
    for (int64_t e = 0; e < ne; ++e) {
        if (edge_kind[e] < 0 || edge_kind[e] >= nkinds || !kinds[edge_kind[e]].is_edge)
            return set_error(ND_ERR_ARGUMENT, "edge %lld: kind index %d is not an edge kind", (long long)(e + 1), edge_kind[e]);
        if (src[e] < 1 || src[e] > nv || dst[e] < 1 || dst[e] > nv)
            return set_error(ND_ERR_ARGUMENT, "edge %lld: endpoint outside 1:%lld", (long long)(e + 1), (long long)nv);
    }
    nd_network* h = new (std::nothrow) nd_network;
    if (!h) return set_error(ND_ERR_ALLOC, "host allocation failed");
    try {
        Network& n = h->net;
        n.nv = nv; n.ne = ne;
        n.src.assign(src, src + ne);
        n.dst.assign(dst, dst + ne);
        n.kinds.assign(kinds, kinds + nkinds);
        n.vkind.assign(vertex_kind, vertex_kind + nv);
        n.ekind.assign(edge_kind, edge_kind + ne);
        int rc = build(n);
        if (rc) { delete h; return rc; }
    } catch (const std::bad_alloc&) {
        delete h;
        return set_error(ND_ERR_ALLOC, "host allocation failed");
    }
    *out = h;
    return ND_OK;
}

void nd_network_destroy(nd_network* nw) { delete nw; }

int nd_network_get_info(const nd_network* h, nd_network_info* info) {
    if (!h || !info) return set_error(ND_ERR_ARGUMENT, "NULL argument");
    const Network& n = h->net;
    info->nv = n.nv; info->ne = n.ne; info->dim = n.dim; info->pdim = n.pdim;
    info->lastidx_out = n.lastidx_out; info->lastidx_aggr = n.lastidx_aggr; info->lastidx_gbuf = n.lastidx_gbuf;
    info->agg_first = n.agg_first; info->agg_last = n.agg_last;
    info->vdepth = n.vdepth; info->edepth = n.edepth;
    info->nvbatches = static_cast<int32_t>(n.vb.size());
    info->nebatches = static_cast<int32_t>(n.eb.size());
    info->lastidx_extbuf = n.lastidx_extbuf;
    return ND_OK;
}

int nd_network_export_ext_ranges(const nd_network* h, int64_t* v_ext, int64_t* e_ext) {
    if (!h) return set_error(ND_ERR_ARGUMENT, "NULL argument");
    const Network& n = h->net;
    if (v_ext)
        for (int64_t v = 0; v < n.nv; ++v) {
            v_ext[2 * v] = n.v_ext_first(v);
            v_ext[2 * v + 1] = n.v_ext_first(v) + n.kinds[n.vkind[v]].extdim - 1;
        }
    if (e_ext)
        for (int64_t e = 0; e < n.ne; ++e) {
            e_ext[2 * e] = n.e_ext_first(e);
            e_ext[2 * e + 1] = n.e_ext_first(e) + n.kinds[n.ekind[e]].extdim - 1;
        }
    return ND_OK;
}

// ExtMap (external_inputs.jl:10-46) with _symidx_to_extidx's checks (:33-46)
int nd_network_set_extmap(nd_network* h, int64_t len, const int64_t* map) {
    if (!h) return set_error(ND_ERR_ARGUMENT, "NULL argument");
    Network& n = h->net;
    if (n.lastidx_extbuf == 0) return set_error(ND_ERR_ARGUMENT, "There are no external inputs.");
    if (len != n.lastidx_extbuf || !map)
        return set_error(ND_ERR_ARGUMENT, "ExtMap needs %lld entries, got %lld", (long long)n.lastidx_extbuf,
                         (long long)len);
    for (int64_t k = 0; k < len; ++k) {
        const int64_t x = map[k];
        if (x > 0) {
            if (x > n.dim) return set_error(ND_ERR_ARGUMENT, "external input %lld: state index %lld outside 1:%lld",
                                            (long long)(k + 1), (long long)x, (long long)n.dim);
        } else if (x < 0) {
            const int64_t o = -x;
            if (o > n.lastidx_out)
                return set_error(ND_ERR_ARGUMENT, "external input %lld: output index %lld outside 1:%lld",
                                 (long long)(k + 1), (long long)o, (long long)n.lastidx_out);
            if (output_state_index(n, o) < 0)
                return set_error(ND_ERR_ARGUMENT, "Cannot resolve external input %lld! Outputs of feed-forward "
                                 "components are not allowed as external inputs.", (long long)(k + 1));
        } else {
            return set_error(ND_ERR_ARGUMENT, "external input %lld: index 0", (long long)(k + 1));
        }
    }
    n.extmap.assign(map, map + len);
    return ND_OK;
}

int nd_network_get_batch(const nd_network* h, int32_t is_edge, int32_t b, nd_batch* out) {
    if (!h || !out) return set_error(ND_ERR_ARGUMENT, "NULL argument");
    const std::vector<Batch>& bs = is_edge ? h->net.eb : h->net.vb;
    if (b < 0 || b >= static_cast<int32_t>(bs.size())) return set_error(ND_ERR_ARGUMENT, "batch %d out of range", b);
    const Batch& x = bs[b];
    out->kind = x.kind; out->pad_ = 0;
    out->count = static_cast<int64_t>(x.indices.size());
    out->indices = x.indices.data();
    out->state_first = x.state_first; out->param_first = x.param_first;
    out->out_first = x.out_first; out->in_first = x.in_first;
    return ND_OK;
}

int nd_network_export_ranges(const nd_network* h, int64_t* vr, int64_t* er) {
    if (!h) return set_error(ND_ERR_ARGUMENT, "NULL argument");
    const Network& n = h->net;
    auto put = [](int64_t* p, int64_t first, int64_t len) { p[0] = first; p[1] = first + len - 1; };
    if (vr)
        for (int64_t v = 0; v < n.nv; ++v) {
            const nd_kind& k = n.kinds[n.vkind[v]];
            int64_t* p = vr + v * 8;
            put(p + 0, n.v_state_first(v), k.dim);
            put(p + 2, n.v_out_first(v), k.outdim_dst);
            put(p + 4, n.v_param_first(v), k.pdim);
            put(p + 6, n.v_aggr_first(v), n.edepth);
        }
    if (er)
        for (int64_t e = 0; e < n.ne; ++e) {
            const nd_kind& k = n.kinds[n.ekind[e]];
            int64_t* p = er + e * 12;
            const int64_t o = n.e_out_first(e), g = n.e_gbuf_first(e);
            put(p + 0, n.e_state_first(e), k.dim);
            put(p + 2, o, k.outdim_src);
            put(p + 4, o + k.outdim_src, k.outdim_dst);
            put(p + 6, n.e_param_first(e), k.pdim);
            put(p + 8, g, n.vdepth);
            put(p + 10, g + n.vdepth, n.vdepth);
        }
    return ND_OK;
}

int nd_network_export_maps(const nd_network* h, int64_t* aggmap, int64_t* gbufmap) {
    if (!h) return set_error(ND_ERR_ARGUMENT, "NULL argument");
    const Network& n = h->net;
    if (aggmap && n.agg_last >= n.agg_first) {
        const int64_t len = n.agg_last - n.agg_first + 1;
        std::fill(aggmap, aggmap + len, int64_t(0));
        for (int64_t a = 0; a < n.lastidx_aggr; ++a)
            for (int64_t q = n.csr_rowptr[a]; q < n.csr_rowptr[a + 1]; ++q)
                aggmap[n.csr_cols[q] - n.agg_first] = a + 1;
    }
    if (gbufmap)
        for (int64_t e = 0; e < n.ne; ++e) {
            const int64_t g = n.e_gbuf_first(e) - 1;
            const int64_t os = n.v_out_first(n.src[e] - 1), od = n.v_out_first(n.dst[e] - 1);
            for (int j = 0; j < n.vdepth; ++j) {
                gbufmap[g + j] = os + j;
                gbufmap[g + n.vdepth + j] = od + j;
            }
        }
    return ND_OK;
}

int nd_network_csr_size(const nd_network* h, int64_t* nnz) {
    if (!h || !nnz) return set_error(ND_ERR_ARGUMENT, "NULL argument");
    *nnz = static_cast<int64_t>(h->net.csr_cols.size());
    return ND_OK;
}

int nd_network_export_csr(const nd_network* h, int64_t* rowptr, int64_t* cols) {
    if (!h) return set_error(ND_ERR_ARGUMENT, "NULL argument");
    const Network& n = h->net;
    if (rowptr) std::copy(n.csr_rowptr.begin(), n.csr_rowptr.end(), rowptr);
    if (cols) std::copy(n.csr_cols.begin(), n.csr_cols.end(), cols);
    return ND_OK;
}

// construct_mass_matrix, construction.jl:409-453.  A component takes part when its
// matrix differs from I (UniformScaling compares its scale only, whatever the size; the
// matrix forms compare entry by entry); the result is diagonal when every taking part
// is (isdiag), and each block is written over the identity.
int nd_network_mass_matrix(const nd_network* h, const nd_mass* vm, const nd_mass* em, int32_t* shape,
                           double* out) {
    if (!h || !shape) return set_error(ND_ERR_ARGUMENT, "NULL argument");
    const Network& n = h->net;
    struct Part { int64_t first; int32_t dim; const nd_mass* m; };
    std::vector<Part> parts;
    bool all_diag = true;
    auto consider = [&](const nd_mass& m, int64_t first, int32_t dim, const char* what, int64_t id) -> int {
        if (m.form < ND_MASS_IDENTITY || m.form > ND_MASS_DENSE)
            return set_error(ND_ERR_ARGUMENT, "%s %lld: unknown mass matrix form %d", what, (long long)id, m.form);
        if (m.form == ND_MASS_IDENTITY) return ND_OK;
        if (!m.m) return set_error(ND_ERR_ARGUMENT, "%s %lld: mass matrix values missing", what, (long long)id);
        bool is_eye = true, diag = true;
        if (m.form == ND_MASS_UNIFORM) {
            is_eye = m.m[0] == 1.0;
        } else if (m.form == ND_MASS_DIAGONAL) {
            for (int32_t i = 0; i < dim; ++i) is_eye = is_eye && m.m[i] == 1.0;
        } else {
            for (int32_t c = 0; c < dim; ++c)
                for (int32_t r = 0; r < dim; ++r) {
                    const double x = m.m[static_cast<int64_t>(c) * dim + r];
                    is_eye = is_eye && x == (r == c ? 1.0 : 0.0);
                    diag = diag && (r == c || x == 0.0);
                }
        }
        if (is_eye) return ND_OK;
        all_diag = all_diag && diag;
        parts.push_back({first, dim, &m});
        return ND_OK;
    };
    if (vm)
        for (int64_t v = 0; v < n.nv; ++v) {
            int rc = consider(vm[v], n.v_state_first(v), n.kinds[n.vkind[v]].dim, "vertex", v + 1);
            if (rc) return rc;
        }
    if (em)
        for (int64_t e = 0; e < n.ne; ++e) {
            int rc = consider(em[e], n.e_state_first(e), n.kinds[n.ekind[e]].dim, "edge", e + 1);
            if (rc) return rc;
        }
    if (parts.empty()) { *shape = ND_MASS_IDENTITY; return ND_OK; }
    *shape = all_diag ? ND_MASS_DIAGONAL : ND_MASS_DENSE;
    if (!out) return ND_OK;
    const int64_t d = n.dim;
    if (all_diag) {
        std::fill(out, out + d, 1.0);
        for (const Part& q : parts)
            for (int32_t i = 0; i < q.dim; ++i) {
                const nd_mass& m = *q.m;
                out[q.first - 1 + i] = m.form == ND_MASS_UNIFORM ? m.m[0]
                                     : m.form == ND_MASS_DIAGONAL ? m.m[i]
                                     : m.m[static_cast<int64_t>(i) * q.dim + i];
            }
        return ND_OK;
    }
    std::fill(out, out + d * d, 0.0);
    for (int64_t i = 0; i < d; ++i) out[i * d + i] = 1.0;
    for (const Part& q : parts) {
        const nd_mass& m = *q.m;
        for (int32_t c = 0; c < q.dim; ++c)
            for (int32_t r = 0; r < q.dim; ++r) {
                double x;
                if (m.form == ND_MASS_DENSE) x = m.m[static_cast<int64_t>(c) * q.dim + r];
                else if (r != c) x = 0.0;
                else x = m.form == ND_MASS_UNIFORM ? m.m[0] : m.m[r];
                out[(q.first - 1 + c) * d + (q.first - 1 + r)] = x;
            }
    }
    return ND_OK;
}

}  // extern "C"
