// Internal declarations shared by the host index layer and the HIP plan.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/nd_hip.h"

namespace nd {

// Last-error slot per calling thread (nd_last_error).
int set_error(int code, const char* fmt, ...);
void clear_error();

// One batch of identical components (ComponentBatch, network_structure.jl:176-198).
// Firsts are 1-based like the reference's BatchStride.first.
struct Batch {
    int32_t kind = 0;
    std::vector<int64_t> indices;    // 1-based vertex / edge ids, ascending
    int64_t state_first = 1;          // statestride.first
    int64_t param_first = 1;          // pstride.first
    int64_t out_first = 1;            // outbufstride.first
    int64_t in_first = 1;             // inbufstride.first (v_aggr or e_gbufr)
    int64_t ext_first = 1;            // extbufstride.first (v_ext or e_ext)
};

// Host-side index layer: what the reference's IndexManager + batches hold,
// stored compactly as batch strides plus per-component (batch, position).
struct Network {
    int64_t nv = 0, ne = 0;
    std::vector<int64_t> src, dst;   // edgevec, 1-based
    std::vector<nd_kind> kinds;
    std::vector<int32_t> vkind, ekind;
    std::vector<Batch> vb, eb;
    std::vector<int32_t> v_batch, e_batch;   // batch of each component
    std::vector<int64_t> v_pos, e_pos;       // 0-based position inside its batch
    int32_t vdepth = 0, edepth = 0;
    int64_t dim = 0, pdim = 0, lastidx_out = 0, lastidx_aggr = 0, lastidx_gbuf = 0;
    int64_t lastidx_extbuf = 0;
    int64_t agg_first = 1, agg_last = 0;
    // ExtMap (external_inputs.jl:10-46): > 0 StateBufIdx (u), < 0 OutBufIdx (o), 1-based
    std::vector<int64_t> extmap;

    // Destination-ordered reduction lists over aggregation indices (0-based rows):
    // row a lists the 1-based output indices summed into aggbuf[a+1], increasing.
    std::vector<int64_t> csr_rowptr;
    std::vector<int64_t> csr_cols;

    // ranges of one vertex (0-based id) -- BatchStride arithmetic, utils.jl:21-29
    int64_t v_state_first(int64_t v) const {
        const Batch& b = vb[v_batch[v]]; return b.state_first + v_pos[v] * kinds[b.kind].dim;
    }
    int64_t v_param_first(int64_t v) const {
        const Batch& b = vb[v_batch[v]]; return b.param_first + v_pos[v] * kinds[b.kind].pdim;
    }
    int64_t v_out_first(int64_t v) const {
        const Batch& b = vb[v_batch[v]]; return b.out_first + v_pos[v] * kinds[b.kind].outdim_dst;
    }
    int64_t v_aggr_first(int64_t v) const {
        const Batch& b = vb[v_batch[v]]; return b.in_first + v_pos[v] * edepth;
    }
    int64_t e_state_first(int64_t e) const {
        const Batch& b = eb[e_batch[e]]; return b.state_first + e_pos[e] * kinds[b.kind].dim;
    }
    int64_t e_param_first(int64_t e) const {
        const Batch& b = eb[e_batch[e]]; return b.param_first + e_pos[e] * kinds[b.kind].pdim;
    }
    // first output index of the src side; the dst side follows after outdim_src
    int64_t e_out_first(int64_t e) const {
        const Batch& b = eb[e_batch[e]]; const nd_kind& k = kinds[b.kind];
        return b.out_first + e_pos[e] * (k.outdim_src + k.outdim_dst);
    }
    int64_t e_gbuf_first(int64_t e) const {
        const Batch& b = eb[e_batch[e]]; return b.in_first + e_pos[e] * 2 * vdepth;
    }
    int64_t v_ext_first(int64_t v) const {
        const Batch& b = vb[v_batch[v]]; return b.ext_first + v_pos[v] * kinds[b.kind].extdim;
    }
    int64_t e_ext_first(int64_t e) const {
        const Batch& b = eb[e_batch[e]]; return b.ext_first + e_pos[e] * kinds[b.kind].extdim;
    }
};

// u-index (0-based) of output-buffer index k (1-based) when that output is a StateMask
// of a state (vertex outputs; ODE-edge StateMask outputs, plain or wrapped), else -1.
// *negated (if given): the output is the negated state (the src side of an
// AntiSymmetric StateMask edge).
int64_t output_state_index(const Network& n, int64_t k, bool* negated = nullptr);

// Vertex output u-index (0-based) of vertex v, output element j: StateMask outputs only.
inline int64_t vertex_out_u_index(const Network& n, int64_t v, int j) {
    const nd_kind& k = n.kinds[n.vkind[v]];
    return n.v_state_first(v) - 1 + (k.mask_dst[j] - 1);
}

}  // namespace nd

struct nd_network {
    nd::Network net;
};
