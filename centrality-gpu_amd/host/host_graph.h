// Host-side structural reduction shared by the C-ABI layer and the betc CLI.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "bcx.h"

namespace bcx {

// One component handed to the GPU (a BCC with more than two vertices).
struct Component {
    int32_t bcc = -1;                 // index in Tarjan emission order
    int32_t n = 0;
    std::vector<int32_t> gid;         // local -> global id
    std::vector<int32_t> row_ptr;     // [n+1]
    std::vector<int32_t> col;         // [arcs]
    std::vector<int32_t> reach;       // [n]
    std::vector<int32_t> chain_ptr;   // [n_chains+1]
    std::vector<int32_t> chain_nodes;
    std::vector<int32_t> chain_ends;  // [2*n_chains]
    int32_t n_free = 0, n_endpoints = 0, n_chain_vertices = 0;

    bcx_component view() const;
};

}  // namespace bcx

struct bcx_host {
    int32_t n = 0;
    int64_t m = 0;
    std::vector<int32_t> uv;          // [2m] file order
    std::vector<int64_t> adj_ptr;     // [n+1]
    std::vector<int32_t> adj;         // [2m] neighbours in file order

    bool decomposed = false;
    bool chains = true;               // use_chains of the decomposition
    bool cached = false;              // loaded from a structure cache (no edge lists)
    std::vector<int64_t> bcc_ptr;     // [n_bcc+1] offsets (in edges) into bcc_uv
    std::vector<int32_t> bcc_uv;      // edges of all BCCs, pop order, global ids
    std::vector<int32_t> bcc_nv;      // vertices per BCC
    std::vector<uint8_t> is_art;      // [n]
    std::vector<int64_t> cut;         // [n]
    std::vector<bcx::Component> comps;
    std::string err;
};

namespace bcx {
// Builds adjacency (file order) from h->uv.  Returns BCX_OK or an error code.
int build_adjacency(bcx_host *h);
int decompose(bcx_host *h, bool use_chains, bool check_input);
}  // namespace bcx
