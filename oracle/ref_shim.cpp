// Harness around the reference's own host sources (src/graph.cpp, src/virtual_graph.cpp),
// compiled unmodified from the reference tree into oracle/_ref/libref.so.  TEST INFRASTRUCTURE
// ONLY: the parity tests call it to pin the loader and the Tigr layout on the reference code.
//
// Defect D2 (src/virtual_graph.cpp:66 leaves outDegreeCounter uninitialised) is made
// deterministic by the zero-filling operator new below, which the library binds to itself
// (-Wl,-Bsymbolic) so every `new` in the reference objects returns zeroed memory; defect D1
// (edges skipped once their slot index reaches num_edges, :84-87) is left as is and the tests
// compare only the slots the reference writes.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>

#include "graph.hpp"
#include "virtual_graph.hpp"

void* operator new(std::size_t n) {
    void* p = std::calloc(1, n ? n : 1);
    if (!p) throw std::bad_alloc();
    return p;
}
void* operator new[](std::size_t n) { return ::operator new(n); }
void operator delete(void* p) noexcept { std::free(p); }
void operator delete[](void* p) noexcept { std::free(p); }
void operator delete(void* p, std::size_t) noexcept { std::free(p); }
void operator delete[](void* p, std::size_t) noexcept { std::free(p); }

extern "C" {

// Reads a text edge list with the reference Graph::ReadGraph (src/graph.cpp:12-113).
// Returns m; copies at most cap edges (raw ids as read) and weights.
long long ref_read_graph(const char* path, int weighted, uint32_t* n_out, int* has_zero,
                         uint32_t* src, uint32_t* dst, uint32_t* w, long long cap) {
    Graph g(std::string(path), weighted != 0);
    g.ReadGraph();
    *n_out = g.num_nodes;
    *has_zero = g.hasZeroID ? 1 : 0;
    const long long m = (long long)g.num_edges;
    for (long long i = 0; i < m && i < cap; ++i) {
        src[i] = g.edges[i].source;
        dst[i] = g.edges[i].end;
        if (weighted && w) w[i] = g.weights[i];
    }
    return m;
}

// Builds the reference Tigr layout (VirtualGraph::MakeGraph, src/virtual_graph.cpp:36-112) for
// a 0-based edge list.  node_pointer: n, edge_list: 2m+n, parts: 2*num_parts (cap_parts pairs).
// Returns num_parts.
long long ref_make_graph(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                         const uint32_t* w, uint32_t* node_pointer, uint32_t* edge_list,
                         uint32_t* parts, long long cap_parts) {
    Graph g(std::string("in-memory.el"), true);
    g.hasZeroID = true;  // ids are already 0-based: no shift (src/virtual_graph.cpp:5-12)
    g.num_nodes = n;
    g.num_edges = (uint)m;
    g.edges.resize(m);
    g.weights.resize(m);
    for (uint64_t i = 0; i < m; ++i) {
        g.edges[i].source = src[i];
        g.edges[i].end = dst[i];
        g.weights[i] = w ? w[i] : 1u;
    }
    VirtualGraph vg(g);
    vg.MakeGraph();
    memcpy(node_pointer, vg.nodePointer, sizeof(uint32_t) * n);
    memcpy(edge_list, vg.edgeList, sizeof(uint32_t) * (2 * m + n));
    const long long P = vg.numParts;
    for (long long i = 0; i < P && i < cap_parts; ++i) {
        parts[2 * i] = vg.partNodePointer[i].node;
        parts[2 * i + 1] = vg.partNodePointer[i].part;
    }
    return P;
}

}  // extern "C"
