// Device-side BLAS construction for large meshes: a linear BVH over Morton codes
// of triangle centroids, boxes fitted bottom-up, subtrees of at most four
// triangles collapsed into leaves, emitted in the node layout of layout.hpp.
//
// Replaces the host SAH build (reference accel/bvh.cpp:131-270) where a scene
// asks for it; closest hits do not depend on the tree (they are the minimum of
// (t, object, prim)), so both builds give the same query results.
#pragma once

#include <hip/hip_runtime_api.h>

#include "layout.hpp"

#include <cstdint>

namespace ab {

// One mesh's build between its two phases. Owns device scratch memory.
class DeviceBlas {
public:
    DeviceBlas() = default;
    ~DeviceBlas();
    DeviceBlas(const DeviceBlas&) = delete;
    DeviceBlas& operator=(const DeviceBlas&) = delete;

    // Phase 1: sorts the n (> 4) triangles (host, in the mesh's kept order) and
    // builds the hierarchy; afterwards nodes() and levels() are known.
    hipError_t prepare(const DevTri* tris, uint32_t n, hipStream_t s);
    // Nodes the emitted tree has, and its levels (a leaf is one level).
    uint32_t nodes() const { return nodes_; }
    uint32_t levels() const { return levels_; }
    // Phase 2: writes nodes() nodes at out_nodes (global index node_base for the
    // root) and the n triangles in leaf order at out_tris (global index tri_base).
    hipError_t emit(DevNode* out_nodes, uint32_t node_base, DevTri* out_tris, uint32_t tri_base, hipStream_t s);

private:
    void* mem_{nullptr};
    uint32_t n_{0};
    uint32_t nodes_{0};
    uint32_t levels_{0};
    int idx_bits_{0};
    // Pointers into mem_.
    DevTri* tris_{nullptr};
    void* boxes_{nullptr};     // blas::Box: triangles [0, n) by index, internal nodes [n, 2n - 1)
    unsigned long long* keys_{nullptr};
    unsigned long long* sorted_{nullptr};
    int32_t* child_{nullptr};  // 2 per internal node: >= 0 internal, < 0 ~leaf
    uint32_t* parent_{nullptr}; // per node as for the boxes; the root's is ~0
    uint32_t* first_{nullptr};  // per internal node: first sorted triangle
    uint32_t* count_{nullptr};  // per internal node: triangles below
    uint32_t* levels_v_{nullptr}; // per internal node: levels of its emitted subtree
    uint32_t* visits_{nullptr};
    uint32_t* keep_{nullptr};   // per internal node: emitted as a node (more than 4 below)
    uint32_t* index_{nullptr};  // exclusive scan of keep_
    uint32_t* bounds_{nullptr}; // 6 order-preserving centroid bounds
    void* scratch_{nullptr};    // sort / scan temporary storage
    size_t scratch_bytes_{0};
};

} // namespace ab
