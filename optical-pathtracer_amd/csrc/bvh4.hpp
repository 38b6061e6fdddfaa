// The device layout of a triangle mesh for batch ray queries and for the transport kernels:
// 4-wide BVH nodes of exactly one 128-byte cache line, and 48-byte triangle records in leaf order.
//
// The reference traverses a binary binned-SAH BVH (src/owe/scene/bvh.cpp:27-337, bvh.hpp:31-78) and
// exports it to its float kernels as 32-byte binary nodes (src/owe/backends/gpu/gpu_scene.cpp:36-66).
// Here the host builds a binary SAH tree and collapses it into 4-child nodes, so one lane's node
// fetch is one cache line holding all four child boxes (structure-of-arrays inside the node), and
// a traversal step tests four boxes instead of two.
#pragma once

#include <cstdint>
#include <vector>

namespace owehip {

struct alignas(16) WideNode {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];  // child boxes; an empty child has lo = +inf, hi = -inf
    uint32_t child[4];  // inner: node index; leaf: kLeafBit | (count - 1) << kCountShift | first slot; kEmpty
    uint32_t meta[4];   // meta[0]: number of non-empty children
};
static_assert(sizeof(WideNode) == 128, "a wide node is one 128-byte cache line");

struct alignas(16) TriRecord {
    float v0[3];
    uint32_t prim;  // the input triangle index
    float v1[3];
    uint32_t face;  // the coplanar face the triangle belongs to (a spawned ray skips its own face)
    float v2[3];
    uint32_t owner; // the boundary (scene) or 0 (a lone mesh)
};
static_assert(sizeof(TriRecord) == 48, "a triangle record is 48 bytes");

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kCountShift = 27;
constexpr uint32_t kFirstMask = (1u << kCountShift) - 1;
constexpr uint32_t kMaxLeaf = 8;  // triangles per leaf child (count - 1 fits the 4-bit field)
constexpr uint32_t kEmptyChild = 0xFFFFFFFFu;

struct Bvh4 {
    std::vector<WideNode> nodes;   // node 0 is the root
    std::vector<TriRecord> tris;   // leaf order
    uint32_t depth = 0;            // levels of wide nodes
    uint32_t stackBound = 0;       // the most entries a traversal stack can hold
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};  // bounds of everything
    double seconds = 0;
};

// Builds the wide BVH of a triangle soup. owner/face may be null (then 0 and the triangle's index).
Bvh4 buildBvh4(const float* positions, uint32_t nVertices, const uint32_t* triangles, uint32_t nTriangles,
               const uint32_t* owner = nullptr, const uint32_t* face = nullptr);

}  // namespace owehip
