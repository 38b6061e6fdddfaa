// TEST INFRASTRUCTURE ONLY — the double-precision CPU oracle (see oracle_core.hpp).
//
// Binned-SAH BVH and triangle-mesh queries restated from the reference:
//   AABB slab test with the far bound widened by 4e-16   src/owe/core/math.hpp:187-205
//   binned SAH (16 bins, widest centroid axis, median
//   fallback, two-pointer partition, leaf rule)           src/owe/scene/bvh.cpp:27-337
//   ordered traversal, 128-entry stack, deferred culling  src/owe/scene/bvh.hpp:31-78
//   mesh closest hit (watertight crossing, plane t)       src/owe/scene/mesh.cpp:12-97
//   any-hit: first crossing in (tmin, tmax) ends the search, as the float kernel's
//   occludedMesh                                          src/owe/backends/gpu/shaders/geometry.slang:577-615
#pragma once

#include <array>
#include <vector>

#include "oracle_core.hpp"

namespace oracle {

struct Box {
    V3 lo{kInf, kInf, kInf}, hi{-kInf, -kInf, -kInf};
    bool valid() const { return lo.x <= hi.x && lo.y <= hi.y && lo.z <= hi.z; }
    void grow(V3 p) {
        lo = vmin(lo, p);
        hi = vmax(hi, p);
    }
    void grow(const Box& b) {
        lo = vmin(lo, b.lo);
        hi = vmax(hi, b.hi);
    }
    V3 center() const { return (lo + hi) * 0.5; }
    double area() const {
        if (!valid()) return 0;
        const V3 e = hi - lo;
        return 2.0 * (e.x * e.y + e.y * e.z + e.z * e.x);
    }
};

// A ray prepared for slab tests: near/far planes picked once from the direction's signs.
struct Slabs {
    double o[3], inv[3];
    bool neg[3];
    Slabs(V3 org, V3 dir) {
        for (int a = 0; a < 3; ++a) {
            o[a] = org[a];
            inv[a] = 1.0 / dir[a];
            neg[a] = inv[a] < 0;
        }
    }
    bool hit(const Box& b, double tmin, double tmax, double& tEnter) const {
        for (int a = 0; a < 3; ++a) {
            const double n = neg[a] ? b.hi[a] : b.lo[a], f = neg[a] ? b.lo[a] : b.hi[a];
            const double t0 = (n - o[a]) * inv[a];
            const double t1 = (f - o[a]) * inv[a] * (1.0 + 4e-16);
            tmin = t0 > tmin ? t0 : tmin;
            tmax = t1 < tmax ? t1 : tmax;
            if (tmin > tmax) return false;
        }
        tEnter = tmin;
        return true;
    }
};

struct Node {
    Box box;
    uint32_t start = 0;  // leaf: first slot in order; inner: right child
    uint32_t count = 0;  // > 0 for a leaf
    uint32_t axis = 0;
};

class Bvh {
public:
    void build(const std::vector<Box>& prims, uint32_t maxLeaf);
    const std::vector<Node>& nodes() const { return nodes_; }
    const std::vector<uint32_t>& order() const { return order_; }

    // visit(slot, tmax&) for candidate slots roughly front to back; visit may shrink tmax.
    // `visits` (optional) counts inner-node box tests and leaf primitives offered.
    template <class F>
    void traverse(V3 o, V3 d, double tmin, double& tmax, F&& visit, uint64_t* boxTests = nullptr,
                  uint64_t* leafPrims = nullptr) const {
        if (nodes_.empty()) return;
        const Slabs rs(o, d);
        struct Entry {
            uint32_t node;
            double t;
        } stack[128];
        int sp = 0;
        uint32_t cur = 0;
        double tEnter;
        if (boxTests) ++*boxTests;
        if (!rs.hit(nodes_[0].box, tmin, tmax, tEnter)) return;
        for (;;) {
            const Node& nd = nodes_[cur];
            bool descend = false;
            if (nd.count > 0) {
                if (leafPrims) *leafPrims += nd.count;
                for (uint32_t i = 0; i < nd.count; ++i) visit(nd.start + i, tmax);
            } else {
                uint32_t a = cur + 1, b = nd.start;
                if (rs.neg[nd.axis]) std::swap(a, b);
                double ta, tb;
                if (boxTests) *boxTests += 2;
                const bool ha = rs.hit(nodes_[a].box, tmin, tmax, ta);
                const bool hb = rs.hit(nodes_[b].box, tmin, tmax, tb);
                if (ha && hb) {
                    if (tb < ta) {
                        std::swap(a, b);
                        std::swap(ta, tb);
                    }
                    if (sp < 128) stack[sp++] = {b, tb};
                    cur = a;
                    descend = true;
                } else if (ha || hb) {
                    cur = ha ? a : b;
                    descend = true;
                }
            }
            if (descend) continue;
            while (sp > 0 && stack[sp - 1].t > tmax) --sp;
            if (sp == 0) break;
            cur = stack[--sp].node;
        }
    }

private:
    std::vector<Node> nodes_;
    std::vector<uint32_t> order_;
};

// A triangle mesh in its own frame with its BVH (leaf ≤ 4).
class Mesh {
public:
    Mesh(std::vector<V3> pos, std::vector<std::array<uint32_t, 3>> tris);
    // Closest crossing in (tmin, tmax); prim = triangle index. skipFace: triangles on the same face
    // as this one are not crossings (a spawned ray's own face); ~0u for none. With `group` (a face id
    // per triangle), every triangle of skipFace's group is skipped instead of its adjacent ones.
    bool closest(V3 o, V3 d, double tmin, double tmax, double& t, uint32_t& prim, uint32_t skipFace = ~0u,
                 uint64_t* boxTests = nullptr, uint64_t* triTests = nullptr, const uint32_t* group = nullptr,
                 uint32_t skipFace2 = ~0u) const;
    bool any(V3 o, V3 d, double tmin, double tmax, uint32_t skipFace = ~0u) const;
    bool bruteClosest(V3 o, V3 d, double tmin, double tmax, double& t, uint32_t& prim) const;
    // Adjacent, exactly coplanar triangles with the same orientation form one face.
    bool sameFace(uint32_t a, uint32_t b) const;
    const Bvh& bvh() const { return bvh_; }
    const std::vector<V3>& positions() const { return pos_; }
    const std::vector<std::array<uint32_t, 3>>& triangles() const { return tris_; }
    Box bounds() const { return bvh_.nodes().empty() ? Box{} : bvh_.nodes()[0].box; }
    double area() const { return area_; }
    const std::vector<double>& areaCdf() const { return cdf_; }
    V3 normalOf(uint32_t prim) const;
    // A triangle's crossing t (false when the line misses it or t is outside (tmin, tmax)).
    bool triangleT(uint32_t prim, V3 o, V3 d, double tmin, double tmax, double& t) const;

private:
    std::vector<V3> pos_;
    std::vector<std::array<uint32_t, 3>> tris_;
    std::vector<double> cdf_;
    double area_ = 0;
    Bvh bvh_;
};

}  // namespace oracle
