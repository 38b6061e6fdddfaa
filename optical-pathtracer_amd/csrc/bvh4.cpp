// Host construction of the 4-wide BVH (see bvh4.hpp). A binary tree is built top-down with a
// binned surface-area heuristic over all three axes (16 bins each), large ranges in parallel on the
// host threads; it is then collapsed into 4-child nodes by repeatedly opening the largest child.
#include "bvh4.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <future>
#include <limits>
#include <stdexcept>
#include <thread>

namespace owehip {

namespace {

constexpr int kBins = 16;
constexpr uint32_t kParallelRange = 1u << 15;

struct FBox {
    float lo[3] = {std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
                   std::numeric_limits<float>::infinity()};
    float hi[3] = {-std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity(),
                   -std::numeric_limits<float>::infinity()};
    void grow(const FBox& b) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    void grow(const float* p) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    bool empty() const { return !(lo[0] <= hi[0] && lo[1] <= hi[1] && lo[2] <= hi[2]); }
    double area() const {
        if (empty()) return 0;
        const double x = double(hi[0]) - lo[0], y = double(hi[1]) - lo[1], z = double(hi[2]) - lo[2];
        return 2 * (x * y + y * z + z * x);
    }
};

struct BNode {
    FBox box;
    uint32_t left = 0, right = 0;
    uint32_t first = 0, count = 0;  // count > 0: leaf over idx[first, first + count)
};

struct Builder {
    const std::vector<FBox>& boxes;
    const std::vector<float>& cen;  // 3 per triangle
    std::vector<uint32_t>& idx;
    std::vector<BNode> nodes;
    std::atomic<uint32_t> used{0};

    Builder(const std::vector<FBox>& b, const std::vector<float>& c, std::vector<uint32_t>& i)
        : boxes(b), cen(c), idx(i), nodes(2 * std::max<size_t>(1, i.size())) {}

    uint32_t alloc() {
        const uint32_t k = used.fetch_add(1);
        if (k >= nodes.size()) throw std::runtime_error("bvh: node pool exhausted");
        return k;
    }

    uint32_t build(uint32_t begin, uint32_t end, int depth) {
        const uint32_t me = alloc();
        BNode& node = nodes[me];
        FBox box, cbox;
        for (uint32_t i = begin; i < end; ++i) {
            box.grow(boxes[idx[i]]);
            cbox.grow(&cen[3 * size_t(idx[i])]);
        }
        node.box = box;
        const uint32_t n = end - begin;
        if (n == 1 || depth > 48) {
            if (n > kMaxLeaf) throw std::runtime_error("bvh: degenerate input deeper than 48 levels");
            node.first = begin;
            node.count = n;
            return me;
        }
        // Best binned split over the three axes.
        double bestCost = std::numeric_limits<double>::infinity();
        int bestAxis = -1, bestBin = -1;
        float scale[3], base[3];
        for (int a = 0; a < 3; ++a) {
            const float ext = cbox.hi[a] - cbox.lo[a];
            base[a] = cbox.lo[a];
            scale[a] = ext > 0 ? float(kBins) / ext : 0.0f;
            if (!(ext > 0)) continue;
            FBox bb[kBins];
            uint32_t bc[kBins] = {};
            for (uint32_t i = begin; i < end; ++i) {
                const uint32_t t = idx[i];
                const int k = std::min(kBins - 1, int((cen[3 * size_t(t) + a] - base[a]) * scale[a]));
                bc[k]++;
                bb[k].grow(boxes[t]);
            }
            double rightArea[kBins];
            uint32_t rightCount[kBins];
            FBox acc;
            uint32_t cnt = 0;
            for (int k = kBins - 1; k >= 1; --k) {
                acc.grow(bb[k]);
                cnt += bc[k];
                rightArea[k] = acc.area();
                rightCount[k] = cnt;
            }
            FBox left;
            uint32_t nl = 0;
            for (int k = 1; k < kBins; ++k) {
                left.grow(bb[k - 1]);
                nl += bc[k - 1];
                if (nl == 0 || rightCount[k] == 0) continue;
                const double cost = left.area() * nl + rightArea[k] * rightCount[k];
                if (cost < bestCost) {
                    bestCost = cost;
                    bestAxis = a;
                    bestBin = k;
                }
            }
        }
        const double parentArea = box.area();
        // Leaf when the split does not pay for itself (traversal step ≈ one triangle test).
        if (n <= 4 && (bestAxis < 0 || (parentArea > 0 && 1.0 + bestCost / parentArea >= double(n)))) {
            node.first = begin;
            node.count = n;
            return me;
        }
        uint32_t mid;
        if (bestAxis < 0) {
            mid = begin + n / 2;  // all centroids coincide: split the range in half
        } else {
            const int a = bestAxis;
            const float b0 = base[a], s = scale[a];
            const int split = bestBin;
            mid = uint32_t(std::partition(idx.begin() + begin, idx.begin() + end,
                                          [&](uint32_t t) {
                                              return std::min(kBins - 1, int((cen[3 * size_t(t) + a] - b0) * s)) < split;
                                          }) -
                           idx.begin());
            if (mid == begin || mid == end) mid = begin + n / 2;
        }
        uint32_t l, r;
        if (n >= kParallelRange) {
            auto fut = std::async(std::launch::async, [&] { return build(begin, mid, depth + 1); });
            r = build(mid, end, depth + 1);
            l = fut.get();
        } else {
            l = build(begin, mid, depth + 1);
            r = build(mid, end, depth + 1);
        }
        nodes[me].left = l;
        nodes[me].right = r;
        return me;
    }
};

struct Collapser {
    const std::vector<BNode>& bn;
    std::vector<WideNode>& out;
    uint32_t maxDepth = 0;

    static void setEmpty(WideNode& w, int i) {
        const float inf = std::numeric_limits<float>::infinity();
        w.lox[i] = w.loy[i] = w.loz[i] = inf;
        w.hix[i] = w.hiy[i] = w.hiz[i] = -inf;
        w.child[i] = kEmptyChild;
    }

    // Returns the wide node's index and its stack bound (the most entries below and at it).
    uint32_t collapse(uint32_t b, uint32_t level, uint32_t& bound) {
        const uint32_t me = uint32_t(out.size());
        out.emplace_back();
        maxDepth = std::max(maxDepth, level + 1);
        uint32_t kids[4];
        int nk = 0;
        if (bn[b].count > 0) {
            kids[nk++] = b;  // a lone leaf under the root
        } else {
            kids[nk++] = bn[b].left;
            kids[nk++] = bn[b].right;
            while (nk < 4) {
                int pick = -1;
                double area = -1;
                for (int i = 0; i < nk; ++i)
                    if (bn[kids[i]].count == 0 && bn[kids[i]].box.area() > area) {
                        area = bn[kids[i]].box.area();
                        pick = i;
                    }
                if (pick < 0) break;
                const uint32_t open = kids[pick];
                kids[pick] = bn[open].left;
                kids[nk++] = bn[open].right;
            }
        }
        WideNode w{};
        for (int i = 0; i < 4; ++i) setEmpty(w, i);
        int inner = 0;
        for (int i = 0; i < nk; ++i) {
            const BNode& c = bn[kids[i]];
            w.lox[i] = c.box.lo[0];
            w.loy[i] = c.box.lo[1];
            w.loz[i] = c.box.lo[2];
            w.hix[i] = c.box.hi[0];
            w.hiy[i] = c.box.hi[1];
            w.hiz[i] = c.box.hi[2];
            if (c.count > 0) {
                if (c.count > kMaxLeaf || c.first > kFirstMask) throw std::runtime_error("bvh: leaf encoding overflow");
                w.child[i] = kLeafBit | ((c.count - 1) << kCountShift) | c.first;
            } else {
                ++inner;
            }
        }
        w.meta[0] = uint32_t(nk);
        out[me] = w;
        uint32_t below = 0;
        for (int i = 0; i < nk; ++i) {
            if (bn[kids[i]].count > 0) continue;
            uint32_t cb = 0;
            const uint32_t ci = collapse(kids[i], level + 1, cb);
            out[me].child[i] = ci;
            below = std::max(below, cb);
        }
        bound = inner > 0 ? uint32_t(inner - 1) + below : 0;
        return me;
    }
};

}  // namespace

Bvh4 buildBvh4(const float* pos, uint32_t nv, const uint32_t* tris, uint32_t nt, const uint32_t* owner,
               const uint32_t* face) {
    const auto t0 = std::chrono::steady_clock::now();
    Bvh4 r;
    if (nt >= (1u << kCountShift)) throw std::runtime_error("bvh: too many triangles");
    std::vector<FBox> boxes(nt);
    std::vector<float> cen(3 * size_t(nt));
    std::vector<uint32_t> idx(nt);
    for (uint32_t i = 0; i < nt; ++i) {
        for (int k = 0; k < 3; ++k)
            if (tris[3 * size_t(i) + k] >= nv) throw std::runtime_error("mesh: triangle index out of range");
        FBox b;
        for (int k = 0; k < 3; ++k) b.grow(&pos[3 * size_t(tris[3 * size_t(i) + k])]);
        boxes[i] = b;
        for (int a = 0; a < 3; ++a) cen[3 * size_t(i) + a] = 0.5f * (b.lo[a] + b.hi[a]);
        idx[i] = i;
    }
    if (nt == 0) {
        WideNode w{};
        for (int i = 0; i < 4; ++i) Collapser::setEmpty(w, i);
        r.nodes.push_back(w);
        r.depth = 1;
        return r;
    }
    Builder b(boxes, cen, idx);
    const uint32_t root = b.build(0, nt, 0);
    b.nodes.resize(b.used.load());
    Collapser c{b.nodes, r.nodes};
    r.nodes.reserve(b.nodes.size() / 2 + 1);
    c.collapse(root, 0, r.stackBound);
    r.depth = c.maxDepth;
    for (int a = 0; a < 3; ++a) {
        r.lo[a] = b.nodes[root].box.lo[a];
        r.hi[a] = b.nodes[root].box.hi[a];
    }
    r.tris.resize(nt);
    for (uint32_t s = 0; s < nt; ++s) {
        const uint32_t t = idx[s];
        TriRecord& q = r.tris[s];
        for (int k = 0; k < 3; ++k) {
            q.v0[k] = pos[3 * size_t(tris[3 * size_t(t)]) + k];
            q.v1[k] = pos[3 * size_t(tris[3 * size_t(t) + 1]) + k];
            q.v2[k] = pos[3 * size_t(tris[3 * size_t(t) + 2]) + k];
        }
        q.prim = t;
        q.face = face ? face[t] : t;
        q.owner = owner ? owner[t] : 0;
    }
    r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return r;
}

}  // namespace owehip
