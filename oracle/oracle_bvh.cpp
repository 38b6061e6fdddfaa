// TEST INFRASTRUCTURE ONLY — the double-precision CPU oracle (see oracle_core.hpp).
// Sequential restatement of the reference's binned-SAH builder (src/owe/scene/bvh.cpp:27-337); the
// reference builds its upper levels in parallel but documents (bvh.cpp:292-295) that the hierarchy
// and primitive order equal a sequential build, which is what this file performs.
#include "oracle_bvh.hpp"

#include <algorithm>

namespace oracle {

namespace {

constexpr int kBins = 16;

struct Prim {
    Box box;
    uint32_t index;
    V3 center() const { return box.center(); }
};

struct Span {
    Box box, cbox;
    uint32_t begin = 0, end = 0;
};

void boundSpan(const std::vector<Prim>& p, Span& s) {
    s.box = s.cbox = Box{};
    for (uint32_t i = s.begin; i < s.end; ++i) {
        s.box.grow(p[i].box);
        s.cbox.grow(p[i].center());
    }
}

// Two-pointer partition: scanning from the left for a record failing `keep` and from the right for
// one passing it, and exchanging them (bvh.cpp:38-49).
template <class Keep>
uint32_t twoPointerPartition(std::vector<Prim>& p, uint32_t lo, uint32_t hi, Keep keep) {
    for (;;) {
        while (lo != hi && keep(p[lo])) ++lo;
        if (lo == hi) return lo;
        --hi;
        while (lo != hi && !keep(p[hi])) --hi;
        if (lo == hi) return lo;
        std::swap(p[lo], p[hi]);
        ++lo;
    }
}

bool splitSpan(std::vector<Prim>& p, const Span& s, uint32_t maxLeaf, int depth, Span& L, Span& R, uint32_t& axisOut) {
    const uint32_t n = s.end - s.begin;
    if (n <= maxLeaf || depth > 60) return false;
    const V3 ext = s.cbox.hi - s.cbox.lo;
    const int axis = (ext.x > ext.y && ext.x > ext.z) ? 0 : (ext.y > ext.z ? 1 : 2);
    if (!(ext[axis] > 0)) return false;

    Box binBox[kBins], binCen[kBins];
    uint32_t binCount[kBins] = {};
    const double lo = s.cbox.lo[axis], scale = kBins / ext[axis];
    auto binOf = [&](const Prim& q) { return std::min(kBins - 1, int((q.center()[axis] - lo) * scale)); };
    for (uint32_t i = s.begin; i < s.end; ++i) {
        const int k = binOf(p[i]);
        binCount[k]++;
        binBox[k].grow(p[i].box);
        binCen[k].grow(p[i].center());
    }
    Box rightBox[kBins];
    uint32_t rightCount[kBins] = {};
    for (int k = kBins - 1; k >= 1; --k) {
        if (k + 1 < kBins) {
            rightBox[k] = rightBox[k + 1];
            rightCount[k] = rightCount[k + 1];
        }
        rightCount[k] += binCount[k];
        if (binCount[k]) rightBox[k].grow(binBox[k]);
    }
    double bestCost = kInf;
    int best = -1;
    Box left;
    uint32_t nl = 0;
    for (int k = 1; k < kBins; ++k) {
        if (binCount[k - 1]) {
            left.grow(binBox[k - 1]);
            nl += binCount[k - 1];
        }
        const uint32_t nr = rightCount[k];
        if (nl == 0 || nr == 0) continue;
        const double cost = left.area() * nl + rightBox[k].area() * nr;
        if (cost < bestCost) {
            bestCost = cost;
            best = k;
        }
    }
    uint32_t mid;
    bool fromBins = false;
    if (best < 0) {
        mid = s.begin + n / 2;
        std::nth_element(p.begin() + s.begin, p.begin() + mid, p.begin() + s.end,
                         [&](const Prim& a, const Prim& b) { return a.center()[axis] < b.center()[axis]; });
    } else {
        if (n <= 16 && bestCost >= s.box.area() * n) return false;
        mid = twoPointerPartition(p, s.begin, s.end, [&](const Prim& q) { return binOf(q) < best; });
        fromBins = mid != s.begin && mid != s.end;
        if (!fromBins) mid = s.begin + n / 2;
    }
    L.begin = s.begin;
    L.end = R.begin = mid;
    R.end = s.end;
    if (fromBins) {
        L.box = L.cbox = R.box = R.cbox = Box{};
        for (int k = 0; k < kBins; ++k) {
            if (!binCount[k]) continue;
            Span& h = k < best ? L : R;
            h.box.grow(binBox[k]);
            h.cbox.grow(binCen[k]);
        }
    } else {
        boundSpan(p, L);
        boundSpan(p, R);
    }
    axisOut = uint32_t(axis);
    return true;
}

// Depth-first layout: a node, its left subtree, then its right subtree (whose root index it keeps).
void buildRec(std::vector<Prim>& p, const Span& s, uint32_t maxLeaf, int depth, std::vector<Node>& out) {
    const uint32_t me = uint32_t(out.size());
    out.emplace_back();
    out[me].box = s.box;
    Span L, R;
    uint32_t axis = 0;
    if (!splitSpan(p, s, maxLeaf, depth, L, R, axis)) {
        out[me].start = s.begin;
        out[me].count = s.end - s.begin;
        return;
    }
    out[me].axis = axis;
    buildRec(p, L, maxLeaf, depth + 1, out);
    out[me].start = uint32_t(out.size());
    buildRec(p, R, maxLeaf, depth + 1, out);
}

}  // namespace

void Bvh::build(const std::vector<Box>& prims, uint32_t maxLeaf) {
    nodes_.clear();
    order_.assign(prims.size(), 0);
    if (prims.empty()) return;
    std::vector<Prim> p(prims.size());
    for (size_t i = 0; i < prims.size(); ++i) p[i] = {prims[i], uint32_t(i)};
    Span root;
    root.begin = 0;
    root.end = uint32_t(p.size());
    boundSpan(p, root);
    nodes_.reserve(2 * prims.size() / std::max(1u, maxLeaf) + 1);
    buildRec(p, root, maxLeaf, 0, nodes_);
    for (size_t i = 0; i < p.size(); ++i) order_[i] = p[i].index;
}

// ---------------------------------------------------------------- Mesh

Mesh::Mesh(std::vector<V3> pos, std::vector<std::array<uint32_t, 3>> tris) : pos_(std::move(pos)), tris_(std::move(tris)) {
    std::vector<Box> boxes(tris_.size());
    cdf_.resize(tris_.size());
    double acc = 0;
    for (size_t i = 0; i < tris_.size(); ++i) {
        const V3 &a = pos_[tris_[i][0]], &b = pos_[tris_[i][1]], &c = pos_[tris_[i][2]];
        Box& bx = boxes[i];
        bx.grow(a);
        bx.grow(b);
        bx.grow(c);
        const double pad = 1e-12 * (1 + maxAbs(bx.hi) + maxAbs(bx.lo));
        bx.lo = bx.lo - V3(pad, pad, pad);
        bx.hi = bx.hi + V3(pad, pad, pad);
        acc += 0.5 * length(cross(b - a, c - a));
        cdf_[i] = acc;
    }
    area_ = acc;
    bvh_.build(boxes, 4);
}

V3 Mesh::normalOf(uint32_t prim) const {
    const auto& t = tris_[prim];
    return normalize(triNormal(pos_[t[0]], pos_[t[1]], pos_[t[2]]));
}

bool Mesh::sameFace(uint32_t a, uint32_t b) const {
    if (a >= tris_.size() || b >= tris_.size()) return false;
    if (a == b) return true;
    const auto &A = tris_[a], &B = tris_[b];
    auto onA = [&](uint32_t v) {
        for (uint32_t u : A)
            if (pos_[v].x == pos_[u].x && pos_[v].y == pos_[u].y && pos_[v].z == pos_[u].z) return true;
        return false;
    };
    int shared = 0;
    for (uint32_t v : B) shared += onA(v) ? 1 : 0;
    if (shared < 2) return false;
    const V3 na = cross(pos_[A[1]] - pos_[A[0]], pos_[A[2]] - pos_[A[0]]);
    if (dot(na, cross(pos_[B[1]] - pos_[B[0]], pos_[B[2]] - pos_[B[0]])) <= 0) return false;
    for (uint32_t v : B)
        if (!onA(v) && dot(na, pos_[v] - pos_[A[0]]) != 0) return false;
    return true;
}

bool Mesh::triangleT(uint32_t prim, V3 o, V3 d, double tmin, double tmax, double& t) const {
    const auto& tri = tris_[prim];
    const V3 &v0 = pos_[tri[0]], &v1 = pos_[tri[1]], &v2 = pos_[tri[2]];
    if (!lineCrossesTri(o, d, v0, v1, v2)) return false;
    const V3 n = normalize(triNormal(v0, v1, v2));
    t = -dot(o - v0, n) / dot(d, n);
    return t > tmin && t < tmax;
}

bool Mesh::closest(V3 o, V3 d, double tmin, double tmax, double& tOut, uint32_t& primOut, uint32_t skipFace,
                   uint64_t* boxTests, uint64_t* triTests, const uint32_t* group, uint32_t skipFace2) const {
    bool found = false;
    double best = tmax;
    const auto& ord = bvh_.order();
    bvh_.traverse(
        o, d, tmin, best,
        [&](uint32_t slot, double& tm) {
            const uint32_t i = ord[slot];
            double t;
            if (!triangleT(i, o, d, tmin, tm, t)) return;
            if (skipFace != ~0u && (group ? group[i] == group[skipFace] : sameFace(skipFace, i))) return;
            if (skipFace2 != ~0u && (group ? group[i] == group[skipFace2] : sameFace(skipFace2, i))) return;
            tm = t;
            found = true;
            tOut = t;
            primOut = i;
        },
        boxTests, triTests);
    return found;
}

bool Mesh::any(V3 o, V3 d, double tmin, double tmax, uint32_t skipFace) const {
    bool hit = false;
    const auto& ord = bvh_.order();
    double tm = tmax;
    // The first accepted crossing ends the search: shrinking tmax to -inf empties the stack.
    bvh_.traverse(o, d, tmin, tm, [&](uint32_t slot, double& m) {
        if (hit) return;
        const uint32_t i = ord[slot];
        double t;
        if (!triangleT(i, o, d, tmin, tmax, t)) return;
        if (skipFace != ~0u && sameFace(skipFace, i)) return;
        hit = true;
        m = -kInf;
    });
    return hit;
}

bool Mesh::bruteClosest(V3 o, V3 d, double tmin, double tmax, double& tOut, uint32_t& primOut) const {
    bool found = false;
    double best = tmax;
    for (uint32_t i = 0; i < tris_.size(); ++i) {
        double t;
        if (!triangleT(i, o, d, tmin, best, t)) continue;
        best = t;
        tOut = t;
        primOut = i;
        found = true;
    }
    return found;
}

}  // namespace oracle
