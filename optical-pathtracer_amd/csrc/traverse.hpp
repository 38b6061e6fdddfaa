// Per-lane traversal of the 4-wide BVH (bvh4.hpp) for one ray: closest hit or any hit.
//
// Semantics follow the reference's mesh query (src/owe/scene/mesh.cpp:73-97 over
// src/owe/scene/bvh.hpp:31-78): a triangle counts when the line crosses it (watertight test) and its
// plane distance t lies strictly inside (tmin, best); the closest wins. The slab test keeps the
// reference float kernels' conservative far bound (shaders/geometry.slang:398-416): widened by a few
// ulps so rounding never culls a box the triangle test would hit.
//
// gfx950 layout: every lane walks its own ray. A node visit is one 128-byte line (eight 16-byte
// loads). Hit leaf children are intersected at once, nearest inner child is entered next, and the
// other hit inner children go on a short stack in LDS laid out [entry][lane] so a wave's pushes and
// pops touch 64 consecutive banks. The stack's capacity is the build's exact bound, so it never
// overflows.
#pragma once

#include "bvh4.hpp"
#include "dmath.hpp"

namespace owehip {

struct RayHit {
    float t;
    uint32_t slot;  // triangle record index (leaf order), kEmptyChild on a miss
};

// skipFace: a spawned ray's own face (its coplanar triangles are not crossings); ~0u for none.
// stack: this lane's first entry in LDS; stride: lanes per block.
// Contacts (world queries only; a triangle record's last word names its boundary): triangles of
// boundary `own` count from tmin, those of every other boundary from tminOther (below zero for a
// spawned ray: a touching surface at its origin); skipFace2 is a second face group that is skipped.
template <bool kAny>
__device__ inline RayHit traverseBvh4(const WideNode* __restrict__ nodes, const TriRecord* __restrict__ tris, f3 o, f3 d,
                                      float tmin, float tmax, uint32_t skipFace, uint32_t* stack, uint32_t stride,
                                      uint32_t own = ~0u, float tminOther = 0.0f, uint32_t skipFace2 = ~0u) {
    const bool contact = own != ~0u;
    const float tlow = contact ? fminf(tmin, tminOther) : tmin;
    const Shear sh = makeShear(d);
    const f3 inv = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    const bool nx = inv.x < 0.0f, ny = inv.y < 0.0f, nz = inv.z < 0.0f;
    RayHit hit{tmax, kEmptyChild};
    uint32_t sp = 0;
    uint32_t cur = 0;
    for (;;) {
        const float4* q = reinterpret_cast<const float4*>(nodes + cur);
        const float4 lx = q[0], hx = q[1], ly = q[2], hy = q[3], lz = q[4], hz = q[5];
        const uint4 ch = reinterpret_cast<const uint4*>(nodes + cur)[6];
        const float lxs[4] = {lx.x, lx.y, lx.z, lx.w}, hxs[4] = {hx.x, hx.y, hx.z, hx.w};
        const float lys[4] = {ly.x, ly.y, ly.z, ly.w}, hys[4] = {hy.x, hy.y, hy.z, hy.w};
        const float lzs[4] = {lz.x, lz.y, lz.z, lz.w}, hzs[4] = {hz.x, hz.y, hz.z, hz.w};
        const uint32_t cs[4] = {ch.x, ch.y, ch.z, ch.w};
        float tn[4];
        bool hitc[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float x0 = ((nx ? hxs[c] : lxs[c]) - o.x) * inv.x, x1 = ((nx ? lxs[c] : hxs[c]) - o.x) * inv.x;
            const float y0 = ((ny ? hys[c] : lys[c]) - o.y) * inv.y, y1 = ((ny ? lys[c] : hys[c]) - o.y) * inv.y;
            const float z0 = ((nz ? hzs[c] : lzs[c]) - o.z) * inv.z, z1 = ((nz ? lzs[c] : hzs[c]) - o.z) * inv.z;
            const float near = fmaxf(fmaxf(tlow, x0), fmaxf(y0, z0));
            float far = fminf(fminf(x1, y1), z1);
            far = far * (far >= 0.0f ? 1.0000004f : 0.9999996f);
            far = fminf(far, hit.t);
            tn[c] = near;
            hitc[c] = near <= far;
        }
        // Leaf children now, in slot order; inner children by entry distance (static indexing only,
        // so the candidates stay in registers).
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (!hitc[c] || !(cs[c] & kLeafBit)) continue;
            const uint32_t first = cs[c] & kFirstMask, count = ((cs[c] >> kCountShift) & 0xFu) + 1;
            for (uint32_t k = 0; k < count; ++k) {
                const uint32_t slot = first + k;
                const float4* tr = reinterpret_cast<const float4*>(tris + slot);
                const float4 a = tr[0], b = tr[1], cc = tr[2];
                if (skipFace != ~0u && __float_as_uint(b.w) == skipFace) continue;
                if (skipFace2 != ~0u && __float_as_uint(b.w) == skipFace2) continue;
                const float tl = contact && __float_as_uint(cc.w) != own ? tminOther : tmin;
                float t;
                if (!hitTriangle(sh, o, d, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(cc.x, cc.y, cc.z), tl, hit.t, t))
                    continue;
                hit.t = t;
                hit.slot = slot;
                if (kAny) return hit;
            }
        }
        // Inner children still nearer than the closest hit found so far.
        float key[4];
        uint32_t nd[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const bool inner = hitc[c] && !(cs[c] & kLeafBit) && tn[c] <= hit.t;
            key[c] = inner ? tn[c] : __builtin_inff();
            nd[c] = inner ? cs[c] : kEmptyChild;
        }
        // Four-key sorting network: (0,1) (2,3) (0,2) (1,3) (1,2).
        auto order = [&](int i, int j) {
            if (key[j] < key[i]) {
                const float tk = key[i];
                key[i] = key[j];
                key[j] = tk;
                const uint32_t tn2 = nd[i];
                nd[i] = nd[j];
                nd[j] = tn2;
            }
        };
        order(0, 1);
        order(2, 3);
        order(0, 2);
        order(1, 3);
        order(1, 2);
        // Empty entries sorted last (key +inf); an inner child entered at +inf sorts among them but
        // keeps its node index, so validity is the index, not the key.
#pragma unroll
        for (int k = 3; k >= 1; --k)
            if (nd[k] != kEmptyChild) stack[(sp++) * stride] = nd[k];
        if (nd[0] != kEmptyChild) {
            cur = nd[0];
            continue;
        }
        if (sp == 0) break;
        cur = stack[(--sp) * stride];
    }
    return hit;
}

// The same traversal one node at a time, for kernels whose lanes take a new ray when theirs ends
// (traceRefillKernel): TravState holds what traverseBvh4 keeps in registers between iterations, and
// stepBvh4 is one iteration of its loop. A ray visits the same nodes in the same order as in
// traverseBvh4, so the hit is the same bit for bit.
struct TravState {
    f3 o, d, inv;
    Shear sh;
    float tmin;
    RayHit hit;
    uint32_t sp, cur;
};

__device__ inline void beginBvh4(TravState& s, f3 o, f3 d, float tmin, float tmax) {
    s.o = o;
    s.d = d;
    s.sh = makeShear(d);
    s.inv = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    s.tmin = tmin;
    s.hit = RayHit{tmax, kEmptyChild};
    s.sp = 0;
    s.cur = 0;
}

// One node of the walk; true when the ray is done (s.hit holds the result).
template <bool kAny>
__device__ inline bool stepBvh4(TravState& s, const WideNode* __restrict__ nodes, const TriRecord* __restrict__ tris,
                                uint32_t* stack, uint32_t stride) {
    const bool nx = s.inv.x < 0.0f, ny = s.inv.y < 0.0f, nz = s.inv.z < 0.0f;
    const float4* q = reinterpret_cast<const float4*>(nodes + s.cur);
    const float4 lx = q[0], hx = q[1], ly = q[2], hy = q[3], lz = q[4], hz = q[5];
    const uint4 ch = reinterpret_cast<const uint4*>(nodes + s.cur)[6];
    const float lxs[4] = {lx.x, lx.y, lx.z, lx.w}, hxs[4] = {hx.x, hx.y, hx.z, hx.w};
    const float lys[4] = {ly.x, ly.y, ly.z, ly.w}, hys[4] = {hy.x, hy.y, hy.z, hy.w};
    const float lzs[4] = {lz.x, lz.y, lz.z, lz.w}, hzs[4] = {hz.x, hz.y, hz.z, hz.w};
    const uint32_t cs[4] = {ch.x, ch.y, ch.z, ch.w};
    float tn[4];
    bool hitc[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float x0 = ((nx ? hxs[c] : lxs[c]) - s.o.x) * s.inv.x, x1 = ((nx ? lxs[c] : hxs[c]) - s.o.x) * s.inv.x;
        const float y0 = ((ny ? hys[c] : lys[c]) - s.o.y) * s.inv.y, y1 = ((ny ? lys[c] : hys[c]) - s.o.y) * s.inv.y;
        const float z0 = ((nz ? hzs[c] : lzs[c]) - s.o.z) * s.inv.z, z1 = ((nz ? lzs[c] : hzs[c]) - s.o.z) * s.inv.z;
        const float near = fmaxf(fmaxf(s.tmin, x0), fmaxf(y0, z0));
        float far = fminf(fminf(x1, y1), z1);
        far = far * (far >= 0.0f ? 1.0000004f : 0.9999996f);
        far = fminf(far, s.hit.t);
        tn[c] = near;
        hitc[c] = near <= far;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        if (!hitc[c] || !(cs[c] & kLeafBit)) continue;
        const uint32_t first = cs[c] & kFirstMask, count = ((cs[c] >> kCountShift) & 0xFu) + 1;
        for (uint32_t k = 0; k < count; ++k) {
            const uint32_t slot = first + k;
            const float4* tr = reinterpret_cast<const float4*>(tris + slot);
            const float4 a = tr[0], b = tr[1], cc = tr[2];
            float t;
            if (!hitTriangle(s.sh, s.o, s.d, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(cc.x, cc.y, cc.z), s.tmin,
                             s.hit.t, t))
                continue;
            s.hit.t = t;
            s.hit.slot = slot;
            if (kAny) return true;
        }
    }
    float key[4];
    uint32_t nd[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const bool inner = hitc[c] && !(cs[c] & kLeafBit) && tn[c] <= s.hit.t;
        key[c] = inner ? tn[c] : __builtin_inff();
        nd[c] = inner ? cs[c] : kEmptyChild;
    }
    auto order = [&](int i, int j) {
        if (key[j] < key[i]) {
            const float tk = key[i];
            key[i] = key[j];
            key[j] = tk;
            const uint32_t tn2 = nd[i];
            nd[i] = nd[j];
            nd[j] = tn2;
        }
    };
    order(0, 1);
    order(2, 3);
    order(0, 2);
    order(1, 3);
    order(1, 2);
#pragma unroll
    for (int k = 3; k >= 1; --k)
        if (nd[k] != kEmptyChild) stack[(s.sp++) * stride] = nd[k];
    if (nd[0] != kEmptyChild) {
        s.cur = nd[0];
        return false;
    }
    if (s.sp == 0) return true;
    s.cur = stack[(--s.sp) * stride];
    return false;
}

}  // namespace owehip
