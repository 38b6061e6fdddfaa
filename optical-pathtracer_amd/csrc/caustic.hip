// Differentiable caustic design: owe_hip_caustic_* of include/owe_hip_caustic.h.
//
// Forward: one thread per particle, each wave over a contiguous range of particles. A particle of
// pass k is stream (hashCombine(seed, i), k): a wavelength importance-sampled over the visible range
// and weighted by the sun's luminous spectrum, and a point on the slab's entry face, stratified over
// square tiles (a wave's particles share a tile or two). It enters at normal incidence (weight 1 − R),
// crosses the slab along +z to the exit surface z = t + h(x, y), refracts out of it (weight 1 − R;
// total internal reflection ends it) and lands on the wall, where its weight spreads over the 3 × 3
// pixels around it with quadratic B-spline weights (continuously differentiable in the landing
// point, so the loss is C¹ in the heights) as 64-bit fixed-point adds — exact, so a pass is
// reproducible.
//
// Backward (path replay): the same particles are drawn again from their seeds and the same float
// operations give the same wall point; the exit computation is also carried through three-component
// dual numbers (∂/∂h, ∂/∂h_x, ∂/∂h_y at the particle's exit point), which with the spline weights
// and ∂L/∂E at the nine pixels gives ∂L/∂(h, h_x, h_y) for this particle. h, h_x and h_y are linear in
// the 4 × 4 heights around the cell (bicubic Hermite patch with central-difference node slopes), so
// the particle adds its share to those sixteen heights: no stored paths. Both passes gather their
// adds (pixels, nodes) in a per-wave table in LDS and send each touched pixel or node to HBM once
// per wave.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "common.hpp"
#include "owe_hip_caustic.h"
#include "transport_dev.hpp"

using namespace owehip;

namespace owehip {
namespace {

constexpr int kCBlock = 256;
constexpr double kSplatUnits = 4294967296.0;  // fixed-point units per unit of irradiance

struct CArgs {
    const float* heights;
    uint32_t nodes, pixels;
    float slabHalf, cell, z0, zWall, wallHalf, pix;
    uint64_t particles, seed, pass;
    DMedium slab, air;
    float kelvin, invY;      // sun temperature; 1 / ∫ B ȳ over 360–830 nm
    float tileArea;          // a tile's area / pixel area: with a tile's particle count, irradiance per unit weight
    uint32_t tiles;          // tiles per side (a power of two)
    unsigned long long* acc;
    float* irradiance;
    const float* gradE;
    float* gradH;
};

// Three-component dual numbers: value and partials with respect to (h, h_x, h_y).
struct D3 {
    float v, d[3];
};
__device__ __forceinline__ D3 mkD(float v, int seed) {
    D3 r{v, {0.0f, 0.0f, 0.0f}};
    if (seed >= 0) r.d[seed] = 1.0f;
    return r;
}
__device__ __forceinline__ float val(float x) { return x; }
__device__ __forceinline__ float val(const D3& x) { return x.v; }
__device__ __forceinline__ D3 operator+(const D3& a, const D3& b) {
    return {a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]}};
}
__device__ __forceinline__ D3 operator-(const D3& a, const D3& b) {
    return {a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]}};
}
__device__ __forceinline__ D3 operator*(const D3& a, const D3& b) {
    return {a.v * b.v, {a.d[0] * b.v + a.v * b.d[0], a.d[1] * b.v + a.v * b.d[1], a.d[2] * b.v + a.v * b.d[2]}};
}
__device__ __forceinline__ D3 operator/(const D3& a, const D3& b) {
    const float q = a.v / b.v, ib = 1.0f / b.v;
    return {q, {(a.d[0] - q * b.d[0]) * ib, (a.d[1] - q * b.d[1]) * ib, (a.d[2] - q * b.d[2]) * ib}};
}
__device__ __forceinline__ D3 operator+(const D3& a, float b) { return {a.v + b, {a.d[0], a.d[1], a.d[2]}}; }
__device__ __forceinline__ D3 operator+(float a, const D3& b) { return b + a; }
__device__ __forceinline__ D3 operator-(const D3& a, float b) { return {a.v - b, {a.d[0], a.d[1], a.d[2]}}; }
__device__ __forceinline__ D3 operator-(float a, const D3& b) { return {a - b.v, {-b.d[0], -b.d[1], -b.d[2]}}; }
__device__ __forceinline__ D3 operator-(const D3& a) { return {-a.v, {-a.d[0], -a.d[1], -a.d[2]}}; }
__device__ __forceinline__ D3 operator*(const D3& a, float s) { return {a.v * s, {a.d[0] * s, a.d[1] * s, a.d[2] * s}}; }
__device__ __forceinline__ D3 operator*(float s, const D3& a) { return a * s; }
__device__ __forceinline__ D3 operator/(const D3& a, float s) { return {a.v / s, {a.d[0] / s, a.d[1] / s, a.d[2] / s}}; }
__device__ __forceinline__ D3 operator/(float a, const D3& b) {
    const float q = a / b.v, ib = 1.0f / b.v;
    return {q, {-q * b.d[0] * ib, -q * b.d[1] * ib, -q * b.d[2] * ib}};
}
__device__ __forceinline__ float sqrtT(float x) { return sqrtf(x); }
__device__ __forceinline__ D3 sqrtT(const D3& x) {
    const float s = sqrtf(x.v), k = s > 0.0f ? 0.5f / s : 0.0f;
    return {s, {x.d[0] * k, x.d[1] * k, x.d[2] * k}};
}

// The exit: refraction out of the slab (index n1) into air (n2) through the surface of height h and
// slope (ax, ay) at (x, y), for a ray travelling +z, and the wall point. Written once for float and
// for dual numbers, without contraction, so the backward pass rounds its values as the forward did.
template <class T>
__device__ __forceinline__ bool exitRay(T h, T ax, T ay, float x, float y, float n1, float n2, float z0, float zWall,
                                        T& xw, T& yw, T& trans) {
#pragma clang fp contract(off)
    const T g2 = ax * ax + ay * ay + 1.0f;
    const T inv = 1.0f / sqrtT(g2);
    const T nx = -(ax * inv), ny = -(ay * inv), nz = inv;  // outward normal; cos θi = n·(+z) = nz
    const float eta = n1 / n2;
    const T k = 1.0f - eta * eta * (1.0f - nz * nz);
    if (!(val(k) >= 0.0f)) return false;  // total internal reflection
    const T m = eta * nz - sqrtT(k);
    const T dx = -(m * nx), dy = -(m * ny), dz = eta - m * nz;
    // Fresnel (optics.hpp:42-54) at cos θi = nz, eta = n2 / n1.
    const float ef = n2 / n1;
    const T s2t = (1.0f - nz * nz) / (ef * ef);
    const T ct = sqrtT(1.0f - s2t);
    const T ra = (ef * nz - ct) / (ef * nz + ct), rb = (nz - ef * ct) / (nz + ef * ct);
    trans = 1.0f - (ra * ra + rb * rb) * 0.5f;
    const T L = zWall - (z0 + h);
    xw = x + dx / dz * L;
    yw = y + dy / dz * L;
    return true;
}

// Hermite bases (surfaces.hpp:190-197) at u for the two ends: P, Q and their derivatives.
__device__ __forceinline__ void hermite(float u, float P[2], float Q[2], float dP[2], float dQ[2]) {
    const float u2 = u * u, u3 = u2 * u;
    P[0] = 2.0f * u3 - 3.0f * u2 + 1.0f;
    P[1] = -2.0f * u3 + 3.0f * u2;
    Q[0] = u3 - 2.0f * u2 + u;
    Q[1] = u3 - u2;
    dP[0] = 6.0f * u2 - 6.0f * u;
    dP[1] = -6.0f * u2 + 6.0f * u;
    dQ[0] = 3.0f * u2 - 4.0f * u + 1.0f;
    dQ[1] = 3.0f * u2 - 2.0f * u;
}

// The patch of the cell holding (x, y): the 4 × 4 node block around it (row = y) and the weights
// with which h, ∂h/∂u and ∂h/∂v (u, v in cells) depend on each node's height.
struct Patch {
    int i0, j0;                                  // node of the block's corner (i − 1, j − 1)
    float wf[4][4], wu[4][4], wv[4][4];
};
__device__ __forceinline__ void patchAt(const CArgs& a, float x, float y, Patch& p) {
    const float fx = (x + a.slabHalf) / a.cell, fy = (y + a.slabHalf) / a.cell;
    const int i = min(max(int(fx), 0), int(a.nodes) - 2), j = min(max(int(fy), 0), int(a.nodes) - 2);
    float P[2], Q[2], dP[2], dQ[2], Pv[2], Qv[2], dPv[2], dQv[2];
    hermite(fx - float(i), P, Q, dP, dQ);
    hermite(fy - float(j), Pv, Qv, dPv, dQv);
    p.i0 = i - 1;
    p.j0 = j - 1;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) p.wf[r][c] = p.wu[r][c] = p.wv[r][c] = 0.0f;
    // Corner (a, b) of the cell is block entry (b + 1, a + 1); its Hermite coefficients are
    // c0 = H, cu = (H[x+1] − H[x−1]) / 2, cv = (H[y+1] − H[y−1]) / 2, cuv = the cross difference / 4.
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int r = b + 1, k = c + 1;
            auto add = [&](int rr, int kk, float s, float f, float fu, float fv) {
                p.wf[rr][kk] += s * f;
                p.wu[rr][kk] += s * fu;
                p.wv[rr][kk] += s * fv;
            };
            add(r, k, 1.0f, P[c] * Pv[b], dP[c] * Pv[b], P[c] * dPv[b]);
            add(r, k + 1, 0.5f, Q[c] * Pv[b], dQ[c] * Pv[b], Q[c] * dPv[b]);
            add(r, k - 1, -0.5f, Q[c] * Pv[b], dQ[c] * Pv[b], Q[c] * dPv[b]);
            add(r + 1, k, 0.5f, P[c] * Qv[b], dP[c] * Qv[b], P[c] * dQv[b]);
            add(r - 1, k, -0.5f, P[c] * Qv[b], dP[c] * Qv[b], P[c] * dQv[b]);
            add(r + 1, k + 1, 0.25f, Q[c] * Qv[b], dQ[c] * Qv[b], Q[c] * dQv[b]);
            add(r + 1, k - 1, -0.25f, Q[c] * Qv[b], dQ[c] * Qv[b], Q[c] * dQv[b]);
            add(r - 1, k + 1, -0.25f, Q[c] * Qv[b], dQ[c] * Qv[b], Q[c] * dQv[b]);
            add(r - 1, k - 1, 0.25f, Q[c] * Qv[b], dQ[c] * Qv[b], Q[c] * dQv[b]);
        }
}
// The global node of block entry (r, c), clamped to the grid (the edge nodes' one-sided slopes).
__device__ __forceinline__ uint32_t nodeOf(const CArgs& a, const Patch& p, int r, int c) {
    const int j = min(max(p.j0 + r, 0), int(a.nodes) - 1), i = min(max(p.i0 + c, 0), int(a.nodes) - 1);
    return uint32_t(j) * a.nodes + uint32_t(i);
}

struct Particle {
    float x, y, weight, n1, n2;  // weight: luminous weight × entry transmittance × area factor
};
// The entry face is cut into tiles × tiles square tiles in Morton order; tile t holds the particles
// [⌊t·N/T⌋, ⌊(t+1)·N/T⌋) (T tiles, N particles), each uniform over its tile. Stratified sampling
// (unbiased: each tile's weight is its area over its own count) that also keeps a wave's
// consecutive particles in one small patch of the slab, so their splats and gradient adds meet the
// same pixels and nodes.
__device__ __forceinline__ uint64_t tileStart(uint64_t t, uint64_t n, uint64_t T) { return t * n / T; }
__device__ __forceinline__ uint32_t mortonHalf(uint32_t v) {  // the even bits of v, packed
    v &= 0x55555555u;
    v = (v | (v >> 1)) & 0x33333333u;
    v = (v | (v >> 2)) & 0x0F0F0F0Fu;
    v = (v | (v >> 4)) & 0x00FF00FFu;
    v = (v | (v >> 8)) & 0x0000FFFFu;
    return v;
}
__device__ __forceinline__ Particle drawParticle(const CArgs& a, uint64_t i) {
    Rng rng;
    rng.seed(hashPair(a.seed, i), a.pass);
    const float ul = rng.uniform();
    const float lambda = sampleVisibleF(ul);
    const float pdf = visiblePdfF(lambda);
    const float u2 = rng.uniform(), u1 = rng.uniform();
    const uint64_t T = uint64_t(a.tiles) * a.tiles;
    uint64_t t = i * T / a.particles;
    while (t + 1 < T && tileStart(t + 1, a.particles, T) <= i) ++t;
    while (t > 0 && tileStart(t, a.particles, T) > i) --t;
    const uint64_t count = tileStart(t + 1, a.particles, T) - tileStart(t, a.particles, T);
    const float side = 2.0f * a.slabHalf / float(a.tiles);
    Particle q;
    q.x = -a.slabHalf + (float(mortonHalf(uint32_t(t))) + u1) * side;
    q.y = -a.slabHalf + (float(mortonHalf(uint32_t(t >> 1))) + u2) * side;
    q.n1 = mediumIndex(a.slab, lambda);
    q.n2 = mediumIndex(a.air, lambda);
    const float lum = planckF(lambda, a.kelvin) * cieXYZ(lambda).y / pdf * a.invY;
    q.weight = lum * (1.0f - fresnelDielectric(1.0f, q.n1 / q.n2)) * (a.tileArea / float(count));
    return q;
}

// A wave's table in LDS of adds keyed by pixel or node: open addressing over kTableSlots (key,
// sum) pairs; an add that finds no slot within kTableProbes goes straight to HBM. flush adds every
// claimed slot to HBM and empties the table (wave-wide, after every lane's adds: the fence keeps
// the compiler from moving LDS accesses across it; one wave's LDS operations complete in order).
constexpr int kTableSlots = 512;
constexpr int kTableProbes = 8;
constexpr uint32_t kTableEmpty = 0xFFFFFFFFu;

template <class V>
struct WaveTable {
    uint32_t* keys;
    V* sums;
    __device__ __forceinline__ void clear(uint32_t lane) const {
        for (uint32_t q = lane; q < uint32_t(kTableSlots); q += 64) {
            keys[q] = kTableEmpty;
            sums[q] = V(0);
        }
    }
    __device__ __forceinline__ void add(V* global, uint32_t key, V v) const {
        uint32_t slot = (key * 2654435761u) >> (32 - 9);
        static_assert(kTableSlots == 512, "the slot hash takes 9 bits");
#pragma unroll 1
        for (int probe = 0; probe < kTableProbes; ++probe) {
            const uint32_t old = atomicCAS(&keys[slot], kTableEmpty, key);
            if (old == kTableEmpty || old == key) {
                atomicAdd(&sums[slot], v);
                return;
            }
            slot = (slot + 1) & uint32_t(kTableSlots - 1);
        }
        atomicAdd(global + key, v);
    }
    __device__ __forceinline__ void flush(V* global, uint32_t lane) const {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        for (uint32_t q = lane; q < uint32_t(kTableSlots); q += 64) {
            const uint32_t key = keys[q];
            if (key != kTableEmpty) {
                atomicAdd(global + key, sums[q]);
                keys[q] = kTableEmpty;
                sums[q] = V(0);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
};

// The contiguous range of particles wave w of the grid takes: ⌈N / waves⌉ rounded up to whole waves.
__device__ __forceinline__ void waveRange(uint64_t n, uint64_t& begin, uint64_t& end) {
    const uint64_t waves = uint64_t(gridDim.x) * (kCBlock / 64);
    const uint64_t w = (uint64_t(blockIdx.x) * kCBlock + threadIdx.x) >> 6;
    const uint64_t per = ((n + waves - 1) / waves + 63) / 64 * 64;
    begin = w * per;
    end = begin + per < n ? begin + per : n;
}

__device__ __forceinline__ void heightsAt(const CArgs& a, const Patch& p, float& h, float& fu, float& fv) {
    h = fu = fv = 0.0f;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float H = a.heights[nodeOf(a, p, r, c)];
            h += p.wf[r][c] * H;
            fu += p.wu[r][c] * H;
            fv += p.wv[r][c] * H;
        }
}

// The splat: a quadratic B-spline over the 3 × 3 pixels about the nearest pixel centre (C¹ in the
// wall point, so the irradiance — and any smooth loss of it — has continuous derivatives): the
// centre pixel (i, j), the offsets t ∈ [−½, ½) from it, and the weights and their derivatives.
__device__ __forceinline__ void splatOf(const CArgs& a, float xw, float yw, int& i, int& j, float wu[3], float wv[3],
                                        float du[3], float dv[3]) {
    const float gu = (xw + a.wallHalf) / a.pix - 0.5f, gv = (yw + a.wallHalf) / a.pix - 0.5f;
    const float fi = floorf(gu + 0.5f), fj = floorf(gv + 0.5f);
    i = int(fmaxf(fminf(fi, 1e8f), -1e8f));
    j = int(fmaxf(fminf(fj, 1e8f), -1e8f));
    const float tu = gu - fi, tv = gv - fj;
    wu[0] = 0.5f * (0.5f - tu) * (0.5f - tu);
    wu[1] = 0.75f - tu * tu;
    wu[2] = 0.5f * (0.5f + tu) * (0.5f + tu);
    wv[0] = 0.5f * (0.5f - tv) * (0.5f - tv);
    wv[1] = 0.75f - tv * tv;
    wv[2] = 0.5f * (0.5f + tv) * (0.5f + tv);
    du[0] = -(0.5f - tu);
    du[1] = -2.0f * tu;
    du[2] = 0.5f + tu;
    dv[0] = -(0.5f - tv);
    dv[1] = -2.0f * tv;
    dv[2] = 0.5f + tv;
}

__global__ __launch_bounds__(kCBlock) void causticForward(CArgs a) {
    __shared__ uint32_t keys[kCBlock / 64][kTableSlots];
    __shared__ unsigned long long sums[kCBlock / 64][kTableSlots];
    const uint32_t lane = __lane_id();
    const WaveTable<unsigned long long> table{keys[threadIdx.x / 64], sums[threadIdx.x / 64]};
    table.clear(lane);
    uint64_t begin, end;
    waveRange(a.particles, begin, end);
    for (uint64_t i = begin + lane; i < end; i += 64) {
        const Particle q = drawParticle(a, i);
        Patch p;
        patchAt(a, q.x, q.y, p);
        float h, fu, fv;
        heightsAt(a, p, h, fu, fv);
        float xw, yw, T2;
        if (!exitRay(h, fu / a.cell, fv / a.cell, q.x, q.y, q.n1, q.n2, a.z0, a.zWall, xw, yw, T2)) continue;
        const float w = q.weight * T2;
        int ci, cj;
        float wu[3], wv[3], du[3], dv[3];
        splatOf(a, xw, yw, ci, cj, wu, wv, du, dv);
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int pi = ci + c - 1, pj = cj + b - 1;
                if (pi < 0 || pj < 0 || pi >= int(a.pixels) || pj >= int(a.pixels)) continue;
                const long long units = llrint(double(w * wu[c] * wv[b]) * kSplatUnits);
                if (units > 0) table.add(a.acc, uint32_t(pj) * a.pixels + uint32_t(pi), (unsigned long long)units);
            }
    }
    table.flush(a.acc, lane);
}

__global__ void causticClear(unsigned long long* acc, size_t n) {
    for (size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += size_t(gridDim.x) * blockDim.x) acc[k] = 0;
}
__global__ void causticResolve(const unsigned long long* acc, float* out, size_t n) {
    for (size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += size_t(gridDim.x) * blockDim.x)
        out[k] = float(double(acc[k]) / kSplatUnits);
}
__global__ void causticZero(float* g, size_t n) {
    for (size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += size_t(gridDim.x) * blockDim.x) g[k] = 0.0f;
}

__global__ __launch_bounds__(kCBlock) void causticBackward(CArgs a) {
    __shared__ uint32_t keys[kCBlock / 64][kTableSlots];
    __shared__ float sums[kCBlock / 64][kTableSlots];
    const uint32_t lane = __lane_id();
    const WaveTable<float> table{keys[threadIdx.x / 64], sums[threadIdx.x / 64]};
    table.clear(lane);
    uint64_t begin, end;
    waveRange(a.particles, begin, end);
    for (uint64_t i = begin + lane; i < end; i += 64) {
        const Particle q = drawParticle(a, i);
        Patch p;
        patchAt(a, q.x, q.y, p);
        float h, fu, fv;
        heightsAt(a, p, h, fu, fv);
        D3 xw, yw, T2;
        if (!exitRay(mkD(h, 0), mkD(fu / a.cell, 1), mkD(fv / a.cell, 2), q.x, q.y, q.n1, q.n2, a.z0, a.zWall, xw, yw, T2))
            continue;
        const float w = q.weight * T2.v;
        int ci, cj;
        float wu[3], wv[3], du[3], dv[3];
        splatOf(a, xw.v, yw.v, ci, cj, wu, wv, du, dv);
        // ∂L/∂(xw, yw, w) from the nine pixels' ∂L/∂E and the spline weights.
        float gx = 0.0f, gy = 0.0f, gw = 0.0f;
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int pi = ci + c - 1, pj = cj + b - 1;
                if (pi < 0 || pj < 0 || pi >= int(a.pixels) || pj >= int(a.pixels)) continue;
                const float g = a.gradE[size_t(pj) * a.pixels + pi];
                gw += g * wu[c] * wv[b];
                gx += g * du[c] * wv[b];
                gy += g * wu[c] * dv[b];
            }
        gx *= w / a.pix;
        gy *= w / a.pix;
        // ∂L/∂(h, h_x, h_y); the weight depends on them through the exit transmittance.
        float gq[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) gq[k] = gx * xw.d[k] + gy * yw.d[k] + gw * q.weight * T2.d[k];
        const float su = gq[1] / a.cell, sv = gq[2] / a.cell;
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const float g = gq[0] * p.wf[r][c] + su * p.wu[r][c] + sv * p.wv[r][c];
                if (g != 0.0f) table.add(a.gradH, nodeOf(a, p, r, c), g);
            }
    }
    table.flush(a.gradH, lane);
}

double planck(double lnm, double kelvin) {
    const double c1 = 1.191042972e8, c2 = 1.438776877e4, l = lnm * 1e-3, l2 = l * l;
    return 1e6 * c1 / (l2 * l2 * l * (std::exp(c2 / (l * kelvin)) - 1));
}
double cieYHost(double l) {
    auto lobe = [](double x, double mu, double s1, double s2) {
        const double t = (x - mu) / (x < mu ? s1 : s2);
        return std::exp(-0.5 * t * t);
    };
    return 0.821 * lobe(l, 568.8, 46.9, 40.5) + 0.286 * lobe(l, 530.9, 16.3, 31.1);
}

DMedium toDevMedium(const owe_medium& m) {
    DMedium d{};
    d.indexKind = m.index_kind;
    d.terms = m.terms;
    d.relAir = m.rel_air;
    for (int k = 0; k < 4; ++k) {
        d.a[k] = float(m.coeff_a[k]);
        d.b[k] = float(m.coeff_b[k]);
    }
    return d;
}

// Tiles per side of the entry face: the largest power of two up to 64 that leaves at least 64
// particles per tile.
uint32_t causticTiles(uint64_t particles) {
    uint32_t g = 1;
    while (g < 64 && uint64_t(4) * g * g * 64 <= particles) g *= 2;
    return g;
}

void checkSettings(const owe_caustic_settings& s) {
    require(s.nodes >= 4 && s.nodes <= 8192, "nodes per side must lie in [4, 8192]");
    require(s.pixels >= 2 && s.pixels <= 8192, "pixels per side must lie in [2, 8192]");
    require(s.particles >= 1 && s.particles <= (uint64_t(1) << 34), "particles must lie in [1, 2^34]");
    require(s.slab_half > 0 && s.thickness > 0 && s.wall_half > 0, "sizes must be positive");
    require(s.wall_z > s.thickness, "the wall must lie beyond the slab");
    require(s.sun_kelvin > 100, "sun temperature must exceed 100 K");
    require(s.medium.index_kind >= OWE_INDEX_CONSTANT && s.medium.index_kind <= OWE_INDEX_CIDDOR, "unknown index model");
}

// The kernels' arguments for settings s (everything but the per-call buffers and pass).
CArgs makeArgs(const owe_caustic_settings& s) {
    CArgs a{};
    a.nodes = s.nodes;
    a.pixels = s.pixels;
    a.slabHalf = float(s.slab_half);
    a.cell = float(2 * s.slab_half / (s.nodes - 1));
    a.z0 = float(s.thickness);
    a.zWall = float(s.wall_z);
    a.wallHalf = float(s.wall_half);
    a.pix = float(2 * s.wall_half / s.pixels);
    a.particles = s.particles;
    a.seed = s.seed;
    a.slab = toDevMedium(s.medium);
    owe_medium air{};
    air.index_kind = OWE_INDEX_CIDDOR;
    a.air = toDevMedium(air);
    a.kelvin = float(s.sun_kelvin);
    // ∫ B ȳ dλ over 360–830 nm (composite Simpson, 0.5 nm).
    const int n = 940;
    const double hh = 470.0 / n;
    double sum = planck(360.0, s.sun_kelvin) * cieYHost(360.0) + planck(830.0, s.sun_kelvin) * cieYHost(830.0);
    for (int k = 1; k < n; ++k) {
        const double l = 360.0 + k * hh;
        sum += planck(l, s.sun_kelvin) * cieYHost(l) * ((k & 1) ? 4 : 2);
    }
    a.invY = float(1.0 / (sum * hh / 3));
    const double pix = 2 * s.wall_half / s.pixels;
    a.tiles = causticTiles(s.particles);
    const double side = 2 * s.slab_half / a.tiles;
    a.tileArea = float(side * side / (pix * pix));
    return a;
}

// The forward kernel's grid: 8 blocks of 256 per compute unit.
int forwardGrid(int device) {
    int cus = 0;
    OWE_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    require(cus > 0, "cannot read the device's compute-unit count");
    return cus * 8;
}

}  // namespace
}  // namespace owehip

struct owe_hip_caustic {
    int device = 0;
    owe_caustic_settings settings{};
    CArgs args{};
    unsigned long long* acc = nullptr;
    int grid = 1024;
};

extern "C" {

int owe_hip_caustic_create(int device, const owe_caustic_settings* settings, owe_hip_caustic** out) {
    owe_hip_caustic* c = nullptr;
    const int st = guard([&] {
        require(settings != nullptr && out != nullptr, "null argument");
        *out = nullptr;
        const owe_caustic_settings& s = *settings;
        checkSettings(s);
        selectDevice(device);
        c = new owe_hip_caustic;
        c->device = device;
        c->settings = s;
        c->grid = forwardGrid(device);
        c->args = makeArgs(s);
        void* p = nullptr;
        OWE_HIP_CHECK(hipMalloc(&p, size_t(s.pixels) * s.pixels * sizeof(unsigned long long)));
        c->acc = static_cast<unsigned long long*>(p);
        c->args.acc = c->acc;
        *out = c;
    });
    if (st != OWE_HIP_OK && c) {
        if (c->acc) (void)hipFree(c->acc);
        delete c;
    }
    return st;
}

int owe_hip_caustic_destroy(owe_hip_caustic* c) {
    return guard([&] {
        if (!c) return;
        (void)hipSetDevice(c->device);
        OWE_HIP_CHECK(hipDeviceSynchronize());
        if (c->acc) OWE_HIP_CHECK(hipFree(c->acc));
        delete c;
    });
}

int owe_hip_caustic_forward(owe_hip_caustic* c, const float* heights, float* irradiance, uint64_t pass, void* stream) {
    return guard([&] {
        require(c != nullptr && heights != nullptr && irradiance != nullptr, "null argument");
        requireDevicePointer(heights, "heights");
        requireDevicePointer(irradiance, "irradiance");
        OWE_HIP_CHECK(hipSetDevice(c->device));
        const hipStream_t s = static_cast<hipStream_t>(stream);
        CArgs a = c->args;
        a.heights = heights;
        a.pass = pass;
        a.irradiance = irradiance;
        const size_t n = size_t(a.pixels) * a.pixels;
        hipLaunchKernelGGL(causticClear, dim3{1024}, dim3{256}, 0, s, c->acc, n);
        hipLaunchKernelGGL(causticForward, dim3{unsigned(c->grid)}, dim3{kCBlock}, 0, s, a);
        hipLaunchKernelGGL(causticResolve, dim3{1024}, dim3{256}, 0, s, c->acc, irradiance, n);
        OWE_HIP_CHECK(hipGetLastError());
    });
}

int owe_hip_caustic_backward(owe_hip_caustic* c, const float* heights, const float* grad_irradiance, float* grad_heights,
                             uint64_t pass, void* stream) {
    return guard([&] {
        require(c != nullptr && heights != nullptr && grad_irradiance != nullptr && grad_heights != nullptr,
                "null argument");
        requireDevicePointer(heights, "heights");
        requireDevicePointer(grad_irradiance, "grad_irradiance");
        requireDevicePointer(grad_heights, "grad_heights");
        OWE_HIP_CHECK(hipSetDevice(c->device));
        const hipStream_t s = static_cast<hipStream_t>(stream);
        CArgs a = c->args;
        a.heights = heights;
        a.pass = pass;
        a.gradE = grad_irradiance;
        a.gradH = grad_heights;
        hipLaunchKernelGGL(causticZero, dim3{1024}, dim3{256}, 0, s, grad_heights, size_t(a.nodes) * a.nodes);
        hipLaunchKernelGGL(causticBackward, dim3{unsigned(c->grid)}, dim3{kCBlock}, 0, s, a);
        OWE_HIP_CHECK(hipGetLastError());
    });
}

// The allocation-free forward. Its fixed-point bins live in `scratch`, one array per device for the
// life of the process (2 MiB: 512 × 512 pixels at most). Host code never passes the array's name to
// a kernel (in host code it names a host-side stand-in, not device memory): the first call on a
// device, with that device current, asks the runtime for the array's device address and keeps it,
// and every later call on that device uses the kept address. Every call clears the bins it uses
// before splatting; calls that overlap on one device share the array and race.
static __device__ __align__(256) unsigned long long scratch[size_t(OWE_CAUSTIC_NOALLOC_MAX_PIXELS)];

namespace {

constexpr int kMaxDevices = 64;

struct ScratchSlot {
    unsigned long long* bins = nullptr;
    int grid = 0;
};

ScratchSlot scratchFor(int device) {
    static std::mutex mu;
    static ScratchSlot slots[kMaxDevices];
    require(device < kMaxDevices, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(mu);
    ScratchSlot& s = slots[device];
    if (!s.bins) {
        void* p = nullptr;
        OWE_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        require(p != nullptr, "the scratch array has no device address");
        s.grid = forwardGrid(device);
        s.bins = static_cast<unsigned long long*>(p);
    }
    return s;
}

}  // namespace

int owe_hip_caustic_forward_noalloc(int device, const owe_caustic_settings* settings, const float* heights,
                                    float* irradiance, uint64_t pass, void* stream) {
    return guard([&] {
        require(settings != nullptr && heights != nullptr && irradiance != nullptr, "null argument");
        const owe_caustic_settings& s = *settings;
        checkSettings(s);
        const size_t n = size_t(s.pixels) * s.pixels;
        if (n > size_t(OWE_CAUSTIC_NOALLOC_MAX_PIXELS))
            fail(OWE_HIP_ERR_INVALID, "the allocation-free forward holds at most " +
                                          std::to_string(OWE_CAUSTIC_NOALLOC_MAX_PIXELS) + " pixels (512 x 512); " +
                                          std::to_string(s.pixels) + " x " + std::to_string(s.pixels) + " asked");
        selectDevice(device);
        requireDevicePointer(heights, "heights", device);
        requireDevicePointer(irradiance, "irradiance", device);
        const ScratchSlot slot = scratchFor(device);
        CArgs a = makeArgs(s);
        a.acc = slot.bins;
        a.heights = heights;
        a.pass = pass;
        a.irradiance = irradiance;
        const hipStream_t st = static_cast<hipStream_t>(stream);
        hipLaunchKernelGGL(causticClear, dim3{1024}, dim3{256}, 0, st, slot.bins, n);
        hipLaunchKernelGGL(causticForward, dim3{unsigned(slot.grid)}, dim3{kCBlock}, 0, st, a);
        hipLaunchKernelGGL(causticResolve, dim3{1024}, dim3{256}, 0, st, slot.bins, irradiance, n);
        OWE_HIP_CHECK(hipGetLastError());
    });
}

}  // extern "C"
