// Batched lens analysis: owe_hip_lens_batch_* of include/owe_hip_lens.h.
//
// Work is a flat list of items — (lens, wavelength, block of `raysPerItem` rays) — drained by a
// persistent grid: each wave takes the next item with one atomic, its lanes trace the block's rays
// through the lens surface after surface, reduce the spot's sums across the wave in double and add
// them to the (lens, wavelength) totals with one atomic per sum; arrived rays also land on the lens's
// through-focus stack of detector grids as 64-bit fixed-point weights, so the irradiance is
// independent of order. The grid adds gather first in a per-wave hash table in LDS (bin → sum),
// which a wave flushes to HBM when its items move to another lens: a spot covers few bins, so most
// of a lens's rays meet a bin already in the table and HBM sees about one atomic per bin touched
// instead of one per ray. Lenses of very different ray counts (4k to 1M) share the launch without
// tails of idle blocks.
//
// Ray–surface roots follow SagSurface (src/owe/scene/geometry.cpp:115-166,
// src/owe/core/shared/surfaces.hpp:13-94) in float32 with the reference's float tolerances; each
// quadric root is then re-solved from the approximate hit point (refineRoot) — in float64 on a
// steep surface (|c|·r_max > 0.5), where the quadric's terms cancel, and throughout instruments with
// mirrors or without focal power (metre paths, high angular magnification), in float elsewhere — and a
// numeric asphere root is polished by Newton steps in float64: the only fp64 of the per-ray path
// (the spot sums are accumulated in double from float per-ray terms). Transmission is the walk's
// deterministic refraction with weight T = 1 − R (transport.cpp:531-595); TIR or a missed aperture
// ends a ray.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "owe_hip_lens.h"
#include "shapes_dev.hpp"
#include "transport_dev.hpp"

using namespace owehip;

namespace owehip {
namespace {

constexpr int kLensBlock = 256;
constexpr double kHistScale = 4294967296.0;  // 2^32 units of weight per unit
// The per-wave bin table: open addressing over kHashSlots (bin, sum) pairs, at most kHashProbes
// probes before an add goes straight to HBM.
constexpr int kHashSlots = 512;
constexpr int kHashProbes = 8;
constexpr uint32_t kHashEmpty = 0xFFFFFFFFu;
// The steepness |c|·r_max above which a quadric root is polished in float64.
constexpr double kSteepPolish = 0.5;

// A lens row: its sag surface (shapes_dev.hpp; a mirror's annulus outside its hole), the vertex
// position along the fixed chart, and whether it reflects.
struct DSurf : DSag {
    float z;
    int32_t reflect;
};

// Parameters differentiated per gradient pass (of up to OWE_LENS_MAX_PARAMS), and the sums a pass
// adds per (lens, wavelength, parameter): ΣA', ΣB' (x, y), ΣA·A', ΣB·B', Σ(A'·B + A·B').
constexpr int kGroup = 4;
constexpr int kGradSums = 7;

struct DLensInfo {
    uint32_t first, count;    // surfaces
    uint64_t rays;            // per wavelength
    uint64_t itemBase;        // first work item
    uint32_t itemsPerLambda;
    uint32_t rings;           // 0: Fibonacci lattice; r: hexapolar rings
    double rEP, zEP, zStart;  // entrance pupil (sampling radius 0.999 rEP) and launch plane
    float detX, detY, detZ, detHalf;  // detector grid centre, plane, half size (0: no grid)
    float detStep;            // through-focus plane spacing along z
    int32_t afocal;           // spot sums of angles (afocalSpot) instead of (A, B) at an image
    int32_t mirrors;          // reflecting rows: their vertex planes obstruct
};

struct LensArgs {
    const DSurf* surf;
    const DLensInfo* lens;
    const float* nAfter;      // [surface × nλ]: index after each surface
    const float* nAmbient;    // [lens × nλ]
    uint32_t nLenses, nLambda, raysPerItem, pixels, planes;
    uint64_t totalItems;
    double dirY, dirZ;        // collimated field direction (0, sin θ, cos θ)
    unsigned long long* next;
    double* moments;          // [lens × nλ × OWE_LENS_MOMENTS]
    unsigned long long* hist; // [lens × planes × pixels²]
    // Gradient passes: parameters [paramBase, paramBase + kGroup) of nParams, each (surface within
    // the lens, field); their sums [lens × nλ × OWE_LENS_MAX_PARAMS × kGradSums].
    uint32_t nParams, paramBase;
    int32_t paramSurface[OWE_LENS_MAX_PARAMS], paramField[OWE_LENS_MAX_PARAMS];
    double* grads;
    int32_t sums;             // a gradient pass also adds the moments and histogram
};

// Pupil point i of n on the entrance pupil of radius rEP, in double: a Fibonacci lattice over the
// disc of radius 0.999 rEP, or with rings > 0 spotAt's hexapolar pattern (analysis.cpp:124-131) —
// point 0 the centre, ring k holding the 6k points after the 1 + 3k(k − 1) before it.
__device__ __forceinline__ void pupilPoint(uint64_t i, uint64_t n, uint32_t rings, double rEP, double& x, double& y) {
    if (rings > 0) {
        if (i == 0) {
            x = y = 0.0;
            return;
        }
        uint64_t k = uint64_t((3.0 + sqrt(12.0 * double(i) - 3.0)) / 6.0);
        while (k > 1 && 1 + 3 * k * (k - 1) > i) --k;
        while (1 + 3 * k * (k + 1) <= i) ++k;
        const uint64_t j = i - (1 + 3 * k * (k - 1)), m = 6 * k;
        const double rr = rEP * double(k) / double(rings) * 0.999;
        const double a = 2.0 * 3.14159265358979323846 * double(j) / double(m);
        x = rr * cos(a);
        y = rr * sin(a);
        return;
    }
    // The lattice in float: the golden-ratio fraction of i from 64-bit fixed point (exact for any
    // i), the radius and angle to float rounding.
    const float rr = 0.999f * float(rEP) * sqrtf((float(i) + 0.5f) / float(n));
    const float f = float(uint32_t((i * 0x9E3779B97F4A7C15ull) >> 40)) * 0x1p-24f;
    float sa, ca;
    sincospif(2.0f * f, &sa, &ca);
    x = double(rr * ca);
    y = double(rr * sa);
}

// Whether the ray o + s d, s in (0, sEnd) (sEnd < 0: unbounded), crosses the vertex-plane annulus
// hole ≤ r ≤ semi-diameter of a mirror of the lens other than surfaces `from` and `to`.
__device__ __forceinline__ bool obstructed(const LensArgs& a, const DLensInfo& L, f3 o, f3 d, float zEnd, uint32_t from,
                                           uint32_t to) {
    for (uint32_t j = 0; j < L.count; ++j) {
        const DSurf& m = a.surf[L.first + j];
        if (!m.reflect || j == from || j == to || d.z == 0.0f) continue;
        if (!(m.z > fminf(o.z, zEnd) && m.z < fmaxf(o.z, zEnd))) continue;
        const float s = (m.z - o.z) / d.z;
        const float x = o.x + d.x * s, y = o.y + d.y * s, r2 = x * x + y * y;
        if (r2 >= m.rMin2 && r2 <= m.rMax2) return true;
    }
    return false;
}

// Derivatives of the ray with respect to the kGroup parameters of a gradient pass.
struct RayGrad {
    float o[3][kGroup], d[3][kGroup];
};

// One ray i of lens L at wavelength w: true when every surface takes it — a refracting row refracts
// it, a mirror reflects it — no mirror body stops it and it leaves along +z; p, d after the last
// surface. With kGrad the derivatives of p and d ride along (launch rays held fixed): at each
// surface the implicit function theorem on F = sag(r²) − z gives the root's derivative
// dt = −(∇F·(do + t dd) + dz_v + F_θ)/F_t, the hit point dp = do + t dd + dt d, the normal's through
// dN of N = (−2g x, −2g y, 1) with g = d sag / d r², and the refracted or reflected direction's by
// the chain rule of its formula.
template <bool kGrad>
__device__ __forceinline__ bool traceLensRay(const LensArgs& a, const DLensInfo& L, uint32_t li, uint32_t w, uint64_t i,
                                             f3& p, f3& d, float& T, RayGrad& G) {
    double x, y;
    pupilPoint(i, L.rays, L.rings, L.rEP, x, y);
    const double s = (L.zStart - L.zEP) / a.dirZ;
    f3 o = mk3(float(x), float(y + a.dirY * s), float(L.zEP + a.dirZ * s));
    d = mk3(0.0f, float(a.dirY), float(a.dirZ));
    T = 1.0f;
    if (kGrad) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int k = 0; k < kGroup; ++k) G.o[c][k] = G.d[c][k] = 0.0f;
    }
    float nPrev = a.nAmbient[li * a.nLambda + w];
    for (uint32_t k = 0; k < L.count; ++k) {
        const DSurf& sf = a.surf[L.first + k];
        f3 q, n;
        float th;
        if (!sagHit(sf, mk3(o.x, o.y, o.z - sf.z), d, 0.0f, __builtin_inff(), false, th, q)) return false;
        if (L.mirrors && obstructed(a, L, o, d, q.z + sf.z, k == 0 ? L.count : k - 1, k)) return false;
        n = sagNormalF(sf, q);
        f3 wv;
        float eta = 1.0f;
        if (sf.reflect) {
            wv = d - n * (2.0f * dot3(d, n));
        } else {
            const float nNext = a.nAfter[(L.first + k) * a.nLambda + w];
            const float R = fresnelDielectric(fabsf(dot3(d, n)), nNext / nPrev);
            if (!refractDir(d, n, nPrev, nNext, wv)) return false;
            T *= 1.0f - R;
            eta = nPrev / nNext;
            nPrev = nNext;
        }
        if (kGrad) {
            // The surface's partials at the local hit point q.
            const float r2 = q.x * q.x + q.y * q.y, c = sf.cf, kk = sf.kf;
            const float S = sqrtf(fmaxf(1e-30f, 1.0f - (1.0f + kk) * c * c * r2));
            float g = c / (2.0f * S), gr = c * (1.0f + kk) * c * c / (4.0f * S * S * S);
            const float Fc = r2 * ((1.0f + S) + (1.0f + kk) * r2 * c * c / S) / ((1.0f + S) * (1.0f + S));
            const float Fk = c * c * c * r2 * r2 / (2.0f * S * (1.0f + S) * (1.0f + S));
            const float gc = 1.0f / (2.0f * S) + (1.0f + kk) * c * c * r2 / (2.0f * S * S * S);
            const float gk = c * c * c * r2 / (4.0f * S * S * S);
            float FA[4], gA[4], rj = 1.0f;  // rj = r2^j
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                FA[j] = rj * r2 * r2;                   // ∂sag/∂A_j = r2^(2+j)
                gA[j] = float(2 + j) * rj * r2;         // ∂g/∂A_j = (2+j) r2^(1+j)
                if (j < sf.nA) {
                    g += sf.A[j] * gA[j];
                    gr += sf.A[j] * float(2 + j) * float(1 + j) * rj;
                }
                rj *= r2;
            }
            const f3 gradF = mk3(2.0f * g * q.x, 2.0f * g * q.y, -1.0f);
            const float Ft = dot3(gradF, d);
            const f3 N = mk3(-2.0f * g * q.x, -2.0f * g * q.y, 1.0f);
            const float invN = rsqrtf(dot3(N, N));
            const float sig = dot3(d, n) < 0.0f ? 1.0f : -1.0f;  // the refraction's facing normal
            f3 W = wv;
            float invW = 1.0f, ci = 0.0f, sq = 0.0f;
            if (!sf.reflect) {
                const f3 nn = n * sig;
                ci = -dot3(d, nn);
                sq = sqrtf(fmaxf(0.0f, 1.0f - eta * eta * (1.0f - ci * ci)));
                W = d * eta + nn * (eta * ci - sq);
                invW = rsqrtf(dot3(W, W));
            }
#pragma unroll
            for (int kp = 0; kp < kGroup; ++kp) {
                const uint32_t pi = a.paramBase + uint32_t(kp);
                float dc = 0.0f, dk = 0.0f, dA[4] = {0.0f, 0.0f, 0.0f, 0.0f}, dzv = 0.0f;
                if (pi < a.nParams) {
                    const uint32_t ps = uint32_t(a.paramSurface[pi]);
                    const int pf = a.paramField[pi];
                    if (ps == k) {
                        if (pf == OWE_LENS_PARAM_CURVATURE) dc = 1.0f;
                        else if (pf == OWE_LENS_PARAM_CONIC) dk = 1.0f;
                        else if (pf >= OWE_LENS_PARAM_A4) dA[pf - OWE_LENS_PARAM_A4] = 1.0f;
                    }
                    if (pf == OWE_LENS_PARAM_THICKNESS && ps < k) dzv = 1.0f;
                }
                const f3 dO = mk3(G.o[0][kp], G.o[1][kp], G.o[2][kp]), dD = mk3(G.d[0][kp], G.d[1][kp], G.d[2][kp]);
                float Fth = Fc * dc + Fk * dk, gth = gc * dc + gk * dk;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    Fth += FA[j] * dA[j];
                    gth += gA[j] * dA[j];
                }
                const f3 dx = dO + dD * th;
                const float dt = -(dot3(gradF, dx) + dzv + Fth) / Ft;
                const f3 dP = dx + d * dt;
                // dN from dg = gr dr² + g_θ and the local point's move.
                const float dg = gr * 2.0f * (q.x * dP.x + q.y * dP.y) + gth;
                const f3 dN = mk3(-2.0f * (dg * q.x + g * dP.x), -2.0f * (dg * q.y + g * dP.y), 0.0f);
                const f3 dn = (dN - n * dot3(n, dN)) * invN;
                f3 dW;
                if (sf.reflect) {
                    dW = dD - (n * (dot3(dD, n) + dot3(d, dn)) + dn * dot3(d, n)) * 2.0f;
                } else {
                    const f3 nn = n * sig, dnn = dn * sig;
                    const float dci = -(dot3(dD, nn) + dot3(d, dnn));
                    const float dsq = sq > 0.0f ? eta * eta * ci * dci / sq : 0.0f;
                    const f3 dWu = dD * eta + nn * (eta * dci - dsq) + dnn * (eta * ci - sq);
                    dW = (dWu - wv * dot3(wv, dWu)) * invW;
                }
                G.o[0][kp] = dP.x;
                G.o[1][kp] = dP.y;
                G.o[2][kp] = dP.z;
                G.d[0][kp] = dW.x;
                G.d[1][kp] = dW.y;
                G.d[2][kp] = dW.z;
            }
        }
        o = mk3(q.x, q.y, q.z + sf.z);
        d = wv;
    }
    if (L.mirrors && obstructed(a, L, o, d, d.z > 0.0f ? __builtin_inff() : -__builtin_inff(), L.count - 1, L.count))
        return false;
    p = o;
    return d.z > 0.0f;
}

__device__ __forceinline__ double waveSumD(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// A wave's bin table in LDS. Keys are bins of the lens the wave is on (plane · pixels² + row ·
// pixels + column); a key is claimed by compare-and-swap, its sum grown by a 64-bit LDS add.
struct BinTable {
    uint32_t* keys;
    unsigned long long* sums;
};

__device__ __forceinline__ void binAdd(BinTable t, unsigned long long* lensHist, uint32_t key, unsigned long long v) {
    uint32_t slot = (key * 2654435761u) >> (32 - 9);
    static_assert(kHashSlots == 512, "the slot hash takes 9 bits");
#pragma unroll 1
    for (int probe = 0; probe < kHashProbes; ++probe) {
        const uint32_t old = atomicCAS(&t.keys[slot], kHashEmpty, key);
        if (old == kHashEmpty || old == key) {
            atomicAdd(&t.sums[slot], v);
            return;
        }
        slot = (slot + 1) & uint32_t(kHashSlots - 1);
    }
    atomicAdd(lensHist + key, v);  // the table is crowded here: straight to HBM
}

// Adds every claimed bin of the table to the lens's grids and empties the table. Wave-wide: every
// lane takes part, after every lane's adds (the fence keeps the compiler from moving LDS accesses
// across it; one wave's LDS operations complete in order).
__device__ __forceinline__ void binFlush(BinTable t, unsigned long long* lensHist, uint32_t lane) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (uint32_t q = lane; q < uint32_t(kHashSlots); q += 64) {
        const uint32_t key = t.keys[q];
        if (key != kHashEmpty) {
            atomicAdd(lensHist + key, t.sums[q]);
            t.keys[q] = kHashEmpty;
            t.sums[q] = 0;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// The persistent batch: a wave takes the next (lens, wavelength, block of rays) item, its lanes
// trace the block, the spot sums (and with kGrad the derivative sums of the pass's parameters) are
// reduced across the wave in double and added with one atomic each; grid adds go through the
// wave's bin table, flushed when the wave's next item is another lens's and before it exits.
template <bool kGrad>
__global__ __launch_bounds__(kLensBlock) void lensKernel(LensArgs a) {
    __shared__ uint32_t tableKeys[kLensBlock / 64][kHashSlots];
    __shared__ unsigned long long tableSums[kLensBlock / 64][kHashSlots];
    const uint32_t lane = __lane_id();
    const bool sums = !kGrad || a.sums;
    const bool grids = sums && a.pixels > 0;
    const BinTable table{tableKeys[threadIdx.x / 64], tableSums[threadIdx.x / 64]};
    for (uint32_t q = lane; q < uint32_t(kHashSlots); q += 64) {
        table.keys[q] = kHashEmpty;
        table.sums[q] = 0;
    }
    const size_t lensBins = size_t(a.planes) * a.pixels * a.pixels;
    uint32_t tableLens = ~0u;  // the lens whose bins the table holds
    for (;;) {
        unsigned long long item = 0;
        if (lane == 0) item = atomicAdd(a.next, 1ull);
        item = __shfl(item, 0);
        if (item >= a.totalItems) break;
        // The lens of this item: the last whose first item is ≤ item.
        uint32_t lo = 0, len = a.nLenses;
        while (len > 0) {
            const uint32_t half = len / 2;
            if (a.lens[lo + half].itemBase <= item) {
                lo += half + 1;
                len -= half + 1;
            } else {
                len = half;
            }
        }
        const uint32_t li = lo - 1;
        const DLensInfo& L = a.lens[li];
        if (grids && li != tableLens) {
            if (tableLens != ~0u) binFlush(table, a.hist + tableLens * lensBins, lane);
            tableLens = li;
        }
        unsigned long long* lensHist = a.hist + size_t(li) * lensBins;
        const uint64_t local = item - L.itemBase;
        const uint32_t w = uint32_t(local / L.itemsPerLambda);
        const uint64_t first = (local % L.itemsPerLambda) * a.raysPerItem;
        double m[OWE_LENS_MOMENTS] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        float gs[kGrad ? kGroup : 1][kGradSums] = {};
        for (uint32_t k = lane; k < a.raysPerItem; k += 64) {
            const uint64_t i = first + k;
            if (i >= L.rays) break;
            f3 p, d;
            float T;
            RayGrad G;
            if (!traceLensRay<kGrad>(a, L, li, w, i, p, d, T, G)) continue;
            // The ray's spot terms in float (B = d_xy / d_z, A at z = 0 or the angles), summed in
            // double.
            const float iz = 1.0f / d.z;
            const float bx = d.x * iz, by = d.y * iz;
            const float ax = L.afocal ? atanf(bx) : p.x - bx * p.z;
            const float ay = L.afocal ? atanf(by) : p.y - by * p.z;
            if (kGrad) {
#pragma unroll
                for (int kp = 0; kp < kGroup; ++kp) {
                    const float dbx = (G.d[0][kp] * d.z - d.x * G.d[2][kp]) * iz * iz;
                    const float dby = (G.d[1][kp] * d.z - d.y * G.d[2][kp]) * iz * iz;
                    float dax, day;
                    if (L.afocal) {
                        dax = dbx / (1.0f + bx * bx);
                        day = dby / (1.0f + by * by);
                    } else {
                        dax = G.o[0][kp] - dbx * p.z - bx * G.o[2][kp];
                        day = G.o[1][kp] - dby * p.z - by * G.o[2][kp];
                    }
                    gs[kp][0] += dax;
                    gs[kp][1] += day;
                    gs[kp][2] += dbx;
                    gs[kp][3] += dby;
                    gs[kp][4] += ax * dax + ay * day;
                    gs[kp][5] += bx * dbx + by * dby;
                    gs[kp][6] += dax * bx + day * by + ax * dbx + ay * dby;
                }
            }
            if (!sums) continue;
            m[0] += 1.0;
            m[1] += double(ax);
            m[2] += double(ay);
            m[3] += double(bx);
            m[4] += double(by);
            m[5] += double(ax * ax + ay * ay);
            m[6] += double(bx * bx + by * by);
            m[7] += double(ax * bx + ay * by);
            m[8] += double(T);
            if (grids && L.detHalf > 0.0f) {
                const unsigned long long v = (unsigned long long)llrint(double(T) * kHistScale);
                const float half = 0.5f * float(a.planes - 1);
                for (uint32_t pl = 0; pl < a.planes; ++pl) {
                    const float z = L.detZ + (float(pl) - half) * L.detStep;
                    const float s = (z - p.z) * iz;
                    const float gx = (p.x + d.x * s - L.detX) / L.detHalf, gy = (p.y + d.y * s - L.detY) / L.detHalf;
                    const float fx = (gx + 1.0f) * 0.5f * float(a.pixels), fy = (1.0f - gy) * 0.5f * float(a.pixels);
                    if (fx >= 0.0f && fy >= 0.0f && fx < float(a.pixels) && fy < float(a.pixels)) {
                        const uint32_t px = min(uint32_t(fx), a.pixels - 1), py = min(uint32_t(fy), a.pixels - 1);
                        binAdd(table, lensHist, (pl * a.pixels + py) * a.pixels + px, v);
                    }
                }
            }
        }
        if (sums) {
            double* dst = a.moments + (size_t(li) * a.nLambda + w) * OWE_LENS_MOMENTS;
#pragma unroll
            for (int q = 0; q < OWE_LENS_MOMENTS; ++q) {
                const double v = waveSumD(m[q]);
                if (lane == 0 && v != 0.0) atomicAdd(dst + q, v);
            }
        }
        if (kGrad) {
            double* dst = a.grads + (size_t(li) * a.nLambda + w) * (OWE_LENS_MAX_PARAMS * kGradSums);
#pragma unroll
            for (int kp = 0; kp < kGroup; ++kp) {
                if (a.paramBase + uint32_t(kp) >= a.nParams) continue;
#pragma unroll
                for (int q = 0; q < kGradSums; ++q) {
                    const double v = waveSumD(double(gs[kp][q]));
                    if (lane == 0 && v != 0.0) atomicAdd(dst + (a.paramBase + kp) * kGradSums + q, v);
                }
            }
        }
    }
    if (grids && tableLens != ~0u) binFlush(table, a.hist + tableLens * lensBins, lane);
}

__global__ void lensBegin(LensArgs a, size_t nMoments, size_t nHist, size_t nGrads) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x, stride = size_t(gridDim.x) * blockDim.x;
    if (i == 0) *a.next = 0;
    for (size_t k = i; k < nMoments; k += stride) a.moments[k] = 0.0;
    for (size_t k = i; k < nHist; k += stride) a.hist[k] = 0;
    for (size_t k = i; k < nGrads; k += stride) a.grads[k] = 0.0;
}

// Resets the item counter between the gradient passes of one run.
__global__ void lensNext(LensArgs a) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.next = 0;
}

__global__ void lensProbe(LensArgs a, uint32_t li, uint32_t w, const uint64_t* index, uint64_t n, double* out) {
    const uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    f3 p, d;
    float T;
    RayGrad G;
    double* r = out + 7 * k;
    if (traceLensRay<false>(a, a.lens[li], li, w, index[k], p, d, T, G)) {
        r[0] = p.x, r[1] = p.y, r[2] = p.z, r[3] = d.x, r[4] = d.y, r[5] = d.z, r[6] = T;
    } else {
        for (int q = 0; q < 6; ++q) r[q] = __builtin_nan("");
        r[6] = 0.0;
    }
}

// ---------------------------------------------------------------- host: media and paraxial optics

double ciddor(double lnm) {
    const double s2 = (1e3 / lnm) * (1e3 / lnm);
    return 1 + 0.05792105 / (238.0185 - s2) + 0.00167917 / (57.362 - s2);
}
// IndexModel::n (src/owe/core/medium.cpp:46-66): absolute index, ×n_air when relative to air.
double mediumIndexHost(const owe_medium& m, double lnm) {
    const double l = lnm * 1e-3, l2 = l * l;
    double n = 1;
    switch (m.index_kind) {
        case OWE_INDEX_CONSTANT: n = m.coeff_a[0]; break;
        case OWE_INDEX_CAUCHY: n = m.coeff_a[0] + m.coeff_a[1] / l2 + m.coeff_a[2] / (l2 * l2); break;
        case OWE_INDEX_SELLMEIER: {
            double s = 1;
            for (int k = 0; k < m.terms; ++k) s += m.coeff_a[k] * l2 / (l2 - m.coeff_b[k]);
            n = std::sqrt(s);
            break;
        }
        case OWE_INDEX_CIDDOR: n = ciddor(lnm); break;
        default: fail(OWE_HIP_ERR_INVALID, "unknown index model");
    }
    return m.rel_air ? n * ciddor(lnm) : n;
}

struct ParaxialOut {
    double efl, bfl, epz, epr, length;
    double lastZ;  // the last vertex along the fixed chart (signed gaps)
    bool afocal;
    double angMag;
};

// paraxialAnalysis, object at infinity (src/owe/scene/prescription.cpp:193-321): the (y, n·u) system
// matrix with a refracting row's power (n' − n) c and a mirror's −2 dir n c (dir the travel along
// the fixed chart, which a mirror reverses; gaps unfolded), EFL −n'/C and BFL −n'A/C or — when C
// vanishes at the reference's scale — afocal with angular magnification D n/n', the stop (marked,
// else the surface that limits a marginal ray) and its image through the rows before it.
ParaxialOut paraxialHost(const owe_lens_surface* s, size_t N, const std::vector<double>& nAfter, double n0) {
    std::vector<double> nBefore(N), phi(N), t(N);
    double nPrev = n0, dir = 1;
    for (size_t i = 0; i < N; ++i) {
        nBefore[i] = nPrev;
        phi[i] = s[i].reflect ? -2 * dir * nAfter[i] * s[i].curvature : (nAfter[i] - nPrev) * s[i].curvature;
        if (s[i].reflect) dir = -dir;
        t[i] = std::fabs(s[i].thickness);
        nPrev = nAfter[i];
    }
    double a = 1, b = 0, c = 0, d = 1;
    for (size_t i = 0; i < N; ++i) {
        c -= phi[i] * a;
        d -= phi[i] * b;
        if (i + 1 < N) {
            const double k = t[i] / nAfter[i];
            a += k * c;
            b += k * d;
        }
    }
    ParaxialOut r{};
    for (size_t i = 0; i + 1 < N; ++i) {
        r.length += t[i];
        r.lastZ += s[i].thickness;
    }
    const double nk = nAfter[N - 1];
    const double scale = 1e-12 * (1 + std::fabs(a) + std::fabs(d));
    r.afocal = std::fabs(c) * std::max(1.0, r.length) < scale * 1e3 || std::fabs(c) < 1e-9;
    if (r.afocal) {
        r.efl = r.bfl = INFINITY;
        r.angMag = d * n0 / nk;
    } else {
        r.efl = -nk / c;
        r.bfl = -nk * a / c;
    }
    int stop = -1;
    for (size_t i = 0; i < N; ++i)
        if (s[i].stop) stop = int(i);
    if (stop < 0) {
        double best = -1, y = 1, nu = 0;
        for (size_t i = 0; i < N; ++i) {
            const double ratio = std::fabs(y) / std::max(s[i].semi_diameter, 1e-12);
            if (ratio > best) {
                best = ratio;
                stop = int(i);
            }
            nu -= phi[i] * y;
            if (i + 1 < N) y += t[i] * nu / nAfter[i];
        }
    }
    const double stopR = s[size_t(stop)].semi_diameter;
    double y = 0, nu = 1e-3 * nBefore[size_t(stop)], yM = stopR, nuM = 0;
    for (int i = stop - 1; i >= 0; --i) {
        y -= t[size_t(i)] * nu / nAfter[size_t(i)];
        yM -= t[size_t(i)] * nuM / nAfter[size_t(i)];
        nu += phi[size_t(i)] * y;
        nuM += phi[size_t(i)] * yM;
    }
    if (stop == 0) {
        r.epz = 0;
        r.epr = stopR;
    } else {
        const double u = nu / n0;
        require(u != 0, "entrance pupil at infinity");
        const double dz = -y / u;
        r.epz = dz;
        r.epr = std::fabs(yM + dz * nuM / n0);
    }
    return r;
}

}  // namespace
}  // namespace owehip

struct owe_hip_lens_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocations;
    std::vector<owe_lens> lenses;
    std::vector<owe_lens_result> paraxial;  // launched/efl/bfl/pupil/last vertex per lens
    std::vector<DLensInfo> info;
    owe_lens_batch_settings settings{};
    LensArgs args{};
    size_t nMoments = 0, nHist = 0, nGrads = 0;
    int computeUnits = 256;
    int blocksPerCU = 4;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double lastMs = 0;
    bool ran = false;
};

namespace owehip {
namespace {

template <class T>
T* lensAlloc(owe_hip_lens_batch* b, size_t n) {
    void* p = nullptr;
    OWE_HIP_CHECK(hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T)));
    b->allocations.push_back(p);
    return static_cast<T*>(p);
}

// The work items of each lens for the current raysPerItem.
void layoutItems(owe_hip_lens_batch* b, uint32_t raysPerItem) {
    uint64_t base = 0;
    for (DLensInfo& L : b->info) {
        L.itemBase = base;
        L.itemsPerLambda = uint32_t((L.rays + raysPerItem - 1) / raysPerItem);
        base += uint64_t(L.itemsPerLambda) * b->args.nLambda;
    }
    OWE_HIP_CHECK(hipMemcpy(const_cast<DLensInfo*>(b->args.lens), b->info.data(), b->info.size() * sizeof(DLensInfo),
                            hipMemcpyHostToDevice));
    b->args.raysPerItem = raysPerItem;
    b->args.totalItems = base;
}

// One run: the plain kernel, or with parameters set one gradient pass per kGroup parameters (the
// first also adding the spot sums and the histogram).
double runOnce(owe_hip_lens_batch* b) {
    hipLaunchKernelGGL(lensBegin, dim3{1024}, dim3{256}, 0, b->stream, b->args, b->nMoments, b->nHist, b->nGrads);
    OWE_HIP_CHECK(hipEventRecord(b->ev[0], b->stream));
    const dim3 grid{unsigned(std::max(1, b->computeUnits * b->blocksPerCU))};
    if (b->args.nParams == 0) {
        hipLaunchKernelGGL(lensKernel<false>, grid, dim3{kLensBlock}, 0, b->stream, b->args);
    } else {
        for (uint32_t base = 0; base < b->args.nParams; base += kGroup) {
            LensArgs a = b->args;
            a.paramBase = base;
            a.sums = base == 0;
            if (base > 0) hipLaunchKernelGGL(lensNext, dim3{1}, dim3{64}, 0, b->stream, a);
            hipLaunchKernelGGL(lensKernel<true>, grid, dim3{kLensBlock}, 0, b->stream, a);
        }
    }
    OWE_HIP_CHECK(hipGetLastError());
    OWE_HIP_CHECK(hipEventRecord(b->ev[1], b->stream));
    OWE_HIP_CHECK(hipEventSynchronize(b->ev[1]));
    float ms = 0;
    OWE_HIP_CHECK(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    b->lastMs = ms;
    b->ran = true;
    return ms;
}

}  // namespace
}  // namespace owehip

extern "C" {

int owe_hip_lens_batch_create(int device, const owe_medium* media, uint32_t n_media, const owe_lens_surface* surfaces,
                              uint32_t n_surfaces, const owe_lens* lenses, uint32_t n_lenses,
                              const owe_lens_batch_settings* settings, owe_hip_lens_batch** out) {
    owe_hip_lens_batch* b = nullptr;
    const int st = guard([&] {
        require(out && media && surfaces && lenses && settings, "null argument");
        *out = nullptr;
        require(n_media > 0 && n_surfaces > 0 && n_lenses > 0, "empty batch");
        const owe_lens_batch_settings& s = *settings;
        require(s.n_wavelengths >= 1 && s.n_wavelengths <= OWE_LENS_MAX_WAVELENGTHS, "1 to 16 wavelengths");
        for (int w = 0; w < s.n_wavelengths; ++w)
            require(s.wavelengths[w] >= 300 && s.wavelengths[w] <= 2000, "wavelengths must lie in [300, 2000] nm");
        require(s.reference_wavelength >= 300 && s.reference_wavelength <= 2000, "reference wavelength out of range");
        require(std::fabs(s.field_deg) < 80, "field angle must lie in (-80, 80) degrees");
        require(s.detector_pixels >= 0 && s.detector_pixels <= 4096, "detector pixels must lie in [0, 4096]");
        require(s.detector_half_size >= 0, "detector half size must be non-negative");
        require(s.focus_planes >= 0 && s.focus_planes <= 16, "focus planes must lie in [0, 16]");
        require(s.focus_step >= 0 && std::isfinite(s.focus_step), "the focus step must be finite and non-negative");
        for (uint32_t i = 0; i < n_surfaces; ++i) {
            const owe_lens_surface& r = surfaces[i];
            const std::string name = "lens surface " + std::to_string(i);
            require(r.medium >= 0 && uint32_t(r.medium) < n_media, name + ": medium out of range");
            require(r.semi_diameter > 0 && std::isfinite(r.semi_diameter), name + ": semi-diameter must be positive");
            require(std::isfinite(r.curvature) && std::isfinite(r.conic) && std::isfinite(r.thickness),
                    name + ": values must be finite");
            require(1 - (1 + r.conic) * r.curvature * r.curvature * r.semi_diameter * r.semi_diameter >= 0,
                    name + ": aperture exceeds the domain of the conic");
            require(r.hole_radius >= 0 && (r.hole_radius == 0 || (r.reflect && r.hole_radius < r.semi_diameter)),
                    name + ": only a mirror has a hole, smaller than its semi-diameter");
        }
        selectDevice(device);
        b = new owe_hip_lens_batch;
        b->device = device;
        b->settings = s;
        hipDeviceProp_t prop;
        OWE_HIP_CHECK(hipGetDeviceProperties(&prop, device));
        b->computeUnits = prop.multiProcessorCount;
        OWE_HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        OWE_HIP_CHECK(hipEventCreate(&b->ev[0]));
        OWE_HIP_CHECK(hipEventCreate(&b->ev[1]));
        const uint32_t nw = uint32_t(s.n_wavelengths);
        std::vector<DSurf> ds;
        std::vector<float> nAfter(size_t(n_surfaces) * nw, 1.0f), nAmb(size_t(n_lenses) * nw, 1.0f);
        std::vector<uint8_t> used(n_surfaces, 0);
        // Lenses whose every root is polished in float64: instruments with mirrors or afocal ones,
        // where metre-long paths and the angular magnification turn a float root's error along
        // the ray (about 1e-7 of the path) into angular errors comparable to the spot itself.
        std::vector<uint8_t> precise;
        const double f = s.field_deg * 3.14159265358979323846 / 180;
        for (uint32_t li = 0; li < n_lenses; ++li) {
            const owe_lens& L = lenses[li];
            const std::string name = "lens " + std::to_string(li);
            require(L.surface_count >= 1 && uint64_t(L.first_surface) + L.surface_count <= n_surfaces,
                    name + ": surfaces out of range");
            require(L.rays >= 1 && L.rays <= (uint64_t(1) << 32), name + ": rays must lie in [1, 2^32]");
            require(L.ambient >= 0 && uint32_t(L.ambient) < n_media, name + ": ambient medium out of range");
            require(L.pupil_rings >= 0 && L.pupil_rings <= 30000, name + ": pupil rings must lie in [0, 30000]");
            require(L.pupil_rings == 0 || L.rays == 1 + 3 * uint64_t(L.pupil_rings) * uint64_t(L.pupil_rings + 1),
                    name + ": with pupil rings r the lens traces 1 + 3 r (r + 1) rays");
            const owe_lens_surface* rows = surfaces + L.first_surface;
            std::vector<double> nRef(L.surface_count);
            int mirrors = 0;
            for (uint32_t k = 0; k < L.surface_count; ++k) {
                require(!used[L.first_surface + k], name + ": surfaces are shared with another lens");
                used[L.first_surface + k] = 1;
                if (rows[k].reflect) {
                    ++mirrors;
                    require(rows[k].medium == (k == 0 ? L.ambient : rows[k - 1].medium),
                            name + ": a mirror reflects into the medium it is met in");
                }
                nRef[k] = mediumIndexHost(media[rows[k].medium], s.reference_wavelength);
                for (uint32_t w = 0; w < nw; ++w)
                    nAfter[size_t(L.first_surface + k) * nw + w] = float(mediumIndexHost(media[rows[k].medium], s.wavelengths[w]));
            }
            for (uint32_t w = 0; w < nw; ++w) nAmb[size_t(li) * nw + w] = float(mediumIndexHost(media[L.ambient], s.wavelengths[w]));
            const ParaxialOut px = paraxialHost(rows, L.surface_count, nRef, mediumIndexHost(media[L.ambient], s.reference_wavelength));
            owe_lens_result r{};
            r.launched = L.rays * nw;
            r.efl = px.efl;
            r.bfl = px.bfl;
            r.entrance_pupil_z = px.epz;
            r.entrance_pupil_radius = px.epr;
            r.last_vertex_z = px.lastZ;
            r.afocal = px.afocal ? 1 : 0;
            r.angular_magnification = px.angMag;
            b->paraxial.push_back(r);
            DLensInfo I{};
            I.first = L.first_surface;
            I.count = L.surface_count;
            I.rays = L.rays;
            I.rings = uint32_t(L.pupil_rings);
            I.rEP = px.epr;
            I.zEP = px.epz;
            I.zStart = -0.05 - 0.1 * px.length;
            // The detector: the paraxial image plane at the reference wavelength, centred on the
            // chief ray's paraxial image height EFL·tan θ.
            I.detZ = float(px.lastZ + px.bfl);
            I.detX = 0.0f;
            I.detY = float(px.efl * std::tan(f));
            I.detHalf = float(s.detector_half_size > 0 ? s.detector_half_size : 5e-3 * std::fabs(px.efl));
            I.detStep = float(s.focus_step > 0 ? s.focus_step : 1e-3 * std::fabs(px.efl));
            I.afocal = px.afocal ? 1 : 0;
            I.mirrors = mirrors;
            if (px.afocal) {  // no image plane: no irradiance grid
                I.detZ = I.detY = I.detHalf = I.detStep = 0.0f;
            }
            b->info.push_back(I);
            precise.push_back(mirrors > 0 || px.afocal);
        }
        ds.resize(n_surfaces);
        for (uint32_t li = 0; li < n_lenses; ++li) {
            const owe_lens& L = lenses[li];
            double z = 0;
            for (uint32_t k = 0; k < L.surface_count; ++k) {
                const owe_lens_surface& r = surfaces[L.first_surface + k];
                DSurf d{};
                static_cast<DSag&>(d) = makeSag(r.curvature, r.conic, r.asphere, r.semi_diameter,
                                                r.reflect ? r.hole_radius : 0.0);
                d.polish64 = precise[li] || std::fabs(r.curvature) * r.semi_diameter > kSteepPolish ? 1 : 0;
                d.z = float(z);
                d.reflect = r.reflect ? 1 : 0;
                ds[L.first_surface + k] = d;
                z += r.thickness;
            }
        }
        b->lenses.assign(lenses, lenses + n_lenses);
        LensArgs& a = b->args;
        DSurf* dsurf = lensAlloc<DSurf>(b, ds.size());
        OWE_HIP_CHECK(hipMemcpy(dsurf, ds.data(), ds.size() * sizeof(DSurf), hipMemcpyHostToDevice));
        a.surf = dsurf;
        a.lens = lensAlloc<DLensInfo>(b, n_lenses);
        float* na = lensAlloc<float>(b, nAfter.size());
        OWE_HIP_CHECK(hipMemcpy(na, nAfter.data(), nAfter.size() * sizeof(float), hipMemcpyHostToDevice));
        a.nAfter = na;
        float* nb = lensAlloc<float>(b, nAmb.size());
        OWE_HIP_CHECK(hipMemcpy(nb, nAmb.data(), nAmb.size() * sizeof(float), hipMemcpyHostToDevice));
        a.nAmbient = nb;
        a.nLenses = n_lenses;
        a.nLambda = nw;
        a.pixels = uint32_t(s.detector_pixels);
        a.planes = uint32_t(std::max(1, s.focus_planes));
        a.dirY = std::sin(f);
        a.dirZ = std::cos(f);
        a.next = lensAlloc<unsigned long long>(b, 1);
        b->nMoments = size_t(n_lenses) * nw * OWE_LENS_MOMENTS;
        a.moments = lensAlloc<double>(b, b->nMoments);
        b->nHist = size_t(n_lenses) * a.planes * a.pixels * a.pixels;
        a.hist = lensAlloc<unsigned long long>(b, b->nHist);
        b->nGrads = 0;
        a.grads = lensAlloc<double>(b, size_t(n_lenses) * nw * OWE_LENS_MAX_PARAMS * kGradSums);
        a.nParams = 0;
        layoutItems(b, 1024);
        *out = b;
    });
    if (st != OWE_HIP_OK && b) {
        for (void* p : b->allocations) (void)hipFree(p);
        for (hipEvent_t e : b->ev)
            if (e) (void)hipEventDestroy(e);
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
    }
    return st;
}

int owe_hip_lens_batch_destroy(owe_hip_lens_batch* b) {
    return guard([&] {
        if (!b) return;
        (void)hipSetDevice(b->device);
        if (b->stream) (void)hipStreamSynchronize(b->stream);
        for (void* p : b->allocations) OWE_HIP_CHECK(hipFree(p));
        for (hipEvent_t e : b->ev)
            if (e) OWE_HIP_CHECK(hipEventDestroy(e));
        if (b->stream) OWE_HIP_CHECK(hipStreamDestroy(b->stream));
        delete b;
    });
}

int owe_hip_lens_batch_run(owe_hip_lens_batch* b) {
    return guard([&] {
        require(b != nullptr, "batch is null");
        OWE_HIP_CHECK(hipSetDevice(b->device));
        runOnce(b);
    });
}

int owe_hip_lens_batch_set_params(owe_hip_lens_batch* b, const owe_lens_param* params, uint32_t n) {
    return guard([&] {
        require(b != nullptr, "batch is null");
        require(n <= OWE_LENS_MAX_PARAMS, "at most OWE_LENS_MAX_PARAMS parameters");
        require(n == 0 || params != nullptr, "null parameters");
        uint32_t minCount = ~0u;
        for (const owe_lens& L : b->lenses) minCount = std::min(minCount, L.surface_count);
        for (uint32_t k = 0; k < n; ++k) {
            require(params[k].surface < minCount, "parameter surface beyond a lens's surfaces");
            require(params[k].field >= OWE_LENS_PARAM_CURVATURE && params[k].field <= OWE_LENS_PARAM_A4 + 3,
                    "unknown parameter field");
            b->args.paramSurface[k] = int32_t(params[k].surface);
            b->args.paramField[k] = params[k].field;
        }
        b->args.nParams = n;
        b->nGrads = n ? size_t(b->lenses.size()) * b->args.nLambda * OWE_LENS_MAX_PARAMS * kGradSums : 0;
        b->ran = false;
    });
}

int owe_hip_lens_batch_gradients(owe_hip_lens_batch* b, double* grad) {
    return guard([&] {
        require(b != nullptr && grad != nullptr, "null argument");
        require(b->ran, "run the batch first");
        require(b->args.nParams > 0, "no parameters are set");
        const size_t nl = b->lenses.size(), nw = b->args.nLambda, np = b->args.nParams;
        std::vector<double> mom(nl * nw * OWE_LENS_MOMENTS), gs(nl * nw * OWE_LENS_MAX_PARAMS * kGradSums);
        OWE_HIP_CHECK(hipMemcpy(mom.data(), b->args.moments, mom.size() * sizeof(double), hipMemcpyDeviceToHost));
        OWE_HIP_CHECK(hipMemcpy(gs.data(), b->args.grads, gs.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t li = 0; li < nl; ++li) {
            double m[OWE_LENS_MOMENTS] = {};
            std::vector<double> g(np * kGradSums, 0.0);
            for (size_t w = 0; w < nw; ++w) {
                for (int q = 0; q < OWE_LENS_MOMENTS; ++q) m[q] += mom[(li * nw + w) * OWE_LENS_MOMENTS + q];
                for (size_t k = 0; k < np * kGradSums; ++k) g[k] += gs[(li * nw + w) * OWE_LENS_MAX_PARAMS * kGradSums + k];
            }
            const double n = m[0];
            for (size_t k = 0; k < np; ++k) {
                double out = NAN;
                const double* G = &g[k * kGradSums];  // ΣA'x, ΣA'y, ΣB'x, ΣB'y, ΣA·A', ΣB·B', Σ(A'·B + A·B')
                if (n >= 3) {
                    const double max = m[1] / n, may = m[2] / n, mbx = m[3] / n, mby = m[4] / n;
                    const double dmax = G[0] / n, dmay = G[1] / n, dmbx = G[2] / n, dmby = G[3] / n;
                    if (b->paraxial[li].afocal) {
                        // v = Σ|A|²/n − 2 c·ΣA/n + |c|², c = atan(mean B): afocalSpot's RMS².
                        const double cx = std::atan(mbx), cy = std::atan(mby);
                        const double dcx = dmbx / (1 + mbx * mbx), dcy = dmby / (1 + mby * mby);
                        const double v = m[5] / n - 2 * (cx * max + cy * may) + cx * cx + cy * cy;
                        const double dv = 2 * G[4] / n - 2 * (dcx * max + dcy * may) - 2 * (cx * dmax + cy * dmay) +
                                          2 * (cx * dcx + cy * dcy);
                        out = v > 0 ? dv / (2 * std::sqrt(v)) : 0.0;
                    } else {
                        // rms² at best focus = var_A − cov²/var_B (focusedSpot).
                        const double va = m[5] / n - (max * max + may * may);
                        const double vb = m[6] / n - (mbx * mbx + mby * mby);
                        const double cv = m[7] / n - (max * mbx + may * mby);
                        const double dva = 2 * G[4] / n - 2 * (max * dmax + may * dmay);
                        const double dvb = 2 * G[5] / n - 2 * (mbx * dmbx + mby * dmby);
                        const double dcv = G[6] / n - (dmax * mbx + dmay * mby + max * dmbx + may * dmby);
                        const double r2 = vb > 0 ? va - cv * cv / vb : va;
                        const double dr2 = vb > 0 ? dva - (2 * cv * dcv * vb - cv * cv * dvb) / (vb * vb) : dva;
                        out = r2 > 0 ? dr2 / (2 * std::sqrt(r2)) : 0.0;
                    }
                }
                grad[li * np + k] = out;
            }
        }
    });
}

int owe_hip_lens_batch_last_ms(const owe_hip_lens_batch* b, double* ms) {
    return guard([&] {
        require(b != nullptr && ms != nullptr, "null argument");
        *ms = b->lastMs;
    });
}

int owe_hip_lens_batch_set_shape(owe_hip_lens_batch* b, int blocks_per_cu, int rays_per_item) {
    return guard([&] {
        require(b != nullptr, "batch is null");
        require(blocks_per_cu >= 0 && blocks_per_cu <= 64, "blocks per compute unit must lie in [0, 64]");
        require(rays_per_item >= 0 && rays_per_item % 64 == 0 && rays_per_item <= (1 << 20),
                "rays per item must be a multiple of 64 up to 2^20");
        OWE_HIP_CHECK(hipSetDevice(b->device));
        if (blocks_per_cu > 0) b->blocksPerCU = blocks_per_cu;
        if (rays_per_item > 0) layoutItems(b, uint32_t(rays_per_item));
    });
}

int owe_hip_lens_batch_autotune(owe_hip_lens_batch* b, double* best_ms) {
    return guard([&] {
        require(b != nullptr, "batch is null");
        OWE_HIP_CHECK(hipSetDevice(b->device));
        // Every shape is timed once; the three fastest three more times each, and the lowest median
        // wins, so one noisy run cannot pick the shape.
        struct Cand {
            int blocks;
            uint32_t rays;
            double ms;
        };
        std::vector<Cand> cands;
        auto run = [&](const Cand& c) {
            b->blocksPerCU = c.blocks;
            layoutItems(b, c.rays);
            return runOnce(b);
        };
        for (int blocks : {1, 2, 4, 8})
            for (uint32_t rays : {256u, 1024u, 4096u, 16384u}) {
                cands.push_back({blocks, rays, 0.0});
                cands.back().ms = run(cands.back());
            }
        std::sort(cands.begin(), cands.end(), [](const Cand& x, const Cand& y) { return x.ms < y.ms; });
        for (size_t i = 0; i < 3 && i < cands.size(); ++i) {
            double t[3];
            for (double& x : t) x = run(cands[i]);
            std::sort(t, t + 3);
            cands[i].ms = t[1];
        }
        std::sort(cands.begin(), cands.begin() + std::min<size_t>(3, cands.size()),
                  [](const Cand& x, const Cand& y) { return x.ms < y.ms; });
        const double best = cands.front().ms;
        b->blocksPerCU = cands.front().blocks;
        layoutItems(b, cands.front().rays);
        b->ran = false;
        if (best_ms) *best_ms = best;
    });
}

int owe_hip_lens_batch_results(owe_hip_lens_batch* b, owe_lens_result* lenses, double* moments, float* irradiance) {
    return guard([&] {
        require(b != nullptr, "batch is null");
        require(b->ran, "no run yet");
        OWE_HIP_CHECK(hipSetDevice(b->device));
        const uint32_t nl = uint32_t(b->lenses.size()), nw = b->args.nLambda, px = b->args.pixels;
        std::vector<double> m(b->nMoments);
        OWE_HIP_CHECK(hipMemcpy(m.data(), b->args.moments, m.size() * sizeof(double), hipMemcpyDeviceToHost));
        if (moments) std::copy(m.begin(), m.end(), moments);
        if (lenses) {
            for (uint32_t li = 0; li < nl; ++li) {
                owe_lens_result r = b->paraxial[li];
                double s[OWE_LENS_MOMENTS] = {0};
                for (uint32_t w = 0; w < nw; ++w)
                    for (int q = 0; q < OWE_LENS_MOMENTS; ++q) s[q] += m[(size_t(li) * nw + w) * OWE_LENS_MOMENTS + q];
                r.arrived = uint64_t(s[0]);
                r.transmitted = s[8] / double(r.launched);
                // focusedSpot (analysis.cpp:85-113) from the sums, or afocalSpot (70-81) from the
                // sums of angles.
                const double zImg = r.last_vertex_z + r.bfl;
                r.apparent_angle = r.angular_rms = NAN;
                if (r.afocal) {
                    r.best_focus_z = r.rms_best = r.rms_paraxial = r.centroid_y = NAN;
                    if (s[0] >= 3) {
                        const double n = s[0], cx = std::atan(s[3] / n), cy = std::atan(s[4] / n);
                        const double v = s[5] / n - 2 * (cx * s[1] + cy * s[2]) / n + cx * cx + cy * cy;
                        r.apparent_angle = cy;
                        r.angular_rms = std::sqrt(std::max(0.0, v));
                    }
                } else if (s[0] >= 3) {
                    const double n = s[0], max_ = s[1] / n, may = s[2] / n, mbx = s[3] / n, mby = s[4] / n;
                    const double varA = s[5] / n - (max_ * max_ + may * may), varB = s[6] / n - (mbx * mbx + mby * mby);
                    const double cov = s[7] / n - (max_ * mbx + may * mby);
                    const double zBest = varB > 0 ? -cov / varB : zImg;
                    auto rms = [&](double z) { return std::sqrt(std::max(0.0, varA + 2 * z * cov + z * z * varB)); };
                    r.best_focus_z = zBest - r.last_vertex_z;
                    r.rms_best = rms(zBest);
                    r.rms_paraxial = rms(zImg);
                    r.centroid_y = may + mby * zImg;
                } else {
                    r.best_focus_z = r.rms_best = r.rms_paraxial = r.centroid_y = NAN;
                }
                lenses[li] = r;
            }
        }
        if (irradiance && px > 0) {
            std::vector<unsigned long long> h(b->nHist);
            OWE_HIP_CHECK(hipMemcpy(h.data(), b->args.hist, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
            const double cosT = b->args.dirZ;
            const size_t lensBins = size_t(b->args.planes) * px * px;
            for (uint32_t li = 0; li < nl; ++li) {
                // Unit irradiance across the beam: each ray carries π r² cos θ / (rays · nλ) of flux.
                const DLensInfo& I = b->info[li];
                float* out = irradiance + size_t(li) * lensBins;
                if (!(I.detHalf > 0)) {
                    std::fill(out, out + lensBins, 0.0f);
                    continue;
                }
                const double pix = 2.0 * I.detHalf / px;
                const double k = 3.14159265358979323846 * I.rEP * I.rEP * 0.999 * 0.999 * cosT /
                                 (double(I.rays) * nw * kHistScale * pix * pix);
                for (size_t q = 0; q < lensBins; ++q) out[q] = float(double(h[size_t(li) * lensBins + q]) * k);
            }
        }
    });
}

int owe_hip_lens_batch_trace(owe_hip_lens_batch* b, uint32_t lens, int wavelength, const uint64_t* index, uint64_t n,
                             double* out) {
    return guard([&] {
        require(b != nullptr, "batch is null");
        require(lens < b->lenses.size(), "lens out of range");
        require(wavelength >= 0 && uint32_t(wavelength) < b->args.nLambda, "wavelength out of range");
        if (n == 0) return;
        require(index != nullptr && out != nullptr, "null buffer");
        for (uint64_t k = 0; k < n; ++k) require(index[k] < b->lenses[lens].rays, "ray index out of range");
        OWE_HIP_CHECK(hipSetDevice(b->device));
        uint64_t* di = nullptr;
        double* dout = nullptr;
        OWE_HIP_CHECK(hipMalloc(&di, n * sizeof(uint64_t)));
        OWE_HIP_CHECK(hipMalloc(&dout, 7 * n * sizeof(double)));
        OWE_HIP_CHECK(hipMemcpy(di, index, n * sizeof(uint64_t), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(lensProbe, dim3{unsigned((n + 255) / 256)}, dim3{256}, 0, b->stream, b->args, lens,
                           uint32_t(wavelength), di, n, dout);
        const hipError_t e = hipGetLastError();
        if (e == hipSuccess) {
            (void)hipStreamSynchronize(b->stream);
            (void)hipMemcpy(out, dout, 7 * n * sizeof(double), hipMemcpyDeviceToHost);
        }
        (void)hipFree(di);
        (void)hipFree(dout);
        OWE_HIP_CHECK(e);
    });
}

}  // extern "C"
