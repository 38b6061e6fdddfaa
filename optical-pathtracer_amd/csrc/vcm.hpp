// Vertex connection and merging (the reference's vcm and bdpt integrators, src/owe/backends/gpu/
// shaders/vcm.slang, after Georgiev, Křivánek, Davidovič & Slusallek 2012, with a light vertex cache
// in place of per-pixel light subpaths): the state the light pass (particle.hip) and the camera pass
// (render.hip) share, and the recursive MIS quantities both ends carry.
//
// One iteration: N_L light subpaths leave the lights as particles do; each non-specular vertex goes
// into its wavelength group's cache with its throughput and Georgiev's dVCM, dVC, dVM, and is
// connected to the eye (light tracing, splatted exactly); with merging the cached vertices enter the
// multi-level hash of the SPPM pass at the level of the merging radius where they lie; then one
// camera path per pixel takes, at every non-specular vertex, next-event estimation, a connection to
// one vertex drawn uniformly from its group's cache (weighted by the cache's size over N_L) and, with
// merging, every cached vertex of its group within the merging radius; emission it meets is weighted
// too. Every strategy is weighted by the balance heuristic over all the others. Participating media
// only attenuate the subpaths here (by their extinction σa + σs): in a scene whose media scatter,
// a camera pass of its own adds the light that reached the eye through a medium scattering vertex.
#pragma once

#include <hip/hip_runtime.h>

#include "sppm.hpp"

namespace owehip {

// Stream keys of this backend: light subpath j of an iteration, the camera path of (pixel, iteration),
// the pixel's group.
constexpr uint64_t kVcmLight = 0x510e527fade682d1ull, kVcmCamera = 0x9b05688c2b3e6c1full,
                   kVcmGroup = 0x1f83d9abfb41bd6bull;

// A cached light vertex (80 B).
struct LightVertex {
    float4 p;     // camera-relative position; w: boundary (bits)
    float4 n;     // geometric normal (toward the boundary's front); w: dVCM
    float4 d;     // direction the light arrived along; w: flags (bits): kLvTerminated | (face + 1) << 1
    float4 beta;  // the subpath's throughput arriving (per wavelength)
    float4 mis;   // x: dVC, y: dVM, z: group (bits), w: the iteration's slot (bits)
};
constexpr uint32_t kLvTerminated = 1u;

// Iterations in flight: the iterations of a pass are independent but for their merging radius and
// wavelength rotation, so up to kVcmSlots of them run in each launch, each in its own slot of the
// cache, the offsets and the merging hash (a 256² image has 65,536 light subpaths per iteration: a
// quarter of a wave per SIMD of the chip; eight in flight fill it).
constexpr uint32_t kVcmSlots = 16;

// Resident waves per SIMD the light and camera kernels are compiled for: their registers are capped
// to fit two (a few spills; measured on a 256² caustic with eight iterations in flight, 0.31 ms per
// iteration at one wave, 0.22 at two, 0.26 at three).
#ifndef OWE_HIP_VCM_WAVES
#define OWE_HIP_VCM_WAVES 2
#endif

// The cache is filled in a deterministic order: light subpath j of slot k's iteration (group g =
// j mod G, i = j / G) is path p = k·N + g·N_L + i; it writes its vertices to staging slots [p·K,
// p·K + n_p) (K = max_depth + 1 bounds n_p) and n_p to pathCount[p]; an exclusive scan of the counts
// gives each path's offset, and its vertices are copied to cache[offset_p, offset_p + n_p). Group g
// of slot k is then the contiguous range [counts[32k + 16 + g], … + counts[32k + 4 + g]) in path
// order — the same cache whatever order the subpaths finish in, so a camera path's connection draw
// picks the same vertex every time, with any number of iterations in flight.
struct VcmDev {
    LightVertex* cache;    // the slots' vertices, slot by slot, group by group, in path order
    LightVertex* staging;  // K slots per light subpath
    uint32_t* pathCount;   // vertices per light subpath (path p = k·N + g·N_L + i)
    uint32_t* offset;      // their exclusive prefix sums; offset[slots · N] = the total
    uint32_t* counts;      // per slot 32: [4 + g] group g's vertex count; [16 + g] its first vertex
    uint32_t* tileSum;     // the scan's per-tile sums ...
    uint32_t* tileOffset;  // ... and their exclusive prefix sums
    uint32_t capacity;     // vertices the cache holds (every subpath's K, in every slot)
    uint32_t pathSlots;    // K
    uint32_t pathsPerGroup;
    uint32_t groups;       // wavelength groups G
    uint32_t slotPaths;    // N = G·N_L: light subpaths per iteration
    uint32_t slots;        // iterations in flight in this launch
    float lightPaths;      // N_L: light subpaths per group and iteration
    uint32_t iteration;    // slot 0's iteration (slot k runs iteration + k); in a slot's view, its own
    float rotation;        // a slot's view: its iteration's base wavelength rotation
    float mergePerMetre;   // a slot's view: merging radius per metre from the eye; 0: no merging (bdpt)
    float mergeFloor;
    uint32_t* head;        // the merging hash, gridMask + 1 buckets per slot: bucket → first vertex ...
    uint32_t* next;        // ... vertex → next vertex of its bucket (kGridNone ends a bucket)
    uint32_t gridMask;
    unsigned long long* film;  // the camera film (32.32 fixed point, three channels per pixel)
    float invNWY;              // 1 / (NW ∫ȳ)
    float mergeBase;           // the merging radius per metre at the first iteration (0: bdpt) ...
    double mergeExponent;      // ... times (i + 1)^mergeExponent at iteration i: (α − 1)/2
    double rotationOffset;     // the wavelength rotation of iteration i: frac(offset + φ⁻¹(i + 1))

    // Slot k's iteration, wavelength rotation, merging radius per metre, counts and hash.
    __device__ __forceinline__ float slotRotation(uint32_t k) const {
        const double u = rotationOffset + 0.6180339887498949 * double(iteration + k + 1u);
        return float(u - floor(u));
    }
    __device__ __forceinline__ float slotPerMetre(uint32_t k) const {
        return mergeBase > 0.0f ? float(double(mergeBase) * pow(double(iteration + k + 1u), mergeExponent)) : 0.0f;
    }
    __device__ __forceinline__ uint32_t* slotHead(uint32_t k) const { return head + size_t(k) * (size_t(gridMask) + 1); }
    // Slot k's view: a copy with those values in place of the launch's.
    __device__ __forceinline__ VcmDev at(uint32_t k) const {
        VcmDev s = *this;
        s.iteration = iteration + k;
        s.rotation = slotRotation(k);
        s.mergePerMetre = slotPerMetre(k);
        s.counts = counts + 32u * k;
        s.head = slotHead(k);
        return s;
    }
};

__device__ __forceinline__ uint32_t vcmPixelGroup(uint32_t pixel, uint32_t iteration, uint32_t groups) {
    return uint32_t(mix64((uint64_t(iteration) << 32 | pixel) ^ kVcmGroup) % uint64_t(groups));
}

// The merging radius at p for a radius per metre perMetre (a slot's).
__device__ __forceinline__ float vcmRadiusAt(const VcmDev& v, float perMetre, f3 p) {
    return fmaxf(perMetre * len3(p), v.mergeFloor);
}
__device__ __forceinline__ float vcmRadius(const VcmDev& v, f3 p) { return vcmRadiusAt(v, v.mergePerMetre, p); }
// η = π r² N_L: merging's weight relative to connections at p (0 without merging), and its inverse.
__device__ __forceinline__ float vcmEtaAt(const VcmDev& v, float perMetre, f3 p) {
    if (!(perMetre > 0.0f)) return 0.0f;
    const float r = vcmRadiusAt(v, perMetre, p);
    return kPiF * r * r * v.lightPaths;
}
__device__ __forceinline__ float vcmEta(const VcmDev& v, f3 p) { return vcmEtaAt(v, v.mergePerMetre, p); }
__device__ __forceinline__ float vcmInvEta(const VcmDev& v, f3 p) {
    const float e = vcmEta(v, p);
    return e > 0.0f ? 1.0f / e : 0.0f;
}

// A subpath's MIS quantities (Georgiev's recursive dVCM, dVC, dVM; vcm.slang:150-178), carried from
// vertex to vertex, in the balance heuristic.
struct VcmMis {
    float dVCM, dVC, dVM;

    // Arriving at a vertex `dist` beyond the last one at |cos| cosIn.
    __device__ __forceinline__ void arrive(float dist, float cosIn) {
        dVCM *= dist * dist;
        dVCM /= cosIn;
        dVC /= cosIn;
        dVM /= cosIn;
    }
    // After the scattering sample s at a vertex of normal n, arrived along d; pdfRev is the
    // reverse sample's pdf (arriving along −s.wi, leaving along −d), eta the vertex's η.
    __device__ __forceinline__ void scatter(f3 n, const SSample& s, float pdfRev, float eta) {
        const float cosOut = fabsf(dot3(n, s.wi));
        if (s.delta) {
            dVCM = 0.0f;
            dVC *= cosOut;
            dVM *= cosOut;
            return;
        }
        dVC = cosOut / s.pdf * (dVC * pdfRev + dVCM + eta);
        dVM = cosOut / s.pdf * (dVM * pdfRev + dVCM * (eta > 0.0f ? 1.0f / eta : 0.0f) + 1.0f);
        dVCM = 1.0f / s.pdf;
    }
};

// The interface a surface at boundary bi presents at wavelength lambda, with normal n.
__device__ __forceinline__ Iface vcmIface(const DScene& sc, uint32_t bi, f3 n, float lambda, float floor) {
    const DBoundary& b = sc.boundaries[bi];
    return Iface{&sc.optics[b.optics], n, regionIndex(sc, b.front, lambda), regionIndex(sc, b.back, lambda), lambda,
                 floor};
}

// f per wavelength of the interface at boundary bi (hero interface `it`) for arrival d and departure
// wi, with the hero's pdf; wavelengths with pdf 0 give 0.
__device__ __forceinline__ bool vcmEval(const DScene& sc, const Iface& it, uint32_t bi, const WL& wl, f3 d, f3 wi,
                                         bool radiance, float f[NW], float& pdf) {
    bool any = false;
    pdf = 0.0f;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        f[i] = 0.0f;
        if (wl.pdf[i] == 0.0f) continue;
        float p;
        f[i] = i == 0 ? evalScatter(sc, it, d, wi, radiance, p)
                      : evalScatter(sc, vcmIface(sc, bi, it.n, wl.lambda[i], it.floor), d, wi, radiance, p);
        if (i == 0) pdf = p;
        any = any || f[i] != 0.0f;
    }
    return any;
}

// Enqueues the light subpaths [0, count) of the slots' iterations (light tracing into the particle
// film's bins, the staging slots filled) on `stream`; count = slots · N.
void launchVcmLight(const DScene& sc, const Settings& set, const VcmDev& v, unsigned long long* lightFilm,
                    unsigned long long* stats, float invNWY, uint32_t count, hipStream_t stream);

}  // namespace owehip
