// Stochastic progressive photon mapping (the reference's sppm integrator, sppm.slang): the state the
// camera pass (render.hip) and the photon pass (particle.hip) share, and the hash grid's arithmetic.
//
// One iteration: every pixel traces one camera path through specular events to its visible point
// (the first non-specular surface) and records it; the points enter a multi-level spatial hash, a
// point of radius r living in the one cell of size 2^level ≥ 2r that holds it; photons leave the
// lights and deposit at every non-specular surface after their first scattering event into each
// point within reach (the 2×2×2 cells about the photon on each level in use cover it); each pixel
// then updates N, r and τ. Deposits are exact integer additions (the particle film's bins), so the
// result does not depend on the order photons arrive in.
#pragma once

#include <hip/hip_runtime.h>

#include "particle.hpp"
#include "shading_dev.hpp"

namespace owehip {

// Visible-point flags (bits 0-2) and the point's wavelength group (bits 8-15).
constexpr uint32_t kVpValid = 1u, kVpTerminated = 2u, kVpFree = 4u;
constexpr uint32_t kGridNone = 0xFFFFFFFFu;

// Stream keys: the camera path of (pixel, iteration), the pixel's wavelength group, photon i.
constexpr uint64_t kSppmCamera = 0x6a09e667f3bcc908ull, kSppmGroup = 0xbb67ae8584caa73bull,
                   kSppmPhoton = 0xa54ff53a5f1d36f1ull;

struct SppmDev {
    // Visible points, one slot per pixel (flags 0: none this iteration).
    float4* vpP;      // position; w: gather radius
    float4* vpN;      // the boundary's geometric normal (toward its front); w: boundary (bits)
    float4* vpD;      // arriving direction; w: flags (bits)
    float4* vpBeta;   // throughput arriving, with the camera weight
    // Per pixel across iterations.
    float* radius;    // gather radius (0 until the pixel's first visible point)
    double2* tauXY;   // τ X, Y
    double2* tauZN;   // τ Z, N (the accumulated fractional photon count)
    // This iteration's deposits per pixel: 3 × kSplatBins exact bins, and M.
    unsigned long long* flux;
    uint32_t* count;
    // The hash: bucket → first point, point → next point in its bucket, [lowest, highest] level in use.
    uint32_t* head;
    uint32_t* next;
    int* levels;
    uint32_t gridMask, groups, pixels, iteration;
    float rotation;      // the iteration's base wavelength rotation
    float alpha;         // radius reduction
    float r0PerMetre;    // initial radius per metre of unfolded camera path
    float r0Floor;
    float invNWY;        // 1 / (NW ∫ȳ)
    unsigned long long* stats;  // += paths, segments, inconsistencies, leaks
    // Photon guiding over the sun's emission disc: counts per cell, their CDF (floats), its total.
    unsigned int* guideCounts;
    float* guideCdf;
    float* guideTotal;
};

// Photon guiding (particle.slang:149-183): kGuideRes² cells over the bounding square of the sun's
// emission disc, in units of its radius. Once the map has counts, kGuideShare of sun photons are
// drawn from it and every sun photon is weighted by the mixture density: the estimate is unchanged
// (the uniform share keeps full support), only its noise falls.
constexpr int kGuideRes = 128, kGuideCells = kGuideRes * kGuideRes;
constexpr float kGuideShare = 0.75f;

// The first cell whose CDF exceeds u.
__device__ __forceinline__ uint32_t guideSample(const float* cdf, float u) {
    uint32_t lo = 0, len = kGuideCells;
    while (len > 0) {
        const uint32_t half = len / 2;
        if (cdf[lo + half] <= u) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo < uint32_t(kGuideCells) ? lo : uint32_t(kGuideCells - 1);
}

// The sun's emission density at (qx, qy) per unit-disc area: the map's mixture with the uniform share.
__device__ __forceinline__ float guidePdf(const float* cdf, float total, float qx, float qy, uint32_t& cell) {
    const float size = 2.0f / float(kGuideRes);
    const int cx = int(floorf((qx + 1.0f) / size)), cy = int(floorf((qy + 1.0f) / size));
    const bool in = cx >= 0 && cy >= 0 && cx < kGuideRes && cy < kGuideRes;
    cell = in ? uint32_t(cy * kGuideRes + cx) : kGridNone;
    const float uniform = qx * qx + qy * qy <= 1.0f ? 1.0f / kPiF : 0.0f;
    if (!(total > 0.0f)) return uniform;
    const float p = in ? cdf[cell] - (cell > 0 ? cdf[cell - 1] : 0.0f) : 0.0f;
    return (in ? kGuideShare * p / (size * size) : 0.0f) + (1.0f - kGuideShare) * uniform;
}

// The level whose cells (size 2^level) are the smallest at least 2r across: exact, from the float's
// exponent (a log2 rounded the wrong way would give cells smaller than the reach).
__device__ __forceinline__ int sppmLevel(float r) {
    int e;
    const float m = frexpf(fmaxf(2.0f * r, 1e-30f), &e);  // 2r = m · 2^e, m in [0.5, 1)
    return m == 0.5f ? e - 1 : e;
}

__device__ __forceinline__ int sppmCell(float x, float size) {
    return int(fminf(fmaxf(floorf(x / size), -2.0e9f), 2.0e9f));
}

__device__ __forceinline__ uint32_t sppmHash(int cx, int cy, int cz, int level, uint32_t mask) {
    const uint32_t h = (uint32_t(cx) * 0x9E3779B1u) ^ (uint32_t(cy) * 0x85EBCA77u) ^ (uint32_t(cz) * 0xC2B2AE3Du) ^
                       (uint32_t(level) * 0x27D4EB2Fu);
    return (h ^ (h >> 15)) & mask;
}

// The wavelengths of group g of an iteration: the base rotation moved by g / (4 G), between the four
// stratified wavelengths of a set, so the groups' union covers the spectrum evenly (sppm.slang:67-76).
__device__ __forceinline__ WL sppmWavelengths(const Settings& set, float rotation, uint32_t g, uint32_t groups) {
    float u = rotation + float(g) / float(4u * groups);
    u -= floorf(u);
    WL w;
    if (set.spectral == OWE_SPECTRAL_UNIFORM1) {
        const float l = set.lambdaMin + (set.lambdaMax - set.lambdaMin) * u;
#pragma unroll
        for (int k = 0; k < NW; ++k) w.lambda[k] = l;
        w.pdf[0] = 1.0f / (set.lambdaMax - set.lambdaMin) / NW;
        w.pdf[1] = w.pdf[2] = w.pdf[3] = 0.0f;
    } else {
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            float up = u + float(k) / NW;
            if (up >= 1.0f) up -= 1.0f;
            w.lambda[k] = sampleVisibleF(up);
            w.pdf[k] = visiblePdfF(w.lambda[k]);
        }
    }
    return w;
}

// Enqueues the iteration's photons [base, base + count) on `stream`.
void launchSppmPhotons(const DScene& sc, const Settings& set, const SppmDev& s, uint64_t base, uint32_t count,
                       uint32_t stackCap, hipStream_t stream);

}  // namespace owehip
