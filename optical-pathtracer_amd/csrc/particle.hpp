// The particle (light-tracing) pass of the renderer: launched by render.hip for the light integrator.
#pragma once

#include <hip/hip_runtime.h>

#include "scene_dev.hpp"

namespace owehip {

struct Settings;

// Exact film splats: each XYZ channel of a pixel is kSplatBins 64-bit integer bins, bin b counting
// units of 2^(kSplatExp0 + 32 b); a float splat adds its 24-bit significand, shifted to its exponent,
// as at most two 32-bit pieces (particle.slang:23-49). Integer sums do not depend on the order
// splats arrive in, so particle images are reproducible bit for bit.
constexpr int kSplatBins = 4;
constexpr int kSplatExp0 = -80;

struct ParticleBatch {
    uint64_t base;              // particle i of the batch is particle base + i of the pass ...
    const uint64_t* index;      // ... or index[i] (device memory) when not null
    uint32_t count, pass;
    unsigned long long* film;   // W·H·3·kSplatBins bins
    double* perParticle;        // when not null: count × 3 XYZ sums per particle instead of the film
    unsigned long long* stats;  // += paths, segments, inconsistencies, leaks
    float invNWY;               // 1 / (NW ∫ȳ)
    // The lane-refill form: particles base + [0, total) of the pass, handed out `chunk` ids at a time
    // from the device counter `next` (zero at launch); a wave starts new particles once `regen` of
    // its lanes are idle.
    uint64_t total;
    unsigned long long* next;
    uint32_t chunk, regen;
};

// Enqueues the batch on `stream`, one thread per particle; stackCap (16, 32 or 48) bounds the BVH
// traversal stack.
void launchParticles(const DScene& sc, const Settings& set, const ParticleBatch& b, uint32_t stackCap,
                     hipStream_t stream);
// Enqueues the pass in the lane-refill form on a grid of `blocks` blocks (the scene's stack bound
// must not exceed 48).
void launchParticlesRefill(const DScene& sc, const Settings& set, const ParticleBatch& b, uint32_t blocks,
                           hipStream_t stream);

}  // namespace owehip
