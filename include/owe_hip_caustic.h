/* Differentiable caustic design on the hip backend: light tracing through a slab whose exit face is a
 * bicubic height field, into an irradiance grid on a wall, and the gradient of any loss of that
 * grid with respect to every height by path-replay backpropagation (the particles are traced again
 * from their seeds; nothing per path is stored).
 *
 * Reference counterparts:
 *   height field: bicubic Hermite patches over a node grid    WaveSurface::height, src/owe/scene/geometry.cpp:251-270,
 *   (value, ∂x, ∂y, ∂xy per node); here the node slopes are  hermiteBases, src/owe/core/shared/surfaces.hpp:190-197
 *   central differences of the heights, so the surface is a
 *   linear function of the heights
 *   particles from a distant (collimated) source              emitDistant, src/owe/backends/gpu/shaders/particle.slang:234-265
 *   deterministic transmission, weight T = 1 − R per face      Tracer::walk PrimaryTransmission, transport.cpp:531-595
 *   Fresnel and Snell at the faces                             core/shared/optics.hpp:14-54
 *   sensor deposit per pixel area                              particle.slang:98-114, 344-350
 *   n(λ) Sellmeier (PMMA)                                      core/medium.cpp:46-66, 135
 * The splat is a quadratic B-spline (each particle spreads its weight over the 3 × 3 pixels about the
 * nearest pixel centre), so the grid is a continuously differentiable function of the particles'
 * wall positions; the bins are 64-bit fixed point, so a pass is bit-reproducible. The reference
 * has no adjoint. */
#ifndef OWE_HIP_CAUSTIC_H
#define OWE_HIP_CAUSTIC_H

#include <stdint.h>

#include "owe_hip_scene.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct owe_caustic_settings {
    uint32_t nodes;       /* heights per side; node (i, j) sits at (−slab_half + i·s, −slab_half + j·s) */
    uint32_t pixels;      /* irradiance grid per side */
    uint64_t particles;   /* per pass */
    uint64_t seed;
    double slab_half;     /* metres: the slab covers [−slab_half, slab_half]² */
    double thickness;     /* metres: entry face at z = 0, exit surface at z = thickness + h(x, y) */
    double wall_z;        /* metres: the wall plane */
    double wall_half;     /* metres: the grid covers [−wall_half, wall_half]² of the wall */
    double sun_kelvin;    /* blackbody temperature of the collimated (+z) sun */
    owe_medium medium;    /* the slab; the surroundings are air (Ciddor) */
} owe_caustic_settings;

typedef struct owe_hip_caustic owe_hip_caustic;

int owe_hip_caustic_create(int device, const owe_caustic_settings* settings, owe_hip_caustic** out);
int owe_hip_caustic_destroy(owe_hip_caustic* caustic);
/* The irradiance of pass `pass` (its particles' streams) through `heights` (nodes² floats, metres,
 * row j = y index, device memory) into `irradiance` (pixels² floats, row = y index from −wall_half,
 * device memory), in units of the incident luminous irradiance. stream: a hipStream_t (null: the
 * default stream); the call returns when the work is enqueued. */
int owe_hip_caustic_forward(owe_hip_caustic* caustic, const float* heights, float* irradiance, uint64_t pass,
                            void* stream);
/* Path replay: the particles of `pass` traced again through `heights`; grad_heights (nodes² floats,
 * device memory, overwritten) receives Σ_pixels grad_irradiance · ∂irradiance/∂height.
 * grad_irradiance: pixels² floats, device memory. */
int owe_hip_caustic_backward(owe_hip_caustic* caustic, const float* heights, const float* grad_irradiance,
                             float* grad_heights, uint64_t pass, void* stream);

/* The largest irradiance grid (pixels², 512 × 512) owe_hip_caustic_forward_noalloc takes. */
#define OWE_CAUSTIC_NOALLOC_MAX_PIXELS 262144u
/* owe_hip_caustic_forward without a handle and without allocating: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under graph capture, serving loops,
 * harnesses that forbid hipMalloc). The same kernels and result, bit for bit, as create + forward
 * with these settings; the fixed-point bins are a library-held array of 2 MiB per device, kept
 * until the process ends. Grids beyond OWE_CAUSTIC_NOALLOC_MAX_PIXELS are refused
 * (OWE_HIP_ERR_INVALID). Calls that overlap on one device share that array and race: serialise
 * them (one stream, or events). */
int owe_hip_caustic_forward_noalloc(int device, const owe_caustic_settings* settings, const float* heights,
                                    float* irradiance, uint64_t pass, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* OWE_HIP_CAUSTIC_H */
