/* Batched lens analysis on the hip backend: many rotationally symmetric sequential lenses traced in
 * one persistent launch, each for its own number of rays at every wavelength of the batch.
 *
 * Reference counterparts:
 *   lens rows (curvature, conic, asphere, thickness,  PrescriptionSurface, src/owe/scene/prescription.hpp:16-34
 *   semi-diameter, medium after, stop)
 *   sag surfaces: conic quadric with refined roots,    SagSurface::intersectQuadric/Numeric, src/owe/scene/geometry.cpp:115-166;
 *   numeric even asphere                               src/owe/core/shared/surfaces.hpp:13-94, 156-171
 *   deterministic transmission through each surface    Tracer::walk (WalkMode::PrimaryTransmission),
 *   (refraction, weight T = 1 − R; TIR or a missed     src/owe/transport/transport.cpp:531-595; traceThrough,
 *   aperture ends the ray)                             src/owe/analysis/analysis.cpp:49-70
 *   paraxial EFL, BFL, entrance pupil                  paraxialAnalysis, src/owe/scene/prescription.cpp:193-321
 *   spot: closed-form best focus and RMS radius        focusedSpot, src/owe/analysis/analysis.cpp:85-113
 *
 * Every ray of lens l at wavelength j starts on the entrance pupil at point i of the lens's pupil
 * pattern (radius 0.999 of the pupil's), collimated at the batch's field angle in
 * the meridional (y, z) plane, from z_start = −0.05 − 0.1·length (analyzeLens,
 * src/owe/analysis/analysis.cpp:174-178). Units are metres; the first vertex sits at z = 0. */
#ifndef OWE_HIP_LENS_H
#define OWE_HIP_LENS_H

#include <stdint.h>

#include "owe_hip_scene.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One surface (PrescriptionSurface, prescription.hpp:16-34): vertex curvature 1/R (0: flat),
 * conic constant, even asphere A4..A10 (SI units), the signed distance to the next vertex along the
 * fixed chart (negative on a leg a mirror has turned back), semi-diameter, the medium after the
 * surface (an index into the batch's media), whether it is the aperture stop, and for a reflecting
 * row (a mirror: the light reflects into the same medium) the radius of its central hole.
 *
 * Mirrors also obstruct: a ray crossing the vertex plane of a mirror other than the two surfaces it
 * travels between is stopped there if it lies on the mirror's body (hole_radius ≤ r ≤
 * semi_diameter) — the bodies buildPrescription places (prescription.cpp:365-388), approximated by
 * their vertex-plane annulus. Spider vanes lie behind a secondary, beyond every launch plane, and
 * are not modelled. */
typedef struct owe_lens_surface {
    double curvature;
    double conic;
    double asphere[4];
    double thickness;
    double semi_diameter;
    int32_t medium;
    int32_t stop;
    int32_t reflect;
    int32_t reserved;
    double hole_radius;
} owe_lens_surface;

/* A lens: surfaces [first_surface, first_surface + surface_count) of the batch, the ambient medium
 * before and after it, and the rays it traces at each wavelength. The entrance pupil is sampled by
 * a Fibonacci disc lattice of `rays` points (pupil_rings = 0), or by spotAt's hexapolar rings
 * (src/owe/analysis/analysis.cpp:115-131): the centre, then 6k points on ring k = 1..pupil_rings
 * at radius 0.999 rEP k / pupil_rings, angle 2π j / 6k — then `rays` must be 1 + 3 r (r + 1). */
typedef struct owe_lens {
    uint32_t first_surface;
    uint32_t surface_count;
    uint64_t rays;
    int32_t ambient;
    int32_t pupil_rings;
} owe_lens;

#define OWE_LENS_MAX_WAVELENGTHS 16

typedef struct owe_lens_batch_settings {
    double field_deg;             /* collimated field angle (degrees) */
    double wavelengths[OWE_LENS_MAX_WAVELENGTHS];  /* nm */
    int32_t n_wavelengths;
    int32_t detector_pixels;      /* side of each lens's square irradiance grid; 0: none */
    double detector_half_size;    /* metres about the chief ray's paraxial image point; 0: 5e-3 × EFL */
    double reference_wavelength;  /* nm: the paraxial quantities and the detector plane */
    /* The through-focus stack: focus_planes grids (1–16; 0 means 1) spaced focus_step metres along
     * z, centred on the paraxial image plane (0: a step of 1e-3 × |EFL|). */
    int32_t focus_planes;
    int32_t reserved;
    double focus_step;
} owe_lens_batch_settings;

/* Per lens, over every wavelength together: focusedSpot over all arrived rays, or for an afocal
 * lens (paraxial power zero at the reference wavelength) afocalSpot (analysis.cpp:70-81): the
 * apparent direction of the mean ray and the angular RMS about it. */
typedef struct owe_lens_result {
    uint64_t launched, arrived;
    double efl, bfl;                 /* paraxial, at the reference wavelength (inf when afocal) */
    double entrance_pupil_z, entrance_pupil_radius;
    double last_vertex_z;
    double best_focus_z;             /* from the last vertex */
    double rms_best, rms_paraxial;   /* RMS spot radius at best focus and at the paraxial image plane */
    double centroid_y;               /* at the paraxial image plane */
    double transmitted;              /* mean weight T over launched rays */
    int32_t afocal;
    int32_t reserved;
    double angular_magnification;    /* afocal: D n0 / nk of the system matrix */
    double apparent_angle, angular_rms;  /* afocal: radians */
} owe_lens_result;

/* The eight sums a spot is computed from, per lens and wavelength, with B = d_xy / d_z of each
 * arrived ray and A = p_xy − B p_z (focused) or A = atan(B) per component (afocal): count, ΣAx, ΣAy,
 * ΣBx, ΣBy, Σ|A|², Σ|B|², ΣA·B; then ΣT. */
#define OWE_LENS_MOMENTS 9

/* Differentiable lens parameters: up to OWE_LENS_MAX_PARAMS (surface, field) pairs, the surface
 * counted within each lens, shared by every lens of the batch. A run then also returns, per lens,
 * the derivative of its spot metric — rms_best when focused, angular_rms when afocal — with respect
 * to each parameter, for the rays the run traced (forward-mode dual numbers through every surface;
 * the launch rays are held fixed). No counterpart in the reference, which has no lens derivatives. */
#define OWE_LENS_MAX_PARAMS 8
#define OWE_LENS_PARAM_CURVATURE 0
#define OWE_LENS_PARAM_CONIC 1
#define OWE_LENS_PARAM_THICKNESS 2  /* the gap after the surface: every later vertex moves */
#define OWE_LENS_PARAM_A4 3         /* A4, A6, A8, A10: 3, 4, 5, 6 */
typedef struct owe_lens_param {
    uint32_t surface;
    int32_t field;
} owe_lens_param;

typedef struct owe_hip_lens_batch owe_hip_lens_batch;

/* Validates and uploads the batch (media as in owe_hip_scene.h; surfaces must be within the conic's
 * domain at their semi-diameter, SagSurface's own check, geometry.cpp:72-77). */
int owe_hip_lens_batch_create(int device, const owe_medium* media, uint32_t n_media, const owe_lens_surface* surfaces,
                              uint32_t n_surfaces, const owe_lens* lenses, uint32_t n_lenses,
                              const owe_lens_batch_settings* settings, owe_hip_lens_batch** out);
int owe_hip_lens_batch_destroy(owe_hip_lens_batch* batch);
/* Traces every ray of every lens at every wavelength once, replacing the previous run's sums. */
int owe_hip_lens_batch_run(owe_hip_lens_batch* batch);
/* Results of the last run: per lens (may be null), the moments per lens and wavelength
 * (n_lenses × n_wavelengths × OWE_LENS_MOMENTS doubles; may be null) and the irradiance grids
 * (n_lenses × focus_planes × pixels × pixels floats, unit incident irradiance across the beam,
 * plane 0 nearest the lens; may be null). */
int owe_hip_lens_batch_results(owe_hip_lens_batch* batch, owe_lens_result* lenses, double* moments, float* irradiance);
/* The persistent launch's shape: resident blocks per compute unit and rays per work item (a
 * multiple of 64); 0 keeps the current value. */
int owe_hip_lens_batch_set_shape(owe_hip_lens_batch* batch, int blocks_per_cu, int rays_per_item);
/* Times runs under candidate shapes, keeps the fastest; best_ms (optional) receives its time. */
int owe_hip_lens_batch_autotune(owe_hip_lens_batch* batch, double* best_ms);
/* Probe for parity: the rays index[i] of lens `lens` at wavelength `wavelength` traced alone. out
 * receives n × 7 doubles: the last surface's point, the outgoing direction and the weight T; a ray
 * that does not arrive has T = 0 and NaN point and direction. */
int owe_hip_lens_batch_trace(owe_hip_lens_batch* batch, uint32_t lens, int wavelength, const uint64_t* index,
                             uint64_t n, double* out);
/* Milliseconds of the last run's persistent kernel (HIP events on the batch's stream). */
int owe_hip_lens_batch_last_ms(const owe_hip_lens_batch* batch, double* ms);
/* Sets the differentiable parameters (n ≤ OWE_LENS_MAX_PARAMS; 0 clears them). */
int owe_hip_lens_batch_set_params(owe_hip_lens_batch* batch, const owe_lens_param* params, uint32_t n);
/* The last run's derivatives: n_lenses × n_params doubles, d(metric)/dθ in SI units. */
int owe_hip_lens_batch_gradients(owe_hip_lens_batch* batch, double* grad);

#ifdef __cplusplus
}
#endif

#endif /* OWE_HIP_LENS_H */
