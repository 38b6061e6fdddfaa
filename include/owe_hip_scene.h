/* owe_hip_scene.h — the flattened scene the hip backend renders: plain structs, plain pointers.
 *
 * It carries what the reference's World, Scene and RenderSettings hold for transport
 * (src/owe/scene/world.hpp:23-344, src/owe/scene/scene.hpp:15-49, src/owe/scene/detector.hpp:44-116):
 * media with n(λ), absorption, scattering and HG g; regions; surface optics; emission; boundaries
 * (a shape in its own frame with a rigid frame→world transform and the regions on its two sides);
 * the environment (sun, uniform sky); one ideal observer. A reference-side adapter fills it from a
 * built owe::Scene; the Python tooling fills it directly. All lengths in metres,
 * wavelengths in nanometres, all values IEEE-754 double (the backend rebases and rounds to float).
 */
#ifndef OWE_HIP_SCENE_H
#define OWE_HIP_SCENE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Spectrum kinds (owe::Spectrum::Kind, core/spectrum.hpp:28). Tabulated: `count` (λ, value) pairs
 * interleaved at table[offset ..], clamped outside, linear inside (spectrum.cpp:135-143). */
#define OWE_SPEC_CONSTANT 0
#define OWE_SPEC_BLACKBODY 1
#define OWE_SPEC_TABULATED 2

typedef struct owe_spectrum {
    int32_t kind;
    int32_t count;
    int32_t offset;
    int32_t reserved;
    double value;  /* constant: the value */
    double kelvin; /* blackbody temperature */
    double scale;  /* multiplies every kind (blackbody: luminance normalisation) */
    double reserved2;
} owe_spectrum;

/* Refractive-index models (owe::IndexModel::Kind, core/medium.hpp:12-38). */
#define OWE_INDEX_CONSTANT 0
#define OWE_INDEX_CAUCHY 1    /* n = A + B/λ² + C/λ⁴, λ in µm: coeff_a[0..2] */
#define OWE_INDEX_SELLMEIER 2 /* n² = 1 + Σ B λ²/(λ² − C): B in coeff_a, C (µm²) in coeff_b */
#define OWE_INDEX_CIDDOR 3    /* standard dry air */

typedef struct owe_medium {
    int32_t index_kind;
    int32_t terms;   /* Sellmeier terms (≤ 4) */
    int32_t rel_air; /* coefficients are relative to air: n = n_rel · n_air(λ) */
    int32_t opaque;
    double coeff_a[4];
    double coeff_b[4];
    owe_spectrum absorption; /* σa(λ) in 1/m */
    owe_spectrum scattering; /* σs(λ) in 1/m */
    double g;                /* Henyey–Greenstein anisotropy */
    double reserved;
} owe_medium;

/* Surface types (owe::SurfaceType, scene/world.hpp:61-70; stained glass is not supported). */
#define OWE_SURF_NULL 0
#define OWE_SURF_DIELECTRIC 1
#define OWE_SURF_DIFFUSE 2
#define OWE_SURF_CONDUCTOR 3
#define OWE_SURF_MIRROR 4
#define OWE_SURF_ABSORBER 5
#define OWE_SURF_DETECTOR 6

typedef struct owe_optics {
    int32_t type;
    int32_t back_absorbs;
    double roughness;          /* GGX α; below 1e-3 the surface is smooth (a delta) */
    owe_spectrum reflectance;  /* diffuse albedo, mirror reflectance, conductor tint */
    owe_spectrum conductor_n;
    owe_spectrum conductor_k;
} owe_optics;

typedef struct owe_emission {
    owe_spectrum radiance; /* Lambertian spectral radiance */
    int32_t front;
    int32_t back;
    int32_t nee;           /* sampled by next-event estimation */
    int32_t reserved;
} owe_emission;

/* Shapes, each in its own frame (scene/geometry.hpp). */
#define OWE_SHAPE_MESH 0     /* triangles [first_triangle, first_triangle + triangle_count) */
#define OWE_SHAPE_PLANE 1    /* z = 0 inside an aperture, normal +z */
#define OWE_SHAPE_SPHERE 2   /* radius a about the origin, normal outward */
#define OWE_SHAPE_CYLINDER 3 /* open cylinder x² + y² = a², params[0] ≤ z ≤ params[1], normal outward
                                (CylinderShape, geometry.hpp:191-209) */
#define OWE_SHAPE_SAG 4      /* z = sag(r) for b ≤ r ≤ a, vertex at the origin, normal +z (SagSurface,
                                geometry.hpp:46-82): params = curvature 1/R, conic k, A4, A6, A8, A10 */
#define OWE_SHAPE_WAVE 5     /* z = h(x, y) over |x| ≤ a, |y| ≤ b, normal +z (WaveSurface,
                                geometry.hpp:90-142): the plane waves [first_wave, first_wave +
                                wave_count) of the scene tapered to zero within params[0] of the
                                edges; params[1] is the node spacing (0: the analytic sum; < 0:
                                automatic, the sum up to six waves, else a bicubic Hermite spline
                                sampled every twelfth of the shortest wavelength) */

#define OWE_APERTURE_DISK 0    /* a = outer radius, b = inner radius */
#define OWE_APERTURE_RECT 1    /* a, b = half sizes along x, y */
#define OWE_APERTURE_ELLIPSE 2 /* a, b = semi-axes */

#define OWE_REGION_NONE 0xFFFFFFFEu

typedef struct owe_boundary {
    int32_t shape;
    int32_t aperture;
    uint32_t front;   /* region on the +normal side */
    uint32_t back;
    int32_t optics;
    int32_t emission; /* -1: none; cylinders, sags and waves do not emit */
    uint32_t first_triangle;
    uint32_t triangle_count;
    double a, b;
    double to_world[12]; /* row-major [R | t]: x_world = R x_local + t (rigid) */
    double params[6];    /* per shape, see OWE_SHAPE_* */
    uint32_t first_wave; /* OWE_SHAPE_WAVE: its records in owe_scene.waves */
    uint32_t wave_count;
} owe_boundary;

/* One plane wave of a liquid surface (PlaneWave, geometry.hpp:84-89): amplitude (m), wave vector
 * (rad/m), phase (rad). */
typedef struct owe_plane_wave {
    double amplitude, kx, ky, phase;
} owe_plane_wave;

#define OWE_SKY_NONE 0
#define OWE_SKY_UNIFORM 1

typedef struct owe_environment {
    int32_t has_sun;
    int32_t sky;
    int32_t reserved[2];
    double sun_dir[3];         /* unit vector toward the sun */
    double sun_angular_radius; /* radians */
    owe_spectrum sun_radiance;
    owe_spectrum sky_radiance;
    double sun_nee_share;      /* fraction of light selections given to the sun (0: by power) */
    double reserved2;
} owe_environment;

#define OWE_DETECTOR_OBSERVER 0
#define OWE_DETECTOR_SENSOR 1

/* The detector a renderer measures with (owe::Detector, scene/detector.hpp:42-72):
 *  - OWE_DETECTOR_OBSERVER: an ideal observer (IdealObserver, detector.hpp:84-116), pinhole when
 *    pupil_radius is 0; it measures radiance.
 *  - OWE_DETECTOR_SENSOR: a rectangular surface sensor (SurfaceSensor, detector.hpp:125-160) on
 *    boundary sensor_boundary (a plane with DETECTOR or DIFFUSE optics), pixels over
 *    [-half_x, half_x] x [-half_y, half_y] of its frame, measuring irradiance on its front face.
 *    Camera paths leave it cosine-distributed, or toward the aim disk with probability aim_share
 *    when has_aim is set; region is ignored (the boundary's front region is used). */
typedef struct owe_observer {
    int32_t width;
    int32_t height;
    uint32_t region;  /* region containing the eye */
    int32_t kind;     /* OWE_DETECTOR_OBSERVER or OWE_DETECTOR_SENSOR */
    double position[3];
    double look_at[3];
    double up[3];
    double fov_y;          /* radians */
    double pupil_radius;
    double focus_distance; /* +inf: focus at infinity */
    double pixel_sigma;    /* Gaussian pixel filter in pixels (0.4 for observers, 0 for sensors by default) */
    /* surface sensor only */
    uint32_t sensor_boundary;
    int32_t flip_x, flip_y; /* readout: which physical corner is pixel (0, 0) */
    int32_t has_aim;
    double half_x, half_y;
    double aim_center[3], aim_normal[3];
    double aim_radius, aim_share;
} owe_observer;

typedef struct owe_scene {
    uint32_t n_media, n_regions, n_optics, n_emissions, n_boundaries, n_vertices, n_triangles, n_table;
    const owe_medium* media;
    const uint32_t* region_medium; /* region → medium; region 0 is the ambient region */
    const owe_optics* optics;
    const owe_emission* emissions;
    const owe_boundary* boundaries;
    const double* vertices;     /* n_vertices × 3, each in its boundary's own frame */
    const uint32_t* triangles;  /* n_triangles × 3 vertex indices */
    const double* table;        /* tabulated spectra */
    owe_environment env;
    owe_observer observer;
    const owe_plane_wave* waves; /* n_waves records for OWE_SHAPE_WAVE boundaries */
    uint32_t n_waves;
    uint32_t reserved;
} owe_scene;

#define OWE_INTEGRATOR_PATH 0
#define OWE_INTEGRATOR_LIGHT 1
#define OWE_INTEGRATOR_SPPM 2
#define OWE_INTEGRATOR_VCM 3
#define OWE_INTEGRATOR_BDPT 4
#define OWE_INTEGRATOR_HYBRID 5

/* Wavelength sampling: the reference's four stratified hero wavelengths importance-sampled over
 * 360–830 nm (core/wavelengths.hpp:36-45), or one wavelength per path uniform over
 * [lambda_min, lambda_max] whose estimate stands for all four strata (pdf/4, as terminateSecondary). */
#define OWE_SPECTRAL_HERO4 0
#define OWE_SPECTRAL_UNIFORM1 1

typedef struct owe_render_settings {
    uint64_t seed;
    int32_t max_depth;
    int32_t rr_depth;
    int32_t integrator;
    int32_t spectral;
    double lambda_min, lambda_max;
    double fresnel_floor;
    /* Stochastic progressive photon mapping (RenderSettings, scene/scene.hpp:27-35): the initial
     * gather radius in pixel footprints, the radius-reduction alpha in (0, 1], photons per pixel and
     * iteration, and wavelength groups per iteration (1–255). A zero selects the reference's
     * default (2, 2/3, 1, 4). Vertex connection and merging reads the same fields as its merging
     * radius, its shrink, its light subpaths per pixel and iteration and its wavelength groups
     * (1–8, default 1: vcmWavelengthGroups, scene.hpp:38); bdpt ignores the radius. The path and
     * light integrators ignore them. */
    double photon_radius, photon_alpha, photons_per_pixel;
    int32_t photon_groups;
    int32_t reserved;
} owe_render_settings;

/* TransportStats (scene/transport_stats.hpp:12-31). */
typedef struct owe_transport_stats {
    uint64_t paths, segments, inconsistencies, leaks;
} owe_transport_stats;

#ifdef __cplusplus
}
#endif

#endif /* OWE_HIP_SCENE_H */
