/* owe_hip — the C ABI of the MI355X (gfx950) transport backend of the Optical World Engine.
 *
 * Every function returns an int status (OWE_HIP_OK on success). On failure a description is kept
 * per thread and returned by owe_hip_last_error(); no C++ exception crosses this boundary. Handles
 * are opaque; each is created and destroyed by the pair of functions named for it. One host thread
 * drives a renderer or mesh at a time (as the reference's Renderer: passes are sequential).
 *
 * Reference interfaces each entry point stands in for (paths into the reference repository):
 *   owe::Backend            src/owe/render/backend.hpp:31-52
 *   owe::Renderer           src/owe/render/renderer.hpp:19-79
 *   owe::MeshShape::intersect, Bvh::traverse   src/owe/scene/mesh.cpp:73-97, src/owe/scene/bvh.hpp:31-78
 *   occludedMesh (any hit)  src/owe/backends/gpu/shaders/geometry.slang:577-615
 *   Tracer::radiance        src/owe/transport/transport.cpp:473-529 (camera paths, per pass)
 */
#ifndef OWE_HIP_H
#define OWE_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "owe_hip_scene.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OWE_HIP_OK 0
#define OWE_HIP_ERR_INVALID 1     /* bad argument or handle (std::invalid_argument-like)          */
#define OWE_HIP_ERR_UNAVAILABLE 2 /* no usable gfx950 device (Backend::available() == false)      */
#define OWE_HIP_ERR_DEVICE 3      /* a HIP runtime call failed                                     */
#define OWE_HIP_ERR_UNSUPPORTED 4 /* the settings ask for something the backend does not implement */

/* The last failure's description on this thread ("" when none). Valid until the next call. */
const char* owe_hip_last_error(void);

/* ------------------------------------------------------------------ backend identity
 * owe::Backend::name/summary/available/devices (backend.hpp:34-43). */

/* "hip". */
const char* owe_hip_backend_name(void);
/* One line: hardware, precision and role. */
const char* owe_hip_backend_summary(void);
/* The integrators a renderer of this backend implements, comma-separated: "path,light,sppm,vcm,bdpt,hybrid"
 * (owe::Backend::integrators, backend.hpp:41-42). */
const char* owe_hip_backend_integrators(void);
/* 1 when a gfx950 device is usable here, else 0 with the reason written to why[0..cap). */
int owe_hip_available(char* why, size_t cap);

typedef struct owe_hip_device_info {
    int index;
    char name[128];   /* marketing name, e.g. "AMD Instinct MI355X" */
    char arch[32];    /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    char type[16];    /* "discrete" (DeviceInfo::type, backend.hpp:23-29) */
    uint64_t hbm_bytes;
    int compute_units;
    int xcds;         /* accelerator complex dies (L2 domains) */
} owe_hip_device_info;

int owe_hip_device_count(int* count);
int owe_hip_device_info_get(int index, owe_hip_device_info* out);

/* ------------------------------------------------------------------ batch ray queries
 * A triangle mesh uploaded to one device with a wide (4-child) BVH built on the host, queried by
 * batches of rays that live in device memory. Stands in for MeshShape::intersect (closest hit) and
 * occludedMesh (any hit) over many rays at once.
 *
 * Ray record: 8 floats {o.x, o.y, o.z, tmin, d.x, d.y, d.z, tmax} (32 bytes); d need not be unit,
 * t is measured in units of |d|. A hit is a crossing with tmin < t < tmax.
 * Closest hit writes prim[i] (the input triangle index, 0xFFFFFFFF on a miss) and t[i] (+inf on a
 * miss). Any hit writes prim[i] = index of some crossed triangle or 0xFFFFFFFF, and t[i] = its t.
 */
typedef struct owe_hip_mesh owe_hip_mesh;

#define OWE_HIP_QUERY_CLOSEST 0
#define OWE_HIP_QUERY_ANY 1

typedef struct owe_hip_mesh_stats {
    uint64_t triangles;
    uint64_t nodes;        /* 4-wide nodes (128 bytes each) */
    uint32_t depth;        /* levels of 4-wide nodes */
    uint32_t stack_bound;  /* the most traversal-stack entries any ray can hold */
    double build_seconds;  /* host build time */
    uint64_t device_bytes; /* nodes + triangles resident on the device */
} owe_hip_mesh_stats;

/* positions: n_vertices × 3 floats; triangles: n_triangles × 3 vertex indices (host memory). */
int owe_hip_mesh_create(int device, const float* positions, uint32_t n_vertices, const uint32_t* triangles,
                        uint32_t n_triangles, owe_hip_mesh** out);
int owe_hip_mesh_destroy(owe_hip_mesh* mesh);
int owe_hip_mesh_stats_get(const owe_hip_mesh* mesh, owe_hip_mesh_stats* out);

/* Queries n rays. rays, prim and t are DEVICE pointers on the mesh's device; stream is a hipStream_t
 * (NULL: the default stream). The call is asynchronous on that stream. */
int owe_hip_trace_batch(owe_hip_mesh* mesh, int mode, const float* rays, uint64_t n, uint32_t* prim, float* t,
                        void* stream);

/* Launch form of owe_hip_trace_batch per query mode (both forms return the same hits bit for bit):
 * OWE_HIP_TRACE_PER_LANE — one lane per ray, a wave lives until its longest ray ends;
 * OWE_HIP_TRACE_REFILL (default) — a persistent grid of blocks_per_cu blocks per compute unit whose
 * waves each own a contiguous range of rays; lanes whose rays ended take new ones together once
 * `refill` lanes of the wave are idle. mode: OWE_HIP_QUERY_CLOSEST, OWE_HIP_QUERY_ANY, or −1 for
 * both; 0 keeps a shape value (blocks_per_cu ≤ 10, refill ≤ 64). No counterpart in the reference:
 * the batch query is this backend's own stand-in for many MeshShape::intersect calls. */
#define OWE_HIP_TRACE_PER_LANE 0
#define OWE_HIP_TRACE_REFILL 1
int owe_hip_mesh_set_launch(owe_hip_mesh* mesh, int mode, int form, int blocks_per_cu, int refill);
int owe_hip_mesh_get_launch(const owe_hip_mesh* mesh, int mode, int* form, int* blocks_per_cu, int* refill);

/* ------------------------------------------------------------------ camera-path renderer
 * owe::Backend::createRenderer (backend.hpp:45-46) and owe::Renderer (renderer.hpp:19-79) over the
 * flattened scene of owe_hip_scene.h. Each pass adds spp samples per pixel of the scene's observer
 * (Renderer::runPass → pass, renderer.hpp:40-46); resolve divides the film by the samples so far.
 * Streams are one PCG32 per (seed, pixel, sample, pass); the pixel's own stream (seed, pixel, pass)
 * gives its hero-wavelength rotation, as CpuRenderer::pass (cpu_backend.cpp:52-94) does per pixel.
 * The scene is copied at creation: the caller's arrays need not outlive the call. */
typedef struct owe_hip_renderer owe_hip_renderer;

/* Integrators: OWE_INTEGRATOR_PATH (camera paths, Tracer::radiance, transport.cpp:473-529),
 * OWE_INTEGRATOR_LIGHT (particles emitted from the lights and the sky, connected to an observer or
 * recorded by a surface sensor: the reference's "light" integrator, particle.slang:356-427, whose
 * passes emit W·H·spp particles, gpu_renderer_passes.cpp:303-311, and whose film is divided by
 * the particles emitted, gpu_renderer_film.cpp:17-20) and OWE_INTEGRATOR_SPPM (stochastic
 * progressive photon mapping, sppm.slang; a pass of spp runs spp iterations, each one camera path
 * per pixel and photons_per_pixel·W·H photons, GpuRenderer::passSppm, gpu_renderer_passes.cpp:259-290;
 * resolve gives L_direct / iterations + τ / (N_emitted π r²)), OWE_INTEGRATOR_VCM and
 * OWE_INTEGRATOR_BDPT (vertex connection and merging, vcm.slang, and the same without merging,
 * GpuRenderer::runVcmIterations, gpu_renderer_passes.cpp:136-199: a pass of spp runs spp iterations,
 * each photons_per_pixel·W·H light subpaths cached and splatted to the eye and one camera path per
 * pixel; resolve gives the camera paths' sum / iterations + the splats / (light subpaths ·
 * iterations; with scattering media the bidirectional paths take media as extinction and a camera
 * pass of spp paths per pixel adds the light that reached the eye through a medium scattering
 * vertex, gpu_renderer.cpp:43-47; a surface sensor is refused) and
 * OWE_INTEGRATOR_HYBRID (a pass of spp runs the camera paths, then W·H·spp particles, with the
 * partition of particle.slang:358-417 and transport.slang:745-906: light tracing owns the paths
 * eye → D → S⁺ → light, camera paths every other; resolve adds both films; observers only). Fails
 * with OWE_HIP_ERR_UNSUPPORTED for any other integrator (renderer.hpp:22-36). */
int owe_hip_renderer_create(int device, const owe_scene* scene, const owe_render_settings* settings,
                            owe_hip_renderer** out);
int owe_hip_renderer_destroy(owe_hip_renderer* renderer);
/* Renderer::runPass(spp): synchronous; spp ≤ 0 does nothing. */
int owe_hip_renderer_pass(owe_hip_renderer* renderer, int spp);
/* Renderer::resolve: width × height × 3 doubles of XYZ (row-major), film ÷ samples per pixel. */
int owe_hip_renderer_resolve(owe_hip_renderer* renderer, double* xyz);
/* The SPPM state per pixel (SppmPixel, sppm.slang:43-46): width × height × 5 doubles of τ (XYZ), N and
 * the gather radius r; the photons emitted per iteration and the iterations so far. Any output may be
 * null. Fails unless the renderer's integrator is OWE_INTEGRATOR_SPPM. */
int owe_hip_renderer_sppm_state(owe_hip_renderer* renderer, double* state, uint64_t* photons_per_iteration,
                                long long* iterations);
/* Renderer::stats: paths, segments, inconsistencies, leaks (transport_stats.hpp:12-31). */
int owe_hip_renderer_stats(const owe_hip_renderer* renderer, owe_transport_stats* out);
/* TransportStats::first (transport_stats.hpp:15-23): the first region inconsistency since creation or
 * reset — a ray reached `boundary` from `ray_region` where the boundary's side holds
 * `boundary_region`; `previous` is the boundary it left; p, origin and direction are in world
 * coordinates; `primitive` is the mesh triangle met (its input index); previous_primitive is not
 * tracked. With several kernels in flight "first" is whichever claimed it first (as in the reference's
 * GPU backend). All fields are 0xFFFFFFFF / 0 and *found is 0 when there was none. */
typedef struct owe_inconsistency {
    uint32_t boundary, ray_region, boundary_region, previous;
    double p[3], origin[3], direction[3];
    uint32_t previous_primitive, primitive;
} owe_inconsistency;
int owe_hip_renderer_first_inconsistency(const owe_hip_renderer* renderer, owe_inconsistency* out, int* found);
/* Renderer::passes() and samplesPerPixel() (renderer.hpp:48-49). */
int owe_hip_renderer_progress(const owe_hip_renderer* renderer, int* passes, long long* samples_per_pixel);
/* Renderer::resetProgress (renderer.hpp:63-69): clears the film, the pass count and the statistics. */
int owe_hip_renderer_reset(owe_hip_renderer* renderer);
/* The XYZ estimate of single camera paths (pixel[i], sample[i]) of a pass of spp samples, without
 * touching the film: the same kernels, for per-path comparison with the oracle. Host arrays; xyz
 * receives n × 3 doubles. */
int owe_hip_renderer_camera_paths(owe_hip_renderer* renderer, int spp, int pass, const uint32_t* pixel,
                                  const uint32_t* sample, uint64_t n, double* xyz);
/* The XYZ each particle index[i] of pass `pass` deposits, summed over its splats (not divided by
 * the particle count), without touching the film: the light integrator's kernel, for per-particle
 * comparison with the oracle. Host arrays; xyz receives n × 3 doubles. Light-integrator renderers only. */
int owe_hip_renderer_particles(owe_hip_renderer* renderer, int pass, const uint64_t* index, uint64_t n, double* xyz);
/* The two forms of the camera-path integrator: the wavefront (one kernel per stage over SoA queues,
 * the default) and the persistent batched form (a grid sized to the chip drains a device queue of
 * path ids, each lane tracing a path to its end in registers with the same device functions). The
 * film is integer, so both give the same image bit for bit. blocks_per_cu and chunk (path ids per
 * lane fetch) shape the persistent launch (0 keeps the current value); waves_per_simd (2 or 3; 0
 * keeps it) selects the persistent kernel compiled for that many resident waves per SIMD (3 trades
 * register spills for latency hiding); refill (1..64; 0 keeps it) is how many of a wave's lanes must
 * have finished their paths before the wave starts new ones together; use_graph replays each pass
 * from a captured HIP graph. No counterpart in the reference (whose GPU backend has one form).
 * A light-integrator renderer has the same two modes for its particle pass: one thread per particle
 * (OWE_HIP_MODE_WAVEFRONT, the reference's particle kernel's form, the default) or the lane-refill
 * form (OWE_HIP_MODE_PERSISTENT: a grid of blocks_per_cu blocks per compute unit, each lane advancing
 * its particle a segment at a time and taking chunk particle ids at a time, refill as above;
 * waves_per_simd and use_graph do not apply). Both give the same image bit for bit. */
#define OWE_HIP_MODE_WAVEFRONT 0
#define OWE_HIP_MODE_PERSISTENT 1
int owe_hip_renderer_set_mode(owe_hip_renderer* renderer, int mode, int blocks_per_cu, int chunk, int use_graph,
                              int waves_per_simd, int refill);
int owe_hip_renderer_get_mode(const owe_hip_renderer* renderer, int* mode, int* blocks_per_cu, int* chunk,
                              int* use_graph, int* waves_per_simd, int* refill);
/* Vertex connection and merging (vcm, bdpt): how many iterations of a pass run together, each in
 * its own slot of the light vertex cache, offsets and merging hash (1 to the number the renderer's
 * caches hold; 0: that number, chosen at creation so that a launch carries enough light subpaths to
 * fill the device, within a 16 GiB cache budget). The image does not depend on it: every iteration
 * fills its cache in subpath order and sums its merges in integers, so any number in flight gives
 * the same image bit for bit. No counterpart in the reference (whose GPU backend runs one iteration
 * after another, gpu_renderer_passes.cpp:136-199). */
int owe_hip_renderer_set_vcm_in_flight(owe_hip_renderer* renderer, int iterations);
int owe_hip_renderer_get_vcm_in_flight(const owe_hip_renderer* renderer, int* iterations, int* most);
/* Times passes of spp samples per pixel under candidate launch shapes of the current form
 * (persistent: waves per SIMD × refill threshold × blocks per compute unit × chunk; wavefront: bounce
 * rounds between host reads of the live counters × paths in flight, a quarter to all of the path
 * slots; a light-integrator renderer in the lane-refill form: refill threshold × blocks per compute
 * unit × chunk), times the five fastest three more times each and keeps the lowest median, and clears the
 * film and statistics again. Only on a renderer without passes. best_ms (optional) receives the
 * winning pass time. With a tuning cache set (below) a recorded shape is applied without timing. */
int owe_hip_renderer_autotune(owe_hip_renderer* renderer, int spp, double* best_ms);

/* The autotuner's per-device tuning cache: a text file of tuned shapes, one line per (device,
 * form, pass size, flattened scene, settings) key. owe_hip_renderer_autotune first looks its key
 * up there (a hit applies the recorded shape without sweeping and reports its recorded time) and
 * appends what a sweep chooses. path NULL or "" disables it; the default is the environment
 * variable OWE_HIP_TUNING_CACHE, else none. Process-wide. */
int owe_hip_set_tuning_cache(const char* path);
/* The wavefront form's tuned bounce rounds per host sync and paths in flight, and whether the last
 * autotune came from the cache. */
int owe_hip_renderer_tuning(const owe_hip_renderer* renderer, int* rounds_per_sync, uint64_t* queue_capacity,
                            int* from_cache);
/* Per-kernel timing with HIP events on the renderer's stream (the reference's OWE_GPU_PROFILE,
 * src/owe/backends/gpu/vulkan_context.cpp:350-422). Enabling clears the totals. */
int owe_hip_renderer_set_timing(owe_hip_renderer* renderer, int enable);
/* Milliseconds and launches per kernel since timing was enabled, in the order generate, intersect,
 * shade, shadow, splat, persistent (6 entries each); a light-integrator renderer counts its particle
 * kernel as shade; a vcm or bdpt renderer its light-subpath kernel as generate and its camera kernel
 * as shade (the medium-only pass's wavefront kernels under their own names). */
int owe_hip_renderer_kernel_times(const owe_hip_renderer* renderer, double* ms, uint64_t* launches);

/* ------------------------------------------------------------------ batches of scenes
 * Many variants of one scene family (perturbed instruments, tolerancing runs) rendered with camera
 * paths as one persistent launch per pass: the launch's waves drain the first scene's pass, then the
 * next one's, each scene's paths from its own id counter into its own film. Each member is a
 * renderer of its scene (created with the settings given, OWE_INTEGRATOR_PATH only), so every
 * member's film equals what its own persistent pass would give, bit for bit. The reference renders
 * one scene per Renderer (renderer.hpp:19-79); a caller there loops over renderers — this is the
 * batching form for that loop, with no counterpart in the reference's interface. */
typedef struct owe_hip_batch owe_hip_batch;
/* Flattens, builds and uploads the count scenes (a few host threads at once). */
int owe_hip_batch_create(int device, const owe_scene* scenes, int32_t count, const owe_render_settings* settings,
                         owe_hip_batch** out);
int owe_hip_batch_destroy(owe_hip_batch* batch);
/* Member index's renderer, owned by the batch (never destroy it): resolve, stats and progress read
 * what the batch's passes left; its launch shape (owe_hip_renderer_set_mode or _autotune on member
 * 0) is the batch's. */
int owe_hip_batch_member(owe_hip_batch* batch, int32_t index, owe_hip_renderer** out);
/* One pass of spp samples per pixel of every member, in one launch; synchronous. */
int owe_hip_batch_pass(owe_hip_batch* batch, int spp);
/* The members' statistics summed. */
int owe_hip_batch_stats(const owe_hip_batch* batch, owe_transport_stats* out);
/* The launch's time by HIP events on the batch's stream; enabling clears the total. */
int owe_hip_batch_set_timing(owe_hip_batch* batch, int enable);
int owe_hip_batch_kernel_time(const owe_hip_batch* batch, double* ms, uint64_t* launches);

#ifdef __cplusplus
}
#endif

#endif /* OWE_HIP_H */
