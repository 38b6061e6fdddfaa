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
 */
#ifndef OWE_HIP_H
#define OWE_HIP_H

#include <stddef.h>
#include <stdint.h>

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

#ifdef __cplusplus
}
#endif

#endif /* OWE_HIP_H */
