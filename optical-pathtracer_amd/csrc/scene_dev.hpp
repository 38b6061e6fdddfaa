// The device-resident scene of the transport kernels: the flattened owe_scene rebased to the
// observer (camera-relative) and rounded to float32, as the reference's float kernels hold it
// (src/owe/backends/gpu/gpu_scene.hpp:1-10: "camera-relative float records").
//
// Layout in HBM: small tables (media, optics, emission, analytic boundaries, lights, spectra) live in
// one read-only block of plain structs; every triangle of every mesh boundary is transformed to
// world space once on the host and goes into one 4-wide BVH (bvh4.hpp) whose records carry the
// owning boundary and the coplanar face id. Analytic shapes (planes, spheres) are few and are tested
// in a loop after the triangle traversal.
#pragma once

#include <cstdint>
#include <vector>

#include "bvh4.hpp"
#include "owe_hip_scene.h"

namespace owehip {

constexpr int NW = 4;
constexpr uint32_t kRegionNone = 0xFFFFFFFEu;
constexpr uint32_t kNoBoundary = 0xFFFFFFFEu;

struct DSpec {
    int32_t kind, count, offset, pad;  // tabulated: (λ, v) pairs at table[offset ..]
    float value, kelvin, scale, pad2;
};

struct DMedium {
    int32_t indexKind, terms, relAir, flags;  // flags: 1 absorbs, 2 scatters
    float a[4], b[4];
    DSpec absorption, scattering;
    float g, pad[3];
};

struct DOptics {
    int32_t type, backAbsorbs, delta, pad;
    float roughness, pad2[3];
    DSpec reflectance, condN, condK;
};

struct DEmission {
    DSpec radiance;
    int32_t front, back, pad[2];
};

// Sag surface (SagSurface) and liquid surface (WaveSurface) parameters; the device code is in
// shapes_dev.hpp.
struct DSag {
    double c, ck;                 // curvature, c (1 + k): the quadric, in double for the polish
    float cf, kf, ckf;            // the same in float
    float rMax2, rMin2, rMax;     // the annulus
    float zMin, zMax;             // the numeric path's slab
    float A[4];                   // even asphere A4..A10
    int32_t nA;                   // 0: quadric path
    int32_t polish64;             // a quadric root is polished in double (else in float)
};

struct DWave {
    float hx, hy, margin;
    float amp, slope;             // |h| ≤ amp, |∇h| ≤ slope
    float dx, dy;                 // sampled node spacing
    uint32_t data;                // first record in the scene's wave data
    uint32_t nx, ny;              // nx = 0: analytic, ny waves; else nx × ny nodes
};

struct DBoundary {
    int32_t shape, aperture;
    uint32_t front, back;
    int32_t optics, emission;
    float a, b;
    float toLocal[12];  // world (camera-relative) → shape frame, row-major [R | t]
    float toWorld[12];
    float lo[3], area;
    float hi[3];
    int32_t light;  // index into the light table, -1 when next-event estimation does not sample it
    float z0, z1;   // cylinder: axial extent
    float depTol;   // numeric sag, wave: a departing ray's own root lies within this of its origin
    int32_t pad;
    DSag sag;
    DWave wave;
};

struct DLight {
    int32_t kind;  // 0 boundary, 1 sun
    uint32_t boundary;
    float pdf, cdf;
};

struct DEnv {
    int32_t hasSun, sky, nLights, pad;
    float sunDir[3], sunOneMinusCos;  // 1 − cos(angular radius), computed in double
    float sunSolidAngle, sunPdf, pad2[2];
    DSpec sunRadiance, skyRadiance;
    // Particles: the sky's share of the emitted power, the radius of the disc distant light leaves
    // from (1.001 × the scene radius) and the scene's centre (gpu_scene.cpp:750-758).
    float skyShare, emitRadius, pad3[2];
    float center[3], pad4;
};

// The detector (owe_observer): an ideal observer at the camera-relative origin, or a surface sensor
// on plane boundary `sensor` (whose centre is then the origin).
struct DObserver {
    int32_t width, height;
    uint32_t region;
    int32_t kind;             // OWE_DETECTOR_OBSERVER or OWE_DETECTOR_SENSOR
    float position[3], tanX;  // position is the origin of the camera-relative frame: (0, 0, 0)
    float fwd[3], tanY;
    float right[3], pupil;
    float up[3], focus;
    float sigma, halfX, halfY, aimShare;
    uint32_t sensor;
    int32_t flipX, flipY, hasAim;
    float normal[3], aimRadius;
    float aimCenter[3], pad;
    float aimNormal[3], pad2;
    float toWorld[12];        // the sensor's frame → camera-relative world
};

// What a kernel needs to read the scene: device pointers and counts.
struct DScene {
    const DMedium* media;
    const uint32_t* regionMedium;
    const DOptics* optics;
    const DEmission* emissions;
    const DBoundary* boundaries;
    const DLight* lights;
    const float* table;
    const float4* waveData;  // wave records (OWE_SHAPE_WAVE)
    const WideNode* nodes;  // null when the scene has no triangles
    const TriRecord* tris;
    uint32_t nBoundaries, nLights, nRegions, stackCap;
    uint32_t contacts;  // two distinct boundaries touch or overlap: rays resolve contacts (intersectContact)
    uint32_t* firstInconsistency;  // the first region inconsistency's record (kFirstWords words), or null
    DEnv env;
    DObserver obs;
};

// Host side: the flattened, rebased, float32 scene and its upload.
struct HostScene {
    std::vector<DMedium> media;
    std::vector<uint32_t> regionMedium;
    std::vector<DOptics> optics;
    std::vector<DEmission> emissions;
    std::vector<DBoundary> boundaries;
    std::vector<DLight> lights;
    std::vector<float> table;
    std::vector<float> waveData;  // 4 floats per record
    Bvh4 bvh;
    bool hasTriangles = false;
    bool contacts = false;  // some pair of distinct boundaries' boxes meet within the contact tolerance
    DEnv env{};
    DObserver obs{};
    double origin[3] = {0, 0, 0};  // world position of the camera-relative frame's origin
    double emittedPower = 0;       // sky + lights: particles carry something when positive
};

// Validates and flattens an owe_scene (throws owehip::Failure on invalid input).
HostScene flattenScene(const owe_scene& s);

// SagSurface's construction (src/owe/scene/geometry.cpp:69-89): trailing zero asphere terms
// dropped, the slab [zMin, zMax] from 513 samples of the annulus padded by 1e-9 + 1e-3 of its
// height.
DSag makeSag(double curvature, double conic, const double asphere[4], double rMax, double rMin);

}  // namespace owehip
