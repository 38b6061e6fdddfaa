// TEST INFRASTRUCTURE ONLY — the double-precision CPU oracle (see oracle_core.hpp).
//
// The world, lights, observer and camera-path transport restated from the reference:
//   spectra, blackbody normalisation, luminance     src/owe/core/spectrum.cpp:30-181
//   n(λ) models, relative-to-air indices            src/owe/core/medium.cpp:46-66
//   plane / sphere / mesh shapes, sampling          src/owe/scene/geometry.cpp:18-27, 391-477; mesh.cpp:12-113
//   light table (by power, sun share)               src/owe/scene/world.cpp:285-377
//   ideal observer (pinhole / pupil, pixel filter)  src/owe/scene/detector.cpp:36-91, detector.hpp:47-62
//   camera path: radiance, NEE/MIS, media, RR       src/owe/transport/transport.cpp:35-529
//   surface scattering                               src/owe/transport/scattering.cpp:65-222
//   per-pixel passes and XYZ film                   src/owe/backends/cpu/cpu_backend.cpp:52-107
//
// World::intersect here is the nearest crossing over all boundaries, without the reference's
// contact and tie machinery (world.cpp:440-550): scenes keep bodies apart. A spawned ray excludes
// only what it provably cannot cross again: its own plane, its own mesh face (coplanar adjacent
// triangles), or — on a sphere — the root at its origin (the other root is −2 o·d / d·d exactly).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../include/owe_hip_scene.h"
#include "oracle_bvh.hpp"
#include "oracle_shapes.hpp"

namespace oracle {

constexpr int NW = 4;
constexpr uint32_t kNone = 0xFFFFFFFEu;

struct Spec4 {
    double v[NW] = {0, 0, 0, 0};
    Spec4() = default;
    explicit Spec4(double x) {
        for (double& a : v) a = x;
    }
    double& operator[](int i) { return v[i]; }
    double operator[](int i) const { return v[i]; }
    Spec4& operator*=(const Spec4& o) {
        for (int i = 0; i < NW; ++i) v[i] *= o.v[i];
        return *this;
    }
    Spec4& operator*=(double s) {
        for (double& a : v) a *= s;
        return *this;
    }
    Spec4& operator+=(const Spec4& o) {
        for (int i = 0; i < NW; ++i) v[i] += o.v[i];
        return *this;
    }
    double maxValue() const {
        double m = v[0];
        for (double a : v) m = std::max(m, a);
        return m;
    }
    bool isZero() const {
        for (double a : v)
            if (a != 0) return false;
        return true;
    }
};
inline Spec4 operator*(Spec4 a, const Spec4& b) { return a *= b; }
inline Spec4 operator*(Spec4 a, double s) { return a *= s; }
inline Spec4 operator+(Spec4 a, const Spec4& b) { return a += b; }

struct Spectrum {
    int kind = OWE_SPEC_CONSTANT;
    double value = 0, kelvin = 0, scale = 1;
    std::vector<double> lam, val;
    double eval(double l) const;
    bool isZero() const { return kind == OWE_SPEC_CONSTANT && (value == 0 || scale == 0); }
    double luminance() const;  // ∫ s ȳ / ∫ ȳ
};

struct Wavelengths {
    double lambda[NW] = {0, 0, 0, 0};
    double pdf[NW] = {0, 0, 0, 0};
    static Wavelengths hero(double u);
    static Wavelengths uniform1(double u, double lo, double hi);
    bool secondaryTerminated() const { return pdf[1] == 0 && pdf[2] == 0 && pdf[3] == 0; }
    void terminateSecondary() {
        if (secondaryTerminated()) return;
        for (int i = 1; i < NW; ++i) pdf[i] = 0;
        pdf[0] /= NW;
    }
    Spec4 eval(const Spectrum& s) const {
        Spec4 r;
        for (int i = 0; i < NW; ++i) r[i] = s.eval(lambda[i]);
        return r;
    }
    void toXYZ(const Spec4& L, double xyz[3]) const;
};

struct IndexModel {
    int kind = OWE_INDEX_CONSTANT, terms = 0;
    bool relAir = false;
    double a[4] = {1, 0, 0, 0}, b[4] = {0, 0, 0, 0};
    double n(double lnm) const;
};

struct Medium {
    IndexModel index;
    Spectrum absorption, scattering;
    double g = 0;
    bool opaque = false;
    bool scatters() const { return !scattering.isZero(); }
    bool absorbs() const { return !absorption.isZero(); }
};

struct Optics {
    int type = OWE_SURF_DIELECTRIC;
    bool backAbsorbs = false;
    double roughness = 0;
    Spectrum reflectance, condN, condK;
    bool isDelta() const {
        return (type == OWE_SURF_DIELECTRIC || type == OWE_SURF_CONDUCTOR) ? roughness < 1e-3 : type == OWE_SURF_MIRROR;
    }
};

struct Emission {
    Spectrum radiance;
    bool front = true, back = false, nee = true;
};

struct Xf {
    double R[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    V3 t;
    V3 point(V3 p) const { return vector(p) + t; }
    V3 vector(V3 v) const {
        return {R[0][0] * v.x + R[0][1] * v.y + R[0][2] * v.z, R[1][0] * v.x + R[1][1] * v.y + R[1][2] * v.z,
                R[2][0] * v.x + R[2][1] * v.y + R[2][2] * v.z};
    }
    Xf inverse() const;
};

struct Boundary {
    int shape = OWE_SHAPE_PLANE, aperture = OWE_APERTURE_RECT;
    uint32_t front = 0, back = 0;
    int optics = 0, emission = -1;
    double a = 0, b = 0;
    Xf toWorld, toLocal;
    Box worldBox;
    std::unique_ptr<Mesh> mesh;
    std::vector<uint32_t> faceGroup;  // mesh: coplanar face id per triangle
    double z0 = 0, z1 = 0;            // cylinder: axial extent
    shapes::Sag sag;                  // sag surface
    std::shared_ptr<shapes::Wave> wave;
    double area() const;
    bool canSample() const { return area() > 0; }
};

struct Ray {
    V3 o, d;
    uint32_t departed = kNone;   // boundary the ray leaves
    uint32_t departedPrim = kNone;
};

struct Hit {
    double t = kInf;
    V3 p, n, pLocal, nLocal;
    uint32_t boundary = kNone, prim = 0;
};

struct Environment {
    bool hasSun = false;
    int sky = OWE_SKY_NONE;
    V3 sunDir{0, 0, 1};
    double sunAngularRadius = 0;
    Spectrum sunRadiance, skyRadiance;
    double sunNeeShare = 0;
    double skyAt(double l) const { return sky == OWE_SKY_UNIFORM ? skyRadiance.eval(l) : 0; }
    double sunAt(V3 dir, double l) const {
        return (hasSun && dot(dir, sunDir) >= std::cos(sunAngularRadius)) ? sunRadiance.eval(l) : 0;
    }
    double sunSolidAngle() const { return 2 * kPi * (1 - std::cos(sunAngularRadius)); }
};

struct Light {
    int kind = 0;  // 0 boundary, 1 sun
    uint32_t boundary = kNone;
    double power = 0;
};

// The detector: an ideal observer (detector.cpp:36-91) or a surface sensor (detector.cpp:100-149).
struct Observer {
    int kind = OWE_DETECTOR_OBSERVER;
    int width = 0, height = 0;
    uint32_t region = 0;
    V3 position, fwd, right, up;
    double tanX = 1, tanY = 1, pupil = 0, focus = kInf, sigma = 0.4;
    // surface sensor
    uint32_t boundary = kNone;
    double halfX = 0, halfY = 0;
    bool flipX = false, flipY = false, hasAim = false;
    V3 aimCenter, aimNormal{0, 0, 1}, normal;
    double aimRadius = 0, aimShare = 1;
    Xf toWorld;
    bool sensor() const { return kind == OWE_DETECTOR_SENSOR; }
    // Pixel filter offset (Detector::filterOffset, detector.hpp:57-62).
    void filterOffset(Pcg& rng, double& fx, double& fy) const;
    bool generate(double px, double py, Pcg& rng, Ray& ray, double& weight) const;
};

class World {
public:
    explicit World(const owe_scene& s);
    // The nearest crossing of `ray` before tmax, with contact resolution when the ray's region is
    // known (region != kNone): a spawned ray meets other boundaries from t > −tol (tol the contact
    // tolerance at its origin), and a crossing met from another region than the ray's, or a detector
    // crossing, gives way to a crossing at the same place (within tol of its point; anywhere ahead for
    // a contact at the ray's own origin) that is met from the ray's region and is not a detector.
    bool intersect(const Ray& ray, double tmax, Hit& hit, uint32_t region = kNone) const;
    // Contact tolerance at a point: 2^-21 (two to four float ulps) of its largest coordinate relative to
    // the rebase origin.
    double contactTol(V3 p) const;
    // The nearest crossing with t in (lo, hi) — the departed boundary's own crossings only for t > max(lo, 0)
    // by its own rules — skipping boundary `skipB` (for a mesh, only the face of triangle `skipPrim`).
    bool intersectRaw(const Ray& ray, double lo, double hi, Hit& hit, uint32_t skipB, uint32_t skipPrim) const;
    Ray spawn(const Hit& h, V3 dir) const { return Ray{h.p, dir, h.boundary, h.prim}; }
    const Medium& mediumOf(uint32_t region) const { return media[regionMedium[region]]; }
    double indexOf(uint32_t region, double l) const { return mediumOf(region).index.n(l); }
    size_t pickLight(double u) const;
    double sceneRadius() const;
    // A point on boundary b seen from ref (solid-angle pdf), and the pdf of a given point.
    bool sampleFrom(const Boundary& b, V3 refLocal, double u1, double u2, double u3, V3& p, V3& n, double& pdfW) const;
    double pdfFrom(const Boundary& b, V3 refLocal, V3 pLocal, V3 nLocal) const;
    // A uniform point of b in its own frame (Shape::sampleArea; geometry.slang:708-752).
    bool sampleArea(const Boundary& b, double u1, double u2, double u3, V3& p, V3& n, uint32_t& prim) const;

    std::vector<Medium> media;
    std::vector<uint32_t> regionMedium;
    std::vector<Optics> optics;
    std::vector<Emission> emissions;
    std::vector<Boundary> boundaries;
    Environment env;
    Observer observer;
    V3 rebase;  // the eye, or the sensor's centre
    // Two distinct boundaries' boxes meet within the contact tolerance: only then do rays resolve
    // contacts (a spawned ray's window below zero, and the region rule).
    bool contacts = false;
    void findContacts();
    Box bounds;
    std::vector<Light> lights;
    std::vector<double> lightPdf, lightCdf;
    std::vector<int> lightOfBoundary;
    // Particles (gpu_scene.cpp:753-758): the sky's share of the emitted power (environment power,
    // world.cpp:352-371, against the lights' total) and the scene's centre.
    double environmentPower = 0, totalLightPower = 0, skyShare = 0;
    V3 sceneCenter() const { return bounds.valid() ? (bounds.lo + bounds.hi) * 0.5 : V3(); }
};

struct Settings {
    uint64_t seed = 1;
    int maxDepth = 64, rrDepth = 6;
    int spectral = OWE_SPECTRAL_HERO4;
    double lambdaMin = 380, lambdaMax = 780;
    double fresnelFloor = 0;
    // SPPM (RenderSettings, scene/scene.hpp:27-35).
    double photonRadius = 2, photonAlpha = 2.0 / 3.0, photonsPerPixel = 1;
    int photonGroups = 4;
    // The hybrid integrator's partition (kFlagPartition, scene.slang:124): light tracing owns the
    // paths eye → D → S⁺ → light, camera paths every other; and camera paths that count only light
    // reaching the eye through a medium scattering vertex (kFlagMediumOnly: the path-traced part of
    // vcm and bdpt in scenes whose media scatter, gpu_renderer.cpp:43-47).
    bool partition = false, mediumOnly = false;
};

// SPPM photon guiding (particle.slang:149-183): a map of kRes² cells over the sun's emission disc's
// bounding square (coordinates in units of the disc's radius) counting the uniform share's photons
// that reached a visible point; once it has counts, kShare of sun photons are drawn from it and every
// sun photon is weighted by the mixture density. The CDF is stored as floats, as the GPU keeps it.
struct PhotonGuide {
    static constexpr int kRes = 128, kCells = kRes * kRes;
    static constexpr double kShare = 0.75;
    std::vector<uint64_t> counts = std::vector<uint64_t>(kCells, 0);
    std::vector<float> cdf = std::vector<float>(kCells, 0.0f);
    double total = 0;
    void build();                              // the CDF of the counts so far
    uint32_t sample(float u) const;            // the first cell whose CDF exceeds u
    double pdf(double qx, double qy, uint32_t& cell) const;  // density per unit-disc area; cell of q or kNone
};

// An SPPM visible point (sppm.slang:33-40): the first non-specular surface a camera path reaches
// through specular events only.
struct VisiblePoint {
    bool valid = false;
    bool terminated = false;  // the path's secondary wavelengths ended on the way
    bool free = false;        // seen directly: no scattering and no absorbing medium on the way
    uint32_t group = 0, boundary = 0;
    V3 p, n, d;               // position, the boundary's geometric normal, arriving direction
    Spec4 beta;               // throughput arriving, with the camera weight
    double pathLength = 0;    // unfolded length from the eye
};

struct Stats {
    uint64_t paths = 0, segments = 0, inconsistencies = 0, leaks = 0;
    void add(const Stats& o) {
        paths += o.paths;
        segments += o.segments;
        inconsistencies += o.inconsistencies;
        leaks += o.leaks;
    }
};

class Tracer {
public:
    Tracer(const World& w, const Settings& s) : w_(w), s_(s) {}
    // Tracer::radiance (transport.cpp:473-529).
    Spec4 radiance(Ray ray, uint32_t region, Wavelengths& wl, Pcg& rng, Stats& st, VisiblePoint* vp = nullptr) const;
    Spec4 transmittance(V3 p, const Ray& start, uint32_t region, V3 dir, double dist, const Wavelengths& wl,
                        uint32_t* finalRegion, uint32_t excludeBoundary) const;
    // One camera sample of pixel (x, y): the draws of cpu_backend.cpp:69-84 from `rng`.
    bool cameraSample(int x, int y, int s, int spp, double rot, Pcg& rng, Stats& st, double xyz[3]) const;

    // Where a particle's contributions go: pixel (px, py) receives XYZ.
    struct Splat {
        virtual void add(int px, int py, const double xyz[3]) = 0;
        virtual ~Splat() = default;
    };
    // One particle of a light-tracing pass (particle_kernel.slang:11-19, particle.slang:356-427):
    // its wavelengths from the first uniform of `rng`, then emission, walk and splats.
    void particle(Pcg& rng, Stats& st, Splat& out) const { particleWalk(nullptr, rng, st, out, nullptr, nullptr); }

    // SPPM (sppm.slang): a photon arriving at a non-specular surface after at least one scattering event.
    struct PhotonSink {
        virtual void gather(const Hit& hit, V3 d, const Spec4& beta, const Wavelengths& wl) = 0;
        bool useful = false;  // the current photon reached a visible point
        virtual ~PhotonSink() = default;
    };
    // The camera pass's path of pixel (x, y) at the iteration's wavelengths `wl` (sppm_camera_kernel):
    // its own radiance (emission met through specular events, direct light at the visible point) in
    // xyz, and the visible point, if any, in vp. False when the sample measures nothing.
    bool sppmCamera(int x, int y, Wavelengths wl, Pcg& rng, Stats& st, double xyz[3], VisiblePoint& vp) const;
    // One photon at the given wavelengths (traceSppmPhoton, sppm.slang:150-201), its sun emission
    // guided by `guide` when given: the guide cell the photon teaches when it proved useful (the sink
    // set `useful`), else kNone.
    uint32_t sppmPhoton(Wavelengths wl, Pcg& rng, Stats& st, PhotonSink& sink, const PhotonGuide* guide) const;
    // A visible point's XYZ for a photon arriving along d with throughput beta at wavelengths wl, once
    // the point has passed the reach, disc, orientation and group tests (gatherPhoton's per-point
    // value, sppm.slang:117-136): β_vp · f_vp(wo, wi) · β_photon, a bound point's share scaled by the
    // groups. False when it is zero.
    bool sppmValue(const VisiblePoint& vp, int groups, V3 d, const Spec4& beta, const Wavelengths& wl,
                   double xyz[3]) const;

    // Vertex connection and merging (the vcm and bdpt integrators, vcm.slang): a light vertex of the
    // iteration's cache, and an iteration's parameters.
    struct VcmVertex {
        V3 p, n, d;  // position, the boundary's geometric normal, the direction the light arrived along
        Spec4 beta;
        double dVCM = 0, dVC = 0, dVM = 0;
        uint32_t boundary = kNone;
        bool terminated = false;  // the subpath left only its hero wavelength
    };
    struct VcmParams {
        double lightPaths = 1;     // N_L, light subpaths per group
        double mergePerMetre = 0;  // merging radius per metre from the eye (0: bdpt)
        uint32_t groups = 1, iteration = 0;
    };
    // The group's cached vertices ordered by x (their x and index in the group), for merging: the
    // vertices within r of p have |x − p.x| ≤ r. Without it merging visits every vertex.
    struct VcmSweep {
        const double* x;
        const uint32_t* index;
    };
    // Light subpath of group `group` (traceLightVcm, vcm.slang:246-292): its non-specular vertices
    // appended to `cache` in order, its connections to the eye (light tracing) to `out`.
    void vcmLight(Wavelengths wl, Pcg& rng, Stats& st, Splat& out, const VcmParams& p, uint32_t group,
                  std::vector<VcmVertex>& cache) const;
    // The camera path of pixel (x, y) (traceCameraVcm, vcm.slang:523-593) against its group's M cached
    // vertices: its XYZ from every strategy but light tracing. False when the sample measures nothing.
    bool vcmCamera(int x, int y, Wavelengths wl, Pcg& rng, Stats& st, const VcmParams& p, const VcmVertex* cache,
                   uint32_t M, double xyz[3], const VcmSweep& sweep = VcmSweep{nullptr, nullptr}) const;

private:
    struct Link;
    struct Emitted;
    double vcmEta(const VcmParams& p, V3 x) const;
    bool linkObserver(V3 x, Pcg& rng, Link& k) const;
    void deposit(Splat& out, int px, int py, const Spec4& v, const Wavelengths& wl) const;
    bool emitParticle(Wavelengths& wl, Pcg& rng, bool connect, Emitted& e, Splat& out,
                      const PhotonGuide* guide = nullptr) const;
    void connectVertex(const Hit& hit, bool surface, const Optics* o, V3 d, double g, uint32_t region,
                       const Spec4& beta, const Wavelengths& wl, Pcg& rng, Splat& out) const;
    struct Path;
    uint32_t particleWalk(const Wavelengths* fixed, Pcg& rng, Stats& st, Splat& out, PhotonSink* sink,
                          const PhotonGuide* guide) const;
    Spec4 directLighting(const Hit& hit, bool surface, const Optics* o, double nFront, double nBack, V3 d,
                         const Wavelengths& wl, Pcg& rng, double g, uint32_t region) const;
    bool scatterAtSurface(const Hit& hit, const Optics& o, uint32_t other, int depth, Wavelengths& wl, Pcg& rng,
                          Path& path) const;
    const World& w_;
    Settings s_;
};

// Stream layouts of a pass: the reference's one stream per (pixel, pass) drawn by every sample in
// turn, or one stream per (pixel, sample, pass) so samples run independently (the hip backend's).
constexpr int kLayoutPixel = 0;
constexpr int kLayoutSample = 1;

}  // namespace oracle
