// The wavefront camera-path renderer: owe_hip_renderer_* of the C ABI.
//
// One pass of spp samples per pixel is a batch of W·H·spp independent paths (one PCG32 stream per
// (pixel, sample, pass); the pixel's own stream gives its hero-wavelength rotation, as
// CpuRenderer::pass, src/owe/backends/cpu/cpu_backend.cpp:52-94). Paths advance one segment per
// round of four kernels over structure-of-arrays state in HBM:
//
//   generate  camera rays, wavelengths, streams                 (detector.cpp:61-91, cpu_backend.cpp:69-81)
//   intersect nearest crossing per live path                    (World::intersect, world.cpp:553-582)
//   shade     media, emission met, sky/sun met, NEE set-up,     (Tracer::radiance body, transport.cpp:481-526)
//             BSDF sample, Russian roulette; survivors are
//             compacted into the next queue by a wave ballot
//   shadow    transmittance of each pending NEE ray             (transport.cpp:237-348)
//
// and a final splat converts each path's spectral estimate to XYZ with its final wavelengths
// (Wavelengths::toXYZ, core/wavelengths.hpp:71-83) and adds it to the film with 64-bit integer
// atomics in 32.32 fixed point, so the film is independent of the order paths finish.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "common.hpp"
#include "particle.hpp"
#include "shading_dev.hpp"
#include "sppm.hpp"
#include "vcm.hpp"

using namespace owehip;

namespace owehip {
namespace {

constexpr int kBlock = 256;
constexpr double kFilmScale = 4294967296.0;  // 2^32 units per unit of XYZ

enum : uint32_t {
    kFlagSpecular = 1u,
    kFlagTerminated = 2u,
    // The hybrid partition's record of a camera path's scattering events (noteScatter,
    // transport.slang:783-791): any event yet, the first one non-delta, a delta event since the last
    // non-delta one, the non-delta events (two bits, saturating at 2); and a medium scattering vertex
    // met (the medium-only pass).
    kFlagEvents = 4u,
    kFlagFirstNonDelta = 8u,
    kFlagDeltaSince = 16u,
    kFlagMediumScattered = 32u,
    kNonDeltaShift = 6,
    kNonDeltaMask = 3u << kNonDeltaShift,
};

__device__ __forceinline__ void noteScatter(uint32_t& flags, bool delta) {
    if (delta) {
        flags |= kFlagDeltaSince;
    } else {
        if (!(flags & kFlagEvents)) flags |= kFlagFirstNonDelta;
        const uint32_t n = min(((flags & kNonDeltaMask) >> kNonDeltaShift) + 1u, 2u);
        flags = (flags & ~(kNonDeltaMask | kFlagDeltaSince)) | (n << kNonDeltaShift);
    }
    flags |= kFlagEvents;
}

// stepPath's ownedByLightTracing (transport.slang:890-891): one non-delta event, the first, then at
// least one delta event — light tracing's path eye → D → S⁺ → light.
__device__ __forceinline__ bool ownedByLightTracing(uint32_t flags) {
    return ((flags & kNonDeltaMask) >> kNonDeltaShift) == 1u && (flags & kFlagFirstNonDelta) && (flags & kFlagDeltaSince);
}

// Structure-of-arrays path state, one slot per path of the current chunk.
struct Paths {
    float4* ray0;     // o.xyz, prevPdf
    float4* ray1;     // d.xyz, camera weight
    float4* beta;     // throughput per wavelength
    float4* L;        // radiance estimate per wavelength
    float4* lam;      // wavelengths
    float4* prevP;    // last scattering point (MIS of emitters met)
    uint4* info;      // pixel, depth | flags << 16, region, departed boundary
    uint32_t* face;   // departed mesh face (~0: none)
    uint2* rng;       // PCG32 state (the increment is the pass's)
    uint4* hit;       // t bits, boundary, face, triangle slot
    float4* shPend;   // pending NEE contribution (β·f·Le·cos·w/pdf), zero when none
    float4* shO;      // shadow ray origin, distance (inf: the sun)
    float4* shD;      // direction, unused
    uint4* shInfo;    // start region, light's region, excluded boundary, departed boundary
    uint32_t* shFace; // departed face
    uint32_t* queue[2];
    uint32_t* count;  // [0], [1] queue lengths; [2] inconsistencies; [3] leaks
    unsigned long long* segments;  // segments traced in the current chunk (sum of live counts)
};

// Whether wavelength i of the path still carries an estimate: the hero always; a secondary until a
// dispersive event terminates it (never, with one wavelength per path). Spectra of wavelengths that
// carry nothing are not evaluated (their throughput is zero).
__device__ __forceinline__ bool carried(const WL& w, int i) { return i == 0 || w.pdf[i] != 0.0f; }

__device__ __forceinline__ WL loadWL(const float4 l, uint32_t flags, const Settings& s) {
    WL w;
    w.lambda[0] = l.x;
    w.lambda[1] = l.y;
    w.lambda[2] = l.z;
    w.lambda[3] = l.w;
    if (s.spectral == OWE_SPECTRAL_UNIFORM1) {
        w.pdf[0] = 1.0f / (s.lambdaMax - s.lambdaMin) / NW;
        w.pdf[1] = w.pdf[2] = w.pdf[3] = 0.0f;
    } else if (flags & kFlagTerminated) {
        w.pdf[0] = visiblePdfF(l.x) / NW;
        w.pdf[1] = w.pdf[2] = w.pdf[3] = 0.0f;
    } else {
#pragma unroll
        for (int i = 0; i < NW; ++i) w.pdf[i] = visiblePdfF(w.lambda[i]);
    }
    return w;
}


// ---------------------------------------------------------------- generate

// A camera path's first state (CpuRenderer::pass per sample, cpu_backend.cpp:66-81, and
// IdealObserver/SurfaceSensor::generate, detector.cpp:61-149), shared by the wavefront and the
// persistent kernels so both trace the same float operations.
struct PathStart {
    f3 o, d;
    float weight;
    float4 lam, beta;
    uint32_t region, departed;
    uint64_t rngState;
};

__device__ __forceinline__ void cameraRay(const DScene& sc, uint32_t x, uint32_t y, Rng& rng, float4 lam, float4 beta,
                                          PathStart& out);

__device__ __forceinline__ void startPath(const DScene& sc, const Settings& set, uint64_t pixel, uint64_t s, uint32_t spp,
                                 uint32_t pass, PathStart& out) {
    const DObserver& ob = sc.obs;
    const uint32_t x = uint32_t(pixel % uint64_t(ob.width)), y = uint32_t(pixel / uint64_t(ob.width));
    Rng pr;
    pr.seed(hashPair(set.seed, pixel), pass);
    const float rot = pr.uniform();
    Rng rng;
    rng.seed(hashPair(hashPair(set.seed, pixel), s), pass);
    float u = (float(s) + rng.uniform()) / float(spp) + rot;
    u -= floorf(u);
    float4 lam;
    float4 beta = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    if (set.spectral == OWE_SPECTRAL_UNIFORM1) {
        const float l = set.lambdaMin + (set.lambdaMax - set.lambdaMin) * u;
        lam = make_float4(l, l, l, l);
        beta = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
    } else {
        float ls[NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            float up = u + float(k) / NW;
            if (up >= 1.0f) up -= 1.0f;
            ls[k] = sampleVisibleF(up);
        }
        lam = make_float4(ls[0], ls[1], ls[2], ls[3]);
    }
    cameraRay(sc, x, y, rng, lam, beta, out);
}

// The camera ray of a pixel from `rng`: position in the pixel, filter offset, then the detector's
// own draws (IdealObserver/SurfaceSensor::generate, detector.cpp:61-149).
__device__ __forceinline__ void cameraRay(const DScene& sc, uint32_t x, uint32_t y, Rng& rng, float4 lam, float4 beta,
                                          PathStart& out) {
    const DObserver& ob = sc.obs;
    const float px = float(x) + rng.uniform(), py = float(y) + rng.uniform();
    float fx = 0.0f, fy = 0.0f;
    if (ob.sigma > 0.0f) {
        const float u2 = rng.uniform(), u1 = rng.uniform();
        const float r = sqrtf(-2.0f * logf(1.0f - u1)), phi = 2.0f * kPiF * u2;
        fx = r * cosf(phi) * ob.sigma;
        fy = r * sinf(phi) * ob.sigma;
    }
    out.lam = lam;
    if (ob.kind == OWE_DETECTOR_SENSOR) {
        // SurfaceSensor::generate (detector.cpp:110-149) without the focus disk: a point of the
        // sensor, a direction over its front (cosine or toward the aim disk), weight cos/pdf.
        float sxp = px + fx, syp = py + fy, weight = 0.0f;
        const float W = float(ob.width), H = float(ob.height);
        f3 p = mk3(0, 0, 0), dir = ld3(ob.normal);
        if (sxp >= 0.0f && syp >= 0.0f && sxp < W && syp < H) {
            if (ob.flipX) sxp = W - sxp;
            if (ob.flipY) syp = H - syp;
            p = xfPoint(ob.toWorld, mk3((2.0f * sxp / W - 1.0f) * ob.halfX, (1.0f - 2.0f * syp / H) * ob.halfY, 0.0f));
            const f3 nrm = ld3(ob.normal);
            if (ob.hasAim) {
                const bool useAim = ob.aimShare >= 1.0f || rng.uniform() < ob.aimShare;
                const float u2 = rng.uniform(), u1 = rng.uniform();
                float dx, dy;
                concentricDisk(u1, u2, dx, dy);
                const Frame af = makeFrame(ld3(ob.aimNormal));
                const f3 w = ld3(ob.aimCenter) + (af.s * dx + af.t * dy) * ob.aimRadius - p;
                dir = w / len3(w);
                if (!useAim) dir = toWorld(makeFrame(nrm), cosineHemisphere(u1, u2));
                const float cosS = dot3(dir, nrm);
                if (cosS > 0.0f) {
                    const float pdfW = ob.aimShare * diskAimPdf(p, dir, ld3(ob.aimCenter), ld3(ob.aimNormal), ob.aimRadius,
                                                                useAim) +
                                       (1.0f - ob.aimShare) * cosS * kInvPiF;
                    if (pdfW > 0.0f) weight = cosS / pdfW;
                }
            } else {
                const float u2 = rng.uniform(), u1 = rng.uniform();
                dir = toWorld(makeFrame(nrm), cosineHemisphere(u1, u2));
                weight = kPiF;
            }
        }
        // A sample the sensor does not measure (weight 0) still takes its slot and adds nothing.
        out.o = p;
        out.d = dir;
        out.weight = weight;
        out.beta = weight > 0.0f ? beta : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        out.region = ob.region;
        out.departed = ob.sensor;
        out.rngState = rng.state;
        return;
    }
    const float sx = (2.0f * (px + fx) / float(ob.width) - 1.0f) * ob.tanX;
    const float sy = (1.0f - 2.0f * (py + fy) / float(ob.height)) * ob.tanY;
    const f3 right = ld3(ob.right), up = ld3(ob.up), fwd = ld3(ob.fwd);
    const f3 aim = right * sx + up * sy + fwd;
    f3 q = mk3(0.0f, 0.0f, 0.0f), dir = norm3(aim);
    if (ob.pupil > 0.0f) {
        const float u2 = rng.uniform(), u1 = rng.uniform();
        float dx, dy;
        concentricDisk(u1, u2, dx, dy);
        q = (right * dx + up * dy) * ob.pupil;
        if (isfinite(ob.focus)) dir = norm3(aim * ob.focus - q);
    }
    out.o = q;
    out.d = dir;
    out.weight = 1.0f;
    out.beta = beta;
    out.region = ob.region;
    out.departed = kNoBoundary;
    out.rngState = rng.state;
}

__global__ __launch_bounds__(kBlock) void generateKernel(DScene sc, Settings set, Paths P, uint64_t base, uint32_t count,
                                                          uint32_t spp, uint32_t pass, const uint32_t* pixelList,
                                                          const uint32_t* sampleList) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    uint64_t pixel, s;
    if (pixelList) {
        pixel = pixelList[i];
        s = sampleList[i];
    } else {
        const uint64_t g = base + i;
        pixel = g / spp;
        s = g % spp;
    }
    PathStart ps;
    startPath(sc, set, pixel, s, spp, pass, ps);
    P.ray0[i] = make_float4(ps.o.x, ps.o.y, ps.o.z, 0.0f);
    P.ray1[i] = make_float4(ps.d.x, ps.d.y, ps.d.z, ps.weight);
    P.beta[i] = ps.beta;
    P.L[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    P.lam[i] = ps.lam;
    P.prevP[i] = make_float4(ps.o.x, ps.o.y, ps.o.z, 0.0f);
    P.info[i] = make_uint4(uint32_t(pixel), kFlagSpecular << 16, ps.region, ps.departed);
    P.face[i] = ~0u;
    P.rng[i] = make_uint2(uint32_t(ps.rngState), uint32_t(ps.rngState >> 32));
    P.queue[0][i] = i;
}

// ---------------------------------------------------------------- intersect

// The live count is read from device memory, so rounds can be enqueued without a host round trip;
// the grid covers a bound on it the host knows (live counts only shrink).
template <int kCap, bool kContacts>
__global__ __launch_bounds__(kBlock) void intersectKernel(DScene sc, Paths P, int q) {
    __shared__ uint32_t stack[kCap * kBlock];
    const uint32_t count = P.count[q];
    if (blockIdx.x == 0 && threadIdx.x == 0 && count > 0) atomicAdd(P.segments, (unsigned long long)count);
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= count) return;
    const uint32_t slot = P.queue[q][k];
    const float4 r0 = P.ray0[slot], r1 = P.ray1[slot];
    const uint4 info = P.info[slot];
    Hit h;
    const bool found = intersectContact<false, kContacts>(sc, mk3(r0.x, r0.y, r0.z), mk3(r1.x, r1.y, r1.z), __builtin_inff(), info.w,
                                               P.face[slot], kNoBoundary, info.z, stack + threadIdx.x, kBlock, h);
    P.hit[slot] = found ? make_uint4(__float_as_uint(h.t), h.boundary, h.face, h.slot)
                        : make_uint4(__float_as_uint(__builtin_inff()), kNoBoundary, 0u, kEmptyChild);
}

// ---------------------------------------------------------------- shade

struct ShadeCtx {
    const DScene& sc;
    const Settings& set;
    Rng rng;
    WL wl;
    float4 beta, L;
    f3 o, d, prevP;
    float prevPdf, weight;
    uint32_t pixel, depth, flags, region, departed, face;
    float4 shPend;
    f3 shO, shD;
    float shDist;
    uint32_t shStart, shLightRegion, shExclude, shDeparted, shFace;
};

// Next-event estimation at point p (sampleLightFrom + directLighting, transport.cpp:35-70, 323-348):
// draws the light, evaluates everything but the transmittance, and leaves a shadow ray.
__device__ __forceinline__ void prepareDirect(ShadeCtx& c, f3 p, bool surface, const Iface* it, uint32_t hitBoundary,
                                     uint32_t hitFace, uint32_t hitSlot, f3 n, float g) {
    const DScene& sc = c.sc;
    if (sc.nLights == 0) return;
    const float ul = c.rng.uniform();
    uint32_t li = 0;
    while (li + 1 < sc.nLights && sc.lights[li].cdf <= ul) ++li;
    const DLight L = sc.lights[li];
    const float u1 = c.rng.uniform(), u2 = c.rng.uniform(), u3 = c.rng.uniform();
    f3 wi;
    float dist = __builtin_inff(), lpdf;
    float Le[NW];
    uint32_t lightRegion = kRegionNone, exclude = kNoBoundary;
    if (L.kind == 1) {
        const Frame f = makeFrame(ld3(sc.env.sunDir));
        wi = toWorld(f, uniformConeOmc(u1, u2, sc.env.sunOneMinusCos));
#pragma unroll
        for (int i = 0; i < NW; ++i) Le[i] = carried(c.wl, i) ? evalSpec(sc.env.sunRadiance, sc.table, c.wl.lambda[i]) : 0.0f;
        lpdf = L.pdf / sc.env.sunSolidAngle;
    } else {
        const DBoundary& b = sc.boundaries[L.boundary];
        f3 pl, nl;
        float pdfW;
        if (!sampleFrom(b, xfPoint(b.toLocal, p), u1, u2, u3, pl, nl, pdfW)) return;
        const f3 ep = xfPoint(b.toWorld, pl), en = norm3(xfVector(b.toWorld, nl));
        const f3 dv = ep - p;
        dist = len3(dv);
        if (!(dist > 0.0f)) return;
        wi = dv / dist;
        const DEmission& e = sc.emissions[b.emission];
        const float cosL = dot3(en, -wi);
        const bool lit = (cosL > 0.0f && e.front) || (cosL < 0.0f && e.back);
#pragma unroll
        for (int i = 0; i < NW; ++i) Le[i] = lit && carried(c.wl, i) ? evalSpec(e.radiance, sc.table, c.wl.lambda[i]) : 0.0f;
        lpdf = L.pdf * pdfW;
        lightRegion = cosL > 0.0f ? b.front : b.back;
        exclude = L.boundary;
    }
    if (!(lpdf > 0.0f)) return;
    bool any = false;
#pragma unroll
    for (int i = 0; i < NW; ++i) any = any || Le[i] != 0.0f;
    if (!any) return;
    float f[NW] = {0, 0, 0, 0};
    float pdfScatter = 0.0f, cosFactor = 1.0f;
    uint32_t start = c.region;
    if (!surface) {
        pdfScatter = hgPhase(dot3(-c.d, wi), g);
#pragma unroll
        for (int i = 0; i < NW; ++i) f[i] = pdfScatter;
    } else {
        const DBoundary& b = sc.boundaries[hitBoundary];
        bool nz = false;
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            if (c.wl.pdf[i] == 0.0f) continue;
            Iface at = *it;
            if (i > 0) {
                at.nFront = regionIndex(sc, b.front, c.wl.lambda[i]);
                at.nBack = regionIndex(sc, b.back, c.wl.lambda[i]);
                at.lambda = c.wl.lambda[i];
            }
            float pdf;
            f[i] = evalScatter(sc, at, c.d, wi, true, pdf);
            if (i == 0) pdfScatter = pdf;
            nz = nz || f[i] != 0.0f;
        }
        if (!nz) return;
        cosFactor = fabsf(dot3(wi, n));
        start = dot3(wi, n) > 0.0f ? b.front : b.back;
    }
    const float w = powerHeuristic(lpdf, pdfScatter) * cosFactor / lpdf;
    c.shPend = make_float4(get(c.beta, 0) * f[0] * Le[0] * w, get(c.beta, 1) * f[1] * Le[1] * w,
                           get(c.beta, 2) * f[2] * Le[2] * w, get(c.beta, 3) * f[3] * Le[3] * w);
    c.shO = p;
    c.shD = wi;
    c.shDist = dist;
    c.shStart = start;
    c.shLightRegion = lightRegion;
    c.shExclude = exclude;
    c.shDeparted = surface ? hitBoundary : kNoBoundary;
    c.shFace = surface && hitSlot != kEmptyChild ? hitFace : ~0u;
}

__device__ __forceinline__ bool survives(ShadeCtx& c) {
    const float q = fminf(1.0f, fmaxf(fmaxf(c.beta.x, c.beta.y), fmaxf(c.beta.z, c.beta.w)));
    if (c.rng.uniform() >= q) return false;
    const float s = 1.0f / q;
    c.beta = make_float4(c.beta.x * s, c.beta.y * s, c.beta.z * s, c.beta.w * s);
    return true;
}

// An SPPM camera path's visible point and what decides it (transport.slang:749-762, 783-791, 846-858).
struct SppmRec {
    bool valid = false;
    bool chromatic = false;  // an absorbing, non-scattering medium on the way
    uint32_t events = 0, nonDelta = 0, boundary = 0, flags = 0;
    float pathLength = 0.0f;  // unfolded length from the eye
    f3 p, n, d;
    float4 beta;
};

// One iteration of Tracer::radiance's loop after the intersection; false when the path ends. With
// kSppm the SPPM camera pass's rule: the first non-specular surface reached through specular events
// only is the visible point; the path takes next-event estimation and one scatter sample there and
// ends at the next matter it meets (light scattered by a medium on that segment is the photons').
template <bool kSppm = false>
__device__ __forceinline__ bool shadeOne(ShadeCtx& c, const uint4 hraw, uint32_t* counters, SppmRec* rec = nullptr) {
    const DScene& sc = c.sc;
    const bool found = hraw.y != kNoBoundary;
    const float t = __uint_as_float(hraw.x);
    const DMedium& med = sc.media[sc.regionMedium[c.region]];
    // sampleMedium (transport.cpp:95-113).
    if (med.flags & 2) {
        float sT[NW];
#pragma unroll
        for (int i = 0; i < NW; ++i)
            sT[i] = evalSpec(med.absorption, sc.table, c.wl.lambda[i]) + evalSpec(med.scattering, sc.table, c.wl.lambda[i]);
        const float tm = -logf(1.0f - c.rng.uniform()) / sT[0];
        const float seg = found ? t : __builtin_inff();
        if (tm < seg) {
            const float pdf0 = sT[0] * expf(-sT[0] * tm);
#pragma unroll
            for (int i = 0; i < NW; ++i)
                setc(c.beta, i, get(c.beta, i) * evalSpec(med.scattering, sc.table, c.wl.lambda[i]) * expf(-sT[i] * tm) / pdf0);
            // scatterInMedium (transport.cpp:379-397).
            const f3 p = c.o + c.d * tm;
            if (kSppm && rec->valid) return false;
            if (int(c.depth) >= c.set.maxDepth) return false;
            c.flags |= kFlagMediumScattered;
            prepareDirect(c, p, false, nullptr, kNoBoundary, 0, kEmptyChild, mk3(0, 0, 0), med.g);
            const float u2 = c.rng.uniform(), u1 = c.rng.uniform();
            f3 wi;
            c.prevPdf = sampleHg(-c.d, med.g, u1, u2, wi);
            if (kSppm) {
                rec->events++;
                rec->nonDelta++;
            }
            c.flags &= ~kFlagSpecular;
            noteScatter(c.flags, false);
            c.prevP = p;
            c.o = p;
            c.d = wi;
            c.departed = kNoBoundary;
            c.face = ~0u;
            return int(c.depth) < c.set.rrDepth || survives(c);
        }
        if (!found) return false;
        const float pass0 = expf(-sT[0] * seg);
#pragma unroll
        for (int i = 0; i < NW; ++i) setc(c.beta, i, get(c.beta, i) * expf(-sT[i] * seg) / pass0);
    } else if (med.flags & 1) {
        if (!found) return false;
#pragma unroll
        for (int i = 0; i < NW; ++i) setc(c.beta, i, get(c.beta, i) * expf(-evalSpec(med.absorption, sc.table, c.wl.lambda[i]) * t));
    }
    if (kSppm && (med.flags & 1) && !(med.flags & 2)) rec->chromatic = true;
    // Light this path may count: under the hybrid partition not what light tracing owns; in the
    // medium-only pass only after a medium scattering vertex (stepPath, transport.slang:868-906).
    const bool counts = !(c.set.mode & kModeMediumOnly) || (c.flags & kFlagMediumScattered);
    const bool owned = (c.set.mode & kModePartition) && ownedByLightTracing(c.flags);
    if (!found) {
        if (c.region == 0) {
            if (owned || !counts) return false;
            // skyMet (transport.cpp:401-417); the sun's disc test in the cancellation-free form
            // 1 − cos θ = |d − s|² / 2 for unit vectors.
            const f3 s = ld3(sc.env.sunDir), dd = c.d - s;
            const bool inSun = sc.env.hasSun && 0.5f * dot3(dd, dd) <= sc.env.sunOneMinusCos;
            float w = 1.0f;
            if (inSun && !(c.flags & kFlagSpecular)) w = powerHeuristic(c.prevPdf, sc.env.sunPdf);
#pragma unroll
            for (int i = 0; i < NW; ++i) {
                if (!carried(c.wl, i)) continue;
                const float sky = sc.env.sky == OWE_SKY_UNIFORM ? evalSpec(sc.env.skyRadiance, sc.table, c.wl.lambda[i]) : 0.0f;
                const float sun = inSun ? evalSpec(sc.env.sunRadiance, sc.table, c.wl.lambda[i]) : 0.0f;
                setc(c.L, i, get(c.L, i) + get(c.beta, i) * (sky + sun * w));
            }
        } else {
            atomicAdd(&counters[3], 1u);
        }
        return false;
    }
    // A boundary met: the point, the normal and the sides (meetBoundary, transport.cpp:121-133).
    const uint32_t bi = hraw.y;
    const DBoundary& b = sc.boundaries[bi];
    f3 p, n, pl = mk3(0, 0, 0), nl = mk3(0, 0, 1);
    if (hraw.w != kEmptyChild) {
        const float4* tr = reinterpret_cast<const float4*>(sc.tris + hraw.w);
        const float4 a = tr[0], bb = tr[1], cc = tr[2];
        n = norm3(triNormal(mk3(a.x, a.y, a.z), mk3(bb.x, bb.y, bb.z), mk3(cc.x, cc.y, cc.z)));
        p = c.o + c.d * t;
    } else {
        const f3 lo = xfPoint(b.toLocal, c.o), ldv = xfVector(b.toLocal, c.d);
        analyticGeometry(sc, b, lo, ldv, t, pl, nl);
        p = xfPoint(b.toWorld, pl);
        n = norm3(xfVector(b.toWorld, nl));
    }
    const bool fromFront = dot3(c.d, n) < 0.0f;
    const uint32_t inc = fromFront ? b.front : b.back, other = fromFront ? b.back : b.front;
    if (inc != c.region) {
        atomicAdd(&counters[2], 1u);
        noteInconsistency(sc, bi, c.region, inc, c.departed, p, c.o, c.d, hraw.w);
        c.region = inc;
    }
    if (kSppm && !rec->valid) rec->pathLength += t;
    const DOptics& o = sc.optics[b.optics];
    // emissionMet (transport.cpp:420-433).
    if (b.emission >= 0 && !owned && counts) {
        const DEmission& e = sc.emissions[b.emission];
        if (fromFront ? e.front : e.back) {
            float w = 1.0f;
            if (!(c.flags & kFlagSpecular) && b.light >= 0) {
                const float pl2 = sc.lights[b.light].pdf * pdfFrom(b, xfPoint(b.toLocal, c.prevP), pl, nl);
                if (pl2 > 0.0f) w = powerHeuristic(c.prevPdf, pl2);
            }
#pragma unroll
            for (int i = 0; i < NW; ++i)
                if (carried(c.wl, i))
                    setc(c.L, i, get(c.L, i) + get(c.beta, i) * evalSpec(e.radiance, sc.table, c.wl.lambda[i]) * w);
        }
    }
    const uint32_t hitFace = hraw.w != kEmptyChild ? hraw.z : ~0u;
    if (o.type == OWE_SURF_NULL) {
        c.region = other;
        c.o = p;
        c.departed = bi;
        c.face = hitFace;
        return int(c.depth) < 4 * c.set.maxDepth;
    }
    if (kSppm && rec->valid) return false;  // the visible point's scatter segment has met matter
    if (o.type == OWE_SURF_ABSORBER || o.type == OWE_SURF_DETECTOR) return false;
    if (int(c.depth) >= c.set.maxDepth) return false;
    // scatterAtSurface (transport.cpp:437-471).
    const float hero = c.wl.lambda[0];
    Iface it{&o, n, regionIndex(sc, b.front, hero), regionIndex(sc, b.back, hero), hero, c.set.fresnelFloor};
    auto at = [&](int i) {
        Iface r = it;
        r.nFront = regionIndex(sc, b.front, c.wl.lambda[i]);
        r.nBack = regionIndex(sc, b.back, c.wl.lambda[i]);
        r.lambda = c.wl.lambda[i];
        return r;
    };
    if (o.type == OWE_SURF_DIELECTRIC && !secondaryTerminated(c.wl)) {
        const float a0 = it.nBack / it.nFront;
#pragma unroll
        for (int i = 1; i < NW; ++i) {
            if (c.wl.pdf[i] == 0.0f) continue;
            const Iface ai = at(i);
            if (fabsf(a0 - ai.nBack / ai.nFront) > 1e-12f * fabsf(a0)) {
                c.flags |= kFlagTerminated;
                c.wl.pdf[0] /= NW;
                c.wl.pdf[1] = c.wl.pdf[2] = c.wl.pdf[3] = 0.0f;
                c.beta.y = c.beta.z = c.beta.w = 0.0f;
                break;
            }
        }
    }
    if (kSppm && !o.delta && rec->nonDelta == 0) {
        rec->valid = true;
        rec->p = p;
        rec->n = n;
        rec->d = c.d;
        rec->boundary = bi;
        rec->beta = c.beta;
        rec->flags = kVpValid | (secondaryTerminated(c.wl) ? kVpTerminated : 0u) |
                     (rec->events == 0 && !rec->chromatic ? kVpFree : 0u);
    }
    if (!o.delta && counts) prepareDirect(c, p, true, &it, bi, hraw.z, hraw.w, n, 0.0f);
    const float u3 = c.rng.uniform(), u2 = c.rng.uniform(), u1 = c.rng.uniform();
    SSample s;
    if (!sampleScatter(sc, it, c.d, u1, u2, u3, true, s) || s.weight <= 0.0f) return false;
    if (kSppm) {
        rec->events++;
        if (!s.delta) rec->nonDelta++;
    }
    c.beta.x *= s.weight;
#pragma unroll
    for (int i = 1; i < NW; ++i)
        if (c.wl.pdf[i] != 0.0f) setc(c.beta, i, get(c.beta, i) * secondaryWeight(sc, at(i), c.d, s, true));
    if (s.transmitted) c.region = other;
    if (s.delta) c.flags |= kFlagSpecular;
    else c.flags &= ~kFlagSpecular;
    noteScatter(c.flags, s.delta);
    c.prevPdf = s.pdf;
    c.prevP = p;
    c.o = p;
    c.d = s.wi;
    c.departed = bi;
    c.face = hitFace;
    if (int(c.depth) >= c.set.rrDepth && !survives(c)) return false;
    return true;
}

__global__ __launch_bounds__(kBlock) void shadeKernel(DScene sc, Settings set, Paths P, int q, uint64_t passInc) {
    const uint32_t count = P.count[q];
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    bool alive = false;
    uint32_t slot = 0;
    if (k < count) {
        slot = P.queue[q][k];
        const float4 r0 = P.ray0[slot], r1 = P.ray1[slot];
        const uint4 info = P.info[slot];
        const uint2 rs = P.rng[slot];
        ShadeCtx c{sc, set};
        c.rng.state = uint64_t(rs.x) | (uint64_t(rs.y) << 32);
        c.rng.inc = passInc;
        c.flags = info.y >> 16;
        c.depth = info.y & 0xFFFFu;
        c.wl = loadWL(P.lam[slot], c.flags, set);
        c.beta = P.beta[slot];
        c.L = P.L[slot];
        c.o = mk3(r0.x, r0.y, r0.z);
        c.prevPdf = r0.w;
        c.d = mk3(r1.x, r1.y, r1.z);
        const float4 pp = P.prevP[slot];
        c.prevP = mk3(pp.x, pp.y, pp.z);
        c.pixel = info.x;
        c.region = info.z;
        c.departed = info.w;
        c.face = P.face[slot];
        c.shPend = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        alive = shadeOne(c, P.hit[slot], P.count);
        P.L[slot] = c.L;
        P.shPend[slot] = c.shPend;
        if (c.shPend.x != 0.0f || c.shPend.y != 0.0f || c.shPend.z != 0.0f || c.shPend.w != 0.0f) {
            P.shO[slot] = make_float4(c.shO.x, c.shO.y, c.shO.z, c.shDist);
            P.shD[slot] = make_float4(c.shD.x, c.shD.y, c.shD.z, 0.0f);
            P.shInfo[slot] = make_uint4(c.shStart, c.shLightRegion, c.shExclude, c.shDeparted);
            P.shFace[slot] = c.shFace;
        }
        // The wavelengths' termination changes the splat; the rest only matters to a live path.
        P.info[slot] = make_uint4(c.pixel, (c.depth + 1) | (c.flags << 16), c.region, c.departed);
        if (alive) {
            P.ray0[slot] = make_float4(c.o.x, c.o.y, c.o.z, c.prevPdf);
            P.ray1[slot] = make_float4(c.d.x, c.d.y, c.d.z, r1.w);
            P.beta[slot] = c.beta;
            P.prevP[slot] = make_float4(c.prevP.x, c.prevP.y, c.prevP.z, 0.0f);
            P.face[slot] = c.face;
            P.rng[slot] = make_uint2(uint32_t(c.rng.state), uint32_t(c.rng.state >> 32));
        }
    }
    // Wave-level compaction of the survivors: one atomic per wave, a prefix count within it.
    const uint64_t mask = __ballot(alive);
    if (mask == 0) return;
    const uint32_t lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t at = 0;
    if (int(lane) == leader) at = atomicAdd(&P.count[1 - q], uint32_t(__popcll(mask)));
    at = __shfl(at, leader);
    if (alive) {
        const uint32_t below = __popcll(mask & ((1ull << lane) - 1ull));
        P.queue[1 - q][at + below] = slot;
    }
}

// ---------------------------------------------------------------- shadow rays

// Tracer::transmittance (transport.cpp:237-285): through null boundaries with Beer–Lambert.
// Tracer::transmittance of one pending NEE ray (transport.cpp:237-285): through null boundaries
// with Beer–Lambert; false when blocked. With `simple` (no null boundaries, no media in the scene)
// any crossing blocks. Shared by the wavefront and the persistent kernels.
template <bool kContacts>
__device__ __forceinline__ bool shadowTransmittance(const DScene& sc, f3 o, f3 d, float remaining, uint32_t region,
                                           uint32_t arrival, uint32_t exclude, uint32_t departed, uint32_t face,
                                           const float4 lam, int simple, uint32_t* stack, float Tr[NW]) {
    const bool infinite = isinf(remaining);
#pragma unroll
    for (int i = 0; i < NW; ++i) Tr[i] = 1.0f;
    for (int guard = 0; guard < 64; ++guard) {
        Hit h;
        const bool found = intersectContact<true, kContacts>(sc, o, d, remaining, departed, face, exclude, region, stack,
                                                             kBlock, h);
        if (simple) {
            // No null boundaries or participating media in the scene: any crossing blocks.
            if (found) return false;
            if (infinite && region != 0) return false;
            break;
        }
        const float seg = found ? fminf(h.t, remaining) : remaining;
        const DMedium& m = sc.media[sc.regionMedium[region]];
        if (m.flags) {
            if (isinf(seg)) return false;
#pragma unroll
            for (int i = 0; i < NW; ++i) {
                const float l = get(lam, i);
                Tr[i] *= expf(-(evalSpec(m.absorption, sc.table, l) + evalSpec(m.scattering, sc.table, l)) * seg);
            }
        }
        if (!found) {
            if (infinite && region != 0) return false;
            if (!infinite && region != arrival) return false;
            break;
        }
        const DBoundary& b = sc.boundaries[h.boundary];
        const bool front = dot3(d, h.n) < 0.0f;
        if ((front ? b.front : b.back) != region) return false;
        if (sc.optics[b.optics].type != OWE_SURF_NULL) return false;
        region = front ? b.back : b.front;
        o = h.p;
        departed = h.boundary;
        face = h.slot != kEmptyChild ? h.face : ~0u;
        if (!infinite) remaining -= h.t;
        if (guard == 63) return false;  // exhausting the guard is not evidence the segment is clear
    }
    return !(simple && !infinite && region != arrival);
}

__device__ __forceinline__ float4 addShadow(float4 L, const float4 pend, const float Tr[NW]) {
    L.x += pend.x * Tr[0];
    L.y += pend.y * Tr[1];
    L.z += pend.z * Tr[2];
    L.w += pend.w * Tr[3];
    return L;
}

template <int kCap, bool kContacts>
__global__ __launch_bounds__(kBlock) void shadowKernel(DScene sc, Settings set, Paths P, int q, int simple) {
    __shared__ uint32_t stack[kCap * kBlock];
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= P.count[q]) return;
    const uint32_t slot = P.queue[q][k];
    const float4 pend = P.shPend[slot];
    if (pend.x == 0.0f && pend.y == 0.0f && pend.z == 0.0f && pend.w == 0.0f) return;
    const float4 so = P.shO[slot], sd = P.shD[slot];
    const uint4 si = P.shInfo[slot];
    float Tr[NW];
    if (!shadowTransmittance<kContacts>(sc, mk3(so.x, so.y, so.z), mk3(sd.x, sd.y, sd.z), so.w, si.x, si.y, si.z, si.w,
                             P.shFace[slot], P.lam[slot], simple, stack + threadIdx.x, Tr))
        return;
    P.L[slot] = addShadow(P.L[slot], pend, Tr);
}

// ---------------------------------------------------------------- splat

__device__ __forceinline__ long long fixedPoint(float v) {
    // |value| stays below 2^30 units of XYZ: far above any sample here, and no overflow of the
    // 64-bit two's-complement sum.
    return llrint(fmin(fmax(double(v) * kFilmScale, -4.6e18), 4.6e18));
}

// Each path's XYZ, then one film atomic per run of same-pixel lanes in the wave: a segmented suffix
// sum of 64-bit integers (exact, so the film stays independent of order) over the lanes' pixels,
// which are contiguous runs because path g belongs to pixel g / spp.
// A path's XYZ (Wavelengths::toXYZ, core/wavelengths.hpp:71-83) from its estimate per wavelength,
// its camera weight and its final wavelengths; false when not finite (the sample is dropped).
__device__ __forceinline__ bool pathXYZ(const float4 L, float weight, const WL& wl, float invNWY, float& X, float& Y, float& Z) {
    X = Y = Z = 0.0f;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        const float val = get(L, k) * weight;
        if (wl.pdf[k] == 0.0f || val == 0.0f) continue;
        const float w = val / wl.pdf[k] * invNWY;
        const f3 c = cieXYZ(wl.lambda[k]);
        X += w * c.x;
        Y += w * c.y;
        Z += w * c.z;
    }
    return isfinite(X + Y + Z);
}

__global__ __launch_bounds__(kBlock) void splatKernel(Settings set, Paths P, uint32_t count, float invNWY,
                                                      unsigned long long* film, double* perPath) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    long long v[3] = {0, 0, 0};
    uint32_t pix = 0xFFFFFFFFu;
    if (i < count) {
        const uint4 info = P.info[i];
        float X, Y, Z;
        const bool ok = pathXYZ(P.L[i], P.ray1[i].w, loadWL(P.lam[i], info.y >> 16, set), invNWY, X, Y, Z);
        if (perPath) {
            perPath[3 * size_t(i)] = ok ? X : 0.0;
            perPath[3 * size_t(i) + 1] = ok ? Y : 0.0;
            perPath[3 * size_t(i) + 2] = ok ? Z : 0.0;
        } else if (ok) {
            pix = info.x;
            v[0] = fixedPoint(X);
            v[1] = fixedPoint(Y);
            v[2] = fixedPoint(Z);
        }
    }
    if (perPath) return;  // uniform across the grid
    const uint32_t lane = __lane_id();
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t po = __shfl_down(pix, off);
        long long o[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) o[c] = __shfl_down(v[c], off);
        if (lane + off < 64 && po == pix) {
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c] += o[c];
        }
    }
    const uint32_t prev = __shfl_up(pix, 1);
    const bool head = lane == 0 || prev != pix;
    if (head && pix != 0xFFFFFFFFu)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (v[c] != 0) atomicAdd(&film[3 * size_t(pix) + c], (unsigned long long)v[c]);
}

__global__ void clearCounts(uint32_t* c, int n) {
    if (threadIdx.x < unsigned(n)) c[threadIdx.x] = 0;
}

// Starts a chunk: queue 0 holds `count` paths, queue 1 none, no segments, inconsistencies or leaks yet.
__global__ void setCount(uint32_t* c, unsigned long long* segments, uint32_t count) {
    if (threadIdx.x < 4) c[threadIdx.x] = threadIdx.x == 0 ? count : 0u;
    if (threadIdx.x == 0) *segments = 0;
}

// ---------------------------------------------------------------- persistent form

// The batched persistent form of the same camera loop: a grid sized to the chip keeps one path per
// lane in registers and advances it one segment per iteration (intersect, shade, shadow — the very
// device functions of the wavefront kernels, so every path's float operations are the same); a lane
// whose path ends splats it and takes the next path id from its chunk, or a new chunk of `chunk`
// consecutive ids from a device counter (one atomic per wave per refill). Splats of consecutive
// paths of one pixel accumulate in the lane and reach the film as one 64-bit atomic per channel.
// The film is integer, so it equals the wavefront's bit for bit.
struct Persist {
    uint64_t total;            // paths of the pass: path g is sample g % spp of pixel g / spp
    uint32_t spp, chunk;
    const uint32_t* pass;      // the pass index, read on the device (a graph replays with a new one)
    unsigned long long* next;  // path ids handed out so far
    unsigned long long* film;
    unsigned long long* segments;
    uint32_t* counters;        // [2] inconsistencies, [3] leaks (as the wavefront's)
    float invNWY;
    int simple;
    uint32_t regen;            // refill a wave's finished lanes once at least this many are idle
};

__global__ void persistBegin(Persist a) {
    if (threadIdx.x == 0) {
        *a.next = 0;
        *a.segments = 0;
    }
    if (threadIdx.x < 4) a.counters[threadIdx.x] = 0;
}

__device__ __forceinline__ void flushFilm(unsigned long long* film, uint32_t pix, long long acc[3]) {
    if (pix == 0xFFFFFFFFu) return;
#pragma unroll
    for (int c = 0; c < 3; ++c)
        if (acc[c] != 0) atomicAdd(&film[3 * size_t(pix) + c], (unsigned long long)acc[c]);
    acc[0] = acc[1] = acc[2] = 0;
}

// The lane loop of the persistent form over one scene's pass (the kernels below run it for one scene
// or for each scene of a batch in turn); returns the segments the lane traced. The wave leaves it
// only when every lane has finished and no path ids of the pass are left.
template <int kCap, bool kContacts>
__device__ __forceinline__ uint32_t persistDrain(const DScene& sc, const Settings& set, const Persist& a,
                                                 uint32_t* st, uint32_t pass) {
    const uint64_t passInc = (uint64_t(pass) << 1) | 1u;
    const uint32_t lane = __lane_id();
    ShadeCtx c{sc, set};
    bool live = false, done = false, pending = false;  // pending: a finished path not yet splatted
    uint64_t id = 0, end = 0;
    uint32_t accPix = 0xFFFFFFFFu, segments = 0;
    long long acc[3] = {0, 0, 0};
    for (;;) {
        // Finished lanes wait until at least a.regen of them (or the whole wave) are idle, then all
        // start new paths together: the wave runs the path set-up code once for many lanes.
        const uint64_t idle = __ballot(!live && !done);
        const bool refill = uint32_t(__popcll(idle)) >= a.regen || __ballot(live) == 0;
        // Finished paths are splatted together at the refill, from the state they ended with.
        if (refill && pending) {
            pending = false;
            float X, Y, Z;
            if (pathXYZ(c.L, c.weight, c.wl, a.invNWY, X, Y, Z)) {
                if (c.pixel != accPix) {
                    flushFilm(a.film, accPix, acc);
                    accPix = c.pixel;
                }
                acc[0] += fixedPoint(X);
                acc[1] += fixedPoint(Y);
                acc[2] += fixedPoint(Z);
            }
        }
        const bool needChunk = refill && !live && !done && id >= end;
        const uint64_t ask = __ballot(needChunk);
        if (ask) {
            const int leader = __ffsll((long long)ask) - 1;
            unsigned long long base = 0;
            if (int(lane) == leader) base = atomicAdd(a.next, (unsigned long long)__popcll(ask) * a.chunk);
            base = __shfl(base, leader);
            if (needChunk) {
                id = base + uint64_t(__popcll(ask & ((1ull << lane) - 1ull))) * a.chunk;
                end = id + a.chunk < a.total ? id + a.chunk : a.total;
                done = id >= a.total;
            }
        }
        if (refill && !live && !done) {
            const uint64_t pixel = id / a.spp;
            PathStart ps;
            startPath(sc, set, pixel, id % a.spp, a.spp, pass, ps);
            c.rng.state = ps.rngState;
            c.rng.inc = passInc;
            c.flags = kFlagSpecular;
            c.depth = 0;
            c.wl = loadWL(ps.lam, c.flags, set);
            c.beta = ps.beta;
            c.L = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            c.o = ps.o;
            c.prevPdf = 0.0f;
            c.d = ps.d;
            c.weight = ps.weight;
            c.prevP = ps.o;
            c.pixel = uint32_t(pixel);
            c.region = ps.region;
            c.departed = ps.departed;
            c.face = ~0u;
            live = true;
            ++id;
        }
        if (__ballot(live) == 0) break;  // every lane of the wave is done: the queue is empty
        if (!live) continue;
        Hit h;
        const bool found = intersectContact<false, kContacts>(sc, c.o, c.d, __builtin_inff(), c.departed, c.face,
                                                              kNoBoundary, c.region, st, kBlock, h);
        ++segments;
        const uint4 hraw = found ? make_uint4(__float_as_uint(h.t), h.boundary, h.face, h.slot)
                                 : make_uint4(__float_as_uint(__builtin_inff()), kNoBoundary, 0u, kEmptyChild);
        c.shPend = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const bool alive = shadeOne(c, hraw, a.counters);
        if (c.shPend.x != 0.0f || c.shPend.y != 0.0f || c.shPend.z != 0.0f || c.shPend.w != 0.0f) {
            float Tr[NW];
            const float4 lam = make_float4(c.wl.lambda[0], c.wl.lambda[1], c.wl.lambda[2], c.wl.lambda[3]);
            if (shadowTransmittance<kContacts>(sc, c.shO, c.shD, c.shDist, c.shStart, c.shLightRegion, c.shExclude, c.shDeparted,
                                    c.shFace, lam, a.simple, st, Tr))
                c.L = addShadow(c.L, c.shPend, Tr);
        }
        c.depth += 1;
        if (!alive) {
            live = false;
            pending = true;
        }
    }
    flushFilm(a.film, accPix, acc);
    return segments;
}

__device__ __forceinline__ void addSegments(unsigned long long* total, uint32_t segments) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) segments += __shfl_xor(segments, off);
    if (__lane_id() == 0 && segments) atomicAdd(total, (unsigned long long)segments);
}

// kWaves: the waves per SIMD the register allocation must allow (2: no spills; 3: more waves to
// hide latency, at the price of some register spills) — the autotuner measures both.
template <int kCap, int kWaves, bool kContacts>
__global__ __launch_bounds__(kBlock, kWaves) void persistentKernel(DScene sc, Settings set, Persist a) {
    __shared__ uint32_t stack[kCap * kBlock];
    addSegments(a.segments, persistDrain<kCap, kContacts>(sc, set, a, stack + threadIdx.x, *a.pass));
}

// A batch of scenes (variants of one family: perturbed instruments, tolerancing runs) in one launch:
// every wave drains scene 0's pass, then scene 1's, and so on — scene v's paths come from its own id
// counter and go to its own film, so each scene's film equals that of its own persistent pass bit for
// bit. The scene a wave works on is uniform across it, so its tables are read through the scalar unit.
struct PersistBatch {
    uint32_t count;                         // scenes
    uint32_t spp, chunk, regen;
    const uint32_t* pass;
    const uint64_t* total;                  // [count] paths of the pass
    const int32_t* simple;                  // [count] plain shadow walks
    unsigned long long* const* film;        // [count]
    unsigned long long* next;               // [count]
    unsigned long long* segments;           // [count]
    uint32_t* counters;                     // [4 count]: [2] inconsistencies, [3] leaks
    float invNWY;
};

__global__ void persistBatchBegin(PersistBatch b) {
    for (uint32_t v = threadIdx.x; v < b.count; v += blockDim.x) {
        b.next[v] = 0;
        b.segments[v] = 0;
        for (int k = 0; k < 4; ++k) b.counters[4 * v + k] = 0;
    }
}

template <int kCap, int kWaves, bool kContacts>
__global__ __launch_bounds__(kBlock, kWaves) void persistentBatchKernel(const DScene* __restrict__ scenes, Settings set,
                                                                       PersistBatch b) {
    __shared__ uint32_t stack[kCap * kBlock];
    const uint32_t pass = *b.pass;
    for (uint32_t v = 0; v < b.count; ++v) {
        Persist a{};
        a.total = b.total[v];
        a.spp = b.spp;
        a.chunk = b.chunk;
        a.regen = b.regen;
        a.next = b.next + v;
        a.film = b.film[v];
        a.counters = b.counters + 4 * v;
        a.invNWY = b.invNWY;
        a.simple = b.simple[v];
        addSegments(b.segments + v, persistDrain<kCap, kContacts>(scenes[v], set, a, stack + threadIdx.x, pass));
    }
}

// ---------------------------------------------------------------- SPPM

// The SPPM camera pass (sppm_camera_kernel): one thread per pixel and one camera path per iteration,
// at the wavelengths of the pixel's group. The path's own radiance goes to the film (one sample per
// pixel per iteration); its visible point, if any, to the pixel's slot with the unfolded path length.
// It reads nothing the previous iteration's photons or update write, so it runs beside them on a
// second stream, into the other of two visible-point buffers. Compiled for two waves per SIMD (with
// contacts its registers would allow one).
template <int kCap, bool kContacts>
__global__ __launch_bounds__(kBlock, 2) void sppmCameraKernel(DScene sc, Settings set, SppmDev s, unsigned long long* film,
                                                           uint32_t* counters, int simple) {
    __shared__ uint32_t stack[kCap * kBlock];
    uint32_t* st = stack + threadIdx.x;
    const uint32_t pixel = blockIdx.x * kBlock + threadIdx.x;
    uint32_t paths = 0, segments = 0;
    if (pixel < s.pixels) {
        const uint32_t group = uint32_t(hashPair(hashPair(set.seed ^ kSppmGroup, pixel), s.iteration) % uint64_t(s.groups));
        ShadeCtx c{sc, set};
        c.wl = sppmWavelengths(set, s.rotation, group, s.groups);
        Rng rng;
        rng.seed(hashPair(set.seed ^ kSppmCamera, pixel), s.iteration);
        const float4 lam = make_float4(c.wl.lambda[0], c.wl.lambda[1], c.wl.lambda[2], c.wl.lambda[3]);
        const float4 beta0 = set.spectral == OWE_SPECTRAL_UNIFORM1 ? make_float4(1.0f, 0.0f, 0.0f, 0.0f)
                                                                   : make_float4(1.0f, 1.0f, 1.0f, 1.0f);
        PathStart ps;
        cameraRay(sc, pixel % sc.obs.width, pixel / sc.obs.width, rng, lam, beta0, ps);
        SppmRec rec;
        c.L = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (ps.weight > 0.0f) {
            c.rng = rng;
            c.flags = kFlagSpecular;
            c.depth = 0;
            c.beta = ps.beta;
            c.o = ps.o;
            c.prevPdf = 0.0f;
            c.d = ps.d;
            c.weight = ps.weight;
            c.prevP = ps.o;
            c.pixel = pixel;
            c.region = ps.region;
            c.departed = ps.departed;
            c.face = ~0u;
            for (bool alive = true; alive; c.depth += 1) {
                Hit h;
                const bool found = intersectContact<false, kContacts>(sc, c.o, c.d, __builtin_inff(), c.departed, c.face,
                                                                      kNoBoundary, c.region, st, kBlock, h);
                ++segments;
                const uint4 hraw = found ? make_uint4(__float_as_uint(h.t), h.boundary, h.face, h.slot)
                                         : make_uint4(__float_as_uint(__builtin_inff()), kNoBoundary, 0u, kEmptyChild);
                c.shPend = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                alive = shadeOne<true>(c, hraw, counters, &rec);
                if (c.shPend.x != 0.0f || c.shPend.y != 0.0f || c.shPend.z != 0.0f || c.shPend.w != 0.0f) {
                    float Tr[NW];
                    const float4 l4 = make_float4(c.wl.lambda[0], c.wl.lambda[1], c.wl.lambda[2], c.wl.lambda[3]);
                    if (shadowTransmittance<kContacts>(sc, c.shO, c.shD, c.shDist, c.shStart, c.shLightRegion, c.shExclude,
                                                       c.shDeparted, c.shFace, l4, simple, st, Tr))
                        c.L = addShadow(c.L, c.shPend, Tr);
                }
            }
            float X, Y, Z;
            if (pathXYZ(c.L, c.weight, c.wl, s.invNWY, X, Y, Z)) {
                // One thread per pixel: the film's words are this thread's alone in this kernel.
                unsigned long long* f = film + 3 * size_t(pixel);
                f[0] += (unsigned long long)fixedPoint(X);
                f[1] += (unsigned long long)fixedPoint(Y);
                f[2] += (unsigned long long)fixedPoint(Z);
            }
            paths = 1;
        }
        // w: the unfolded path length; the hash fill turns it into the pixel's radius.
        const float w = ps.weight;
        s.vpP[pixel] = make_float4(rec.p.x, rec.p.y, rec.p.z, rec.pathLength);
        s.vpN[pixel] = make_float4(rec.n.x, rec.n.y, rec.n.z, __uint_as_float(rec.boundary));
        s.vpD[pixel] = make_float4(rec.d.x, rec.d.y, rec.d.z, __uint_as_float(rec.valid ? rec.flags | (group << 8) : 0u));
        s.vpBeta[pixel] = make_float4(rec.beta.x * w, rec.beta.y * w, rec.beta.z * w, rec.beta.w * w);
    }
    // One pair of counter atomics per wave.
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        paths += __shfl_xor(paths, off);
        segments += __shfl_xor(segments, off);
    }
    if (__lane_id() == 0 && paths) {
        atomicAdd(&s.stats[0], (unsigned long long)paths);
        atomicAdd(&s.stats[1], (unsigned long long)segments);
    }
}

__global__ void sppmClearGrid(SppmDev s) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= s.gridMask) s.head[i] = kGridNone;
    if (i == 0) {
        s.levels[0] = 0x7FFFFFFF;
        s.levels[1] = int(0x80000001u);
    }
}

// The visible points enter the hash (sppm_grid_kernel): each in the one cell of its level holding it,
// with the pixel's radius (set at its first visible point from the unfolded path length).
// Launched with whole waves (blocks of 256): every lane takes part in the level reduction.
__global__ void sppmGridKernel(SppmDev s) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    int lo = 0x7FFFFFFF, hi = int(0x80000001u);
    if (i < s.pixels && (__float_as_uint(s.vpD[i].w) & kVpValid)) {
        float4 p = s.vpP[i];
        float r = s.radius[i];
        if (r == 0.0f) {
            r = fmaxf(s.r0PerMetre * p.w, s.r0Floor);
            s.radius[i] = r;
        }
        p.w = r;
        s.vpP[i] = p;
        const int level = sppmLevel(r);
        const float size = ldexpf(1.0f, level);
        const uint32_t h = sppmHash(sppmCell(p.x, size), sppmCell(p.y, size), sppmCell(p.z, size), level, s.gridMask);
        s.next[i] = atomicExch(&s.head[h], i);
        lo = hi = level;
    }
    // The levels in use: one pair of atomics per wave rather than per point.
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, __shfl_xor(lo, off));
        hi = max(hi, __shfl_xor(hi, off));
    }
    if (__lane_id() == 0 && lo <= hi) {
        atomicMin(&s.levels[0], lo);
        atomicMax(&s.levels[1], hi);
    }
}

// The photon guide's CDF from its counts so far (sppm_guide_kernel): one block of 1024 threads, 16
// cells each. The counts are staged through LDS in coalesced reads, the per-thread sums scanned
// (Hillis–Steele), each cell's exact integer prefix normalised in double and stored as a float, and
// the CDF written back coalesced. The counts are stable here: the previous iteration's photons have
// finished and this one's have not started.
constexpr uint32_t kGuideThreads = 1024;
__global__ __launch_bounds__(kGuideThreads) void sppmGuideKernel(SppmDev s) {
    __shared__ uint32_t cells[kGuideCells];  // counts, then the CDF's float bits (64 KiB)
    __shared__ unsigned long long part[kGuideThreads];
    const uint32_t t = threadIdx.x, per = kGuideCells / kGuideThreads, base = t * per;
    for (uint32_t i = t; i < uint32_t(kGuideCells); i += kGuideThreads) cells[i] = s.guideCounts[i];
    __syncthreads();
    unsigned long long sum = 0;
    for (uint32_t i = 0; i < per; ++i) sum += cells[base + i];
    part[t] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < kGuideThreads; off <<= 1) {
        const unsigned long long v = t >= off ? part[t - off] : 0ull;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    const unsigned long long total = part[kGuideThreads - 1];
    const double tot = double(total);
    unsigned long long run = part[t] - sum;  // exclusive prefix of this thread's cells
    for (uint32_t i = 0; i < per; ++i) {
        run += cells[base + i];
        cells[base + i] = __float_as_uint(total > 0 ? float(double(run) / tot) : 0.0f);
    }
    __syncthreads();
    for (uint32_t i = t; i < uint32_t(kGuideCells); i += kGuideThreads) s.guideCdf[i] = __uint_as_float(cells[i]);
    if (t == 0) *s.guideTotal = float(total);
}

// The progressive update (sppm_update_kernel): N' = N + αM, r' = r·√(N'/(N+M)), τ' = (τ + Φ)·N'/(N+M);
// then the iteration's deposits are cleared. A pixel no photon reached keeps its state.
__global__ void sppmUpdateKernel(SppmDev s) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.pixels) return;
    const uint32_t M = s.count[i];
    if (M == 0) return;
    unsigned long long* bins = s.flux + size_t(i) * (3 * kSplatBins);
    double phi[3];
    for (int k = 0; k < 3; ++k) {
        double v = 0.0;
        for (int b = kSplatBins - 1; b >= 0; --b) v += ldexp(double(bins[k * kSplatBins + b]), kSplatExp0 + 32 * b);
        phi[k] = v;
    }
    double2 xy = s.tauXY[i], zn = s.tauZN[i];
    const double N = zn.y, Nn = N + double(s.alpha) * double(M), ratio = Nn / (N + double(M));
    xy.x = (xy.x + phi[0]) * ratio;
    xy.y = (xy.y + phi[1]) * ratio;
    zn.x = (zn.x + phi[2]) * ratio;
    zn.y = Nn;
    s.tauXY[i] = xy;
    s.tauZN[i] = zn;
    s.radius[i] = float(double(s.radius[i]) * sqrt(ratio));
    for (int k = 0; k < 3 * kSplatBins; ++k) bins[k] = 0ull;
    s.count[i] = 0u;
}

}  // namespace
}  // namespace owehip

// ---------------------------------------------------------------- vertex connection and merging

// The light subpaths' offsets in the cache, in three launches: per-tile sums of the vertex counts
// (kScanTile consecutive subpaths per block, read coalesced), an exclusive scan of the tile sums by
// one block, then each tile's own scan plus its tile's offset, with the subpath's staged vertices
// copied to their place in the cache.
constexpr uint32_t kScanBlock = 256, kScanPer = 4, kScanTile = kScanBlock * kScanPer;

// Exclusive prefix of x over the block's threads (thread order), and the block's total.
__device__ __forceinline__ uint32_t blockExclusiveSum(uint32_t x, uint32_t& total) {
    __shared__ uint32_t waveTotal[kScanBlock / 64];
    const uint32_t lane = __lane_id(), wave = threadIdx.x / 64;
    uint32_t inc = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(inc, off);
        if (lane >= uint32_t(off)) inc += y;
    }
    if (lane == 63) waveTotal[wave] = inc;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
#pragma unroll
    for (uint32_t w = 0; w < kScanBlock / 64; ++w) {
        const uint32_t t = waveTotal[w];
        if (w < wave) before += t;
        total += t;
    }
    return before + inc - x;
}

__global__ __launch_bounds__(kScanBlock) void vcmTileSumKernel(const uint32_t* count, uint32_t n, uint32_t* tileSum) {
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanPer; ++k)
        if (base + k < n) sum += count[base + k];
    uint32_t total;
    blockExclusiveSum(sum, total);
    if (threadIdx.x == 0) tileSum[blockIdx.x] = total;
}

// Exclusive scan of n values by one block of kScanThreads threads, each over a contiguous run;
// offset[n] receives the total.
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void vcmScanKernel(const uint32_t* count, uint32_t* offset, uint32_t n) {
    __shared__ uint32_t part[kScanThreads];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (n + kScanThreads - 1) / kScanThreads;
    const uint32_t lo = min(n, t * per), hi = min(n, lo + per);
    uint32_t sum = 0;
    for (uint32_t k = lo; k < hi; ++k) sum += count[k];
    part[t] = sum;
    __syncthreads();
    for (uint32_t step = 1; step < kScanThreads; step <<= 1) {  // inclusive Hillis–Steele scan of the runs
        const uint32_t add = t >= step ? part[t - step] : 0u;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    for (uint32_t k = lo; k < hi; ++k) {
        offset[k] = run;
        run += count[k];
    }
    if (t == kScanThreads - 1) offset[n] = part[t];
}

// Each subpath's offset (its tile's offset plus its place in the tile) and its staged vertices to
// their place in the cache; n = slots · N subpaths; offset[n] receives the total.
__global__ __launch_bounds__(kScanBlock) void vcmScanCompactKernel(VcmDev v, uint32_t n) {
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    uint32_t c[kScanPer], sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanPer; ++k) {
        c[k] = base + k < n ? v.pathCount[base + k] : 0u;
        sum += c[k];
    }
    uint32_t total;
    uint32_t o = v.tileOffset[blockIdx.x] + blockExclusiveSum(sum, total);
#pragma unroll
    for (uint32_t k = 0; k < kScanPer; ++k) {
        if (base + k >= n) break;
        v.offset[base + k] = o;
        const LightVertex* from = v.staging + size_t(base + k) * v.pathSlots;
        for (uint32_t i = 0; i < c[k]; ++i) v.cache[o + i] = from[i];
        o += c[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanBlock - 1) v.offset[n] = o;
}

// Group g's range of slot k's cache: thread 8k + g.
__global__ void vcmRangesKernel(VcmDev v) {
    const uint32_t k = threadIdx.x / 8, g = threadIdx.x % 8;
    if (k >= v.slots || g >= v.groups) return;
    const uint32_t p = k * v.slotPaths + g * v.pathsPerGroup;
    const uint32_t first = v.offset[p], last = v.offset[p + v.pathsPerGroup];
    v.counts[32 * k + 4 + g] = last - first;
    v.counts[32 * k + 16 + g] = first;
}

// The cached light vertices into their slot's merging hash, each at the level of the merging radius
// where it lies (vcm_grid_kernel.slang).
// A grid of at most kGridBlocks blocks strides over the vertices the slots hold (a few percent of the
// cache's capacity, which bounds them).
constexpr uint32_t kGridBlocks = 2048;
__global__ void vcmGridKernel(VcmDev v) {
    const uint32_t total = min(v.offset[v.slots * v.slotPaths], v.capacity);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const LightVertex& y = v.cache[i];
        const float4 q = y.p;
        const uint32_t slot = __float_as_uint(y.mis.w);
        if (slot >= v.slots) continue;  // cannot happen: the light kernel writes its own slot
        const VcmDev s = v.at(slot);
        const f3 p = mk3(q.x, q.y, q.z);
        const int level = sppmLevel(vcmRadius(s, p));
        const float size = ldexpf(1.0f, level);
        const uint32_t h = sppmHash(sppmCell(p.x, size), sppmCell(p.y, size), sppmCell(p.z, size), level, v.gridMask);
        v.next[i] = atomicExch(&s.head[h], i);
    }
}

// XYZ of a contribution per wavelength (Wavelengths::toXYZ), with the wavelengths the joined subpaths
// share: when the light side left only its hero, only the hero counts (joinedXYZ, vcm.slang:319-324).
__device__ __forceinline__ f3 vcmXYZ(const WL& camera, bool lightTerminated, const float v[NW], float invNWY) {
    WL w = camera;
    if (lightTerminated && !secondaryTerminated(w)) {
        w.pdf[0] /= NW;
        w.pdf[1] = w.pdf[2] = w.pdf[3] = 0.0f;
    }
    f3 c = mk3(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        if (w.pdf[k] == 0.0f || v[k] == 0.0f) continue;
        c = c + cieXYZ(w.lambda[k]) * (v[k] / w.pdf[k] * invNWY);
    }
    return c;
}

// The camera side of an iteration's light emission pdfs (vcm.slang:105-143).
__device__ __forceinline__ float vcmAreaEmissionPdf(const DScene& sc, const DBoundary& b, uint32_t li, float cosL) {
    const DEmission& e = sc.emissions[b.emission];
    const float side = e.front && e.back ? 0.5f : 1.0f;
    return (1.0f - sc.env.skyShare) * sc.lights[li].pdf / b.area * side * fabsf(cosL) * kInvPiF;
}
__device__ __forceinline__ float vcmDistantEmissionPdf(const DScene& sc, float sel, float pdfDir) {
    const float R = sc.env.emitRadius;
    return sel * pdfDir / (kPiF * R * R);
}

struct VcmCamera {
    const DScene& sc;
    const Settings& set;
    const VcmDev& v;          // the launch's (its slot's own values below)
    float perMetre;           // the slot's merging radius per metre from the eye (0: bdpt)
    const uint32_t* head;     // the slot's merging hash
    Rng rng;
    WL wl;
    float beta[NW];
    VcmMis mis;
    uint32_t* stack;
    uint32_t group, M, base;  // the pixel's group, its cache's vertex count and first vertex
    f3 L;               // XYZ so far but merging's
    long long Lm[3];    // merging's XYZ in film units: integer sums, whatever order the hash visits in
    bool bad;           // a merge gave a non-finite value (the sample is dropped, as for L)
};

// A non-specular camera vertex: what every strategy at it shares.
struct VcmVertex {
    Hit h;
    Iface it;
    f3 d;
    float eta;
    uint32_t face;
};

template <bool kContacts>
__device__ __forceinline__ bool vcmWalk(const VcmCamera& c, f3 o, f3 dir, float dist, uint32_t start, uint32_t departed,
                                        uint32_t face, uint32_t exclude, uint32_t arrival, float Tr[NW]) {
    return walkTransmittance<kContacts>(c.sc, o, dir, dist, start, departed, face, exclude, arrival, c.wl.lambda, c.stack,
                                        kBlock, Tr);
}

// Next-event estimation (s = 1; cameraNextEvent, vcm.slang:402-425): the light sampled as the path
// tracer samples it, weighted against the BSDF sample and the light subpaths that leave it.
template <bool kContacts>
__device__ __forceinline__ void vcmNextEvent(VcmCamera& c, const VcmVertex& x) {
    const DScene& sc = c.sc;
    if (sc.nLights == 0) return;
    const float ul = c.rng.uniform();
    uint32_t li = 0;
    while (li + 1 < sc.nLights && sc.lights[li].cdf <= ul) ++li;
    const DLight Lt = sc.lights[li];
    const float u1 = c.rng.uniform(), u2 = c.rng.uniform(), u3 = c.rng.uniform();
    f3 wi;
    float dist = __builtin_inff(), lpdf, cosAtLight = 1.0f, emissionPdf;
    float Le[NW];
    uint32_t arrival = kRegionNone, exclude = kNoBoundary;
    if (Lt.kind == 1) {
        const Frame f = makeFrame(ld3(sc.env.sunDir));
        wi = toWorld(f, uniformConeOmc(u1, u2, sc.env.sunOneMinusCos));
#pragma unroll
        for (int i = 0; i < NW; ++i) Le[i] = c.wl.pdf[i] != 0.0f ? evalSpec(sc.env.sunRadiance, sc.table, c.wl.lambda[i]) : 0.0f;
        lpdf = Lt.pdf / sc.env.sunSolidAngle;
        emissionPdf = vcmDistantEmissionPdf(sc, (1.0f - sc.env.skyShare) * Lt.pdf, 1.0f / sc.env.sunSolidAngle);
    } else {
        const DBoundary& b = sc.boundaries[Lt.boundary];
        f3 pl, nl;
        float pdfW;
        if (!sampleFrom(b, xfPoint(b.toLocal, x.h.p), u1, u2, u3, pl, nl, pdfW)) return;
        const f3 ep = xfPoint(b.toWorld, pl), en = norm3(xfVector(b.toWorld, nl));
        const f3 dv = ep - x.h.p;
        dist = len3(dv);
        if (!(dist > 0.0f)) return;
        wi = dv / dist;
        const DEmission& e = sc.emissions[b.emission];
        const float cosL = dot3(en, -wi);
        const bool lit = (cosL > 0.0f && e.front) || (cosL < 0.0f && e.back);
#pragma unroll
        for (int i = 0; i < NW; ++i) Le[i] = lit && c.wl.pdf[i] != 0.0f ? evalSpec(e.radiance, sc.table, c.wl.lambda[i]) : 0.0f;
        lpdf = Lt.pdf * pdfW;
        cosAtLight = fabsf(cosL);
        emissionPdf = vcmAreaEmissionPdf(sc, b, li, cosL);
        arrival = cosL > 0.0f ? b.front : b.back;
        exclude = Lt.boundary;
    }
    if (!(lpdf > 0.0f) || !(cosAtLight > 0.0f)) return;
    float f[NW], pdfFwd;
    if (!vcmEval(sc, x.it, x.h.boundary, c.wl, x.d, wi, true, f, pdfFwd)) return;
    float pdfRev;
    evalScatter(sc, x.it, -wi, -x.d, true, pdfRev);
    const float cosToLight = fabsf(dot3(x.h.n, wi));
    const float wLight = pdfFwd / lpdf;
    const float wCamera = emissionPdf * cosToLight / (lpdf * cosAtLight) * (x.eta + c.mis.dVCM + c.mis.dVC * pdfRev);
    const float w = 1.0f / (wLight + 1.0f + wCamera);
    const DBoundary& hb = sc.boundaries[x.h.boundary];
    const uint32_t start = dot3(wi, x.h.n) > 0.0f ? hb.front : hb.back;
    float Tr[NW];
    if (!vcmWalk<kContacts>(c, x.h.p, wi, dist, start, x.h.boundary, x.face, exclude, arrival, Tr)) return;
    float val[NW];
    const float s = cosToLight * w / lpdf;
#pragma unroll
    for (int i = 0; i < NW; ++i) val[i] = c.beta[i] * f[i] * Le[i] * Tr[i] * s;
    c.L = c.L + vcmXYZ(c.wl, false, val, c.v.invNWY);
}

// A connection to one light vertex y drawn uniformly from the group's cache of M (s, t ≥ 2;
// cameraConnect, vcm.slang:428-462).
template <bool kContacts>
__device__ __forceinline__ void vcmConnect(VcmCamera& c, const VcmVertex& x) {
    const DScene& sc = c.sc;
    const VcmDev& v = c.v;
    if (c.M == 0) return;
    const uint32_t pick = min(uint32_t(c.rng.uniform() * float(c.M)), c.M - 1);
    const LightVertex y = v.cache[c.base + pick];
    const f3 yp = mk3(y.p.x, y.p.y, y.p.z), yn = mk3(y.n.x, y.n.y, y.n.z), yd = mk3(y.d.x, y.d.y, y.d.z);
    const f3 dv = yp - x.h.p;
    const float dist2 = dot3(dv, dv);
    if (!(dist2 > 0.0f)) return;
    const float dist = sqrtf(dist2);
    const f3 dir = dv / dist;
    const uint32_t yflags = __float_as_uint(y.d.w), yb = __float_as_uint(y.p.w);
    const bool lightTerminated = (yflags & kLvTerminated) != 0u;
    WL wj = c.wl;
    if (lightTerminated && !secondaryTerminated(wj)) {
        wj.pdf[0] /= NW;
        wj.pdf[1] = wj.pdf[2] = wj.pdf[3] = 0.0f;
    }
    float fc[NW], pdfCamFwd;
    if (!vcmEval(sc, x.it, x.h.boundary, wj, x.d, dir, true, fc, pdfCamFwd)) return;
    const Iface hl = vcmIface(sc, yb, yn, wj.lambda[0], c.set.fresnelFloor);
    float fl[NW], pdfLightFwd;
    if (!vcmEval(sc, hl, yb, wj, yd, -dir, false, fl, pdfLightFwd)) return;
    float pdfCamRev, pdfLightRev;
    evalScatter(sc, x.it, -dir, -x.d, true, pdfCamRev);
    evalScatter(sc, hl, dir, -yd, true, pdfLightRev);
    const float cosX = fabsf(dot3(x.h.n, dir)), cosY = fabsf(dot3(yn, dir));
    const float camPdfA = pdfCamFwd * cosY / dist2, lightPdfA = pdfLightFwd * cosX / dist2;
    const float wLight = camPdfA * (vcmEtaAt(v, c.perMetre, yp) + y.n.w + y.mis.x * pdfLightRev);
    const float wCamera = lightPdfA * (x.eta + c.mis.dVCM + c.mis.dVC * pdfCamRev);
    const float w = 1.0f / (wLight + 1.0f + wCamera);
    const DBoundary& xb = sc.boundaries[x.h.boundary];
    const DBoundary& lb = sc.boundaries[yb];
    const uint32_t start = dot3(dir, x.h.n) > 0.0f ? xb.front : xb.back;
    const uint32_t arrival = dot3(-dir, yn) > 0.0f ? lb.front : lb.back;
    // The walk stops 1e-5 of the segment short of y (its own surface is the far end).
    float Tr[NW];
    if (!walkTransmittance<kContacts>(sc, x.h.p, dir, dist * (1.0f - 1e-5f), start, x.h.boundary, x.face, kNoBoundary,
                                      arrival, wj.lambda, c.stack, kBlock, Tr))
        return;
    const float g = cosX * cosY / dist2 * w * float(c.M) / v.lightPaths;
    const float yb4[NW] = {y.beta.x, y.beta.y, y.beta.z, y.beta.w};
    float val[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) val[i] = c.beta[i] * fc[i] * yb4[i] * fl[i] * Tr[i] * g;
    c.L = c.L + vcmXYZ(c.wl, lightTerminated, val, c.v.invNWY);
}

// Merging with the group's light vertices within the merging radius of the vertex (cameraMerge,
// vcm.slang:465-520).
__device__ __forceinline__ void vcmMerge(VcmCamera& c, const VcmVertex& x) {
    const DScene& sc = c.sc;
    const VcmDev& v = c.v;
    if (!(c.perMetre > 0.0f)) return;
    const f3 p = x.h.p;
    const float r = vcmRadiusAt(v, c.perMetre, p), r2 = r * r;
    const float invEta = 1.0f / x.eta;
    const float dx = len3(p);
    const float near = dx > 0.0f ? fmaxf(dx - r, 0.0f) / dx : 0.0f, far = dx > 0.0f ? (dx + r) / dx : 1.0f;
    // The levels light vertices within reach live on: their radii span those at the nearest and
    // farthest distance from the eye in reach (almost always one level; at most the top four are
    // searched, which leaves out only vertices within a merging radius of the eye itself).
    const int levelHi = sppmLevel(vcmRadiusAt(v, c.perMetre, p * far));
    const int levelLo = max(sppmLevel(vcmRadiusAt(v, c.perMetre, p * near)), levelHi - 3);
    for (int level = levelLo; level <= levelHi; ++level) {
        const float size = ldexpf(1.0f, level);
        const f3 q = p - mk3(0.5f, 0.5f, 0.5f) * size;
        const int bx = sppmCell(q.x, size), by = sppmCell(q.y, size), bz = sppmCell(q.z, size);
        for (int c8 = 0; c8 < 8; ++c8) {
            const int cx = bx + (c8 & 1), cy = by + ((c8 >> 1) & 1), cz = bz + ((c8 >> 2) & 1);
            for (uint32_t i = c.head[sppmHash(cx, cy, cz, level, v.gridMask)]; i != kGridNone; i = v.next[i]) {
                const float4 vp = v.cache[i].p;
                const f3 yp = mk3(vp.x, vp.y, vp.z);
                const f3 dp = yp - p;
                if (dot3(dp, dp) > r2) continue;
                if (sppmLevel(vcmRadiusAt(v, c.perMetre, yp)) != level || sppmCell(yp.x, size) != cx || sppmCell(yp.y, size) != cy ||
                    sppmCell(yp.z, size) != cz)
                    continue;
                const LightVertex y = v.cache[i];
                const f3 yn = mk3(y.n.x, y.n.y, y.n.z);
                if (fabsf(dot3(dp, x.h.n)) > 0.25f * r || dot3(yn, x.h.n) < 0.5f) continue;
                if (__float_as_uint(y.mis.z) != c.group) continue;
                const f3 wi = -mk3(y.d.x, y.d.y, y.d.z);
                const bool lightTerminated = (__float_as_uint(y.d.w) & kLvTerminated) != 0u;
                WL wj = c.wl;
                if (lightTerminated && !secondaryTerminated(wj)) {
                    wj.pdf[0] /= NW;
                    wj.pdf[1] = wj.pdf[2] = wj.pdf[3] = 0.0f;
                }
                float fc[NW], pdfFwd;
                if (!vcmEval(sc, x.it, x.h.boundary, wj, x.d, wi, true, fc, pdfFwd)) continue;
                float pdfRev;
                evalScatter(sc, x.it, -wi, -x.d, true, pdfRev);
                const float wLight = y.n.w * invEta + y.mis.y * pdfFwd;
                const float wCamera = c.mis.dVCM * invEta + c.mis.dVM * pdfRev;
                const float w = 1.0f / (wLight + 1.0f + wCamera);
                const float yb4[NW] = {y.beta.x, y.beta.y, y.beta.z, y.beta.w};
                float val[NW];
#pragma unroll
                for (int k = 0; k < NW; ++k) val[k] = c.beta[k] * fc[k] * yb4[k] * (w * invEta);
                const f3 m = vcmXYZ(c.wl, lightTerminated, val, c.v.invNWY);
                if (!isfinite(m.x + m.y + m.z)) {
                    c.bad = true;
                    continue;
                }
                c.Lm[0] += fixedPoint(m.x);
                c.Lm[1] += fixedPoint(m.y);
                c.Lm[2] += fixedPoint(m.z);
            }
        }
    }
}

// One camera path of pixel `pixel` in slot k's iteration (thread k·pixels + pixel; traceCameraVcm,
// vcm.slang:523-593); its contributions, every strategy but light tracing, go to the camera film.
template <bool kContacts>
__global__ __launch_bounds__(kBlock, OWE_HIP_VCM_WAVES) void vcmCameraKernel(DScene sc, Settings set, VcmDev vAll, unsigned long long* stats) {
    __shared__ uint32_t stack[48 * kBlock];
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    const DObserver& ob = sc.obs;
    const uint32_t pixels = uint32_t(ob.width) * uint32_t(ob.height);
    uint32_t nPaths = 0, nSegments = 0, nInconsistent = 0, nLeaks = 0;
    if (t < pixels * vAll.slots) {
        // The slot's own values are computed here, not copied with the rest of the launch's record,
        // which stays a kernel argument.
        const uint32_t slot = t / pixels, pixel = t - slot * pixels;
        const VcmDev& v = vAll;
        const uint32_t iteration = v.iteration + slot;
        const uint32_t* counts = v.counts + 32u * slot;
        VcmCamera c{sc, set, v, v.slotPerMetre(slot), v.slotHead(slot)};
        c.Lm[0] = c.Lm[1] = c.Lm[2] = 0;
        c.bad = false;
        c.stack = stack + threadIdx.x;
        c.rng.seed(hashPair(set.seed ^ kVcmCamera, pixel), iteration);
        c.group = vcmPixelGroup(pixel, iteration, v.groups);
        c.wl = sppmWavelengths(set, v.slotRotation(slot), c.group, v.groups);
        c.M = counts[4 + c.group];
        c.base = counts[16 + c.group];
        if (c.base > v.capacity || c.M > v.capacity - c.base) c.M = 0;  // cannot happen: the cache holds every vertex
        c.L = mk3(0.0f, 0.0f, 0.0f);
        PathStart ps;
        cameraRay(sc, pixel % uint32_t(ob.width), pixel / uint32_t(ob.width), c.rng, make_float4(0, 0, 0, 0),
                  make_float4(1, 1, 1, 1), ps);
        const f3 fwd = ld3(ob.fwd);
        const float cosCam = dot3(ps.d, fwd);
        const float W = float(ob.width), H = float(ob.height);
        const float pixelArea = (2.0f * ob.tanX / W) * (2.0f * ob.tanY / H);
        if (ps.weight > 0.0f && cosCam > 0.0f) {
            ++nPaths;
            const float cameraPdfW = 1.0f / (pixelArea * cosCam * cosCam * cosCam);
            c.mis = VcmMis{v.lightPaths / cameraPdfW, 0.0f, 0.0f};
#pragma unroll
            for (int i = 0; i < NW; ++i) c.beta[i] = c.wl.pdf[i] != 0.0f ? ps.weight : 0.0f;
            f3 o = ps.o, d = ps.d, prevP = ps.o;
            uint32_t region = ps.region, departed = ps.departed, face = ~0u;
            float segment = 0.0f;
            int pathLength = 0;
            const int maxDepth = set.maxDepth, rrDepth = set.rrDepth;
            for (int depth = 0;; ++depth) {
                Hit h;  // with its geometry: every strategy below reads the point and the normal
                const bool found = intersectContact<true, kContacts>(sc, o, d, __builtin_inff(), departed, face, kNoBoundary,
                                                                     region, c.stack, kBlock, h);
                ++nSegments;
                if (found && !(isfinite(h.p.x + h.p.y + h.p.z) && isfinite(h.n.x + h.n.y + h.n.z))) break;
                const DMedium& med = sc.media[sc.regionMedium[region]];
                if (med.flags) {  // extinction only: light scattered in a medium is the medium-only pass's
                    if (!found) break;
#pragma unroll
                    for (int i = 0; i < NW; ++i) {
                        const float l = c.wl.lambda[i];
                        c.beta[i] *= expf(-(evalSpec(med.absorption, sc.table, l) + evalSpec(med.scattering, sc.table, l)) * h.t);
                    }
                }
                if (!found) {
                    if (region == 0) {
                        // cameraSkyMet (vcm.slang:360-376): the sky and the sun, weighted after a vertex.
                        const f3 sd = ld3(sc.env.sunDir), dd = d - sd;
                        const bool inSun = sc.env.hasSun && 0.5f * dot3(dd, dd) <= sc.env.sunOneMinusCos;
                        float wSky = 1.0f, wSun = 1.0f;
                        if (pathLength > 0) {
                            wSky = 1.0f / (1.0f + vcmDistantEmissionPdf(sc, sc.env.skyShare, 1.0f / (4.0f * kPiF)) * c.mis.dVC);
                            float sunLightPdf = 0.0f;
                            for (uint32_t li = 0; li < sc.nLights; ++li)
                                if (sc.lights[li].kind == 1) sunLightPdf = sc.lights[li].pdf;
                            if (inSun && sc.env.sunSolidAngle > 0.0f)
                                wSun = 1.0f / (1.0f + sunLightPdf / sc.env.sunSolidAngle * c.mis.dVCM +
                                               vcmDistantEmissionPdf(sc, (1.0f - sc.env.skyShare) * sunLightPdf,
                                                                     1.0f / sc.env.sunSolidAngle) * c.mis.dVC);
                        }
                        float val[NW];
#pragma unroll
                        for (int i = 0; i < NW; ++i) {
                            const float sky = sc.env.sky == OWE_SKY_UNIFORM && c.wl.pdf[i] != 0.0f
                                                  ? evalSpec(sc.env.skyRadiance, sc.table, c.wl.lambda[i]) : 0.0f;
                            const float sun = inSun && c.wl.pdf[i] != 0.0f ? evalSpec(sc.env.sunRadiance, sc.table, c.wl.lambda[i]) : 0.0f;
                            val[i] = c.beta[i] * (sky * wSky + sun * wSun);
                        }
                        c.L = c.L + vcmXYZ(c.wl, false, val, v.invNWY);
                    } else {
                        ++nLeaks;
                    }
                    break;
                }
                const uint32_t bi = h.boundary;
                const DBoundary& b = sc.boundaries[bi];
                const bool fromFront = dot3(d, h.n) < 0.0f;
                const uint32_t inc = fromFront ? b.front : b.back, other = fromFront ? b.back : b.front;
                if (inc != region) {
                    ++nInconsistent;
                    noteInconsistency(sc, bi, region, inc, departed, h.p, o, d, h.slot);
                    region = inc;
                }
                segment += h.t;
                const DOptics& o2 = sc.optics[b.optics];
                const uint32_t hitFace = h.slot != kEmptyChild ? h.face : ~0u;
                if (o2.type == OWE_SURF_NULL) {
                    region = other;
                    o = h.p;
                    departed = bi;
                    face = hitFace;
                    if (depth >= 4 * maxDepth) break;
                    continue;
                }
                const float cosIn = fabsf(dot3(h.n, d));
                const float dist = segment;
                c.mis.arrive(dist, cosIn);
                segment = 0.0f;
                ++pathLength;
                // cameraEmissionMet (vcm.slang:380-399).
                if (b.emission >= 0) {
                    const DEmission& e = sc.emissions[b.emission];
                    if (fromFront ? e.front : e.back) {
                        float w = 1.0f;
                        if (pathLength > 1 && b.light >= 0) {
                            f3 pl, nl;
                            const f3 lo = xfPoint(b.toLocal, o), ld = xfVector(b.toLocal, d);
                            analyticGeometry(sc, b, lo, ld, h.t, pl, nl);
                            const float directPdfA = sc.lights[b.light].pdf * pdfFrom(b, xfPoint(b.toLocal, prevP), pl, nl) *
                                                     cosIn / (dist * dist);
                            const float emissionPdfW = vcmAreaEmissionPdf(sc, b, uint32_t(b.light), cosIn);
                            w = 1.0f / (1.0f + directPdfA * c.mis.dVCM + emissionPdfW * c.mis.dVC);
                        }
                        float val[NW];
#pragma unroll
                        for (int i = 0; i < NW; ++i)
                            val[i] = c.wl.pdf[i] != 0.0f ? c.beta[i] * evalSpec(e.radiance, sc.table, c.wl.lambda[i]) * w : 0.0f;
                        c.L = c.L + vcmXYZ(c.wl, false, val, v.invNWY);
                    }
                }
                if (o2.type == OWE_SURF_ABSORBER || o2.type == OWE_SURF_DETECTOR) break;
                if (depth >= maxDepth) break;
                const float hero = c.wl.lambda[0];
                const Iface it{&o2, h.n, regionIndex(sc, b.front, hero), regionIndex(sc, b.back, hero), hero,
                               set.fresnelFloor};
                // terminateIfDispersive (transport.slang:639-648).
                if (o2.type == OWE_SURF_DIELECTRIC && !secondaryTerminated(c.wl)) {
                    const float a0 = it.nBack / it.nFront;
#pragma unroll
                    for (int i = 1; i < NW; ++i) {
                        if (c.wl.pdf[i] == 0.0f) continue;
                        const Iface ai = vcmIface(sc, bi, h.n, c.wl.lambda[i], set.fresnelFloor);
                        if (fabsf(a0 - ai.nBack / ai.nFront) > 1e-12f * fabsf(a0)) {
                            c.wl.pdf[0] /= NW;
                            c.wl.pdf[1] = c.wl.pdf[2] = c.wl.pdf[3] = 0.0f;
                            c.beta[1] = c.beta[2] = c.beta[3] = 0.0f;
                            break;
                        }
                    }
                }
                const float eta = vcmEtaAt(v, c.perMetre, h.p);
                if (!o2.delta) {
                    const VcmVertex x{h, it, d, eta, hitFace};
                    vcmNextEvent<kContacts>(c, x);
                    vcmConnect<kContacts>(c, x);
                    vcmMerge(c, x);
                }
                const float u3 = c.rng.uniform(), u2 = c.rng.uniform(), u1 = c.rng.uniform();
                SSample s;
                if (!sampleScatter(sc, it, d, u1, u2, u3, true, s) || s.weight <= 0.0f) break;
                c.beta[0] *= s.weight;
#pragma unroll
                for (int i = 1; i < NW; ++i)
                    if (c.wl.pdf[i] != 0.0f)
                        c.beta[i] *= secondaryWeight(sc, vcmIface(sc, bi, h.n, c.wl.lambda[i], set.fresnelFloor), d, s, true);
                float pdfRev = 0.0f;
                if (!s.delta) evalScatter(sc, it, -s.wi, -d, true, pdfRev);
                c.mis.scatter(h.n, s, pdfRev, eta);
                if (s.transmitted) region = other;
                prevP = h.p;
                o = h.p;
                d = s.wi;
                departed = bi;
                face = hitFace;
                if (depth >= rrDepth) {
                    float q = 0.0f;
#pragma unroll
                    for (int i = 0; i < NW; ++i) q = fmaxf(q, c.beta[i]);
                    q = fminf(1.0f, q);
                    if (c.rng.uniform() >= q) break;
#pragma unroll
                    for (int i = 0; i < NW; ++i) c.beta[i] /= q;
                }
            }
            if (isfinite(c.L.x + c.L.y + c.L.z) && !c.bad) {
                unsigned long long* fp = v.film + 3 * size_t(pixel);
                const long long a = fixedPoint(c.L.x) + c.Lm[0], bb = fixedPoint(c.L.y) + c.Lm[1],
                                cc = fixedPoint(c.L.z) + c.Lm[2];
                if (a) atomicAdd(&fp[0], (unsigned long long)a);
                if (bb) atomicAdd(&fp[1], (unsigned long long)bb);
                if (cc) atomicAdd(&fp[2], (unsigned long long)cc);
            }
        }
    }
    const uint32_t n[4] = {nPaths, nSegments, nInconsistent, nLeaks};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint32_t t = n[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
        if (__lane_id() == 0 && t) atomicAdd(&stats[k], (unsigned long long)t);
    }
}

// ---------------------------------------------------------------- the renderer object

struct owe_hip_renderer {
    int device = 0;
    HostScene host;
    DScene dev{};
    Settings set{};
    std::vector<void*> allocations;
    Paths P{};
    uint32_t capacity = 0;  // path slots
    uint32_t queueCapacity = 0;  // wavefront paths in flight per chunk (≤ capacity; tuned)
    bool tunedFromCache = false;
    unsigned long long* film = nullptr;
    double* perPath = nullptr;
    uint32_t* pathPixels = nullptr;
    uint32_t* pathSamples = nullptr;
    uint32_t listCapacity = 0;
    int passes = 0;
    long long totalSpp = 0;
    int integrator = OWE_INTEGRATOR_PATH;
    // Light integrator: the exact splat bins, the pass's counters, and the particles emitted so far.
    unsigned long long* lightFilm = nullptr;
    unsigned long long* particleStats = nullptr;
    unsigned long long* particleNext = nullptr;  // the lane-refill particle form's id counter
    double particleCount = 0;
    uint64_t* particleIndex = nullptr;
    double* perParticle = nullptr;
    uint32_t particleListCapacity = 0;
    // SPPM: the device state, photons per iteration and iterations so far.
    owehip::SppmDev sppm{};
    float4* vpBuf[2][4] = {};           // two visible-point buffers: P, N, D, beta
    hipStream_t sppmStream = nullptr;   // the camera pass's stream
    hipEvent_t camDone = nullptr, photonsDone[2] = {nullptr, nullptr};
    hipEvent_t cleared = nullptr;       // the SPPM call's counters are zero (recorded on `stream`)
    uint64_t photons = 0;
    uint64_t sppmIterations = 0;
    owe_transport_stats stats{};
    // Vertex connection and merging: the light vertex cache, its counters and merging hash, light
    // subpaths per iteration (all groups), the merging radius per metre at the first iteration, and
    // the iterations so far.
    owehip::VcmDev vcm{};
    uint64_t vcmLightPaths = 0;
    uint32_t vcmGridSize = 0;
    float vcmRadiusPerMetre = 0.0f;
    float vcmAlpha = 2.0f / 3.0f;
    uint64_t vcmIterations = 0;
    uint32_t vcmSlotsMax = 1, vcmInFlight = 1;  // iterations the caches hold, and run per launch
    bool vcmMediumPass = false;  // scattering media: a medium-only camera pass beside each vcm pass
    uint32_t* firstInconsistency = nullptr;  // the first region inconsistency's record (kFirstWords words)
    hipStream_t stream = nullptr;
    uint32_t stackCap = 16;
    bool simpleShadows = true;
    float invNWY = 0;
    // Optional per-kernel timing with HIP events on the renderer's stream: generate, intersect,
    // shade, shadow, splat.
    int roundsPerSync = 4;     // bounce rounds enqueued between host reads of the live counters
    bool timing = false;
    hipEvent_t ev[2][6] = {};
    double kernelMs[6] = {0, 0, 0, 0, 0, 0};
    uint64_t kernelLaunches[6] = {0, 0, 0, 0, 0, 0};
    // The persistent form: mode, launch shape, its device counters and the captured pass graph.
    int mode = OWE_HIP_MODE_WAVEFRONT;
    int computeUnits = 256;
    int blocksPerCU = 2;
    int persistWaves = 2;
    uint32_t regen = 1;
    uint32_t chunk = 4;
    bool useGraph = false;
    unsigned long long* persistNext = nullptr;
    unsigned long long* persistSegments = nullptr;
    uint32_t* persistCounters = nullptr;
    uint32_t* passSlot = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    uint64_t graphTotal = 0;  // the pass size the graph was captured for
};

// A batch of renderers whose persistent passes run as one launch (owe_hip_batch_*).
struct owe_hip_batch {
    int device = 0;
    std::vector<owe_hip_renderer*> members;
    std::vector<void*> allocations;
    hipStream_t stream = nullptr;
    owehip::PersistBatch args{};
    owehip::DScene* scenes = nullptr;
    uint32_t stackCap = 1;
    bool contacts = false;
    int passes = 0;
    bool timing = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double kernelMs = 0;
    uint64_t launches = 0;
};

namespace owehip {
namespace {

template <class T>
T* devAlloc(owe_hip_renderer* r, size_t count) {
    void* p = nullptr;
    OWE_HIP_CHECK(hipMalloc(&p, std::max<size_t>(1, count) * sizeof(T)));
    r->allocations.push_back(p);
    return static_cast<T*>(p);
}
template <class T>
const T* upload(owe_hip_renderer* r, const std::vector<T>& v) {
    T* p = devAlloc<T>(r, v.size());
    if (!v.empty()) OWE_HIP_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
}

void allocPaths(owe_hip_renderer* r, uint32_t n) {
    Paths& P = r->P;
    P.ray0 = devAlloc<float4>(r, n);
    P.ray1 = devAlloc<float4>(r, n);
    P.beta = devAlloc<float4>(r, n);
    P.L = devAlloc<float4>(r, n);
    P.lam = devAlloc<float4>(r, n);
    P.prevP = devAlloc<float4>(r, n);
    P.info = devAlloc<uint4>(r, n);
    P.face = devAlloc<uint32_t>(r, n);
    P.rng = devAlloc<uint2>(r, n);
    P.hit = devAlloc<uint4>(r, n);
    P.shPend = devAlloc<float4>(r, n);
    P.shO = devAlloc<float4>(r, n);
    P.shD = devAlloc<float4>(r, n);
    P.shInfo = devAlloc<uint4>(r, n);
    P.shFace = devAlloc<uint32_t>(r, n);
    P.queue[0] = devAlloc<uint32_t>(r, n);
    P.queue[1] = devAlloc<uint32_t>(r, n);
    P.count = devAlloc<uint32_t>(r, 4);
    P.segments = devAlloc<unsigned long long>(r, 1);
    r->capacity = n;
}

dim3 gridFor(uint64_t n) { return dim3{unsigned((n + kBlock - 1) / kBlock)}; }

enum { kGenerate = 0, kIntersect = 1, kShade = 2, kShadow = 3, kSplat = 4, kPersistent = 5 };
constexpr int kTimers = 6;

// Brackets a launch with events when timing is on; the elapsed time is collected after the next
// stream synchronisation (collectTimes).
template <class F>
void timed(owe_hip_renderer* r, int which, F&& launch) {
    if (r->timing) OWE_HIP_CHECK(hipEventRecord(r->ev[0][which], r->stream));
    launch();
    if (r->timing) {
        OWE_HIP_CHECK(hipEventRecord(r->ev[1][which], r->stream));
        r->kernelLaunches[which]++;
    }
}
void collectTimes(owe_hip_renderer* r, std::initializer_list<int> which) {
    if (!r->timing) return;
    for (int k : which) {
        float ms = 0;
        OWE_HIP_CHECK(hipEventElapsedTime(&ms, r->ev[0][k], r->ev[1][k]));
        r->kernelMs[k] += ms;
    }
}

// One segment of every live path; the shade kernel's PCG increment is the pass's ((pass << 1) | 1).
template <int kCap, bool kContacts>
void launchBounce(owe_hip_renderer* r, int q, uint32_t upper, uint64_t passInc) {
    // upper: the live count at the host's last read, a bound on every later round's.
    const dim3 grid = gridFor(upper);
    timed(r, kIntersect, [&] {
        hipLaunchKernelGGL((intersectKernel<kCap, kContacts>), grid, dim3{kBlock}, 0, r->stream, r->dev, r->P, q);
    });
    timed(r, kShade, [&] {
        hipLaunchKernelGGL(shadeKernel, grid, dim3{kBlock}, 0, r->stream, r->dev, r->set, r->P, q, passInc);
    });
    timed(r, kShadow, [&] {
        hipLaunchKernelGGL((shadowKernel<kCap, kContacts>), grid, dim3{kBlock}, 0, r->stream, r->dev, r->set, r->P, q,
                           int(r->simpleShadows));
    });
    // The consumed queue's counter is cleared for the round after next (a kernel, not a memset).
    hipLaunchKernelGGL(clearCounts, dim3{1}, dim3{64}, 0, r->stream, r->P.count + q, 1);
}

// Runs count paths: slots [0, count) with global path ids base + slot, or the (pixel, sample) lists.
void runPaths(owe_hip_renderer* r, uint64_t base, uint32_t count, uint32_t spp, uint32_t pass, bool lists,
              double* perPath) {
    if (!r->P.count) allocPaths(r, r->capacity);
    const uint64_t passInc = (uint64_t(pass) << 1) | 1u;
    timed(r, kGenerate, [&] {
        hipLaunchKernelGGL(generateKernel, gridFor(count), dim3{kBlock}, 0, r->stream, r->dev, r->set, r->P, base, count,
                           spp, pass, lists ? r->pathPixels : nullptr, lists ? r->pathSamples : nullptr);
    });
    OWE_HIP_CHECK(hipGetLastError());
    bool generated = true;
    r->stats.paths += count;
    const uint64_t segmentsBefore = r->stats.segments;
    // Rounds are enqueued in groups without waiting: each kernel reads its live count on the device.
    // The host reads the counters back once per group to stop, and the bound it knew (`upper`, the
    // live count at the last read) sizes the grids.
    uint32_t upper = count;
    uint32_t hostCounts[4] = {count, 0, 0, 0};
    int q = 0;
    const int maxRounds = 4 * r->set.maxDepth + 4;
    // Every round raises each live path's depth by one and no path outlives depth 4·maxDepth (null
    // crossings); the round limit only bounds the loop should that ever fail.
    hipLaunchKernelGGL(setCount, dim3{1}, dim3{64}, 0, r->stream, r->P.count, r->P.segments, count);
    for (int round = 0; upper > 0 && round <= maxRounds;) {
        const int group = round < 2 ? 1 : r->roundsPerSync;
        for (int g = 0; g < group && round <= maxRounds; ++g, ++round) {
            if (r->dev.contacts) {
                if (r->stackCap <= 16) launchBounce<16, true>(r, q, upper, passInc);
                else if (r->stackCap <= 32) launchBounce<32, true>(r, q, upper, passInc);
                else launchBounce<48, true>(r, q, upper, passInc);
            } else {
                if (r->stackCap <= 16) launchBounce<16, false>(r, q, upper, passInc);
                else if (r->stackCap <= 32) launchBounce<32, false>(r, q, upper, passInc);
                else launchBounce<48, false>(r, q, upper, passInc);
            }
            OWE_HIP_CHECK(hipGetLastError());
            q = 1 - q;
            if (r->timing) {
                OWE_HIP_CHECK(hipStreamSynchronize(r->stream));
                if (generated) collectTimes(r, {kGenerate});
                generated = false;
                collectTimes(r, {kIntersect, kShade, kShadow});
            }
        }
        OWE_HIP_CHECK(hipMemcpyAsync(hostCounts, r->P.count, sizeof hostCounts, hipMemcpyDeviceToHost, r->stream));
        OWE_HIP_CHECK(hipStreamSynchronize(r->stream));
        if (generated) collectTimes(r, {kGenerate});
        generated = false;
        upper = hostCounts[q];
    }
    r->stats.inconsistencies += hostCounts[2];
    r->stats.leaks += hostCounts[3];
    {
        unsigned long long segs = 0;
        OWE_HIP_CHECK(hipMemcpyAsync(&segs, r->P.segments, sizeof segs, hipMemcpyDeviceToHost, r->stream));
        OWE_HIP_CHECK(hipStreamSynchronize(r->stream));
        r->stats.segments = segmentsBefore + segs;
    }
    timed(r, kSplat, [&] {
        hipLaunchKernelGGL(splatKernel, gridFor(count), dim3{kBlock}, 0, r->stream, r->set, r->P, count, r->invNWY, r->film,
                           perPath);
    });
    OWE_HIP_CHECK(hipGetLastError());
    if (r->timing) {
        OWE_HIP_CHECK(hipStreamSynchronize(r->stream));
        if (generated) collectTimes(r, {kGenerate});
        collectTimes(r, {kSplat});
    }
}

__global__ void setPass(uint32_t* slot, uint32_t pass) {
    if (threadIdx.x == 0) *slot = pass;
}

Persist persistArgs(owe_hip_renderer* r, uint64_t total, uint32_t spp) {
    Persist a{};
    a.total = total;
    a.spp = spp;
    a.chunk = r->chunk;
    a.regen = r->regen;
    a.pass = r->passSlot;
    a.next = r->persistNext;
    a.film = r->film;
    a.segments = r->persistSegments;
    a.counters = r->persistCounters;
    a.invNWY = r->invNWY;
    a.simple = int(r->simpleShadows);
    return a;
}

void launchPersistent(owe_hip_renderer* r, const Persist& a) {
    hipLaunchKernelGGL(persistBegin, dim3{1}, dim3{64}, 0, r->stream, a);
    const dim3 grid{unsigned(std::max(1, r->computeUnits * r->blocksPerCU))};
    auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, grid, dim3{kBlock}, 0, r->stream, r->dev, r->set, a); };
    // One LDS stack size (48 entries, every scene's bound) for the persistent kernels: registers hold
    // them to 2 or 3 waves per SIMD, i.e. 2 or 3 blocks per compute unit, and 3 × 48 KiB of stacks
    // fit its 160 KiB of LDS, so a smaller stack would not let more waves in.
    const bool c = r->dev.contacts != 0u;
    if (r->persistWaves == 3) c ? go(persistentKernel<48, 3, true>) : go(persistentKernel<48, 3, false>);
    else c ? go(persistentKernel<48, 2, true>) : go(persistentKernel<48, 2, false>);
}

void launchPersistentBatch(owe_hip_batch* b, const PersistBatch& a, int waves, int blocksPerCU, int computeUnits) {
    hipLaunchKernelGGL(persistBatchBegin, dim3{1}, dim3{256}, 0, b->stream, a);
    const dim3 grid{unsigned(std::max(1, computeUnits * blocksPerCU))};
    auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, grid, dim3{kBlock}, 0, b->stream, b->scenes, b->members[0]->set, a); };
    const bool c = b->contacts;  // the persistent form's one stack size, as launchPersistent
    if (waves == 3) c ? go(persistentBatchKernel<48, 3, true>) : go(persistentBatchKernel<48, 3, false>);
    else c ? go(persistentBatchKernel<48, 2, true>) : go(persistentBatchKernel<48, 2, false>);
}

template <class T>
T* batchAlloc(owe_hip_batch* b, size_t count) {
    void* p = nullptr;
    OWE_HIP_CHECK(hipMalloc(&p, std::max<size_t>(1, count) * sizeof(T)));
    b->allocations.push_back(p);
    return static_cast<T*>(p);
}
template <class T>
T* batchUpload(owe_hip_batch* b, const std::vector<T>& v) {
    T* p = batchAlloc<T>(b, v.size());
    if (!v.empty()) OWE_HIP_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
}

void dropGraph(owe_hip_renderer* r) {
    if (r->graphExec) OWE_HIP_CHECK(hipGraphExecDestroy(r->graphExec));
    if (r->graph) OWE_HIP_CHECK(hipGraphDestroy(r->graph));
    r->graphExec = nullptr;
    r->graph = nullptr;
    r->graphTotal = 0;
}

// One pass of the persistent form: the pass index goes to the device, then the counters are cleared
// and the persistent grid drains the pass's path ids (both in one captured graph when enabled; the
// graph holds kernels only, no memset nodes).
void runPersistent(owe_hip_renderer* r, uint32_t spp, uint32_t pass) {
    const uint64_t total = uint64_t(r->host.obs.width) * uint64_t(r->host.obs.height) * uint64_t(spp);
    const Persist a = persistArgs(r, total, spp);
    hipLaunchKernelGGL(setPass, dim3{1}, dim3{64}, 0, r->stream, r->passSlot, pass);
    timed(r, kPersistent, [&] {
        if (!r->useGraph) {
            launchPersistent(r, a);
            return;
        }
        if (r->graphExec && r->graphTotal != total) dropGraph(r);
        if (!r->graphExec) {
            OWE_HIP_CHECK(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
            launchPersistent(r, a);
            OWE_HIP_CHECK(hipStreamEndCapture(r->stream, &r->graph));
            OWE_HIP_CHECK(hipGraphInstantiate(&r->graphExec, r->graph, nullptr, nullptr, 0));
            r->graphTotal = total;
        }
        OWE_HIP_CHECK(hipGraphLaunch(r->graphExec, r->stream));
    });
    OWE_HIP_CHECK(hipGetLastError());
    unsigned long long segs = 0;
    uint32_t counters[4] = {0, 0, 0, 0};
    OWE_HIP_CHECK(hipMemcpyAsync(&segs, r->persistSegments, sizeof segs, hipMemcpyDeviceToHost, r->stream));
    OWE_HIP_CHECK(hipMemcpyAsync(counters, r->persistCounters, sizeof counters, hipMemcpyDeviceToHost, r->stream));
    OWE_HIP_CHECK(hipStreamSynchronize(r->stream));
    if (r->timing) collectTimes(r, {kPersistent});
    r->stats.paths += total;
    r->stats.segments += segs;
    r->stats.inconsistencies += counters[2];
    r->stats.leaks += counters[3];
}

// One light-tracing pass of `total` particles (W·H·ppp), in launches of at most `capacity`; with a
// particle list (device) and perParticle, the listed particles only, summed per particle.
void runParticles(owe_hip_renderer* r, uint64_t total, uint32_t pass, const uint64_t* list = nullptr,
                  double* perParticle = nullptr) {
    if (!list) r->particleCount += double(total);
    if (!(r->host.emittedPower > 0)) return;  // nothing emits: the particle estimate is zero
    OWE_HIP_CHECK(hipMemsetAsync(r->particleStats, 0, 4 * sizeof(unsigned long long), r->stream));
    if (r->mode == OWE_HIP_MODE_PERSISTENT && !list && !perParticle) {
        // The lane-refill form: the whole pass in one launch on a grid sized to the chip.
        ParticleBatch b{};
        b.total = total;
        b.pass = pass;
        b.film = r->lightFilm;
        b.stats = r->particleStats;
        b.invNWY = r->invNWY;
        b.next = r->particleNext;
        b.chunk = r->chunk;
        b.regen = r->regen;
        OWE_HIP_CHECK(hipMemsetAsync(r->particleNext, 0, sizeof(unsigned long long), r->stream));
        timed(r, kShade, [&] {
            launchParticlesRefill(r->dev, r->set, b, uint32_t(std::max(1, r->computeUnits * r->blocksPerCU)), r->stream);
        });
    } else {
        // One thread per particle, in launches of at most `capacity` particles.
        for (uint64_t base = 0; base < total; base += r->capacity) {
            ParticleBatch b{};
            b.base = list ? 0 : base;
            b.index = list ? list + base : nullptr;
            b.count = uint32_t(std::min<uint64_t>(r->capacity, total - base));
            b.pass = pass;
            b.film = r->lightFilm;
            b.perParticle = perParticle ? perParticle + 3 * base : nullptr;
            b.stats = r->particleStats;
            b.invNWY = r->invNWY;
            timed(r, kShade, [&] { launchParticles(r->dev, r->set, b, r->stackCap, r->stream); });
        }
    }
    unsigned long long st[4];
    OWE_HIP_CHECK(hipMemcpyAsync(st, r->particleStats, sizeof st, hipMemcpyDeviceToHost, r->stream));
    OWE_HIP_CHECK(hipStreamSynchronize(r->stream));
    if (r->timing) collectTimes(r, {kShade});
    if (list) return;
    r->stats.paths += st[0];
    r->stats.segments += st[1];
    r->stats.inconsistencies += st[2];
    r->stats.leaks += st[3];
}

// An iteration's base wavelength rotation: a golden-ratio sequence offset by the seed (the rotation of
// GpuRenderer::wavelengthRotation, gpu_renderer_passes.cpp:52-56, from this backend's seed hash).
double sppmRotation(uint64_t seed, uint64_t iteration) {
    const double offset = double(uint32_t(mix64(seed) >> 32)) / 4294967296.0;
    const double u = offset + 0.6180339887498949 * double(iteration + 1);
    return u - std::floor(u);
}

template <int kCap>
void launchSppmCamera(owe_hip_renderer* r) {
    const dim3 grid{(r->sppm.pixels + kBlock - 1) / kBlock}, block{kBlock};
    if (r->dev.contacts)
        hipLaunchKernelGGL((sppmCameraKernel<kCap, true>), grid, block, 0, r->sppmStream, r->dev, r->set, r->sppm,
                           r->film, r->persistCounters, r->simpleShadows ? 1 : 0);
    else
        hipLaunchKernelGGL((sppmCameraKernel<kCap, false>), grid, block, 0, r->sppmStream, r->dev, r->set, r->sppm,
                           r->film, r->persistCounters, r->simpleShadows ? 1 : 0);
    OWE_HIP_CHECK(hipGetLastError());
}

// Iterations of SPPM (GpuRenderer::passSppm, gpu_renderer_passes.cpp:259-290): per iteration the hash
// is cleared, then the camera pass, the hash fill, the photons and the update run in stream order.
// The call's statistics and counters are cleared on the render stream, and the camera stream waits
// for that clear before its first pass (its other waits are on events of earlier passes, recorded
// before the clear), so no camera-pass count lands in a counter that is then zeroed.
// Vertex connection and merging: per iteration the light subpaths stage their vertices (and splat to
// the eye), the scan and compaction order them into the cache, the cached vertices enter the merging
// hash (vcm), then one camera path per pixel (GpuRenderer::runVcmIterations,
// gpu_renderer_passes.cpp:136-199). The cache holds every vertex a subpath can leave, so it never
// overflows.
void runVcm(owe_hip_renderer* r, int iterations) {
    VcmDev v = r->vcm;
    const bool merging = r->vcmRadiusPerMetre > 0.0f;
    const uint32_t N = v.slotPaths;
    const uint32_t pixels = uint32_t(r->host.obs.width) * uint32_t(r->host.obs.height);
    OWE_HIP_CHECK(hipMemsetAsync(r->particleStats, 0, 4 * sizeof(unsigned long long), r->stream));
    for (int it = 0; it < iterations;) {
        const uint32_t S = std::min<uint32_t>(r->vcmInFlight, uint32_t(iterations - it));
        const uint64_t first = r->vcmIterations + uint64_t(it);
        v.iteration = uint32_t(first);
        v.slots = S;
        v.rotationOffset = double(uint32_t(mix64(r->set.seed) >> 32)) / 4294967296.0;  // as sppmRotation
        v.mergeBase = merging ? r->vcmRadiusPerMetre : 0.0f;
        v.mergeExponent = (double(r->vcmAlpha) - 1.0) / 2.0;
        OWE_HIP_CHECK(hipMemsetAsync(v.counts, 0, 32 * S * sizeof(uint32_t), r->stream));
        if (merging)
            OWE_HIP_CHECK(hipMemsetAsync(v.head, 0xFF, size_t(S) * r->vcmGridSize * sizeof(uint32_t), r->stream));
        const uint32_t paths = S * N;
        if (r->host.emittedPower > 0) {
            const uint32_t tiles = (paths + kScanTile - 1) / kScanTile;
            timed(r, kGenerate,
                  [&] { launchVcmLight(r->dev, r->set, v, r->lightFilm, r->particleStats, r->invNWY, paths, r->stream); });
            hipLaunchKernelGGL(vcmTileSumKernel, dim3{tiles}, dim3{kScanBlock}, 0, r->stream, v.pathCount, paths, v.tileSum);
            hipLaunchKernelGGL(vcmScanKernel, dim3{1}, dim3{kScanThreads}, 0, r->stream, v.tileSum, v.tileOffset, tiles);
            hipLaunchKernelGGL(vcmScanCompactKernel, dim3{tiles}, dim3{kScanBlock}, 0, r->stream, v, paths);
            hipLaunchKernelGGL(vcmRangesKernel, dim3{1}, dim3{8 * kVcmSlots}, 0, r->stream, v);
            if (merging) {
                const uint64_t n = uint64_t(paths) * v.pathSlots;
                const unsigned blocks = unsigned(std::min<uint64_t>(kGridBlocks, (n + 255) / 256));
                hipLaunchKernelGGL(vcmGridKernel, dim3{blocks}, dim3{256}, 0, r->stream, v);
            }
        }
        timed(r, kShade, [&] {
            if (r->dev.contacts)
                hipLaunchKernelGGL((vcmCameraKernel<true>), gridFor(S * pixels), dim3{kBlock}, 0, r->stream, r->dev, r->set,
                                   v, r->particleStats);
            else
                hipLaunchKernelGGL((vcmCameraKernel<false>), gridFor(S * pixels), dim3{kBlock}, 0, r->stream, r->dev, r->set,
                                   v, r->particleStats);
        });
        OWE_HIP_CHECK(hipGetLastError());
        if (r->timing) {
            OWE_HIP_CHECK(hipStreamSynchronize(r->stream));
            if (r->host.emittedPower > 0) collectTimes(r, {kGenerate});
            collectTimes(r, {kShade});
        }
        it += int(S);
    }
    unsigned long long st[4];
    OWE_HIP_CHECK(hipMemcpyAsync(st, r->particleStats, sizeof st, hipMemcpyDeviceToHost, r->stream));
    OWE_HIP_CHECK(hipStreamSynchronize(r->stream));
    r->vcmIterations += uint64_t(iterations);
    r->particleCount += double(r->vcmLightPaths) * double(iterations);
    r->stats.paths += st[0];
    r->stats.segments += st[1];
    r->stats.inconsistencies += st[2];
    r->stats.leaks += st[3];
}

void runSppm(owe_hip_renderer* r, int iterations) {
    SppmDev& q = r->sppm;
    OWE_HIP_CHECK(hipMemsetAsync(q.stats, 0, 4 * sizeof(unsigned long long), r->stream));
    OWE_HIP_CHECK(hipMemsetAsync(r->persistCounters, 0, 4 * sizeof(uint32_t), r->stream));
    OWE_HIP_CHECK(hipEventRecord(r->cleared, r->stream));
    OWE_HIP_CHECK(hipStreamWaitEvent(r->sppmStream, r->cleared, 0));
    const dim3 block{256}, pixelGrid{(q.pixels + 255) / 256}, gridGrid{(q.gridMask + 256) / 256};
    for (int k = 0; k < iterations; ++k) {
        q.iteration = uint32_t(r->sppmIterations);
        q.rotation = float(sppmRotation(r->set.seed, r->sppmIterations));
        const int b = int(r->sppmIterations & 1u);
        q.vpP = r->vpBuf[b][0];
        q.vpN = r->vpBuf[b][1];
        q.vpD = r->vpBuf[b][2];
        q.vpBeta = r->vpBuf[b][3];
        // The camera pass on its own stream, once the photons that last read buffer b are done: it
        // overlaps the previous iteration's photons and update.
        OWE_HIP_CHECK(hipStreamWaitEvent(r->sppmStream, r->photonsDone[b], 0));
        if (r->stackCap <= 16) launchSppmCamera<16>(r);
        else if (r->stackCap <= 32) launchSppmCamera<32>(r);
        else launchSppmCamera<48>(r);
        OWE_HIP_CHECK(hipEventRecord(r->camDone, r->sppmStream));
        hipLaunchKernelGGL(sppmClearGrid, gridGrid, block, 0, r->stream, q);
        OWE_HIP_CHECK(hipStreamWaitEvent(r->stream, r->camDone, 0));
        hipLaunchKernelGGL(sppmGridKernel, pixelGrid, block, 0, r->stream, q);
        hipLaunchKernelGGL(sppmGuideKernel, dim3{1}, dim3{kGuideThreads}, 0, r->stream, q);
        if (r->host.emittedPower > 0)
            for (uint64_t base = 0; base < r->photons; base += r->capacity)
                launchSppmPhotons(r->dev, r->set, q, base, uint32_t(std::min<uint64_t>(r->capacity, r->photons - base)),
                                  r->stackCap, r->stream);
        hipLaunchKernelGGL(sppmUpdateKernel, pixelGrid, block, 0, r->stream, q);
        OWE_HIP_CHECK(hipGetLastError());
        OWE_HIP_CHECK(hipEventRecord(r->photonsDone[b], r->stream));
        r->sppmIterations++;
    }
    OWE_HIP_CHECK(hipStreamSynchronize(r->sppmStream));
    unsigned long long st[4];
    uint32_t counters[4];
    OWE_HIP_CHECK(hipMemcpyAsync(st, q.stats, sizeof st, hipMemcpyDeviceToHost, r->stream));
    OWE_HIP_CHECK(hipMemcpyAsync(counters, r->persistCounters, sizeof counters, hipMemcpyDeviceToHost, r->stream));
    OWE_HIP_CHECK(hipStreamSynchronize(r->stream));
    // Photons of a scene where nothing emits are still emitted (and carry nothing).
    const uint64_t silent = r->host.emittedPower > 0 ? 0 : r->photons * uint64_t(iterations);
    r->stats.paths += st[0] + silent;
    r->stats.segments += st[1];
    r->stats.inconsistencies += st[2] + counters[2];
    r->stats.leaks += st[3] + counters[3];
}

}  // namespace
}  // namespace owehip

using namespace owehip;

extern "C" {

int owe_hip_renderer_create(int device, const owe_scene* scene, const owe_render_settings* settings,
                            owe_hip_renderer** out) {
    owe_hip_renderer* r = nullptr;
    const int st = guard([&] {
        require(out != nullptr && scene != nullptr && settings != nullptr, "null argument");
        *out = nullptr;
        const owe_render_settings& s = *settings;
        require(s.max_depth >= 0 && s.rr_depth >= 0, "depths must be non-negative");
        require(s.max_depth <= 4096, "max depth above 4096 is not supported (depth is kept in 16 bits)");
        if (s.integrator != OWE_INTEGRATOR_PATH && s.integrator != OWE_INTEGRATOR_LIGHT && s.integrator != OWE_INTEGRATOR_SPPM &&
            s.integrator != OWE_INTEGRATOR_VCM && s.integrator != OWE_INTEGRATOR_BDPT && s.integrator != OWE_INTEGRATOR_HYBRID)
            fail(OWE_HIP_ERR_UNSUPPORTED,
                 "unknown integrator (this backend implements path, light, sppm, vcm, bdpt and hybrid)");
        const bool vcmLike = s.integrator == OWE_INTEGRATOR_VCM || s.integrator == OWE_INTEGRATOR_BDPT;
        if (vcmLike) {
            require(s.photon_radius >= 0 && s.photon_alpha >= 0 && s.photons_per_pixel >= 0 && s.photon_groups >= 0,
                    "photon settings must not be negative");
            require((s.photon_alpha > 0 ? s.photon_alpha : 2.0 / 3.0) <= 1.0, "photon_alpha must lie in (0, 1]");
            require(s.photon_groups <= 8, "vcm wavelength groups must be 1 to 8");
        }
        const double photonRadius = s.photon_radius > 0 ? s.photon_radius : 2.0;
        const double photonAlpha = s.photon_alpha > 0 ? s.photon_alpha : 2.0 / 3.0;
        const double photonsPerPixel = s.photons_per_pixel > 0 ? s.photons_per_pixel : 1.0;
        const int photonGroups = s.photon_groups > 0 ? s.photon_groups : 4;
        if (s.integrator == OWE_INTEGRATOR_SPPM) {
            require(s.photon_radius >= 0 && s.photon_alpha >= 0 && s.photons_per_pixel >= 0 && s.photon_groups >= 0,
                    "photon settings must not be negative");
            require(photonAlpha <= 1.0, "photon_alpha must lie in (0, 1]");
            require(photonGroups <= 255, "photon wavelength groups must be 1 to 255");
        }
        require(s.spectral == OWE_SPECTRAL_HERO4 || s.spectral == OWE_SPECTRAL_UNIFORM1, "unknown spectral mode");
        if (s.spectral == OWE_SPECTRAL_UNIFORM1)
            require(s.lambda_min >= 360.0 && s.lambda_max <= 830.0 && s.lambda_min < s.lambda_max,
                    "the uniform wavelength range must lie inside [360, 830] nm");
        require(s.fresnel_floor >= 0 && s.fresnel_floor < 0.5, "fresnel floor must lie in [0, 0.5)");
        selectDevice(device);
        r = new owe_hip_renderer;
        r->device = device;
        r->integrator = s.integrator;
        r->host = flattenScene(*scene);
        const HostScene& h = r->host;
        if (h.bvh.stackBound > 48)
            fail(OWE_HIP_ERR_UNSUPPORTED, "scene hierarchy needs a traversal stack of " + std::to_string(h.bvh.stackBound) +
                                              " entries; the kernels hold at most 48");
        if (vcmLike) {
            if (h.obs.kind != OWE_DETECTOR_OBSERVER)
                fail(OWE_HIP_ERR_UNSUPPORTED, "vcm and bdpt render for an ideal observer (the reference's GPU backend too)");
            // With scattering media, vcm and bdpt take the paths without a medium scattering vertex
            // (media as extinction) and camera paths, in a pass of their own, all the others
            // (gpu_renderer.cpp:43-47, 171).
            for (const DMedium& m : h.media) r->vcmMediumPass = r->vcmMediumPass || (m.flags & 2) != 0;
        }
        if (s.integrator == OWE_INTEGRATOR_HYBRID && h.obs.kind != OWE_DETECTOR_OBSERVER)
            fail(OWE_HIP_ERR_UNSUPPORTED,
                 "the hybrid integrator needs a virtual observer; use path or light for surface sensors (renderer.hpp:34-35)");
        r->stackCap = std::max<uint32_t>(1, h.bvh.stackBound);
        OWE_HIP_CHECK(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
        DScene& d = r->dev;
        d.media = upload(r, h.media);
        d.regionMedium = upload(r, h.regionMedium);
        d.optics = upload(r, h.optics);
        d.emissions = upload(r, h.emissions);
        d.boundaries = upload(r, h.boundaries);
        d.lights = upload(r, h.lights);
        d.table = upload(r, h.table);
        d.waveData = reinterpret_cast<const float4*>(upload(r, h.waveData));
        if (h.hasTriangles) {
            d.nodes = upload(r, h.bvh.nodes);
            d.tris = upload(r, h.bvh.tris);
        }
        d.nBoundaries = uint32_t(h.boundaries.size());
        d.nLights = uint32_t(h.lights.size());
        d.nRegions = uint32_t(h.regionMedium.size());
        d.stackCap = r->stackCap;
        d.contacts = h.contacts ? 1u : 0u;
        r->firstInconsistency = devAlloc<uint32_t>(r, kFirstWords);
        OWE_HIP_CHECK(hipMemset(r->firstInconsistency, 0, kFirstWords * sizeof(uint32_t)));
        d.firstInconsistency = r->firstInconsistency;
        d.env = h.env;
        d.obs = h.obs;
        bool anyNull = false, anyMedium = false;
        for (const DBoundary& b : h.boundaries) anyNull = anyNull || h.optics[size_t(b.optics)].type == OWE_SURF_NULL;
        for (const DMedium& m : h.media) anyMedium = anyMedium || m.flags != 0;
        r->simpleShadows = !anyNull && !anyMedium;
        r->set.seed = s.seed;
        r->set.maxDepth = s.max_depth;
        r->set.rrDepth = s.rr_depth;
        r->set.spectral = s.spectral;
        r->set.lambdaMin = float(s.lambda_min);
        r->set.lambdaMax = float(s.lambda_max);
        r->set.fresnelFloor = float(s.fresnel_floor);
        r->set.mode = s.integrator == OWE_INTEGRATOR_HYBRID ? kModePartition : r->vcmMediumPass ? kModeMediumOnly : 0;
        // 1 / (NW · ∫ȳ) with ∫ȳ by composite Simpson over [360, 830] nm at 0.5 nm (core/spectrum.cpp:11-23).
        {
            auto cy = [](double l) {
                auto lobe = [](double x, double mu, double s1, double s2) {
                    const double t = (x - mu) / (x < mu ? s1 : s2);
                    return std::exp(-0.5 * t * t);
                };
                return 0.821 * lobe(l, 568.8, 46.9, 40.5) + 0.286 * lobe(l, 530.9, 16.3, 31.1);
            };
            const int n = 940;
            const double hh = 470.0 / n;
            double sum = cy(360.0) + cy(830.0);
            for (int i = 1; i < n; ++i) sum += cy(360.0 + i * hh) * ((i & 1) ? 4 : 2);
            r->invNWY = float(1.0 / (NW * (sum * hh / 3)));
        }
        const uint64_t pixels = uint64_t(h.obs.width) * uint64_t(h.obs.height);
        r->film = devAlloc<unsigned long long>(r, 3 * pixels);
        OWE_HIP_CHECK(hipMemset(r->film, 0, 3 * pixels * sizeof(unsigned long long)));
        if (r->integrator == OWE_INTEGRATOR_PATH || r->integrator == OWE_INTEGRATOR_HYBRID) {
            // The wavefront form's path slots are allocated at its first pass (runPaths): a renderer
            // that only runs the persistent form, or only as a batch member, never holds them. The
            // hybrid integrator's particles (same capacity per launch) splat to the particle film.
            r->capacity = uint32_t(std::min<uint64_t>(pixels * 64, 1u << 22));
            if (r->integrator == OWE_INTEGRATOR_HYBRID) {
                r->lightFilm = devAlloc<unsigned long long>(r, 3 * kSplatBins * pixels);
                OWE_HIP_CHECK(hipMemset(r->lightFilm, 0, 3 * kSplatBins * pixels * sizeof(unsigned long long)));
                r->particleStats = devAlloc<unsigned long long>(r, 4);
                r->particleNext = devAlloc<unsigned long long>(r, 1);
            }
            hipDeviceProp_t prop;
            OWE_HIP_CHECK(hipGetDeviceProperties(&prop, device));
            r->computeUnits = prop.multiProcessorCount;
            r->persistNext = devAlloc<unsigned long long>(r, 1);
            r->persistSegments = devAlloc<unsigned long long>(r, 1);
            r->persistCounters = devAlloc<uint32_t>(r, 4);
            r->passSlot = devAlloc<uint32_t>(r, 1);
        } else if (r->integrator == OWE_INTEGRATOR_SPPM) {
            require(pixels <= (1u << 30), "SPPM keeps one visible point per pixel: at most 2^30 pixels");
            SppmDev& q = r->sppm;
            q.pixels = uint32_t(pixels);
            for (auto& buf : r->vpBuf)
                for (float4*& a : buf) a = devAlloc<float4>(r, pixels);
            // Default priority, like the renderer's own stream.
            OWE_HIP_CHECK(hipStreamCreateWithFlags(&r->sppmStream, hipStreamNonBlocking));
            OWE_HIP_CHECK(hipEventCreateWithFlags(&r->camDone, hipEventDisableTiming));
            OWE_HIP_CHECK(hipEventCreateWithFlags(&r->cleared, hipEventDisableTiming));
            for (hipEvent_t& e : r->photonsDone) {
                OWE_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                OWE_HIP_CHECK(hipEventRecord(e, r->stream));
            }
            q.radius = devAlloc<float>(r, pixels);
            q.tauXY = devAlloc<double2>(r, pixels);
            q.tauZN = devAlloc<double2>(r, pixels);
            q.flux = devAlloc<unsigned long long>(r, pixels * 3 * kSplatBins);
            q.count = devAlloc<uint32_t>(r, pixels);
            q.next = devAlloc<uint32_t>(r, pixels);
            uint64_t gridSize = 1024;
            while (gridSize < 2 * pixels) gridSize <<= 1;
            q.head = devAlloc<uint32_t>(r, gridSize);
            q.gridMask = uint32_t(gridSize - 1);
            q.levels = devAlloc<int>(r, 2);
            q.stats = devAlloc<unsigned long long>(r, 4);
            q.guideCounts = devAlloc<unsigned int>(r, kGuideCells);
            q.guideCdf = devAlloc<float>(r, kGuideCells);
            q.guideTotal = devAlloc<float>(r, 1);
            OWE_HIP_CHECK(hipMemset(q.guideCounts, 0, kGuideCells * sizeof(unsigned int)));
            q.groups = uint32_t(photonGroups);
            q.alpha = float(photonAlpha);
            q.invNWY = r->invNWY;
            // The initial radius (GpuRenderer::resetSppm, gpu_renderer_passes.cpp:238-257): photon_radius
            // pixel footprints at the distance the path travelled — a pixel's angle for an observer, a
            // sensor pixel seen from its aim disc's centre; a sensor without one takes a fixed radius.
            double pixelAngle = 0, floorRadius = 0;
            const DObserver& ob = h.obs;
            if (ob.kind == OWE_DETECTOR_OBSERVER) {
                pixelAngle = 2.0 * double(ob.tanY) / double(ob.height);
            } else {
                const double pix = 2.0 * double(ob.halfY) / double(ob.height);
                const double dx = double(ob.aimCenter[0]) - ob.toWorld[3], dy = double(ob.aimCenter[1]) - ob.toWorld[7],
                             dz = double(ob.aimCenter[2]) - ob.toWorld[11];
                const double dist = ob.hasAim ? std::sqrt(dx * dx + dy * dy + dz * dz) : 0.0;
                if (dist > 0) pixelAngle = pix / dist;
                else floorRadius = photonRadius * pix;
            }
            q.r0PerMetre = float(photonRadius * pixelAngle);
            q.r0Floor = float(std::max(floorRadius, 1e-9));
            r->photons = std::max<uint64_t>(1, uint64_t(std::llround(double(pixels) * photonsPerPixel)));
            r->capacity = uint32_t(std::min<uint64_t>(r->photons, 1u << 22));  // photons per launch
            r->persistCounters = devAlloc<uint32_t>(r, 4);
            for (void* p : {static_cast<void*>(q.radius), static_cast<void*>(q.count)})
                OWE_HIP_CHECK(hipMemset(p, 0, pixels * sizeof(uint32_t)));
            OWE_HIP_CHECK(hipMemset(q.tauXY, 0, pixels * sizeof(double2)));
            OWE_HIP_CHECK(hipMemset(q.tauZN, 0, pixels * sizeof(double2)));
            OWE_HIP_CHECK(hipMemset(q.flux, 0, pixels * 3 * kSplatBins * sizeof(unsigned long long)));
        } else if (vcmLike) {
            // Light subpaths per iteration (photons_per_pixel per pixel, a multiple of the groups),
            // a cache every subpath's vertices fit in (one per scattering event up to max_depth + 1),
            // the merging hash, and the merging radius: photon_radius pixel footprints per metre from
            // the eye, shrinking as i^((α − 1)/2) (gpu_renderer_passes.cpp:136-147).
            const uint32_t G = uint32_t(s.photon_groups > 0 ? s.photon_groups : 1);
            const uint64_t want = std::max<uint64_t>(1, uint64_t(std::llround(double(pixels) * photonsPerPixel)));
            r->vcmLightPaths = (want + G - 1) / G * G;
            require(r->vcmLightPaths <= (1u << 30), "vcm: at most 2^30 light subpaths per iteration");
            const uint64_t slots = uint64_t(s.max_depth) + 1;
            const uint64_t vertices = r->vcmLightPaths * slots;  // per iteration
            const uint64_t budget = uint64_t(16) << 30;          // cache + staging, every slot
            if (2 * vertices * sizeof(LightVertex) > budget || vertices >= 0xFFFFFFFFull)
                fail(OWE_HIP_ERR_UNSUPPORTED, "vcm: the light vertex cache for this many subpaths and this depth would "
                                              "exceed 16 GiB (lower photons_per_pixel or max_depth)");
            // Iterations in flight: enough light subpaths per launch for sixteen waves per SIMD of
            // the chip (2^20 on 256 compute units: eight rounds of its resident waves), within the
            // budget and the 32-bit vertex indices.
            hipDeviceProp_t prop;
            OWE_HIP_CHECK(hipGetDeviceProperties(&prop, device));
            const uint64_t fill = uint64_t(prop.multiProcessorCount) * 4 * 16 * 64;
            uint64_t inFlight = std::min<uint64_t>(kVcmSlots, (fill + r->vcmLightPaths - 1) / r->vcmLightPaths);
            inFlight = std::min<uint64_t>(inFlight, budget / (2 * vertices * sizeof(LightVertex)));
            inFlight = std::min<uint64_t>(inFlight, 0xFFFFFFFEull / vertices);
            inFlight = std::max<uint64_t>(inFlight, 1);
            r->vcmSlotsMax = uint32_t(inFlight);
            r->vcmInFlight = uint32_t(inFlight);
            const uint64_t allPaths = inFlight * r->vcmLightPaths, allVertices = inFlight * vertices;
            const uint64_t tiles = (allPaths + kScanTile - 1) / kScanTile;
            VcmDev& v = r->vcm;
            v.cache = devAlloc<LightVertex>(r, allVertices);
            v.staging = devAlloc<LightVertex>(r, allVertices);
            v.pathCount = devAlloc<uint32_t>(r, allPaths);
            v.offset = devAlloc<uint32_t>(r, allPaths + 1);
            v.counts = devAlloc<uint32_t>(r, 32 * kVcmSlots);
            v.tileSum = devAlloc<uint32_t>(r, tiles);
            v.tileOffset = devAlloc<uint32_t>(r, tiles + 1);
            v.capacity = uint32_t(allVertices);
            v.pathSlots = uint32_t(slots);
            v.pathsPerGroup = uint32_t(r->vcmLightPaths / G);
            v.groups = G;
            v.slotPaths = uint32_t(r->vcmLightPaths);
            v.slots = 1;
            v.lightPaths = float(r->vcmLightPaths / G);
            uint64_t gridSize = 1024;
            while (gridSize < std::min<uint64_t>(2 * vertices, uint64_t(1) << 28)) gridSize <<= 1;
            r->vcmGridSize = uint32_t(gridSize);
            v.head = devAlloc<uint32_t>(r, gridSize * inFlight);
            v.next = devAlloc<uint32_t>(r, allVertices);
            v.gridMask = uint32_t(gridSize - 1);
            v.film = r->film;
            v.invNWY = r->invNWY;
            v.mergeFloor = 1e-9f;
            r->vcmRadiusPerMetre = s.integrator == OWE_INTEGRATOR_VCM
                                       ? float(photonRadius * 2.0 * double(h.obs.tanY) / double(h.obs.height)) : 0.0f;
            r->vcmAlpha = float(photonAlpha);
            r->lightFilm = devAlloc<unsigned long long>(r, 3 * kSplatBins * pixels);
            OWE_HIP_CHECK(hipMemset(r->lightFilm, 0, 3 * kSplatBins * pixels * sizeof(unsigned long long)));
            r->particleStats = devAlloc<unsigned long long>(r, 4);
            if (r->vcmMediumPass) r->capacity = uint32_t(std::min<uint64_t>(pixels * 64, 1u << 22));  // paths per launch
        } else {
            r->capacity = uint32_t(std::min<uint64_t>(pixels * 64, 1u << 22));  // particles per launch
            r->lightFilm = devAlloc<unsigned long long>(r, 3 * kSplatBins * pixels);
            OWE_HIP_CHECK(hipMemset(r->lightFilm, 0, 3 * kSplatBins * pixels * sizeof(unsigned long long)));
            r->particleStats = devAlloc<unsigned long long>(r, 4);
            r->particleNext = devAlloc<unsigned long long>(r, 1);
            hipDeviceProp_t prop;
            OWE_HIP_CHECK(hipGetDeviceProperties(&prop, device));
            r->computeUnits = prop.multiProcessorCount;
            r->blocksPerCU = 8;
            r->chunk = 4;
            r->regen = 32;
        }
        *out = r;
    });
    if (st != OWE_HIP_OK && r) {
        for (void* p : r->allocations) (void)hipFree(p);
        if (r->stream) (void)hipStreamDestroy(r->stream);
        if (r->sppmStream) (void)hipStreamDestroy(r->sppmStream);
        if (r->camDone) (void)hipEventDestroy(r->camDone);
        if (r->cleared) (void)hipEventDestroy(r->cleared);
        for (hipEvent_t e : r->photonsDone)
            if (e) (void)hipEventDestroy(e);
        delete r;
    }
    return st;
}

int owe_hip_renderer_destroy(owe_hip_renderer* r) {
    return guard([&] {
        if (!r) return;
        (void)hipSetDevice(r->device);
        if (r->stream) (void)hipStreamSynchronize(r->stream);
        if (r->sppmStream) (void)hipStreamSynchronize(r->sppmStream);
        dropGraph(r);
        for (void* p : r->allocations) OWE_HIP_CHECK(hipFree(p));
        for (auto& row : r->ev)
            for (hipEvent_t& e : row)
                if (e) OWE_HIP_CHECK(hipEventDestroy(e));
        if (r->sppmStream) OWE_HIP_CHECK(hipStreamDestroy(r->sppmStream));
        if (r->camDone) OWE_HIP_CHECK(hipEventDestroy(r->camDone));
        if (r->cleared) OWE_HIP_CHECK(hipEventDestroy(r->cleared));
        for (hipEvent_t e : r->photonsDone)
            if (e) OWE_HIP_CHECK(hipEventDestroy(e));
        if (r->stream) OWE_HIP_CHECK(hipStreamDestroy(r->stream));
        delete r;
    });
}

int owe_hip_renderer_set_timing(owe_hip_renderer* r, int enable) {
    return guard([&] {
        require(r != nullptr, "renderer is null");
        OWE_HIP_CHECK(hipSetDevice(r->device));
        if (enable && !r->ev[0][0])
            for (auto& row : r->ev)
                for (hipEvent_t& e : row) OWE_HIP_CHECK(hipEventCreate(&e));
        r->timing = enable != 0;
        for (int k = 0; k < kTimers; ++k) {
            r->kernelMs[k] = 0;
            r->kernelLaunches[k] = 0;
        }
    });
}

int owe_hip_renderer_kernel_times(const owe_hip_renderer* r, double* ms, uint64_t* launches) {
    return guard([&] {
        require(r != nullptr && ms != nullptr && launches != nullptr, "null argument");
        for (int k = 0; k < kTimers; ++k) {
            ms[k] = r->kernelMs[k];
            launches[k] = r->kernelLaunches[k];
        }
    });
}

int owe_hip_renderer_pass(owe_hip_renderer* r, int spp) {
    return guard([&] {
        require(r != nullptr, "renderer is null");
        if (spp <= 0) return;
        OWE_HIP_CHECK(hipSetDevice(r->device));
        const uint64_t total = uint64_t(r->host.obs.width) * uint64_t(r->host.obs.height) * uint64_t(spp);
        if (r->integrator == OWE_INTEGRATOR_LIGHT) {
            runParticles(r, total, uint32_t(r->passes));
            r->passes++;
            r->totalSpp += spp;
            return;
        }
        if (r->integrator == OWE_INTEGRATOR_VCM || r->integrator == OWE_INTEGRATOR_BDPT) {  // spp iterations
            runVcm(r, spp);
            if (r->vcmMediumPass) {
                // The medium-only camera pass, spp paths per pixel (GpuRenderer::pass,
                // gpu_renderer.cpp:171), into the same camera film.
                for (uint64_t base = 0; base < total; base += r->capacity) {
                    const uint32_t count = uint32_t(std::min<uint64_t>(r->capacity, total - base));
                    runPaths(r, base, count, uint32_t(spp), uint32_t(r->passes), false, nullptr);
                }
                OWE_HIP_CHECK(hipStreamSynchronize(r->stream));
            }
            r->passes++;
            r->totalSpp += spp;
            return;
        }
        if (r->integrator == OWE_INTEGRATOR_SPPM) {  // a pass of spp iterations
            runSppm(r, spp);
            r->passes++;
            r->totalSpp += spp;
            return;
        }
        // path, and hybrid's camera paths (its particles follow, GpuRenderer::pass,
        // gpu_renderer.cpp:166-172).
        if (r->mode == OWE_HIP_MODE_PERSISTENT) {
            runPersistent(r, uint32_t(spp), uint32_t(r->passes));
        } else {
            const uint32_t cap = r->queueCapacity ? r->queueCapacity : r->capacity;
            for (uint64_t base = 0; base < total; base += cap) {
                const uint32_t count = uint32_t(std::min<uint64_t>(cap, total - base));
                runPaths(r, base, count, uint32_t(spp), uint32_t(r->passes), false, nullptr);
            }
        }
        OWE_HIP_CHECK(hipStreamSynchronize(r->stream));
        if (r->integrator == OWE_INTEGRATOR_HYBRID) runParticles(r, total, uint32_t(r->passes));
        r->passes++;
        r->totalSpp += spp;
    });
}

namespace {
// Per pixel: τ (XYZ), N and r.
void sppmState(owe_hip_renderer* r, double* out) {
    const SppmDev& q = r->sppm;
    const size_t pixels = q.pixels;
    std::vector<double2> xy(pixels), zn(pixels);
    std::vector<float> rad(pixels);
    OWE_HIP_CHECK(hipMemcpyAsync(xy.data(), q.tauXY, pixels * sizeof(double2), hipMemcpyDeviceToHost, r->stream));
    OWE_HIP_CHECK(hipMemcpyAsync(zn.data(), q.tauZN, pixels * sizeof(double2), hipMemcpyDeviceToHost, r->stream));
    OWE_HIP_CHECK(hipMemcpyAsync(rad.data(), q.radius, pixels * sizeof(float), hipMemcpyDeviceToHost, r->stream));
    OWE_HIP_CHECK(hipStreamSynchronize(r->stream));
    for (size_t i = 0; i < pixels; ++i) {
        out[5 * i + 0] = xy[i].x;
        out[5 * i + 1] = xy[i].y;
        out[5 * i + 2] = zn[i].x;
        out[5 * i + 3] = zn[i].y;
        out[5 * i + 4] = double(rad[i]);
    }
}
}  // namespace

int owe_hip_renderer_sppm_state(owe_hip_renderer* r, double* state, uint64_t* photons_per_iteration,
                                long long* iterations) {
    return guard([&] {
        require(r != nullptr, "renderer is null");
        require(r->integrator == OWE_INTEGRATOR_SPPM, "SPPM state needs an sppm-integrator renderer");
        OWE_HIP_CHECK(hipSetDevice(r->device));
        if (state) sppmState(r, state);
        if (photons_per_iteration) *photons_per_iteration = r->photons;
        if (iterations) *iterations = (long long)r->sppmIterations;
    });
}

int owe_hip_renderer_resolve(owe_hip_renderer* r, double* xyz) {
    return guard([&] {
        require(r != nullptr && xyz != nullptr, "null argument");
        OWE_HIP_CHECK(hipSetDevice(r->device));
        const size_t n = 3 * size_t(r->host.obs.width) * size_t(r->host.obs.height);
        std::vector<long long> film(n);
        OWE_HIP_CHECK(hipMemcpyAsync(film.data(), r->film, n * sizeof(long long), hipMemcpyDeviceToHost, r->stream));
        OWE_HIP_CHECK(hipStreamSynchronize(r->stream));
        const double norm = r->totalSpp > 0 ? 1.0 / (kFilmScale * double(r->totalSpp)) : 0.0;
        for (size_t i = 0; i < n; ++i) xyz[i] = r->integrator != OWE_INTEGRATOR_LIGHT ? double(film[i]) * norm : 0.0;
        if (r->integrator == OWE_INTEGRATOR_SPPM && r->sppmIterations > 0) {
            // L = L_direct / iterations + τ / (N_emitted π r²) (sppm.slang:19-20).
            const size_t pixels = n / 3;
            std::vector<double> state(5 * pixels);
            sppmState(r, state.data());
            const double emitted = double(r->photons) * double(r->sppmIterations);
            for (size_t i = 0; i < pixels; ++i) {
                const double rr = state[5 * i + 4];
                if (!(rr > 0)) continue;
                const double k = 1.0 / (emitted * 3.141592653589793 * rr * rr);
                for (int c = 0; c < 3; ++c) xyz[3 * i + size_t(c)] += state[5 * i + size_t(c)] * k;
            }
        }
        if (r->lightFilm) {
            // Particle splats ÷ particles emitted (gpu_renderer_film.cpp:17-20): each channel's bins
            // are integers in units of 2^(kSplatExp0 + 32 b).
            std::vector<unsigned long long> bins(n * kSplatBins);
            OWE_HIP_CHECK(hipMemcpyAsync(bins.data(), r->lightFilm, bins.size() * sizeof(unsigned long long),
                                         hipMemcpyDeviceToHost, r->stream));
            OWE_HIP_CHECK(hipStreamSynchronize(r->stream));
            const double lightNorm = r->particleCount > 0 ? 1.0 / r->particleCount : 0.0;
            for (size_t i = 0; i < n; ++i) {
                double v = 0;
                for (int b = kSplatBins - 1; b >= 0; --b) v += std::ldexp(double(bins[i * kSplatBins + b]), kSplatExp0 + 32 * b);
                xyz[i] += v * lightNorm;
            }
        }
    });
}

int owe_hip_renderer_stats(const owe_hip_renderer* r, owe_transport_stats* out) {
    return guard([&] {
        require(r != nullptr && out != nullptr, "null argument");
        *out = r->stats;
    });
}

int owe_hip_renderer_first_inconsistency(const owe_hip_renderer* r, owe_inconsistency* out, int* found) {
    return guard([&] {
        require(r != nullptr && out != nullptr && found != nullptr, "null argument");
        OWE_HIP_CHECK(hipSetDevice(r->device));
        uint32_t w[kFirstWords];
        OWE_HIP_CHECK(hipStreamSynchronize(r->stream));
        OWE_HIP_CHECK(hipMemcpy(w, r->firstInconsistency, sizeof w, hipMemcpyDeviceToHost));
        *out = owe_inconsistency{};
        out->boundary = out->ray_region = out->boundary_region = out->previous = 0xFFFFFFFFu;
        out->previous_primitive = out->primitive = 0xFFFFFFFFu;
        *found = w[0] != 0u ? 1 : 0;
        if (!*found) return;
        auto none = [](uint32_t v) { return v == kNoBoundary || v == kRegionNone ? 0xFFFFFFFFu : v; };
        out->boundary = w[1];
        out->ray_region = none(w[2]);
        out->boundary_region = none(w[3]);
        out->previous = none(w[4]);
        for (int k = 0; k < 3; ++k) {  // camera-relative floats → world coordinates
            out->p[k] = double(__builtin_bit_cast(float, w[5 + k])) + r->host.origin[k];
            out->origin[k] = double(__builtin_bit_cast(float, w[8 + k])) + r->host.origin[k];
            out->direction[k] = double(__builtin_bit_cast(float, w[11 + k]));
        }
        out->previous_primitive = w[14];
        out->primitive = w[15];
    });
}

int owe_hip_renderer_progress(const owe_hip_renderer* r, int* passes, long long* samples_per_pixel) {
    return guard([&] {
        require(r != nullptr, "renderer is null");
        if (passes) *passes = r->passes;
        if (samples_per_pixel) *samples_per_pixel = r->totalSpp;
    });
}

int owe_hip_renderer_reset(owe_hip_renderer* r) {
    return guard([&] {
        require(r != nullptr, "renderer is null");
        OWE_HIP_CHECK(hipSetDevice(r->device));
        const size_t n = 3 * size_t(r->host.obs.width) * size_t(r->host.obs.height);
        OWE_HIP_CHECK(hipMemsetAsync(r->film, 0, n * sizeof(unsigned long long), r->stream));
        OWE_HIP_CHECK(hipMemsetAsync(r->firstInconsistency, 0, kFirstWords * sizeof(uint32_t), r->stream));
        if (r->lightFilm)
            OWE_HIP_CHECK(hipMemsetAsync(r->lightFilm, 0, n * kSplatBins * sizeof(unsigned long long), r->stream));
        if (r->integrator == OWE_INTEGRATOR_SPPM) {
            const SppmDev& q = r->sppm;
            const size_t pixels = q.pixels;
            OWE_HIP_CHECK(hipMemsetAsync(q.radius, 0, pixels * sizeof(float), r->stream));
            OWE_HIP_CHECK(hipMemsetAsync(q.count, 0, pixels * sizeof(uint32_t), r->stream));
            OWE_HIP_CHECK(hipMemsetAsync(q.tauXY, 0, pixels * sizeof(double2), r->stream));
            OWE_HIP_CHECK(hipMemsetAsync(q.tauZN, 0, pixels * sizeof(double2), r->stream));
            OWE_HIP_CHECK(hipMemsetAsync(q.flux, 0, pixels * 3 * kSplatBins * sizeof(unsigned long long), r->stream));
            OWE_HIP_CHECK(hipMemsetAsync(q.guideCounts, 0, kGuideCells * sizeof(unsigned int), r->stream));
            r->sppmIterations = 0;
        }
        r->vcmIterations = 0;
        OWE_HIP_CHECK(hipStreamSynchronize(r->stream));
        r->passes = 0;
        r->totalSpp = 0;
        r->particleCount = 0;
        r->stats = {};
    });
}

int owe_hip_renderer_camera_paths(owe_hip_renderer* r, int spp, int pass, const uint32_t* pixel, const uint32_t* sample,
                                  uint64_t n, double* xyz) {
    return guard([&] {
        require(r != nullptr, "renderer is null");
        require(r->integrator == OWE_INTEGRATOR_PATH || r->integrator == OWE_INTEGRATOR_HYBRID,
                "camera paths need a path- or hybrid-integrator renderer");
        require(spp > 0 && pass >= 0, "spp must be positive and pass non-negative");
        if (n == 0) return;
        require(pixel != nullptr && sample != nullptr && xyz != nullptr, "null buffer");
        const uint64_t pixels = uint64_t(r->host.obs.width) * uint64_t(r->host.obs.height);
        for (uint64_t i = 0; i < n; ++i)
            require(pixel[i] < pixels && sample[i] < uint32_t(spp), "pixel or sample index out of range");
        OWE_HIP_CHECK(hipSetDevice(r->device));
        if (r->listCapacity < r->capacity) {
            r->pathPixels = devAlloc<uint32_t>(r, r->capacity);
            r->pathSamples = devAlloc<uint32_t>(r, r->capacity);
            r->perPath = devAlloc<double>(r, 3 * size_t(r->capacity));
            r->listCapacity = r->capacity;
        }
        for (uint64_t base = 0; base < n; base += r->capacity) {
            const uint32_t count = uint32_t(std::min<uint64_t>(r->capacity, n - base));
            OWE_HIP_CHECK(hipMemcpyAsync(r->pathPixels, pixel + base, count * sizeof(uint32_t), hipMemcpyHostToDevice, r->stream));
            OWE_HIP_CHECK(hipMemcpyAsync(r->pathSamples, sample + base, count * sizeof(uint32_t), hipMemcpyHostToDevice, r->stream));
            runPaths(r, 0, count, uint32_t(spp), uint32_t(pass), true, r->perPath);
            OWE_HIP_CHECK(hipMemcpyAsync(xyz + 3 * base, r->perPath, 3 * size_t(count) * sizeof(double), hipMemcpyDeviceToHost,
                                         r->stream));
            OWE_HIP_CHECK(hipStreamSynchronize(r->stream));
        }
    });
}

int owe_hip_renderer_particles(owe_hip_renderer* r, int pass, const uint64_t* index, uint64_t n, double* xyz) {
    return guard([&] {
        require(r != nullptr, "renderer is null");
        require(r->integrator == OWE_INTEGRATOR_LIGHT || r->integrator == OWE_INTEGRATOR_HYBRID,
                "particles need a light- or hybrid-integrator renderer");
        require(pass >= 0, "pass must be non-negative");
        if (n == 0) return;
        require(index != nullptr && xyz != nullptr, "null buffer");
        OWE_HIP_CHECK(hipSetDevice(r->device));
        if (r->particleListCapacity < r->capacity) {
            r->particleIndex = devAlloc<uint64_t>(r, r->capacity);
            r->perParticle = devAlloc<double>(r, 3 * size_t(r->capacity));
            r->particleListCapacity = r->capacity;
        }
        std::fill(xyz, xyz + 3 * n, 0.0);
        for (uint64_t base = 0; base < n; base += r->capacity) {
            const uint32_t count = uint32_t(std::min<uint64_t>(r->capacity, n - base));
            OWE_HIP_CHECK(hipMemcpyAsync(r->particleIndex, index + base, count * sizeof(uint64_t), hipMemcpyHostToDevice,
                                         r->stream));
            OWE_HIP_CHECK(hipMemsetAsync(r->perParticle, 0, 3 * size_t(count) * sizeof(double), r->stream));
            runParticles(r, count, uint32_t(pass), r->particleIndex, r->perParticle);
            OWE_HIP_CHECK(hipMemcpyAsync(xyz + 3 * base, r->perParticle, 3 * size_t(count) * sizeof(double),
                                         hipMemcpyDeviceToHost, r->stream));
            OWE_HIP_CHECK(hipStreamSynchronize(r->stream));
        }
    });
}

int owe_hip_renderer_set_mode(owe_hip_renderer* r, int mode, int blocks_per_cu, int chunk, int use_graph,
                              int waves_per_simd, int refill) {
    return guard([&] {
        require(r != nullptr, "renderer is null");
        require(mode == OWE_HIP_MODE_WAVEFRONT || mode == OWE_HIP_MODE_PERSISTENT, "unknown mode");
        require(mode == OWE_HIP_MODE_WAVEFRONT || r->integrator == OWE_INTEGRATOR_PATH || r->integrator == OWE_INTEGRATOR_LIGHT ||
                    r->integrator == OWE_INTEGRATOR_HYBRID,
                "the persistent forms trace camera paths or particles (not SPPM or VCM iterations)");
        require(blocks_per_cu >= 0 && blocks_per_cu <= 64, "blocks per compute unit must lie in [0, 64]");
        require(chunk >= 0 && chunk <= 1024, "chunk must lie in [0, 1024]");
        require(waves_per_simd == 0 || waves_per_simd == 2 || waves_per_simd == 3, "waves per SIMD must be 0, 2 or 3");
        require(refill >= 0 && refill <= 64, "refill threshold must lie in [0, 64]");
        OWE_HIP_CHECK(hipSetDevice(r->device));
        r->mode = mode;
        if (blocks_per_cu > 0) r->blocksPerCU = blocks_per_cu;
        if (chunk > 0) r->chunk = uint32_t(chunk);
        if (waves_per_simd > 0) r->persistWaves = waves_per_simd;
        if (refill > 0) r->regen = uint32_t(refill);
        r->useGraph = use_graph != 0;
        dropGraph(r);
    });
}

int owe_hip_renderer_set_vcm_in_flight(owe_hip_renderer* r, int iterations) {
    return guard([&] {
        require(r != nullptr, "renderer is null");
        require(r->integrator == OWE_INTEGRATOR_VCM || r->integrator == OWE_INTEGRATOR_BDPT,
                "iterations in flight apply to the vcm and bdpt integrators");
        require(iterations >= 0 && uint32_t(iterations) <= r->vcmSlotsMax,
                "vcm: iterations in flight must lie in [0, " + std::to_string(r->vcmSlotsMax) +
                    "] (the caches this renderer holds)");
        r->vcmInFlight = iterations == 0 ? r->vcmSlotsMax : uint32_t(iterations);
    });
}

int owe_hip_renderer_get_vcm_in_flight(const owe_hip_renderer* r, int* iterations, int* most) {
    return guard([&] {
        require(r != nullptr, "renderer is null");
        require(r->integrator == OWE_INTEGRATOR_VCM || r->integrator == OWE_INTEGRATOR_BDPT,
                "iterations in flight apply to the vcm and bdpt integrators");
        if (iterations) *iterations = int(r->vcmInFlight);
        if (most) *most = int(r->vcmSlotsMax);
    });
}

int owe_hip_renderer_get_mode(const owe_hip_renderer* r, int* mode, int* blocks_per_cu, int* chunk, int* use_graph,
                              int* waves_per_simd, int* refill) {
    return guard([&] {
        require(r != nullptr, "renderer is null");
        if (mode) *mode = r->mode;
        if (blocks_per_cu) *blocks_per_cu = r->blocksPerCU;
        if (chunk) *chunk = int(r->chunk);
        if (use_graph) *use_graph = r->useGraph ? 1 : 0;
        if (waves_per_simd) *waves_per_simd = r->persistWaves;
        if (refill) *refill = int(r->regen);
    });
}

}  // extern "C"

namespace owehip {
namespace {

// ---------------------------------------------------------------- the autotuner's tuning cache

// A tuned launch shape: the form, the persistent kernel's blocks per CU, path ids per fetch,
// waves per SIMD and refill threshold, the wavefront's bounce rounds per host sync and paths in
// flight (0: keep), and the pass time it was chosen at.
struct TuningRecord {
    int mode = 0, blocksPerCU = 0, chunk = 0, waves = 0, regen = 0, rounds = 0;
    uint32_t queue = 0;
    double ms = 0;
};

void applyTuning(owe_hip_renderer* r, const TuningRecord& t) {
    if (t.mode == OWE_HIP_MODE_PERSISTENT) {
        r->blocksPerCU = t.blocksPerCU;
        r->chunk = uint32_t(t.chunk);
        r->persistWaves = t.waves;
        r->regen = uint32_t(t.regen);
    } else {
        if (t.rounds > 0) r->roundsPerSync = t.rounds;
        if (t.queue > 0) r->queueCapacity = std::min(t.queue, r->capacity);
    }
}

uint64_t fnv1a(uint64_t h, const void* data, size_t n) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; ++i) h = (h ^ p[i]) * 0x100000001b3ull;
    return h;
}
template <class T>
uint64_t fnvVec(uint64_t h, const std::vector<T>& v) {
    const uint64_t n = v.size();
    h = fnv1a(h, &n, sizeof n);
    return v.empty() ? h : fnv1a(h, v.data(), v.size() * sizeof(T));
}

// The cache key: the device (architecture, compute units), the form, the pass size and the scene
// as flattened for the device (every table the kernels read, the hierarchy's size), the settings
// that change the work (integrator, spectral mode, depths) — not the seed.
std::string tuningKey(const owe_hip_renderer* r, int spp) {
    hipDeviceProp_t p;
    OWE_HIP_CHECK(hipGetDeviceProperties(&p, r->device));
    uint64_t h = 0xcbf29ce484222325ull;
    h = fnv1a(h, p.gcnArchName, std::strlen(p.gcnArchName));
    const int32_t head[] = {p.multiProcessorCount, r->mode, spp, r->host.obs.width, r->host.obs.height,
                            r->integrator, r->set.spectral, r->set.maxDepth, r->set.rrDepth,
                            int32_t(r->host.contacts), int32_t(r->simpleShadows), int32_t(r->stackCap),
                            int32_t(r->capacity)};
    h = fnv1a(h, head, sizeof head);
    h = fnvVec(h, r->host.media);
    h = fnvVec(h, r->host.regionMedium);
    h = fnvVec(h, r->host.optics);
    h = fnvVec(h, r->host.emissions);
    h = fnvVec(h, r->host.boundaries);
    h = fnvVec(h, r->host.lights);
    h = fnvVec(h, r->host.table);
    h = fnvVec(h, r->host.waveData);
    h = fnvVec(h, r->host.bvh.nodes);
    h = fnvVec(h, r->host.bvh.tris);
    h = fnv1a(h, &r->host.obs, sizeof r->host.obs);
    char buf[40];
    std::snprintf(buf, sizeof buf, "owe-hip-tune-1-%016llx", static_cast<unsigned long long>(h));
    return buf;
}

std::mutex& tuningMutex() {
    static std::mutex m;
    return m;
}
// The cache file: set by owe_hip_set_tuning_cache, else $OWE_HIP_TUNING_CACHE, else none.
std::string& tuningPath() {
    static std::string path = [] {
        const char* e = std::getenv("OWE_HIP_TUNING_CACHE");
        return std::string(e ? e : "");
    }();
    return path;
}

// One line per record: key mode blocks chunk waves regen rounds queue ms; the last line of a key wins.
bool tuningLookup(const std::string& key, TuningRecord& out) {
    std::lock_guard<std::mutex> lock(tuningMutex());
    if (tuningPath().empty()) return false;
    std::FILE* f = std::fopen(tuningPath().c_str(), "r");
    if (!f) return false;
    char k[64];
    TuningRecord t;
    bool found = false;
    while (std::fscanf(f, "%63s %d %d %d %d %d %d %u %lf", k, &t.mode, &t.blocksPerCU, &t.chunk, &t.waves, &t.regen,
                       &t.rounds, &t.queue, &t.ms) == 9)
        if (key == k) {
            out = t;
            found = true;
        }
    std::fclose(f);
    return found && (out.mode == OWE_HIP_MODE_PERSISTENT || out.mode == OWE_HIP_MODE_WAVEFRONT);
}

void tuningStore(const std::string& key, const TuningRecord& t) {
    std::lock_guard<std::mutex> lock(tuningMutex());
    if (tuningPath().empty()) return;
    std::FILE* f = std::fopen(tuningPath().c_str(), "a");
    if (!f) return;  // an unwritable cache only costs the next run its sweep
    std::fprintf(f, "%s %d %d %d %d %d %d %u %.6f\n", key.c_str(), t.mode, t.blocksPerCU, t.chunk, t.waves, t.regen,
                 t.rounds, t.queue, t.ms);
    std::fclose(f);
}

}  // namespace
}  // namespace owehip

extern "C" {

int owe_hip_set_tuning_cache(const char* path) {
    return guard([&] {
        std::lock_guard<std::mutex> lock(tuningMutex());
        tuningPath() = path ? path : "";
    });
}

int owe_hip_renderer_tuning(const owe_hip_renderer* r, int* rounds_per_sync, uint64_t* queue_capacity, int* from_cache) {
    return guard([&] {
        require(r != nullptr, "renderer is null");
        if (rounds_per_sync) *rounds_per_sync = r->roundsPerSync;
        if (queue_capacity) *queue_capacity = r->queueCapacity ? r->queueCapacity : r->capacity;
        if (from_cache) *from_cache = r->tunedFromCache ? 1 : 0;
    });
}

int owe_hip_renderer_autotune(owe_hip_renderer* r, int spp, double* best_ms) {
    return guard([&] {
        require(r != nullptr, "renderer is null");
        require(spp > 0, "spp must be positive");
        require(r->integrator == OWE_INTEGRATOR_PATH ||
                    (r->integrator == OWE_INTEGRATOR_LIGHT && r->mode == OWE_HIP_MODE_PERSISTENT),
                "the autotuner tunes the camera-path forms and the particle pass's lane-refill form");
        require(r->passes == 0, "autotune a fresh (or reset) renderer: its passes would enter the film");
        OWE_HIP_CHECK(hipSetDevice(r->device));
        const std::string key = tuningKey(r, spp);
        TuningRecord rec;
        r->tunedFromCache = false;
        if (tuningLookup(key, rec)) {
            applyTuning(r, rec);
            r->tunedFromCache = true;
            if (best_ms) *best_ms = rec.ms;
            return;
        }
        hipEvent_t e0, e1;
        OWE_HIP_CHECK(hipEventCreate(&e0));
        OWE_HIP_CHECK(hipEventCreate(&e1));
        const bool wasTiming = r->timing;
        r->timing = false;
        auto timePass = [&](auto&& run) {
            OWE_HIP_CHECK(hipEventRecord(e0, r->stream));
            run();
            OWE_HIP_CHECK(hipEventRecord(e1, r->stream));
            OWE_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0;
            OWE_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            return double(ms);
        };
        const uint64_t total = uint64_t(r->host.obs.width) * uint64_t(r->host.obs.height) * uint64_t(spp);
        // Every candidate is timed once; the five fastest are timed three more times each and the
        // one with the lowest median wins, so a single noisy pass cannot pick the shape.
        std::vector<TuningRecord> cands;
        const bool graph = r->useGraph;
        auto run = [&](const TuningRecord& c) {
            applyTuning(r, c);
            if (r->integrator == OWE_INTEGRATOR_LIGHT) return timePass([&] { runParticles(r, total, 0u); });
            if (r->mode == OWE_HIP_MODE_PERSISTENT) {
                hipLaunchKernelGGL(setPass, dim3{1}, dim3{64}, 0, r->stream, r->passSlot, 0u);
                return timePass([&] { launchPersistent(r, persistArgs(r, total, uint32_t(spp))); });
            }
            const uint32_t cap = r->queueCapacity;
            return timePass([&] {
                for (uint64_t base = 0; base < total; base += cap)
                    runPaths(r, base, uint32_t(std::min<uint64_t>(cap, total - base)), uint32_t(spp), 0u, false, nullptr);
            });
        };
        if (r->integrator == OWE_INTEGRATOR_LIGHT) {
            // Refill threshold × blocks per compute unit × particle ids per lane fetch.
            for (int g : {1, 16, 32, 48})
                for (int b : {2, 4, 8, 16})
                    for (int c : {1, 4, 16}) cands.push_back({r->mode, b, c, r->persistWaves, g, 0, 0, 0.0});
        } else if (r->mode == OWE_HIP_MODE_PERSISTENT) {
            // Register budget (waves per SIMD) × refill threshold × blocks per compute unit × path
            // ids per lane fetch.
            r->useGraph = false;
            for (int w : {2, 3})
                for (int g : {1, 16, 32, 48})
                    for (int b : {1, 2, 3, 4, 8})
                        for (int c : {1, 4, 16}) cands.push_back({r->mode, b, c, w, g, 0, 0, 0.0});
        } else {
            // Bounce rounds enqueued between host reads of the live counters × paths in flight.
            for (int k : {1, 2, 4, 8})
                for (uint32_t q : {r->capacity / 4, r->capacity / 2, r->capacity})
                    if (q >= 1024) cands.push_back({r->mode, r->blocksPerCU, int(r->chunk), r->persistWaves, int(r->regen), k, q, 0.0});
        }
        for (TuningRecord& c : cands) c.ms = run(c);
        std::sort(cands.begin(), cands.end(), [](const TuningRecord& a, const TuningRecord& b) { return a.ms < b.ms; });
        const size_t finalists = std::min<size_t>(5, cands.size());
        for (size_t i = 0; i < finalists; ++i) {
            double t[3];
            for (double& x : t) x = run(cands[i]);
            std::sort(t, t + 3);
            cands[i].ms = t[1];
        }
        std::sort(cands.begin(), cands.begin() + finalists,
                  [](const TuningRecord& a, const TuningRecord& b) { return a.ms < b.ms; });
        rec = cands.front();
        applyTuning(r, rec);
        r->useGraph = graph;
        dropGraph(r);
        tuningStore(key, rec);
        const double best = rec.ms;
        OWE_HIP_CHECK(hipEventDestroy(e0));
        OWE_HIP_CHECK(hipEventDestroy(e1));
        r->timing = wasTiming;
        // The tuning passes leave nothing behind.
        const size_t n = 3 * size_t(r->host.obs.width) * size_t(r->host.obs.height);
        OWE_HIP_CHECK(hipMemsetAsync(r->film, 0, n * sizeof(unsigned long long), r->stream));
        if (r->lightFilm)
            OWE_HIP_CHECK(hipMemsetAsync(r->lightFilm, 0, n * kSplatBins * sizeof(unsigned long long), r->stream));
        r->particleCount = 0;
        OWE_HIP_CHECK(hipStreamSynchronize(r->stream));
        r->stats = {};
        for (int k = 0; k < kTimers; ++k) {
            r->kernelMs[k] = 0;
            r->kernelLaunches[k] = 0;
        }
        if (best_ms) *best_ms = best;
    });
}

// ---------------------------------------------------------------- batches of scenes

static void destroyBatch(owe_hip_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (void* p : b->allocations) (void)hipFree(p);
    for (hipEvent_t e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    for (owe_hip_renderer* m : b->members)
        if (m) (void)owe_hip_renderer_destroy(m);
    delete b;
}

int owe_hip_batch_create(int device, const owe_scene* scenes, int32_t count, const owe_render_settings* settings,
                         owe_hip_batch** out) {
    owe_hip_batch* b = nullptr;
    const int st = guard([&] {
        require(out != nullptr && scenes != nullptr && settings != nullptr, "null argument");
        *out = nullptr;
        require(count > 0, "a batch needs at least one scene");
        require(settings->integrator == OWE_INTEGRATOR_PATH, "a batch renders camera paths (the path integrator)");
        selectDevice(device);
        b = new owe_hip_batch;
        b->device = device;
        b->members.assign(size_t(count), nullptr);
        // The scenes are flattened, their hierarchies built and uploaded by a few host threads.
        std::vector<int> status(size_t(count), OWE_HIP_OK);
        std::vector<std::string> why(static_cast<size_t>(count));
        std::atomic<int> next{0};
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const int threads = int(std::min<unsigned>(unsigned(count), std::min(8u, hw)));
        auto work = [&] {
            for (;;) {
                const int i = next.fetch_add(1);
                if (i >= count) break;
                status[size_t(i)] = owe_hip_renderer_create(device, &scenes[i], settings, &b->members[size_t(i)]);
                if (status[size_t(i)] != OWE_HIP_OK) why[size_t(i)] = owe_hip_last_error();
            }
        };
        std::vector<std::thread> pool;
        for (int k = 0; k < threads; ++k) pool.emplace_back(work);
        for (std::thread& t : pool) t.join();
        for (int i = 0; i < count; ++i)
            if (status[size_t(i)] != OWE_HIP_OK) fail(status[size_t(i)], "scene " + std::to_string(i) + ": " + why[size_t(i)]);
        OWE_HIP_CHECK(hipSetDevice(device));
        OWE_HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        std::vector<DScene> ds;
        std::vector<int32_t> simple;
        std::vector<unsigned long long*> films;
        for (owe_hip_renderer* m : b->members) {
            ds.push_back(m->dev);
            simple.push_back(m->simpleShadows ? 1 : 0);
            films.push_back(m->film);
            b->stackCap = std::max(b->stackCap, m->stackCap);
            b->contacts = b->contacts || m->dev.contacts != 0u;
        }
        b->scenes = batchUpload(b, ds);
        PersistBatch& a = b->args;
        a.count = uint32_t(count);
        a.pass = batchAlloc<uint32_t>(b, 1);
        a.total = batchAlloc<uint64_t>(b, size_t(count));
        a.simple = batchUpload(b, simple);
        a.film = batchUpload(b, films);
        a.next = batchAlloc<unsigned long long>(b, size_t(count));
        a.segments = batchAlloc<unsigned long long>(b, size_t(count));
        a.counters = batchAlloc<uint32_t>(b, 4 * size_t(count));
        a.invNWY = b->members[0]->invNWY;
        *out = b;
    });
    if (st != OWE_HIP_OK) destroyBatch(b);
    return st;
}

int owe_hip_batch_destroy(owe_hip_batch* b) {
    return guard([&] { destroyBatch(b); });
}

int owe_hip_batch_member(owe_hip_batch* b, int32_t index, owe_hip_renderer** out) {
    return guard([&] {
        require(b != nullptr && out != nullptr, "null argument");
        require(index >= 0 && size_t(index) < b->members.size(), "member index out of range");
        *out = b->members[size_t(index)];
    });
}

int owe_hip_batch_pass(owe_hip_batch* b, int spp) {
    return guard([&] {
        require(b != nullptr, "batch is null");
        if (spp <= 0) return;
        OWE_HIP_CHECK(hipSetDevice(b->device));
        const size_t n = b->members.size();
        const owe_hip_renderer* lead = b->members[0];
        std::vector<uint64_t> total(n);
        for (size_t v = 0; v < n; ++v)
            total[v] = uint64_t(b->members[v]->host.obs.width) * uint64_t(b->members[v]->host.obs.height) * uint64_t(spp);
        PersistBatch a = b->args;
        a.spp = uint32_t(spp);
        a.chunk = lead->chunk;
        a.regen = lead->regen;
        OWE_HIP_CHECK(hipMemcpyAsync(const_cast<uint64_t*>(a.total), total.data(), n * sizeof(uint64_t),
                                     hipMemcpyHostToDevice, b->stream));
        hipLaunchKernelGGL(setPass, dim3{1}, dim3{64}, 0, b->stream, const_cast<uint32_t*>(a.pass), uint32_t(b->passes));
        if (b->timing) OWE_HIP_CHECK(hipEventRecord(b->ev[0], b->stream));
        launchPersistentBatch(b, a, lead->persistWaves, lead->blocksPerCU, lead->computeUnits);
        if (b->timing) OWE_HIP_CHECK(hipEventRecord(b->ev[1], b->stream));
        OWE_HIP_CHECK(hipGetLastError());
        std::vector<unsigned long long> segs(n);
        std::vector<uint32_t> counters(4 * n);
        OWE_HIP_CHECK(hipMemcpyAsync(segs.data(), a.segments, n * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                     b->stream));
        OWE_HIP_CHECK(hipMemcpyAsync(counters.data(), a.counters, 4 * n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                     b->stream));
        OWE_HIP_CHECK(hipStreamSynchronize(b->stream));
        if (b->timing) {
            float ms = 0;
            OWE_HIP_CHECK(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
            b->kernelMs += ms;
            b->launches++;
        }
        for (size_t v = 0; v < n; ++v) {
            owe_hip_renderer* m = b->members[v];
            m->stats.paths += total[v];
            m->stats.segments += segs[v];
            m->stats.inconsistencies += counters[4 * v + 2];
            m->stats.leaks += counters[4 * v + 3];
            m->passes++;
            m->totalSpp += spp;
        }
        b->passes++;
    });
}

int owe_hip_batch_stats(const owe_hip_batch* b, owe_transport_stats* out) {
    return guard([&] {
        require(b != nullptr && out != nullptr, "null argument");
        owe_transport_stats s{};
        for (const owe_hip_renderer* m : b->members) {
            s.paths += m->stats.paths;
            s.segments += m->stats.segments;
            s.inconsistencies += m->stats.inconsistencies;
            s.leaks += m->stats.leaks;
        }
        *out = s;
    });
}

int owe_hip_batch_set_timing(owe_hip_batch* b, int enable) {
    return guard([&] {
        require(b != nullptr, "batch is null");
        OWE_HIP_CHECK(hipSetDevice(b->device));
        if (enable && !b->ev[0])
            for (hipEvent_t& e : b->ev) OWE_HIP_CHECK(hipEventCreate(&e));
        b->timing = enable != 0;
        b->kernelMs = 0;
        b->launches = 0;
    });
}

int owe_hip_batch_kernel_time(const owe_hip_batch* b, double* ms, uint64_t* launches) {
    return guard([&] {
        require(b != nullptr && ms != nullptr && launches != nullptr, "null argument");
        *ms = b->kernelMs;
        *launches = b->launches;
    });
}

}  // extern "C"
