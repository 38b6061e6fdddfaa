// Particles (light tracing) in float32: one thread per particle, emitted by power from the sun, a
// uniform sky or an emitting plane/sphere, walked through the world in importance transport, and
// splatted exactly into fixed-point film bins.
//
//   stream and wavelengths of particle i       particle_kernel.slang:11-19 (salt 0x5bd1e995)
//   emission: distant disc, area + cosine       particle.slang:184-344 (without photon guiding)
//   observer link (pupil, focus, pixel filter)  particle.slang:59-95
//   vertex connection, emitter seen directly    particle.slang:118-138, 267-279
//   surface sensor readout                      particle.slang:98-114
//   the walk: media, null crossings, sensor,    particle.slang:356-427
//   non-delta connections, dispersion, scatter,
//   roulette against the event's throughput
//   exact splats                                particle.slang:23-49
//
// Connections walk from the vertex toward the eye (the reference walks from the eye toward the
// vertex); both visit the same crossings, and the region reached must be the eye's. A distant
// particle's first hit is found again from just before it, where float is precise
// (refineDistantHit, particle.slang:143-147).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "particle.hpp"
#include "shading_dev.hpp"
#include "sppm.hpp"
#include "vcm.hpp"

namespace owehip {
namespace {

constexpr int kPBlock = 128;
constexpr int kRefillStack = 48;  // the refill kernel's LDS stack: every scene's bound

struct Link {
    f3 q, dir;
    int px, py;
    float factor;
};

struct PCtx {
    const DScene& sc;
    const Settings& set;
    const ParticleBatch& pb;
    uint32_t* stack;
    Rng rng;
    WL wl;
    float beta[NW];
    float sum[3];
    uint32_t segments, inconsistencies, leaks;
    uint32_t guideCell;  // SPPM: the guide cell a sun photon of the uniform share teaches
    bool useful;         // SPPM: the photon reached a visible point
};

__device__ __forceinline__ void addFixedPoint(unsigned long long* bins, float v) {
    if (!(v > 0.0f) || isinf(v)) return;  // zero, negative or not finite: nothing to add
    const uint32_t bits = __float_as_uint(v);
    const int e = int((bits >> 23) & 0xFFu);
    if (e == 0) return;  // subnormal
    uint32_t mant = (bits & 0x7FFFFFu) | 0x800000u;
    int p = e - 150 - kSplatExp0;  // the significand's lowest bit, above bin 0's
    if (p < 0) {
        if (p <= -24) return;
        mant >>= uint32_t(-p);
        p = 0;
    }
    const int b = p >> 5, r = p & 31;
    if (b >= kSplatBins) return;
    const uint64_t shifted = uint64_t(mant) << r;
    atomicAdd(&bins[b], (unsigned long long)(shifted & 0xFFFFFFFFull));
    const uint64_t high = shifted >> 32;
    if (high != 0 && b + 1 < kSplatBins) atomicAdd(&bins[b + 1], (unsigned long long)high);
}

__device__ __forceinline__ void deposit(PCtx& c, int px, int py, const float v[NW]) {
    float X = 0.0f, Y = 0.0f, Z = 0.0f;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        if (c.wl.pdf[k] == 0.0f || v[k] == 0.0f) continue;
        const float w = v[k] / c.wl.pdf[k] * c.pb.invNWY;
        const f3 m = cieXYZ(c.wl.lambda[k]);
        X += w * m.x;
        Y += w * m.y;
        Z += w * m.z;
    }
    if (c.pb.perParticle) {
        const float xyz[3] = {X, Y, Z};
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (xyz[k] > 0.0f && !isinf(xyz[k])) c.sum[k] += xyz[k];
        return;
    }
    unsigned long long* bins = c.pb.film + (size_t(py) * size_t(c.sc.obs.width) + size_t(px)) * (3 * kSplatBins);
    addFixedPoint(bins, X);
    addFixedPoint(bins + kSplatBins, Y);
    addFixedPoint(bins + 2 * kSplatBins, Z);
}

__device__ __forceinline__ void filterOffset(PCtx& c, float& fx, float& fy) {
    fx = fy = 0.0f;
    const float sigma = c.sc.obs.sigma;
    if (!(sigma > 0.0f)) return;
    const float u2 = c.rng.uniform(), u1 = c.rng.uniform();
    const float r = sqrtf(-2.0f * logf(1.0f - u1)), phi = 2.0f * kPiF * u2;
    fx = r * cosf(phi) * sigma;
    fy = r * sinf(phi) * sigma;
}

// The eye sits at the camera-relative origin.
__device__ __forceinline__ bool linkObserver(PCtx& c, f3 x, Link& k) {
    const DObserver& ob = c.sc.obs;
    const f3 right = ld3(ob.right), up = ld3(ob.up), fwd = ld3(ob.fwd);
    k.q = mk3(0.0f, 0.0f, 0.0f);
    if (ob.pupil > 0.0f) {
        const float u2 = c.rng.uniform(), u1 = c.rng.uniform();
        float dx, dy;
        concentricDisk(u1, u2, dx, dy);
        k.q = (right * dx + up * dy) * ob.pupil;
    }
    const f3 w = x - k.q;
    const float dist2 = dot3(w, w);
    if (!(dist2 > 0.0f)) return false;
    const f3 om = w / sqrtf(dist2);
    const float cc = dot3(om, fwd);
    if (cc <= 1e-9f) return false;
    float sx, sy;
    if (ob.pupil > 0.0f && isfinite(ob.focus)) {
        const f3 F = k.q + om * (ob.focus / cc);
        sx = dot3(F, right) / ob.focus;
        sy = dot3(F, up) / ob.focus;
    } else {
        sx = dot3(om, right) / cc;
        sy = dot3(om, up) / cc;
    }
    float ox, oy;
    filterOffset(c, ox, oy);
    const float W = float(ob.width), H = float(ob.height);
    const float fx = (sx / ob.tanX + 1.0f) * 0.5f * W + ox, fy = (1.0f - sy / ob.tanY) * 0.5f * H + oy;
    if (!(fx >= 0.0f && fy >= 0.0f && fx < W && fy < H)) return false;
    k.px = min(int(fx), ob.width - 1);
    k.py = min(int(fy), ob.height - 1);
    const float pixelArea = (2.0f * ob.tanX / W) * (2.0f * ob.tanY / H);
    k.factor = 1.0f / (dist2 * pixelArea * cc * cc * cc);
    k.dir = norm3(k.q - x);
    return true;
}

// A surface vertex (interface `it` at hit boundary bi, face) or a medium point (phase g) joined to
// the eye, value β·f·|cos|·W·Tr.
template <bool kContacts>
__device__ __forceinline__ void connectVertex(PCtx& c, f3 p, bool surface, const Iface* it, uint32_t bi, uint32_t face, f3 n,
                                     f3 d, float g, uint32_t region) {
    Link k;
    if (!linkObserver(c, p, k)) return;
    const DScene& sc = c.sc;
    float f[NW] = {0.0f, 0.0f, 0.0f, 0.0f};
    float cosX = 1.0f;
    uint32_t start = region, departed = kNoBoundary, dface = ~0u;
    if (!surface) {
        const float ph = hgPhase(dot3(-d, k.dir), g);
#pragma unroll
        for (int i = 0; i < NW; ++i) f[i] = ph;
    } else {
        const DBoundary& b = sc.boundaries[bi];
        bool any = false;
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
            f[i] = evalScatter(sc, at, d, k.dir, false, pdf);
            any = any || f[i] != 0.0f;
        }
        if (!any) return;
        cosX = fabsf(dot3(k.dir, n));
        start = dot3(k.dir, n) > 0.0f ? b.front : b.back;
        departed = bi;
        dface = face;
    }
    float Tr[NW];
    if (!walkTransmittance<kContacts>(sc, p, k.dir, len3(k.q - p), start, departed, dface, kNoBoundary, sc.obs.region,
                           c.wl.lambda, c.stack, kPBlock, Tr))
        return;
    float v[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) v[i] = c.beta[i] * f[i] * Tr[i] * (cosX * k.factor);
    deposit(c, k.px, k.py, v);
}

struct PRay {
    f3 o, d;
    uint32_t region, departed, face;
    bool distant;
};

// What vertex connection and merging needs of an emission (Emitted, particle.slang:192-202): the
// pdf of the emitted position and direction with the choice of the light, next-event estimation's
// pdf of a distant direction, the |cos| at an area emitter (1 for distant light), and the area
// emitter's light index, point and normal in its own frame.
struct EmitInfo {
    bool distant;
    float emissionPdfW, directPdfW, usedCos;
    uint32_t light;
    f3 pL, nL;
};

// emitParticle (particle.slang:306-325): the sky or a light by power, then its emission.
template <bool kContacts, bool kGuide = false>
__device__ __forceinline__ bool emitParticle(PCtx& c, bool observer, PRay& ray, const SppmDev* g = nullptr,
                                             EmitInfo* info = nullptr) {
    const DScene& sc = c.sc;
    const float pSky = sc.env.skyShare;
    const bool fromSky = c.rng.uniform() < pSky;
    uint32_t li = 0;
    if (!fromSky) {
        if (sc.nLights == 0) return false;
        const float ul = c.rng.uniform();
        while (li + 1 < sc.nLights && sc.lights[li].cdf <= ul) ++li;
    }
    float Le[NW];
    if (fromSky || sc.lights[li].kind == 1) {
        // emitDistant (particle.slang:234-265): a direction, then a point of a disc facing it.
        const float sel = fromSky ? pSky : (1.0f - pSky) * sc.lights[li].pdf;
        const float R = sc.env.emitRadius;
        const float u2 = c.rng.uniform(), u1 = c.rng.uniform();
        f3 toward;
        float pdfDir;
        if (fromSky) {
            toward = uniformSphere(u1, u2);
            pdfDir = 1.0f / (4.0f * kPiF);
#pragma unroll
            for (int i = 0; i < NW; ++i)
                Le[i] = sc.env.sky == OWE_SKY_UNIFORM && (i == 0 || c.wl.pdf[i] != 0.0f)
                            ? evalSpec(sc.env.skyRadiance, sc.table, c.wl.lambda[i]) : 0.0f;
        } else {
            toward = toWorld(makeFrame(ld3(sc.env.sunDir)), uniformConeOmc(u1, u2, sc.env.sunOneMinusCos));
            pdfDir = 1.0f / sc.env.sunSolidAngle;
#pragma unroll
            for (int i = 0; i < NW; ++i)
                Le[i] = i == 0 || c.wl.pdf[i] != 0.0f ? evalSpec(sc.env.sunRadiance, sc.table, c.wl.lambda[i]) : 0.0f;
        }
        const f3 dir = -toward;
        const Frame disc = makeFrame(dir);
        float qx, qy, pdfUnit = 1.0f / kPiF;  // the point's density per unit-disc area
        if (kGuide && !fromSky && *g->guideTotal > 0.0f) {
            // sampleEmissionDisc (particle.slang:205-232): the map's share, or uniform over the disc.
            const bool guided = c.rng.uniform() < kGuideShare;
            const float v2 = c.rng.uniform(), v1 = c.rng.uniform();
            if (guided) {
                const uint32_t cell = guideSample(g->guideCdf, v1);
                const float size = 2.0f / float(kGuideRes);
                qx = (float(cell % kGuideRes) + v2) * size - 1.0f;
                qy = (float(cell / kGuideRes) + c.rng.uniform()) * size - 1.0f;
            } else {
                concentricDisk(v1, v2, qx, qy);
            }
            uint32_t cell;
            pdfUnit = guidePdf(g->guideCdf, *g->guideTotal, qx, qy, cell);
            if (!(pdfUnit > 0.0f)) return false;
            if (!guided) c.guideCell = cell;  // the uniform share teaches the map
        } else {
            const float v2 = c.rng.uniform(), v1 = c.rng.uniform();
            concentricDisk(v1, v2, qx, qy);
            if (kGuide && !fromSky) {
                uint32_t cell;
                guidePdf(g->guideCdf, 0.0f, qx, qy, cell);
                c.guideCell = cell;  // before the map has learnt: every sun photon teaches it
            }
        }
        ray.o = ld3(sc.env.center) - dir * (1.01f * R) + disc.s * (qx * R) + disc.t * (qy * R);
        ray.d = dir;
        ray.region = 0;
        ray.departed = kNoBoundary;
        ray.face = ~0u;
        ray.distant = true;
        const float w = kGuide ? R * R / (sel * pdfDir * pdfUnit) : kPiF * R * R / (sel * pdfDir);
#pragma unroll
        for (int i = 0; i < NW; ++i) c.beta[i] = Le[i] * w;
        if (info) {
            info->distant = true;
            info->emissionPdfW = 1.0f / w;
            info->directPdfW = fromSky ? 0.0f : sc.lights[li].pdf / sc.env.sunSolidAngle;
            info->usedCos = 1.0f;
            info->light = kNoBoundary;
        }
    } else {
        // emitFromArea (particle.slang:281-304): a point, a side, a cosine-distributed direction.
        const float sel = (1.0f - pSky) * sc.lights[li].pdf;
        const uint32_t bi = sc.lights[li].boundary;
        const DBoundary& b = sc.boundaries[bi];
        const DEmission& em = sc.emissions[b.emission];
#pragma unroll
        for (int i = 0; i < NW; ++i)
            Le[i] = i == 0 || c.wl.pdf[i] != 0.0f ? evalSpec(em.radiance, sc.table, c.wl.lambda[i]) : 0.0f;
        const float u3 = c.rng.uniform(), u2 = c.rng.uniform(), u1 = c.rng.uniform();
        (void)u3;  // a mesh's triangle choice; meshes are not in the light table
        f3 pl, nl;
        if (!sampleArea(b, u1, u2, pl, nl)) return false;
        const f3 p = xfPoint(b.toWorld, pl), n = norm3(xfVector(b.toWorld, nl));
        const float pdfA = 1.0f / b.area;
        if (observer) {
            // areaEmitterToObserver (particle.slang:267-279): the emitter seen directly.
            Link k;
            if (linkObserver(c, p, k)) {
                const float cosL = dot3(n, k.dir);
                if ((cosL > 0.0f && em.front) || (cosL < 0.0f && em.back)) {
                    float Tr[NW];
                    if (walkTransmittance<kContacts>(sc, p, k.dir, len3(k.q - p), cosL > 0.0f ? b.front : b.back, bi, ~0u,
                                          kNoBoundary, sc.obs.region, c.wl.lambda, c.stack, kPBlock, Tr)) {
                        const float s = fabsf(cosL) * k.factor / (sel * pdfA);
                        float v[NW];
#pragma unroll
                        for (int i = 0; i < NW; ++i) v[i] = Le[i] * Tr[i] * s;
                        deposit(c, k.px, k.py, v);
                    }
                }
            }
        }
        float sidePdf = 1.0f;
        bool front = em.front != 0;
        if (em.front && em.back) {
            front = c.rng.uniform() < 0.5f;
            sidePdf = 0.5f;
        }
        const float c2 = c.rng.uniform(), c1 = c.rng.uniform();
        ray.o = p;
        ray.d = toWorld(makeFrame(front ? n : -n), cosineHemisphere(c1, c2));
        ray.region = front ? b.front : b.back;
        ray.departed = bi;
        ray.face = ~0u;
        ray.distant = false;
        const float w = kPiF / (sel * pdfA * sidePdf);
#pragma unroll
        for (int i = 0; i < NW; ++i) c.beta[i] = Le[i] * w;
        if (info) {
            info->distant = false;
            info->usedCos = fabsf(dot3(n, ray.d));
            info->emissionPdfW = sel * pdfA * sidePdf * info->usedCos * kInvPiF;
            info->directPdfW = 0.0f;
            info->light = li;
            info->pL = pl;
            info->nL = nl;
        }
    }
    bool any = c.beta[0] != 0.0f;
#pragma unroll
    for (int i = 1; i < NW; ++i) {
        if (c.wl.pdf[i] == 0.0f) c.beta[i] = 0.0f;
        any = any || c.beta[i] != 0.0f;
    }
    return any;
}

__device__ __forceinline__ bool roulette(PCtx& c, float q) {
    if (c.rng.uniform() >= q) return false;
    const float s = 1.0f / q;
#pragma unroll
    for (int i = 0; i < NW; ++i) c.beta[i] *= s;
    return true;
}

// An SPPM photon arriving at a non-specular surface point along d (gatherPhoton, sppm.slang:88-148):
// adds its contribution to every visible point within reach, on a disc in the point's tangent plane
// with a matching surface orientation, whose wavelengths agree with the photon's.
__device__ __forceinline__ void sppmGather(PCtx& c, const SppmDev& s, const Hit& h, f3 d, uint32_t group) {
    const DScene& sc = c.sc;
    const int lo = s.levels[0], hi = s.levels[1];
    if (lo > hi) return;  // no visible points this iteration
    const f3 wi = -d;
    for (int level = lo; level <= hi; ++level) {
        const float size = ldexpf(1.0f, level);
        const int bx = sppmCell(h.p.x - 0.5f * size, size), by = sppmCell(h.p.y - 0.5f * size, size),
                  bz = sppmCell(h.p.z - 0.5f * size, size);
        for (int k = 0; k < 8; ++k) {
            const int cx = bx + (k & 1), cy = by + ((k >> 1) & 1), cz = bz + ((k >> 2) & 1);
            for (uint32_t i = s.head[sppmHash(cx, cy, cz, level, s.gridMask)]; i != kGridNone; i = s.next[i]) {
                const float4 vp = s.vpP[i];
                const float r = vp.w;
                const f3 dp = h.p - mk3(vp.x, vp.y, vp.z);
                if (dot3(dp, dp) > r * r) continue;
                // The point's own cell (a bucket is shared by hash collisions): counted once.
                if (sppmLevel(r) != level || sppmCell(vp.x, size) != cx || sppmCell(vp.y, size) != cy ||
                    sppmCell(vp.z, size) != cz)
                    continue;
                const float4 vn = s.vpN[i];
                const f3 n = mk3(vn.x, vn.y, vn.z);
                if (fabsf(dot3(dp, n)) > 0.25f * r || dot3(n, h.n) < 0.5f) continue;
                const float4 vd = s.vpD[i];
                const uint32_t flags = __float_as_uint(vd.w);
                const bool free = (flags & kVpFree) != 0;
                if (!free && ((flags >> 8) & 0xFFu) != group) continue;  // wavelengths must agree
                WL w = c.wl;  // the photon's (a free point is evaluated at them)
                if (!free && (flags & kVpTerminated) && !secondaryTerminated(w)) {
                    w.pdf[0] /= NW;
                    w.pdf[1] = w.pdf[2] = w.pdf[3] = 0.0f;
                }
                const float4 vb = s.vpBeta[i];
                const DBoundary& b = sc.boundaries[__float_as_uint(vn.w)];
                const DOptics& o = sc.optics[b.optics];
                const f3 wo = mk3(vd.x, vd.y, vd.z);
                float X = 0.0f, Y = 0.0f, Z = 0.0f;
                bool any = false;
#pragma unroll
                for (int j = 0; j < NW; ++j) {
                    if (w.pdf[j] == 0.0f) continue;
                    const float l = w.lambda[j];
                    const Iface it{&o, n, regionIndex(sc, b.front, l), regionIndex(sc, b.back, l), l, c.set.fresnelFloor};
                    float pdf;
                    const float f = evalScatter(sc, it, wo, wi, true, pdf);
                    // A point bound to its group sees 1/groups of the photons: its share is scaled back up.
                    const float v = (free ? vb.x : get(vb, j) * float(s.groups)) * f * c.beta[j];
                    if (v == 0.0f) continue;
                    any = true;
                    const float wt = v / w.pdf[j] * s.invNWY;
                    const f3 cm = cieXYZ(l);
                    X += wt * cm.x;
                    Y += wt * cm.y;
                    Z += wt * cm.z;
                }
                if (!any || !isfinite(X + Y + Z)) continue;
                unsigned long long* bins = s.flux + size_t(i) * (3 * kSplatBins);
                addFixedPoint(bins, X);
                addFixedPoint(bins + kSplatBins, Y);
                addFixedPoint(bins + 2 * kSplatBins, Z);
                atomicAdd(&s.count[i], 1u);
                c.useful = true;
            }
        }
    }
}

// A particle between segments: its ray, the segments and scattering events so far (SPPM: direct
// light is the camera pass's, so photons gather only after their first event).
struct PState {
    PRay ray;
    int depth, events;
    int nonDelta, deltas;  // non-specular and specular events (the hybrid partition)
};

// One segment of the walk: returns whether the particle goes on. traceParticle runs it to the end in
// one thread; the lane-refill kernel advances every lane of a wave one segment at a time.
template <bool kContacts, bool kSppm>
__device__ __forceinline__ bool particleSegment(PCtx& c, PState& st, const SppmDev* sp, uint32_t group) {
    const DScene& sc = c.sc;
    const DObserver& ob = sc.obs;
    const bool observer = !kSppm && ob.kind == OWE_DETECTOR_OBSERVER;
    const int maxDepth = c.set.maxDepth, rrDepth = c.set.rrDepth;
    PRay& ray = st.ray;
    int& events = st.events;
    const int depth = st.depth;
    // The hybrid partition (particle.slang:358-417): connect only at the first non-specular vertex,
    // after at least one specular event, and end there.
    const bool partition = !kSppm && (c.set.mode & kModePartition) != 0;
    const bool mayConnect = !partition || (st.nonDelta == 0 && st.deltas >= 1);
    {
        Hit h;
        bool found = intersectContact<true, kContacts>(sc, ray.o, ray.d, __builtin_inff(), ray.departed, ray.face, kNoBoundary, ray.region,
                                      c.stack, kPBlock, h);
        if (found && depth == 0 && ray.distant) {
            const float back = 1e-5f * (1.0f + fmaxf(fabsf(h.p.x), fmaxf(fabsf(h.p.y), fabsf(h.p.z))));
            Hit again;
            if (intersectWorld(sc, h.p - ray.d * back, ray.d, 4.0f * back, kNoBoundary, ~0u, kNoBoundary, c.stack,
                               kPBlock, again))
                h = again;
        }
        c.segments++;
        // sampleMedium (transport.slang:544-562).
        const DMedium& med = sc.media[sc.regionMedium[ray.region]];
        if (med.flags & 2) {
            float sT[NW];
#pragma unroll
            for (int i = 0; i < NW; ++i)
                sT[i] = evalSpec(med.absorption, sc.table, c.wl.lambda[i]) + evalSpec(med.scattering, sc.table, c.wl.lambda[i]);
            const float tm = -logf(1.0f - c.rng.uniform()) / sT[0];
            const float seg = found ? h.t : __builtin_inff();
            if (tm < seg) {
                const float pdf0 = sT[0] * expf(-sT[0] * tm);
#pragma unroll
                for (int i = 0; i < NW; ++i)
                    c.beta[i] *= evalSpec(med.scattering, sc.table, c.wl.lambda[i]) * expf(-sT[i] * tm) / pdf0;
                const f3 p = ray.o + ray.d * tm;
                if (observer && mayConnect)
                    connectVertex<kContacts>(c, p, false, nullptr, kNoBoundary, ~0u, mk3(0, 0, 0), ray.d, med.g, ray.region);
                ++st.nonDelta;
                if (partition) return false;
                if (depth >= maxDepth) return false;
                const float u2 = c.rng.uniform(), u1 = c.rng.uniform();
                f3 wi;
                sampleHg(-ray.d, med.g, u1, u2, wi);
                ray.o = p;
                ray.d = wi;
                ray.departed = kNoBoundary;
                ray.face = ~0u;
                ++events;
                if (depth >= rrDepth && !roulette(c, 0.95f)) return false;
                return ++st.depth, true;
            }
            if (!found) return false;
            const float pass0 = expf(-sT[0] * seg);
#pragma unroll
            for (int i = 0; i < NW; ++i) c.beta[i] *= expf(-sT[i] * seg) / pass0;
        } else if (med.flags & 1) {
            if (!found) return false;
#pragma unroll
            for (int i = 0; i < NW; ++i) c.beta[i] *= expf(-evalSpec(med.absorption, sc.table, c.wl.lambda[i]) * h.t);
        }
        if (!found) {
            if (ray.region != 0) c.leaks++;
            return false;
        }
        // meetBoundary (transport.slang:600-615).
        const uint32_t bi = h.boundary;
        const DBoundary& b = sc.boundaries[bi];
        const bool fromFront = dot3(ray.d, h.n) < 0.0f;
        const uint32_t inc = fromFront ? b.front : b.back, other = fromFront ? b.back : b.front;
        if (inc != ray.region) {
            c.inconsistencies++;
            noteInconsistency(sc, bi, ray.region, inc, ray.departed, h.p, ray.o, ray.d, h.slot);
            ray.region = inc;
        }
        const DOptics& o = sc.optics[b.optics];
        const uint32_t hitFace = h.slot != kEmptyChild ? h.face : ~0u;
        if (o.type == OWE_SURF_NULL) {
            ray.region = other;
            ray.o = h.p;
            ray.departed = bi;
            ray.face = hitFace;
            if (depth >= 4 * maxDepth) return false;
            return ++st.depth, true;
        }
        if (!kSppm && !observer && fromFront && bi == ob.sensor) {
            // sensorPixel (particle.slang:98-114): the flux per pixel area.
            const float W = float(ob.width), H = float(ob.height);
            float fx = (h.pLocal.x / ob.halfX + 1.0f) * 0.5f * W;
            float fy = (1.0f - h.pLocal.y / ob.halfY) * 0.5f * H;
            if (ob.flipX) fx = W - fx;
            if (ob.flipY) fy = H - fy;
            float ox, oy;
            filterOffset(c, ox, oy);
            fx += ox;
            fy += oy;
            if (fx >= 0.0f && fy >= 0.0f && fx < W && fy < H) {
                const float inv = 1.0f / ((2.0f * ob.halfX / W) * (2.0f * ob.halfY / H));
                float v[NW];
#pragma unroll
                for (int i = 0; i < NW; ++i) v[i] = c.beta[i] * inv;
                deposit(c, min(int(fx), ob.width - 1), min(int(fy), ob.height - 1), v);
            }
        }
        if (o.type == OWE_SURF_DETECTOR || o.type == OWE_SURF_ABSORBER) return false;
        const float hero = c.wl.lambda[0];
        const Iface it{&o, h.n, regionIndex(sc, b.front, hero), regionIndex(sc, b.back, hero), hero, c.set.fresnelFloor};
        if (!o.delta && observer && mayConnect)
            connectVertex<kContacts>(c, h.p, true, &it, bi, hitFace, h.n, ray.d, 0.0f, ray.region);
        if (kSppm && !o.delta && events >= 1) sppmGather(c, *sp, h, ray.d, group);
        if (!o.delta) {
            ++st.nonDelta;
            if (partition) return false;
        }
        if (depth >= maxDepth) return false;
        auto at = [&](int i) {
            Iface r = it;
            r.nFront = regionIndex(sc, b.front, c.wl.lambda[i]);
            r.nBack = regionIndex(sc, b.back, c.wl.lambda[i]);
            r.lambda = c.wl.lambda[i];
            return r;
        };
        // terminateIfDispersive (transport.slang:639-648).
        if (o.type == OWE_SURF_DIELECTRIC && !secondaryTerminated(c.wl)) {
            const float a0 = it.nBack / it.nFront;
#pragma unroll
            for (int i = 1; i < NW; ++i) {
                if (c.wl.pdf[i] == 0.0f) continue;
                const Iface ai = at(i);
                if (fabsf(a0 - ai.nBack / ai.nFront) > 1e-12f * fabsf(a0)) {
                    c.wl.pdf[0] /= NW;
                    c.wl.pdf[1] = c.wl.pdf[2] = c.wl.pdf[3] = 0.0f;
                    c.beta[1] = c.beta[2] = c.beta[3] = 0.0f;
                    break;
                }
            }
        }
        // scatterAtSurface in importance mode (transport.slang:671-684).
        const float u3 = c.rng.uniform(), u2 = c.rng.uniform(), u1 = c.rng.uniform();
        SSample s;
        if (!sampleScatter(sc, it, ray.d, u1, u2, u3, false, s) || s.weight <= 0.0f) return false;
        c.beta[0] *= s.weight;
#pragma unroll
        for (int i = 1; i < NW; ++i)
            if (c.wl.pdf[i] != 0.0f) c.beta[i] *= secondaryWeight(sc, at(i), ray.d, s, false);
        if (s.transmitted) ray.region = other;
        ray.o = h.p;
        ray.d = s.wi;
        ray.departed = bi;
        ray.face = hitFace;
        ++events;
        if (s.delta) ++st.deltas;
        // Particles carry flux: roulette against the event's own throughput.
        if (depth >= rrDepth && s.weight < 1.0f && !roulette(c, s.weight)) return false;
    }
    ++st.depth;
    return true;
}

// traceParticle (particle.slang:356-427), with the hybrid partition when the settings ask for it
// (no direct view of an emitter then); with kSppm an SPPM photon (traceSppmPhoton,
// sppm.slang:150-201): no observer connections or sensor splats, a gather at every non-specular
// surface after the first scattering event.
template <bool kContacts, bool kSppm = false>
__device__ __forceinline__ void traceParticle(PCtx& c, const SppmDev* sp = nullptr, uint32_t group = 0) {
    const bool observer = !kSppm && c.sc.obs.kind == OWE_DETECTOR_OBSERVER && !(c.set.mode & kModePartition);
    PState st;
    st.depth = 0;
    st.events = st.nonDelta = st.deltas = 0;
    if (!emitParticle<kContacts, kSppm>(c, observer, st.ray, sp)) return;
    while (particleSegment<kContacts, kSppm>(c, st, sp, group)) {
    }
}


__device__ __forceinline__ uint32_t waveSum(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Particle `index` of pass `pass`: its own stream (seed, index, pass) and its wavelengths.
__device__ __forceinline__ void startParticle(PCtx& c, const Settings& set, uint64_t index, uint32_t pass) {
    c.rng.seed(hashPair(set.seed ^ 0x5bd1e995ull, index), pass);
    const float u = c.rng.uniform();
    if (set.spectral == OWE_SPECTRAL_UNIFORM1) {
        const float l = set.lambdaMin + (set.lambdaMax - set.lambdaMin) * u;
#pragma unroll
        for (int k = 0; k < NW; ++k) c.wl.lambda[k] = l;
        c.wl.pdf[0] = 1.0f / (set.lambdaMax - set.lambdaMin) / NW;
        c.wl.pdf[1] = c.wl.pdf[2] = c.wl.pdf[3] = 0.0f;
    } else {
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            float up = u + float(k) / NW;
            if (up >= 1.0f) up -= 1.0f;
            c.wl.lambda[k] = sampleVisibleF(up);
            c.wl.pdf[k] = visiblePdfF(c.wl.lambda[k]);
        }
    }
}

template <int kCap, bool kContacts>
__global__ __launch_bounds__(kPBlock) void particleKernel(DScene sc, Settings set, ParticleBatch pb) {
    __shared__ uint32_t stack[kCap * kPBlock];
    const uint32_t i = blockIdx.x * kPBlock + threadIdx.x;
    PCtx c{sc, set, pb, stack + threadIdx.x};
    c.sum[0] = c.sum[1] = c.sum[2] = 0.0f;
    c.segments = c.inconsistencies = c.leaks = 0;
    const bool live = i < pb.count;
    if (live) {
        startParticle(c, set, pb.index ? pb.index[i] : pb.base + i, pb.pass);
        traceParticle<kContacts>(c);
        if (pb.perParticle) {
#pragma unroll
            for (int k = 0; k < 3; ++k) pb.perParticle[3 * size_t(i) + k] = c.sum[k];
        }
    }
    const uint32_t n[4] = {waveSum(live ? 1u : 0u), waveSum(c.segments), waveSum(c.inconsistencies), waveSum(c.leaks)};
    if (__lane_id() == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (n[k]) atomicAdd(&pb.stats[k], (unsigned long long)n[k]);
}

// The lane-refill form of a particle pass: a grid sized to the chip, each lane advancing its particle
// one segment per iteration with particleSegment — the device code particleKernel runs, so every
// particle deposits the same splats. A lane whose particle has ended waits until `regen` lanes of
// its wave are idle; then they start the next particles of their chunks together, so the wave runs
// the emission code once for many lanes and stays full instead of idling until its longest walk
// ends. Particle ids come in chunks of `chunk` from one device counter (one atomic per wave per
// refill). The bins are integers, so the image equals particleKernel's bit for bit.
template <bool kContacts>
__global__ __launch_bounds__(kPBlock) void particleRefillKernel(DScene sc, Settings set, ParticleBatch pb) {
    __shared__ uint32_t stack[kRefillStack * kPBlock];
    PCtx c{sc, set, pb, stack + threadIdx.x};
    c.sum[0] = c.sum[1] = c.sum[2] = 0.0f;
    c.segments = c.inconsistencies = c.leaks = 0;
    const bool observer = sc.obs.kind == OWE_DETECTOR_OBSERVER && !(set.mode & kModePartition);  // emitter seen directly
    const uint32_t lane = __lane_id();
    PState st;
    st.depth = st.events = st.nonDelta = st.deltas = 0;
    bool live = false, done = false;
    uint64_t id = 0, end = 0;
    uint32_t started = 0;
    for (;;) {
        const uint64_t idle = __ballot(!live && !done);
        const bool refill = uint32_t(__popcll(idle)) >= pb.regen || __ballot(live) == 0;
        const bool needChunk = refill && !live && !done && id >= end;
        const uint64_t ask = __ballot(needChunk);
        if (ask) {
            const int leader = __ffsll((long long)ask) - 1;
            unsigned long long base = 0;
            if (int(lane) == leader) base = atomicAdd(pb.next, (unsigned long long)__popcll(ask) * pb.chunk);
            base = __shfl(base, leader);
            if (needChunk) {
                id = base + uint64_t(__popcll(ask & ((1ull << lane) - 1ull))) * pb.chunk;
                end = id + pb.chunk < pb.total ? id + pb.chunk : pb.total;
                done = id >= pb.total;
            }
        }
        if (refill && !live && !done) {
            startParticle(c, set, pb.base + id, pb.pass);
            st.depth = 0;
            st.events = st.nonDelta = st.deltas = 0;
            live = emitParticle<kContacts>(c, observer, st.ray);
            ++started;
            ++id;
        }
        if (__ballot(live || !done) == 0) break;  // every lane has finished and no ids are left
        if (!live) continue;
        live = particleSegment<kContacts, false>(c, st, nullptr, 0u);
    }
    const uint32_t n[4] = {waveSum(started), waveSum(c.segments), waveSum(c.inconsistencies), waveSum(c.leaks)};
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (n[k]) atomicAdd(&pb.stats[k], (unsigned long long)n[k]);
}

// ---------------------------------------------------------------- vertex connection and merging

// Into the light subpath's staging slots (cacheLightVertex, vcm.slang:199-227): vertex n of path p.
__device__ __forceinline__ void vcmCache(const VcmDev& v, const Hit& h, f3 d, const float beta[NW], const VcmMis& mis,
                                         bool terminated, uint32_t group, uint32_t slot, uint32_t p, uint32_t& n) {
    if (n >= v.pathSlots) return;  // cannot happen: one cached vertex per scattering event up to max_depth
    LightVertex& y = v.staging[size_t(p) * v.pathSlots + n++];
    const uint32_t face = h.slot != kEmptyChild ? h.face : ~0u;
    y.p = make_float4(h.p.x, h.p.y, h.p.z, __uint_as_float(h.boundary));
    y.n = make_float4(h.n.x, h.n.y, h.n.z, mis.dVCM);
    y.d = make_float4(d.x, d.y, d.z, __uint_as_float((terminated ? kLvTerminated : 0u) | ((face + 1u) << 1)));
    y.beta = make_float4(beta[0], beta[1], beta[2], beta[3]);
    y.mis = make_float4(mis.dVC, mis.dVM, __uint_as_float(group), __uint_as_float(slot));
}

// A light vertex to the eye (light tracing, t = 1; lightVertexToEye, vcm.slang:229-244): only to the
// pixels of this path's group this iteration, scaled by G (each pixel takes 1/G of the light paths).
template <bool kContacts>
__device__ __forceinline__ void vcmToEye(PCtx& c, const VcmDev& v, const Hit& h, const Iface& it, f3 d,
                                         const VcmMis& mis, uint32_t group) {
    Link k;
    if (!linkObserver(c, h.p, k)) return;
    if (vcmPixelGroup(uint32_t(k.py) * uint32_t(c.sc.obs.width) + uint32_t(k.px), v.iteration, v.groups) != group) return;
    const DScene& sc = c.sc;
    float f[NW], pdfFwd;
    if (!vcmEval(sc, it, h.boundary, c.wl, d, k.dir, false, f, pdfFwd)) return;
    float pdfRev;
    evalScatter(sc, it, -k.dir, -d, true, pdfRev);
    const float cosX = fabsf(dot3(h.n, k.dir));
    const float cameraPdfA = k.factor * cosX;
    const float wLight = cameraPdfA / v.lightPaths * (vcmEta(v, h.p) + mis.dVCM + mis.dVC * pdfRev);
    const float w = 1.0f / (wLight + 1.0f);
    const DBoundary& b = sc.boundaries[h.boundary];
    const uint32_t start = dot3(k.dir, h.n) > 0.0f ? b.front : b.back;
    float Tr[NW];
    if (!walkTransmittance<kContacts>(sc, h.p, k.dir, len3(k.q - h.p), start, h.boundary,
                                      h.slot != kEmptyChild ? h.face : ~0u, kNoBoundary, sc.obs.region, c.wl.lambda,
                                      c.stack, kPBlock, Tr))
        return;
    float val[NW];
    const float s = cosX * k.factor * w * float(v.groups);
#pragma unroll
    for (int i = 0; i < NW; ++i) val[i] = c.beta[i] * f[i] * Tr[i] * s;
    deposit(c, k.px, k.py, val);
}

// One light subpath (traceLightVcm, vcm.slang:246-292): its non-specular vertices cached and
// connected to the eye.
template <bool kContacts>
__device__ __forceinline__ void vcmLightPath(PCtx& c, const VcmDev& v, uint32_t group, uint32_t slot, uint32_t p,
                                             uint32_t& n) {
    const DScene& sc = c.sc;
    const int maxDepth = c.set.maxDepth, rrDepth = c.set.rrDepth;
    PRay ray;
    EmitInfo em;
    if (!emitParticle<kContacts>(c, false, ray, nullptr, &em) || !(em.emissionPdfW > 0.0f)) return;
    VcmMis mis{0.0f, 0.0f, 0.0f};
    float segment = 0.0f;  // distance since the last scattering vertex (null boundaries do not scatter)
    bool first = true;
    for (int depth = 0;; ++depth) {
        Hit h;
        bool found = intersectContact<true, kContacts>(sc, ray.o, ray.d, __builtin_inff(), ray.departed, ray.face,
                                                       kNoBoundary, ray.region, c.stack, kPBlock, h);
        if (found && depth == 0 && ray.distant) {
            const float back = 1e-5f * (1.0f + fmaxf(fabsf(h.p.x), fmaxf(fabsf(h.p.y), fabsf(h.p.z))));
            Hit again;
            if (intersectWorld(sc, h.p - ray.d * back, ray.d, 4.0f * back, kNoBoundary, ~0u, kNoBoundary, c.stack,
                               kPBlock, again))
                h = again;
        }
        c.segments++;
        const DMedium& med = sc.media[sc.regionMedium[ray.region]];
        if (med.flags) {  // extinction only: light scattered in a medium is the medium-only camera pass's
            if (!found) return;
#pragma unroll
            for (int i = 0; i < NW; ++i) {
                const float l = c.wl.lambda[i];
                c.beta[i] *= expf(-(evalSpec(med.absorption, sc.table, l) + evalSpec(med.scattering, sc.table, l)) * h.t);
            }
        }
        if (!found) {
            if (ray.region != 0) c.leaks++;
            return;
        }
        const uint32_t bi = h.boundary;
        const DBoundary& b = sc.boundaries[bi];
        const bool fromFront = dot3(ray.d, h.n) < 0.0f;
        const uint32_t inc = fromFront ? b.front : b.back, other = fromFront ? b.back : b.front;
        if (inc != ray.region) {
            c.inconsistencies++;
            noteInconsistency(sc, bi, ray.region, inc, ray.departed, h.p, ray.o, ray.d, h.slot);
            ray.region = inc;
        }
        segment += h.t;
        const DOptics& o = sc.optics[b.optics];
        const uint32_t hitFace = h.slot != kEmptyChild ? h.face : ~0u;
        if (o.type == OWE_SURF_NULL) {
            ray.region = other;
            ray.o = h.p;
            ray.departed = bi;
            ray.face = hitFace;
            if (depth >= 4 * maxDepth) return;
            continue;
        }
        if (o.type == OWE_SURF_DETECTOR || o.type == OWE_SURF_ABSORBER) return;
        const f3 d = ray.d;
        if (first) {
            // firstLightVertexMis (vcm.slang:185-197): next-event estimation would have chosen the
            // emitter point from here.
            mis.dVC = em.usedCos / em.emissionPdfW;
            if (em.distant) {
                mis.dVCM = em.directPdfW / em.emissionPdfW;
            } else {
                const DBoundary& lb = sc.boundaries[sc.lights[em.light].boundary];
                const float directPdfW = sc.lights[em.light].pdf * pdfFrom(lb, xfPoint(lb.toLocal, h.p), em.pL, em.nL);
                mis.dVCM = directPdfW * em.usedCos / em.emissionPdfW;
            }
            mis.dVM = mis.dVC * vcmInvEta(v, h.p);
        }
        mis.arrive(first ? 1.0f : segment, fabsf(dot3(h.n, d)));
        first = false;
        segment = 0.0f;
        const float hero = c.wl.lambda[0];
        const Iface it{&o, h.n, regionIndex(sc, b.front, hero), regionIndex(sc, b.back, hero), hero, c.set.fresnelFloor};
        if (!o.delta) {
            vcmCache(v, h, d, c.beta, mis, secondaryTerminated(c.wl), group, slot, p, n);
            vcmToEye<kContacts>(c, v, h, it, d, mis, group);
        }
        if (depth >= maxDepth) return;
        // terminateIfDispersive (transport.slang:639-648).
        if (o.type == OWE_SURF_DIELECTRIC && !secondaryTerminated(c.wl)) {
            const float a0 = it.nBack / it.nFront;
#pragma unroll
            for (int i = 1; i < NW; ++i) {
                if (c.wl.pdf[i] == 0.0f) continue;
                const Iface ai = vcmIface(sc, bi, h.n, c.wl.lambda[i], c.set.fresnelFloor);
                if (fabsf(a0 - ai.nBack / ai.nFront) > 1e-12f * fabsf(a0)) {
                    c.wl.pdf[0] /= NW;
                    c.wl.pdf[1] = c.wl.pdf[2] = c.wl.pdf[3] = 0.0f;
                    c.beta[1] = c.beta[2] = c.beta[3] = 0.0f;
                    break;
                }
            }
        }
        const float u3 = c.rng.uniform(), u2 = c.rng.uniform(), u1 = c.rng.uniform();
        SSample s;
        if (!sampleScatter(sc, it, d, u1, u2, u3, false, s) || s.weight <= 0.0f) return;
        c.beta[0] *= s.weight;
#pragma unroll
        for (int i = 1; i < NW; ++i)
            if (c.wl.pdf[i] != 0.0f)
                c.beta[i] *= secondaryWeight(sc, vcmIface(sc, bi, h.n, c.wl.lambda[i], c.set.fresnelFloor), d, s, false);
        float pdfRev = 0.0f;
        if (!s.delta) evalScatter(sc, it, -s.wi, -d, true, pdfRev);
        mis.scatter(h.n, s, pdfRev, vcmEta(v, h.p));
        if (s.transmitted) ray.region = other;
        ray.o = h.p;
        ray.d = s.wi;
        ray.departed = bi;
        ray.face = hitFace;
        ray.distant = false;
        if (depth >= rrDepth && s.weight < 1.0f && !roulette(c, s.weight)) return;
    }
}

// One thread per light subpath j of slot k's iteration (thread k·N + j): its own stream (seed, j,
// iteration) and the wavelengths of its group, j mod G (vcm_light_kernel.slang).
template <bool kContacts>
__global__ __launch_bounds__(kPBlock, OWE_HIP_VCM_WAVES) void vcmLightKernel(DScene sc, Settings set, VcmDev v, ParticleBatch pb) {
    __shared__ uint32_t stack[kRefillStack * kPBlock];
    const uint32_t t = blockIdx.x * kPBlock + threadIdx.x;
    PCtx c{sc, set, pb, stack + threadIdx.x};
    c.sum[0] = c.sum[1] = c.sum[2] = 0.0f;
    c.segments = c.inconsistencies = c.leaks = 0;
    const bool live = t < pb.count;
    if (live) {
        const uint32_t k = t / v.slotPaths, j = t - k * v.slotPaths;
        const VcmDev s = v.at(k);
        c.rng.seed(hashPair(set.seed ^ kVcmLight, j), s.iteration);
        const uint32_t group = j % v.groups;
        const uint32_t p = k * v.slotPaths + group * v.pathsPerGroup + j / v.groups;
        c.wl = sppmWavelengths(set, s.rotation, group, v.groups);
        uint32_t n = 0;
        vcmLightPath<kContacts>(c, s, group, k, p, n);
        v.pathCount[p] = n;
    }
    const uint32_t n[4] = {waveSum(live ? 1u : 0u), waveSum(c.segments), waveSum(c.inconsistencies), waveSum(c.leaks)};
    if (__lane_id() == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (n[k]) atomicAdd(&pb.stats[k], (unsigned long long)n[k]);
}

// One thread per SPPM photon (sppm_photon_kernel): photon i of the iteration has its own stream
// (seed, i, iteration) and the wavelengths of its group, i mod G.
template <int kCap, bool kContacts>
__global__ __launch_bounds__(kPBlock) void sppmPhotonKernel(DScene sc, Settings set, SppmDev s, uint64_t base,
                                                            uint32_t count) {
    __shared__ uint32_t stack[kCap * kPBlock];
    const uint32_t i = blockIdx.x * kPBlock + threadIdx.x;
    const ParticleBatch none{};
    PCtx c{sc, set, none, stack + threadIdx.x};
    c.sum[0] = c.sum[1] = c.sum[2] = 0.0f;
    c.segments = c.inconsistencies = c.leaks = 0;
    const bool live = i < count;
    if (live) {
        const uint64_t index = base + i;
        c.rng.seed(hashPair(set.seed ^ kSppmPhoton, index), s.iteration);
        const uint32_t group = uint32_t(index % uint64_t(s.groups));
        c.wl = sppmWavelengths(set, s.rotation, group, s.groups);
        c.guideCell = kGridNone;
        c.useful = false;
        traceParticle<kContacts, true>(c, &s, group);
        if (c.useful && c.guideCell != kGridNone) atomicAdd(&s.guideCounts[c.guideCell], 1u);
    }
    const uint32_t n[4] = {waveSum(live ? 1u : 0u), waveSum(c.segments), waveSum(c.inconsistencies), waveSum(c.leaks)};
    if (__lane_id() == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (n[k]) atomicAdd(&s.stats[k], (unsigned long long)n[k]);
}

}  // namespace

void launchSppmPhotons(const DScene& sc, const Settings& set, const SppmDev& s, uint64_t base, uint32_t count,
                       uint32_t stackCap, hipStream_t stream) {
    if (count == 0) return;
    const dim3 grid{(count + kPBlock - 1) / kPBlock}, block{kPBlock};
    if (sc.contacts) {
        if (stackCap <= 16) hipLaunchKernelGGL((sppmPhotonKernel<16, true>), grid, block, 0, stream, sc, set, s, base, count);
        else if (stackCap <= 32) hipLaunchKernelGGL((sppmPhotonKernel<32, true>), grid, block, 0, stream, sc, set, s, base, count);
        else hipLaunchKernelGGL((sppmPhotonKernel<48, true>), grid, block, 0, stream, sc, set, s, base, count);
    } else {
        if (stackCap <= 16) hipLaunchKernelGGL((sppmPhotonKernel<16, false>), grid, block, 0, stream, sc, set, s, base, count);
        else if (stackCap <= 32) hipLaunchKernelGGL((sppmPhotonKernel<32, false>), grid, block, 0, stream, sc, set, s, base, count);
        else hipLaunchKernelGGL((sppmPhotonKernel<48, false>), grid, block, 0, stream, sc, set, s, base, count);
    }
    OWE_HIP_CHECK(hipGetLastError());
}

void launchParticles(const DScene& sc, const Settings& set, const ParticleBatch& b, uint32_t stackCap,
                     hipStream_t stream) {
    if (b.count == 0) return;
    const dim3 grid{(b.count + kPBlock - 1) / kPBlock}, block{kPBlock};
    if (sc.contacts) {
        if (stackCap <= 16) hipLaunchKernelGGL((particleKernel<16, true>), grid, block, 0, stream, sc, set, b);
        else if (stackCap <= 32) hipLaunchKernelGGL((particleKernel<32, true>), grid, block, 0, stream, sc, set, b);
        else hipLaunchKernelGGL((particleKernel<48, true>), grid, block, 0, stream, sc, set, b);
    } else {
        if (stackCap <= 16) hipLaunchKernelGGL((particleKernel<16, false>), grid, block, 0, stream, sc, set, b);
        else if (stackCap <= 32) hipLaunchKernelGGL((particleKernel<32, false>), grid, block, 0, stream, sc, set, b);
        else hipLaunchKernelGGL((particleKernel<48, false>), grid, block, 0, stream, sc, set, b);
    }
    OWE_HIP_CHECK(hipGetLastError());
}

void launchVcmLight(const DScene& sc, const Settings& set, const VcmDev& v, unsigned long long* lightFilm,
                    unsigned long long* stats, float invNWY, uint32_t count, hipStream_t stream) {
    if (count == 0) return;
    ParticleBatch pb{};
    pb.count = count;
    pb.film = lightFilm;
    pb.stats = stats;
    pb.invNWY = invNWY;
    const dim3 grid{(count + kPBlock - 1) / kPBlock}, block{kPBlock};
    if (sc.contacts) hipLaunchKernelGGL((vcmLightKernel<true>), grid, block, 0, stream, sc, set, v, pb);
    else hipLaunchKernelGGL((vcmLightKernel<false>), grid, block, 0, stream, sc, set, v, pb);
    OWE_HIP_CHECK(hipGetLastError());
}

void launchParticlesRefill(const DScene& sc, const Settings& set, const ParticleBatch& b, uint32_t blocks,
                           hipStream_t stream) {
    if (b.total == 0) return;
    const dim3 grid{std::max(1u, blocks)}, block{kPBlock};
    if (sc.contacts) hipLaunchKernelGGL((particleRefillKernel<true>), grid, block, 0, stream, sc, set, b);
    else hipLaunchKernelGGL((particleRefillKernel<false>), grid, block, 0, stream, sc, set, b);
    OWE_HIP_CHECK(hipGetLastError());
}

}  // namespace owehip
