// Surface scattering of both transport directions in float32: the interface at a hit, BSDF sampling
// and evaluation, the spectral weights of the secondary wavelengths, and emitter sampling. Camera
// paths call it with radiance = true, particles with radiance = false (importance transport: no
// n² factor at refraction; the rough dielectric's adjoint form).
//   sampleScatter / evalScatter / secondaryScatterWeight    src/owe/transport/scattering.cpp:65-222
//   Shape::sampleFrom / pdfFrom (sphere cone, area → solid) src/owe/scene/geometry.cpp:391-477,
//                                                            src/owe/core/shared/warps.hpp:104-113
//   sampleArea (uniform points of an emitter)               src/owe/backends/gpu/shaders/geometry.slang:708-752
#pragma once

#include "transport_dev.hpp"

namespace owehip {

// What every transport kernel is told of a render besides the scene.
// Settings::mode bits: the hybrid integrator's partition (light tracing owns the paths eye → D →
// S⁺ → light, camera paths every other; kFlagPartition, scene.slang:124), and camera paths that
// count only light reaching them through a medium scattering vertex (kFlagMediumOnly, the path
// pass beside vcm and bdpt in scenes whose media scatter, gpu_renderer.cpp:43-47, 171).
constexpr int kModePartition = 1, kModeMediumOnly = 2;

struct Settings {
    uint64_t seed;
    int maxDepth, rrDepth, spectral, mode;
    float lambdaMin, lambdaMax, fresnelFloor, pad2;
};

__device__ __forceinline__ float get(const float4& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w)); }
__device__ __forceinline__ void setc(float4& v, int i, float x) {
    if (i == 0) v.x = x;
    else if (i == 1) v.y = x;
    else if (i == 2) v.z = x;
    else v.w = x;
}

__device__ __forceinline__ float regionIndex(const DScene& sc, uint32_t region, float l) {
    return mediumIndex(sc.media[sc.regionMedium[region]], l);
}

struct Iface {
    const DOptics* o;
    f3 n;
    float nFront, nBack, lambda, floor;
};

struct SSample {
    f3 wi;
    float weight, pdf;
    bool delta, transmitted;
};

__device__ __forceinline__ Cx conductorEta(const DScene& sc, const DOptics& o, float lambda, float nMed) {
    const float inv = 1.0f / nMed;
    return {evalSpec(o.condN, sc.table, lambda) * inv, evalSpec(o.condK, sc.table, lambda) * inv};
}

__device__ __forceinline__ bool specularReflection(const Frame& f, f3 wo, float weight, SSample& s) {
    s.wi = toWorld(f, mk3(-wo.x, -wo.y, wo.z));
    s.weight = weight;
    s.delta = true;
    return weight > 0.0f;
}

// sampleScatter (src/owe/transport/scattering.cpp:65-157).
__device__ __forceinline__ bool sampleScatter(const DScene& sc, const Iface& it, f3 d, float uc, float u1, float u2, bool radiance,
                                     SSample& s) {
    const DOptics& o = *it.o;
    const Frame frame = makeFrame(it.n);
    const f3 wo = toLocal(frame, -d);
    const bool fromFront = wo.z > 0.0f;
    s = SSample{d, 0.0f, 0.0f, false, false};
    switch (o.type) {
        case OWE_SURF_DIELECTRIC: {
            const float eta = it.nBack / it.nFront;
            if (o.roughness >= 1e-3f && eta != 1.0f) {
                Lobe l;
                if (!sampleRoughDielectric(ggxAlpha(o.roughness), wo, eta, uc, u1, u2, radiance, l)) return false;
                s.wi = toWorld(frame, l.wi);
                s.weight = l.weight;
                s.pdf = l.pdf;
                s.transmitted = l.transmitted;
                return true;
            }
            const float ni = fromFront ? it.nFront : it.nBack, nt = fromFront ? it.nBack : it.nFront;
            const float R = fresnelDielectric(wo.z, eta);
            const float pR = it.floor > 0.0f && R < 1.0f ? clampf(R, it.floor, 1.0f - it.floor) : R;
            s.delta = true;
            if (uc < pR) {
                specularReflection(frame, wo, R / pR, s);
                return true;
            }
            if (!refractDir(d, it.n, ni, nt, s.wi)) return false;
            s.weight = (1.0f - R) / (1.0f - pR) * (radiance ? sqf(ni / nt) : 1.0f);
            s.transmitted = true;
            return true;
        }
        case OWE_SURF_DIFFUSE: {
            if (!fromFront && o.backAbsorbs) return false;
            const float albedo = evalSpec(o.reflectance, sc.table, it.lambda);
            if (albedo <= 0.0f) return false;
            f3 w = cosineHemisphere(u1, u2);
            if (!fromFront) w.z = -w.z;
            s.wi = toWorld(frame, w);
            s.pdf = fabsf(w.z) * kInvPiF;
            s.weight = albedo;
            return s.pdf > 0.0f;
        }
        case OWE_SURF_MIRROR:
            if (!fromFront && o.backAbsorbs) return false;
            return specularReflection(frame, wo, evalSpec(o.reflectance, sc.table, it.lambda), s);
        case OWE_SURF_CONDUCTOR: {
            if (!fromFront && o.backAbsorbs) return false;
            const Cx eta = conductorEta(sc, o, it.lambda, fromFront ? it.nFront : it.nBack);
            const float tint = evalSpec(o.reflectance, sc.table, it.lambda);
            const f3 w = fromFront ? wo : mk3(wo.x, wo.y, -wo.z);
            if (o.roughness < 1e-3f) return specularReflection(frame, wo, fresnelConductor(w.z, eta) * tint, s);
            Lobe l;
            if (!sampleRoughConductor(ggxAlpha(o.roughness), w, eta, tint, u1, u2, l)) return false;
            if (!fromFront) l.wi.z = -l.wi.z;
            s.wi = toWorld(frame, l.wi);
            s.weight = l.weight;
            s.pdf = l.pdf;
            return true;
        }
        default: return false;
    }
}

// evalScatter (scattering.cpp:159-188).
__device__ __forceinline__ float evalScatter(const DScene& sc, const Iface& it, f3 d, f3 wiW, bool radiance, float& pdf) {
    pdf = 0.0f;
    const DOptics& o = *it.o;
    if (o.delta) return 0.0f;
    const Frame frame = makeFrame(it.n);
    f3 wo = toLocal(frame, -d), wi = toLocal(frame, wiW);
    const bool fromFront = wo.z > 0.0f;
    switch (o.type) {
        case OWE_SURF_DIFFUSE:
            if (!fromFront && o.backAbsorbs) return 0.0f;
            if (!sameHemi(wo, wi)) return 0.0f;
            pdf = fabsf(wi.z) * kInvPiF;
            return evalSpec(o.reflectance, sc.table, it.lambda) * kInvPiF;
        case OWE_SURF_CONDUCTOR: {
            if (!fromFront && o.backAbsorbs) return 0.0f;
            if (!sameHemi(wo, wi)) return 0.0f;
            if (!fromFront) {
                wo.z = -wo.z;
                wi.z = -wi.z;
            }
            const Cx eta = conductorEta(sc, o, it.lambda, fromFront ? it.nFront : it.nBack);
            return evalRoughConductor(ggxAlpha(o.roughness), wo, wi, eta, evalSpec(o.reflectance, sc.table, it.lambda), pdf);
        }
        case OWE_SURF_DIELECTRIC:
            return evalRoughDielectric(ggxAlpha(o.roughness), wo, wi, it.nBack / it.nFront, radiance, pdf);
        default: return 0.0f;
    }
}

__device__ __forceinline__ float secondaryWeight(const DScene& sc, const Iface& it, f3 d, const SSample& s, bool radiance) {
    const DOptics& o = *it.o;
    if (!s.delta) {
        float pdf;
        const float f = evalScatter(sc, it, d, s.wi, radiance, pdf);
        return s.pdf > 0.0f ? f * fabsf(dot3(s.wi, it.n)) / s.pdf : 0.0f;
    }
    if (o.type == OWE_SURF_DIELECTRIC) return s.weight;
    if (o.type == OWE_SURF_MIRROR) return evalSpec(o.reflectance, sc.table, it.lambda);
    if (o.type == OWE_SURF_CONDUCTOR) {
        const float nMed = dot3(d, it.n) < 0.0f ? it.nFront : it.nBack;
        return fresnelConductor(fabsf(dot3(d, it.n)), conductorEta(sc, o, it.lambda, nMed)) *
               evalSpec(o.reflectance, sc.table, it.lambda);
    }
    return 0.0f;
}

// Solid-angle pdf of a light point on analytic boundary b seen from ref (Shape::pdfFrom).
__device__ __forceinline__ float sphereConeSize(f3 ref, float R) {
    const float dc2 = dot3(ref, ref);
    if (dc2 <= R * R * (1.0f + 1e-6f)) return 0.0f;
    const float s2 = R * R / dc2;
    return s2 / (1.0f + sqrt0f(1.0f - s2));
}
__device__ __forceinline__ bool areaToSolidAngle(f3 ref, f3 p, f3 n, float area, float& pdfW) {
    pdfW = 0.0f;
    const f3 w = p - ref;
    const float d2 = dot3(w, w);
    if (!(d2 > 0.0f)) return false;
    const float cosL = fabsf(dot3(n, w)) / sqrtf(d2);
    if (cosL <= 1e-7f) return false;
    pdfW = d2 / (cosL * area);
    return true;
}
__device__ __forceinline__ float pdfFrom(const DBoundary& b, f3 refLocal, f3 pLocal, f3 nLocal) {
    if (b.shape == OWE_SHAPE_SPHERE) {
        const float omc = sphereConeSize(refLocal, b.a);
        if (omc > 0.0f) return 1.0f / (2.0f * kPiF * omc);
    }
    float pdfW;
    return b.area > 0.0f && areaToSolidAngle(refLocal, pLocal, nLocal, b.area, pdfW) ? pdfW : 0.0f;
}
__device__ __forceinline__ bool sampleFrom(const DBoundary& b, f3 ref, float u1, float u2, float u3, f3& p, f3& n, float& pdfW) {
    if (b.shape == OWE_SHAPE_SPHERE) {
        const float omc = sphereConeSize(ref, b.a);
        if (omc > 0.0f) {
            const f3 axis = -ref / sqrtf(dot3(ref, ref));
            const Frame f = makeFrame(axis);
            const f3 wl = uniformConeOmc(u1, u2, omc);
            const f3 w = f.s * wl.x + f.t * wl.y + axis * wl.z;
            float t0, t1;
            if (sphereRoots(ref, w, b.a, t0, t1) && (t0 > 0.0f || t1 > 0.0f)) p = ref + w * (t0 > 0.0f ? t0 : t1);
            else p = norm3(ref + w * dot3(-ref, w)) * b.a;
            n = norm3(p);
            pdfW = 1.0f / (2.0f * kPiF * omc);
            return true;
        }
        n = uniformSphere(u1, u2);
        p = n * b.a;
        return areaToSolidAngle(ref, p, n, b.area, pdfW);
    }
    n = mk3(0, 0, 1);
    if (b.aperture == OWE_APERTURE_DISK) {
        const float r = sqrtf(b.b * b.b + u1 * (b.a * b.a - b.b * b.b)), phi = 2.0f * kPiF * u2;
        p = mk3(r * cosf(phi), r * sinf(phi), 0.0f);
    } else if (b.aperture == OWE_APERTURE_RECT) {
        p = mk3((2.0f * u1 - 1.0f) * b.a, (2.0f * u2 - 1.0f) * b.b, 0.0f);
    } else {
        const float r = sqrtf(u1), phi = 2.0f * kPiF * u2;
        p = mk3(b.a * r * cosf(phi), b.b * r * sinf(phi), 0.0f);
    }
    return areaToSolidAngle(ref, p, n, b.area, pdfW);
}

// The solid-angle density at p of directions toward a uniform point of a disc (diskAimPdf,
// core/shared/optics.hpp:248-256); generatingDisk: wi is known to come from the disc.
__device__ __forceinline__ float diskAimPdf(f3 p, f3 wi, f3 center, f3 normal, float radius, bool generatingDisk) {
    if (!(radius > 0.0f)) return 0.0f;
    const float dn = dot3(wi, normal);
    if (fabsf(dn) < 1e-12f) return 0.0f;
    const float t = dot3(center - p, normal) / dn;
    const f3 delta = p + wi * t - center;
    if (t <= 0.0f || (!generatingDisk && dot3(delta, delta) > radius * radius)) return 0.0f;
    return t * t / (fabsf(dn) * kPiF * radius * radius);
}

// A uniform point of emitter b in its own frame (sampleArea, geometry.slang:708-752): planes by
// aperture (annulus, rectangle, ellipse), spheres over their whole surface. Meshes are not sampled.
__device__ __forceinline__ bool sampleArea(const DBoundary& b, float u1, float u2, f3& p, f3& n) {
    n = mk3(0, 0, 1);
    if (b.shape == OWE_SHAPE_SPHERE) {
        n = uniformSphere(u1, u2);
        p = n * b.a;
        return true;
    }
    if (b.shape != OWE_SHAPE_PLANE) return false;
    if (b.aperture == OWE_APERTURE_DISK) {
        const float r = sqrtf(b.b * b.b + u1 * (b.a * b.a - b.b * b.b)), phi = 2.0f * kPiF * u2;
        p = mk3(r * cosf(phi), r * sinf(phi), 0.0f);
    } else if (b.aperture == OWE_APERTURE_RECT) {
        p = mk3((2.0f * u1 - 1.0f) * b.a, (2.0f * u2 - 1.0f) * b.b, 0.0f);
    } else {
        const float r = sqrtf(u1), phi = 2.0f * kPiF * u2;
        p = mk3(b.a * r * cosf(phi), b.b * r * sinf(phi), 0.0f);
    }
    return true;
}

// Tracer::transmittance (transport.cpp:237-285): from o along unit d to distance dist (infinity: out
// of the world, which it must leave from the ambient region), through null boundaries with
// Beer–Lambert on the way; zero at matter. The walk starts in `region`, leaving boundary `departed`
// (mesh face `face`, ~0 for none) and never meets `exclude` (the far end's own boundary). For a
// finite segment the region it arrives in must be `arrival`. The traversal stack is stack[k·stride]. Tr holds the transmittance per
// wavelength on return; false when the segment is blocked.
template <bool kContacts = true>
__device__ __forceinline__ bool walkTransmittance(const DScene& sc, f3 o, f3 d, float dist, uint32_t region, uint32_t departed,
                                         uint32_t face, uint32_t exclude, uint32_t arrival, const float lam[NW],
                                         uint32_t* stack, uint32_t stride, float Tr[NW]) {
    float remaining = dist;
    const bool infinite = isinf(remaining);
#pragma unroll
    for (int i = 0; i < NW; ++i) Tr[i] = 1.0f;
    for (int guard = 0; guard < 64; ++guard) {
        Hit h;
        const bool found = intersectContact<true, kContacts>(sc, o, d, remaining, departed, face, exclude, region, stack,
                                                             stride, h);
        const float seg = found ? fminf(h.t, remaining) : remaining;
        const DMedium& m = sc.media[sc.regionMedium[region]];
        if (m.flags) {
            if (isinf(seg)) return false;
#pragma unroll
            for (int i = 0; i < NW; ++i)
                Tr[i] *= expf(-(evalSpec(m.absorption, sc.table, lam[i]) + evalSpec(m.scattering, sc.table, lam[i])) * seg);
        }
        if (!found) return infinite ? region == 0 : region == arrival;
        const DBoundary& b = sc.boundaries[h.boundary];
        const bool front = dot3(d, h.n) < 0.0f;
        if ((front ? b.front : b.back) != region) return false;
        if (sc.optics[b.optics].type != OWE_SURF_NULL) return false;
        region = front ? b.back : b.front;
        o = h.p;
        departed = h.boundary;
        face = h.slot != kEmptyChild ? h.face : ~0u;
        if (!infinite) remaining -= h.t;
    }
    return false;  // exhausting the guard is not evidence the segment is clear
}

}  // namespace owehip
