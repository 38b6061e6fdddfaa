// Device functions of the camera-path transport in float32: random numbers, spectra, n(λ), interface
// optics, sampling warps, world intersection and shadow transmittance.
//
// Each restates the reference's double-precision formula (the same ones the reference compiles to
// float for its own kernels through src/owe/core/shared/*.hpp):
//   PCG32 stream, 24-bit float uniform of the first word    core/shared/warps.hpp:23-29, shaders/common.slang:30-47
//   Draw2/Draw3 order (last argument drawn first)            core/sampling.hpp:37-51
//   Fresnel, Snell, conductor, GGX lobes                     core/shared/optics.hpp:14-256
//   CIE fit, λ sampling, Planck, Sellmeier/Cauchy/Ciddor     core/shared/spectra.hpp:13-61
//   warps, MIS, Henyey–Greenstein                            core/shared/warps.hpp:31-160
//   shadow walk through null boundaries                      transport/transport.cpp:237-285
// Self-intersection is excluded structurally, not by offsets: a spawned ray skips its own plane and
// its own mesh face (coplanar group), and on a sphere takes the root other than its origin's.
#pragma once

#include <hip/hip_runtime.h>

#include "dmath.hpp"
#include "scene_dev.hpp"
#include "shapes_dev.hpp"
#include "traverse.hpp"

namespace owehip {

constexpr float kPiF = 3.14159265358979323846f;
constexpr float kInvPiF = 0.31830988618379067154f;

// ---------------------------------------------------------------- random numbers

__host__ __device__ inline uint64_t mix64(uint64_t v) {
    v = (v ^ (v >> 31)) * 0x7fb5d329728ea185ULL;
    v = (v ^ (v >> 27)) * 0x81dadef4bc2dd44dULL;
    return v ^ (v >> 33);
}
__host__ __device__ inline uint64_t hashPair(uint64_t a, uint64_t b) { return mix64(a ^ (mix64(b) + 0x9e3779b97f4a7c15ULL)); }

struct Rng {
    uint64_t state, inc;
    __device__ uint32_t next() {
        const uint64_t s = state;
        state = s * 6364136223846793005ULL + inc;
        const uint32_t xs = uint32_t(((s >> 18) ^ s) >> 27);
        const uint32_t r = uint32_t(s >> 59);
        return (xs >> r) | (xs << ((32u - r) & 31u));
    }
    __device__ void seed(uint64_t s, uint64_t stream) {
        state = 0;
        inc = (stream << 1) | 1u;
        next();
        state += s;
        next();
    }
    // The double reference's two words per uniform: the first gives 24 bits, the second is skipped.
    __device__ float uniform() {
        const uint32_t a = next();
        next();
        return float(a >> 8) * (1.0f / 16777216.0f);
    }
};

// ---------------------------------------------------------------- small math

__device__ __forceinline__ float sqf(float x) { return x * x; }
__device__ __forceinline__ float sqrt0f(float x) { return sqrtf(fmaxf(0.0f, x)); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(hi, fmaxf(lo, x)); }
__device__ __forceinline__ f3 reflect3(f3 v, f3 n) { return v - n * (2.0f * dot3(v, n)); }
__device__ __forceinline__ f3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }
__device__ __forceinline__ f3 xfPoint(const float* m, f3 p) {
    return mk3(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
               m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}
__device__ __forceinline__ f3 xfVector(const float* m, f3 v) {
    return mk3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
               m[8] * v.x + m[9] * v.y + m[10] * v.z);
}

struct Frame {
    f3 s, t, n;
};
__device__ __forceinline__ Frame makeFrame(f3 n) {
    const float sg = copysignf(1.0f, n.z);
    const float a = -1.0f / (sg + n.z);
    const float b = n.x * n.y * a;
    return {mk3(1.0f + sg * n.x * n.x * a, sg * b, -sg * n.x), mk3(b, sg + n.y * n.y * a, -n.y), n};
}
__device__ __forceinline__ f3 toLocal(const Frame& f, f3 v) { return mk3(dot3(v, f.s), dot3(v, f.t), dot3(v, f.n)); }
__device__ __forceinline__ f3 toWorld(const Frame& f, f3 v) { return f.s * v.x + f.t * v.y + f.n * v.z; }

__device__ __forceinline__ void concentricDisk(float u1, float u2, float& x, float& y) {
    const float a = 2.0f * u1 - 1.0f, b = 2.0f * u2 - 1.0f;
    x = y = 0.0f;
    if (a == 0.0f && b == 0.0f) return;
    float r, phi;
    if (fabsf(a) > fabsf(b)) {
        r = a;
        phi = (kPiF / 4.0f) * (b / a);
    } else {
        r = b;
        phi = kPiF / 2.0f - (kPiF / 4.0f) * (a / b);
    }
    x = r * cosf(phi);
    y = r * sinf(phi);
}
__device__ __forceinline__ f3 cosineHemisphere(float u1, float u2) {
    float x, y;
    concentricDisk(u1, u2, x, y);
    return mk3(x, y, sqrt0f(1.0f - x * x - y * y));
}
__device__ __forceinline__ f3 uniformSphere(float u1, float u2) {
    const float z = 1.0f - 2.0f * u1, r = sqrt0f(1.0f - z * z), phi = 2.0f * kPiF * u2;
    return mk3(r * cosf(phi), r * sinf(phi), z);
}
// A direction in the cone of 1 − cos = omc about +z, uniform (cancellation-free for tiny cones).
__device__ __forceinline__ f3 uniformConeOmc(float u1, float u2, float omc) {
    const float o = u1 * omc, c = 1.0f - o, s = sqrt0f(o * (2.0f - o)), phi = 2.0f * kPiF * u2;
    return mk3(cosf(phi) * s, sinf(phi) * s, c);
}
__device__ __forceinline__ float powerHeuristic(float f, float g) {
    const float a = f * f, b = g * g;
    if (isinf(a)) return 1.0f;
    if (a + b == 0.0f) return 0.0f;
    return a / (a + b);
}
__device__ __forceinline__ float hgPhase(float c, float g) {
    const float d = 1.0f + g * g + 2.0f * g * c;
    return (1.0f - g * g) / (4.0f * kPiF * d * sqrtf(fmaxf(d, 1e-30f)));
}
__device__ __forceinline__ float sampleHg(f3 wo, float g, float u1, float u2, f3& wi) {
    float c;
    if (fabsf(g) < 1e-3f) c = 1.0f - 2.0f * u1;
    else c = -1.0f / (2.0f * g) * (1.0f + g * g - sqf((1.0f - g * g) / (1.0f + g - 2.0f * g * u1)));
    const float s = sqrt0f(1.0f - c * c), phi = 2.0f * kPiF * u2;
    const Frame f = makeFrame(wo);
    wi = f.s * (s * cosf(phi)) + f.t * (s * sinf(phi)) + wo * c;
    return hgPhase(dot3(wo, wi), g);
}

// ---------------------------------------------------------------- spectra

__device__ __forceinline__ float planckF(float lnm, float kelvin) {
    const float c1 = 1.191042972e8f, c2 = 1.438776877e4f;
    const float l = lnm * 1e-3f, l2 = l * l;
    return 1e6f * c1 / (l2 * l2 * l * (expf(c2 / (l * kelvin)) - 1.0f));
}
__device__ __forceinline__ float evalSpec(const DSpec& s, const float* table, float l) {
    if (s.kind == OWE_SPEC_CONSTANT) return s.scale * s.value;
    if (s.kind == OWE_SPEC_BLACKBODY) return s.scale * planckF(l, s.kelvin);
    const float* t = table + s.offset;
    const int n = s.count;
    if (l <= t[0]) return s.scale * t[1];
    if (l >= t[2 * (n - 1)]) return s.scale * t[2 * (n - 1) + 1];
    int i = 1;
    while (i < n - 1 && t[2 * i] <= l) ++i;
    const float u = (l - t[2 * (i - 1)]) / (t[2 * i] - t[2 * (i - 1)]);
    return s.scale * (t[2 * (i - 1) + 1] * (1.0f - u) + t[2 * i + 1] * u);
}
__device__ __forceinline__ float cieLobe(float x, float mu, float s1, float s2) {
    const float t = (x - mu) / (x < mu ? s1 : s2);
    return expf(-0.5f * t * t);
}
__device__ __forceinline__ f3 cieXYZ(float l) {
    return mk3(1.056f * cieLobe(l, 599.8f, 37.9f, 31.0f) + 0.362f * cieLobe(l, 442.0f, 16.0f, 26.7f) -
                   0.065f * cieLobe(l, 501.1f, 20.4f, 26.2f),
               0.821f * cieLobe(l, 568.8f, 46.9f, 40.5f) + 0.286f * cieLobe(l, 530.9f, 16.3f, 31.1f),
               1.217f * cieLobe(l, 437.0f, 11.8f, 36.0f) + 0.681f * cieLobe(l, 459.0f, 26.0f, 13.8f));
}
__device__ __forceinline__ float sampleVisibleF(float u) {
    return clampf(538.0f - 138.888889f * atanhf(0.85691062f - 1.82750197f * u), 360.0f, 830.0f);
}
__device__ __forceinline__ float visiblePdfF(float l) {
    if (l < 360.0f || l > 830.0f) return 0.0f;
    return 0.0039398042f / sqf(coshf(0.0072f * (l - 538.0f)));
}
__device__ __forceinline__ float ciddorAirF(float l2) {
    const float s2 = 1.0f / l2;
    return 1.0f + 0.05792105f / (238.0185f - s2) + 0.00167917f / (57.362f - s2);
}
__device__ __forceinline__ float mediumIndex(const DMedium& m, float lnm) {
    const float l = lnm * 1e-3f, l2 = l * l;
    float v = 1.0f;
    if (m.indexKind == OWE_INDEX_CONSTANT) v = m.a[0];
    else if (m.indexKind == OWE_INDEX_CAUCHY) v = m.a[0] + m.a[1] / l2 + m.a[2] / (l2 * l2);
    else if (m.indexKind == OWE_INDEX_SELLMEIER) {
        float s = 1.0f;
        for (int i = 0; i < m.terms; ++i) s += m.a[i] * l2 / (l2 - m.b[i]);
        v = sqrtf(s);
    } else v = ciddorAirF(l2);
    return m.relAir ? v * ciddorAirF(l2) : v;
}

// Four wavelengths and their pdfs; pdf 0 marks a terminated secondary (core/wavelengths.hpp:14-84).
struct WL {
    float lambda[NW];
    float pdf[NW];
};
__device__ __forceinline__ bool secondaryTerminated(const WL& w) { return w.pdf[1] == 0.0f && w.pdf[2] == 0.0f && w.pdf[3] == 0.0f; }

// ---------------------------------------------------------------- interface optics

__device__ __forceinline__ bool refractDir(f3 d, f3 n, float ni, float nt, f3& out) {
    const f3 nn = dot3(d, n) < 0.0f ? n : -n;
    const float ci = -dot3(d, nn), eta = ni / nt;
    const float k = 1.0f - eta * eta * (1.0f - ci * ci);
    out = d;
    if (k < 0.0f) return false;
    out = norm3(d * eta + nn * (eta * ci - sqrtf(k)));
    return true;
}
__device__ __forceinline__ float fresnelDielectric(float cosI, float eta) {
    cosI = clampf(cosI, -1.0f, 1.0f);
    if (cosI < 0.0f) {
        eta = 1.0f / eta;
        cosI = -cosI;
    }
    const float s2t = (1.0f - cosI * cosI) / (eta * eta);
    if (s2t >= 1.0f) return 1.0f;
    const float ct = sqrt0f(1.0f - s2t);
    const float a = (eta * cosI - ct) / (eta * cosI + ct), b = (cosI - eta * ct) / (cosI + eta * ct);
    return (a * a + b * b) * 0.5f;
}
struct Cx {
    float re, im;
};
__device__ __forceinline__ Cx cmul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ Cx cdiv(Cx a, Cx b) {
    const float d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
__device__ __forceinline__ Cx csqrtx(Cx z) {
    const float r = sqrtf(z.re * z.re + z.im * z.im);
    return {sqrtf(fmaxf(0.0f, (r + z.re) * 0.5f)), copysignf(sqrtf(fmaxf(0.0f, (r - z.re) * 0.5f)), z.im)};
}
__device__ __forceinline__ float fresnelConductor(float cosI, Cx eta) {
    cosI = clampf(cosI, 0.0f, 1.0f);
    const Cx s2t = cdiv({1.0f - cosI * cosI, 0.0f}, cmul(eta, eta));
    const Cx ct = csqrtx({1.0f - s2t.re, -s2t.im});
    const Cx ec{eta.re * cosI, eta.im * cosI}, ect = cmul(eta, ct);
    const Cx a = cdiv({ec.re - ct.re, ec.im - ct.im}, {ec.re + ct.re, ec.im + ct.im});
    const Cx b = cdiv({cosI - ect.re, -ect.im}, {cosI + ect.re, ect.im});
    return (a.re * a.re + a.im * a.im + b.re * b.re + b.im * b.im) * 0.5f;
}
__device__ __forceinline__ bool sameHemi(f3 a, f3 b) { return a.z * b.z > 0.0f; }
__device__ __forceinline__ bool refractLocal(f3 wi, f3 n, float eta, float& etap, f3& wt) {
    float ci = dot3(n, wi);
    etap = 1.0f;
    wt = wi;
    if (ci < 0.0f) {
        eta = 1.0f / eta;
        ci = -ci;
        n = -n;
    }
    const float s2t = fmaxf(0.0f, 1.0f - ci * ci) / (eta * eta);
    if (s2t >= 1.0f) return false;
    const float ct = sqrt0f(1.0f - s2t);
    wt = -wi / eta + n * (ci / eta - ct);
    etap = eta;
    return true;
}
__device__ __forceinline__ float ggxAlpha(float r) { return fmaxf(r, 1e-4f); }
__device__ __forceinline__ float ggxD(float a, f3 m) {
    const float c2 = m.z * m.z;
    if (c2 <= 0.0f) return 0.0f;
    const float c4 = c2 * c2;
    if (c4 < 1e-16f) return 0.0f;
    const float e = (1.0f - c2) / c2 / (a * a);
    return 1.0f / (kPiF * a * a * c4 * sqf(1.0f + e));
}
__device__ __forceinline__ float ggxLambda(float a, f3 w) {
    const float c2 = w.z * w.z;
    if (c2 <= 0.0f) return 0.0f;
    return (sqrtf(1.0f + a * a * ((1.0f - c2) / c2)) - 1.0f) * 0.5f;
}
__device__ __forceinline__ float ggxG1(float a, f3 w) { return 1.0f / (1.0f + ggxLambda(a, w)); }
__device__ __forceinline__ float ggxG(float a, f3 wo, f3 wi) { return 1.0f / (1.0f + ggxLambda(a, wo) + ggxLambda(a, wi)); }
__device__ __forceinline__ float ggxPdf(float a, f3 w, f3 m) {
    if (w.z == 0.0f) return 0.0f;
    return ggxG1(a, w) / fabsf(w.z) * ggxD(a, m) * fabsf(dot3(w, m));
}
__device__ __forceinline__ f3 ggxSampleWm(float a, f3 w, float u1, float u2) {
    f3 wh = norm3(mk3(a * w.x, a * w.y, w.z));
    if (wh.z < 0.0f) wh = -wh;
    const f3 t1 = wh.z < 0.99999f ? norm3(cross3(mk3(0, 0, 1), wh)) : mk3(1, 0, 0);
    const f3 t2 = cross3(wh, t1);
    const float r = sqrtf(u1), phi = 2.0f * kPiF * u2;
    const float px = r * cosf(phi);
    float py = r * sinf(phi);
    const float h = sqrtf(fmaxf(0.0f, 1.0f - px * px)), s = (1.0f + wh.z) * 0.5f;
    py = (1.0f - s) * h + s * py;
    const float pz = sqrt0f(1.0f - px * px - py * py);
    const f3 nh = t1 * px + t2 * py + wh * pz;
    return norm3(mk3(a * nh.x, a * nh.y, fmaxf(1e-6f, nh.z)));
}
struct Lobe {
    f3 wi;
    float weight, pdf;
    bool transmitted;
};
__device__ __forceinline__ bool sampleRoughDielectric(float a, f3 wo, float eta, float uc, float u1, float u2, bool radiance, Lobe& s) {
    const f3 wm = ggxSampleWm(a, wo, u1, u2);
    const float R = fresnelDielectric(dot3(wo, wm), eta), T = 1.0f - R;
    s = Lobe{wo, 0.0f, 0.0f, false};
    if (uc < R / (R + T)) {
        s.wi = reflect3(-wo, wm);
        if (!sameHemi(wo, s.wi)) return false;
        s.pdf = ggxPdf(a, wo, wm) / (4.0f * fabsf(dot3(wo, wm))) * R / (R + T);
        const float f = ggxD(a, wm) * ggxG(a, wo, s.wi) * R / (4.0f * fabsf(s.wi.z) * fabsf(wo.z));
        s.weight = f * fabsf(s.wi.z) / s.pdf;
    } else {
        float etap;
        f3 wi;
        if (!refractLocal(wo, wm, eta, etap, wi)) return false;
        if (sameHemi(wo, wi) || wi.z == 0.0f) return false;
        s.wi = wi;
        s.transmitted = true;
        const float den = sqf(dot3(wi, wm) + dot3(wo, wm) / etap);
        s.pdf = ggxPdf(a, wo, wm) * (fabsf(dot3(wi, wm)) / den) * T / (R + T);
        float ft = T * ggxD(a, wm) * ggxG(a, wo, wi) * fabsf(dot3(wi, wm) * dot3(wo, wm) / (wi.z * wo.z * den));
        if (radiance) ft /= sqf(etap);
        s.weight = ft * fabsf(wi.z) / s.pdf;
    }
    return isfinite(s.weight) && s.pdf > 0.0f;
}
__device__ __forceinline__ float evalRoughDielectric(float a, f3 wo, f3 wi, float eta, bool radiance, float& pdf) {
    pdf = 0.0f;
    const float co = wo.z, ci = wi.z;
    if (co == 0.0f || ci == 0.0f) return 0.0f;
    const bool refl = co * ci > 0.0f;
    float etap = 1.0f;
    if (!refl) etap = co > 0.0f ? eta : 1.0f / eta;
    f3 wm = wi * etap + wo;
    if (dot3(wm, wm) == 0.0f) return 0.0f;
    wm = norm3(wm);
    if (wm.z < 0.0f) wm = -wm;
    if (dot3(wm, wi) * ci < 0.0f || dot3(wm, wo) * co < 0.0f) return 0.0f;
    const float R = fresnelDielectric(dot3(wo, wm), eta), T = 1.0f - R;
    if (refl) {
        pdf = ggxPdf(a, wo, wm) / (4.0f * fabsf(dot3(wo, wm))) * R / (R + T);
        return ggxD(a, wm) * ggxG(a, wo, wi) * R / fabsf(4.0f * ci * co);
    }
    const float den = sqf(dot3(wi, wm) + dot3(wo, wm) / etap);
    pdf = ggxPdf(a, wo, wm) * fabsf(dot3(wi, wm)) / den * T / (R + T);
    float ft = ggxD(a, wm) * T * ggxG(a, wo, wi) * fabsf(dot3(wi, wm) * dot3(wo, wm) / (den * ci * co));
    if (radiance) ft /= sqf(etap);
    return ft;
}
__device__ __forceinline__ bool sampleRoughConductor(float a, f3 w, Cx eta, float tint, float u1, float u2, Lobe& s) {
    const f3 wm = ggxSampleWm(a, w, u1, u2);
    s = Lobe{reflect3(-w, wm), 0.0f, 0.0f, false};
    const float F = fresnelConductor(fabsf(dot3(w, wm)), eta) * tint;
    if (s.wi.z <= 0.0f) return false;
    s.pdf = ggxPdf(a, w, wm) / (4.0f * fabsf(dot3(w, wm)));
    if (!(s.pdf > 0.0f)) return false;
    const float f = ggxD(a, wm) * F * ggxG(a, w, s.wi) / (4.0f * s.wi.z * w.z);
    s.weight = f * s.wi.z / s.pdf;
    return isfinite(s.weight);
}
__device__ __forceinline__ float evalRoughConductor(float a, f3 wo, f3 wi, Cx eta, float tint, float& pdf) {
    pdf = 0.0f;
    f3 wm = wo + wi;
    if (dot3(wm, wm) == 0.0f) return 0.0f;
    wm = norm3(wm);
    const float F = fresnelConductor(fabsf(dot3(wo, wm)), eta) * tint;
    pdf = ggxPdf(a, wo, wm) / (4.0f * fabsf(dot3(wo, wm)));
    return ggxD(a, wm) * F * ggxG(a, wo, wi) / (4.0f * wi.z * wo.z);
}

// ---------------------------------------------------------------- world intersection

struct Hit {
    float t;
    uint32_t boundary;  // kNoBoundary: none
    uint32_t face;      // mesh: the triangle's face group; analytic: 0
    uint32_t slot;      // mesh: triangle record index
    f3 p, n;            // world (camera-relative) point and unit normal toward the front region
    f3 pLocal, nLocal;  // in the boundary's own frame (analytic shapes)
};

__device__ __forceinline__ bool inAperture(int kind, float a, float b, f3 p) {
    if (kind == OWE_APERTURE_DISK) {
        const float r2 = p.x * p.x + p.y * p.y;
        return r2 <= a * a && r2 >= b * b;
    }
    if (kind == OWE_APERTURE_RECT) return fabsf(p.x) <= a && fabsf(p.y) <= b;
    return sqf(p.x / a) + sqf(p.y / b) <= 1.0f;
}

__device__ __forceinline__ bool boxHit(const DBoundary& b, f3 o, f3 inv, float tmax) {
    const float x0 = (b.lo[0] - o.x) * inv.x, x1 = (b.hi[0] - o.x) * inv.x;
    const float y0 = (b.lo[1] - o.y) * inv.y, y1 = (b.hi[1] - o.y) * inv.y;
    const float z0 = (b.lo[2] - o.z) * inv.z, z1 = (b.hi[2] - o.z) * inv.z;
    const float n = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
    float f = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1));
    f *= 1.0000004f;
    return n <= fminf(f, tmax);
}

// Sphere roots in the geometric form (core/shared/surfaces.hpp:98-117), ascending.
__device__ __forceinline__ bool sphereRoots(f3 o, f3 d, float R, float& t0, float& t1) {
    const float dd = dot3(d, d);
    const float b = dot3(o, d) / dd;
    const f3 perp = o - d * b;
    const float disc = (R * R - dot3(perp, perp)) / dd;
    t0 = t1 = 0.0f;
    if (disc < 0.0f) return false;
    const float c = (dot3(o, o) - R * R) / dd;
    const float q = -b - copysignf(sqrtf(disc), b);
    t0 = q;
    t1 = q != 0.0f ? c / q : q;
    if (t0 > t1) {
        const float tt = t0;
        t0 = t1;
        t1 = tt;
    }
    return true;
}

// Out-of-line forms of the large analytic-shape solvers for the world walk: one copy of each in a
// kernel instead of one per inlined walk (a camera kernel walks for its path and for its shadow
// ray, with and without contacts), which keeps the walk's hot code — BVH, planes, spheres — small
// for the instruction cache. The lens batch inlines sagHit itself.
static __device__ __noinline__ bool sagHitOut(const DSag& s, f3 o, f3 d, float tmin, float tmax, bool departed, float& t,
                                       f3& p) {
    return sagHit(s, o, d, tmin, tmax, departed, t, p);
}
static __device__ __noinline__ bool waveHitOut(const DWave& w, const float4* data, f3 o, f3 d, float tmin, float tmax,
                                        bool departed, float depTol, float& t) {
    return waveHit(w, data, o, d, tmin, tmax, departed, depTol, t);
}
static __device__ __noinline__ bool cylinderHitOut(float R, float z0, float z1, f3 o, f3 d, float tmin, float tmax,
                                            bool departed, float& t) {
    return cylinderHit(R, z0, z1, o, d, tmin, tmax, departed, t);
}
static __device__ __noinline__ f3 sagNormalOut(const DSag& s, f3 p) { return sagNormalF(s, p); }
static __device__ __noinline__ f3 waveNormalOut(const DWave& w, const float4* data, float x, float y) {
    return waveNormalF(w, data, x, y);
}

// The nearest crossing in (0, tmax). departed/departedFace: the spawning boundary and (mesh) face;
// exclude: an analytic boundary the segment ends on (a sampled light point). kGeometry = false
// leaves the point and normals unset (a caller that keeps only t, boundary, face and slot).
// The local point and normal of an analytic boundary met at distance t along the local ray.
__device__ __forceinline__ void analyticGeometry(const DScene& sc, const DBoundary& b, f3 lo, f3 ldv, float t, f3& pl,
                                                 f3& nl) {
    pl = lo + ldv * t;
    if (b.shape == OWE_SHAPE_PLANE) {
        pl.z = 0.0f;
        nl = mk3(0, 0, 1);
    } else if (b.shape == OWE_SHAPE_SPHERE) {
        nl = norm3(pl / b.a);
    } else if (b.shape == OWE_SHAPE_CYLINDER) {
        nl = norm3(mk3(pl.x, pl.y, 0.0f));
    } else if (b.shape == OWE_SHAPE_SAG) {
        nl = sagNormalOut(b.sag, pl);
    } else {
        nl = waveNormalOut(b.wave, sc.waveData, pl.x, pl.y);
    }
}

template <bool kGeometry = true, bool kContact = false>
__device__ __forceinline__ bool intersectWorld(const DScene& sc, f3 o, f3 d, float tmax, uint32_t departed, uint32_t departedFace,
                                      uint32_t exclude, uint32_t* stack, uint32_t stride, Hit& h, float lo = 0.0f,
                                      uint32_t skipB = kNoBoundary, uint32_t skipFace = ~0u) {
    // Crossings count for t in (lo, tmax): the departed boundary's own from max(lo, 0) by its own
    // rules, every other boundary's from lo (below zero for a spawned ray among touching bodies).
    // skipB (an analytic boundary) and skipFace (a mesh face group) are not crossings.
    const float loOwn = fmaxf(lo, 0.0f);
    h.boundary = kNoBoundary;
    h.t = tmax;
    if (sc.nodes) {
        const RayHit rh = kContact ? traverseBvh4<false>(sc.nodes, sc.tris, o, d, loOwn, tmax, departedFace, stack, stride,
                                                         departed, lo, skipFace)
                                   : traverseBvh4<false>(sc.nodes, sc.tris, o, d, loOwn, tmax, departedFace, stack, stride);
        if (rh.slot != kEmptyChild) {
            const float4* tr = reinterpret_cast<const float4*>(sc.tris + rh.slot);
            const float4 a = tr[0], b = tr[1], c = tr[2];
            h.t = rh.t;
            h.slot = rh.slot;
            h.face = __float_as_uint(b.w);
            h.boundary = __float_as_uint(c.w);
            if (kGeometry) h.n = norm3(triNormal(mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(c.x, c.y, c.z)));
        }
    }
    const f3 inv = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    // Only (t, boundary) travel through the loop; the winner's point and normal are rebuilt after it.
    uint32_t analytic = kNoBoundary;
    for (uint32_t i = 0; i < sc.nBoundaries; ++i) {
        const DBoundary& b = sc.boundaries[i];
        if (b.shape == OWE_SHAPE_MESH || i == exclude || i == skipB) continue;
        if (!boxHit(b, o, inv, h.t)) continue;
        const f3 lo3 = xfPoint(b.toLocal, o), ldv = xfVector(b.toLocal, d);
        const bool own = i == departed;
        const float tmin = own ? loOwn : lo;
        float t;
        if (b.shape == OWE_SHAPE_PLANE) {
            if (own || ldv.z == 0.0f) continue;
            t = -lo3.z / ldv.z;
            if (!(t > tmin && t < h.t)) continue;
            f3 pl = lo3 + ldv * t;
            pl.z = 0.0f;
            if (!inAperture(b.aperture, b.a, b.b, pl)) continue;
        } else if (b.shape == OWE_SHAPE_SPHERE) {
            if (own) {
                t = -2.0f * dot3(lo3, ldv) / dot3(ldv, ldv);  // the root other than the origin's own
                if (!(t > tmin && t < h.t)) continue;
            } else {
                float t0, t1;
                if (!sphereRoots(lo3, ldv, b.a, t0, t1)) continue;
                t = t0 > tmin ? t0 : t1;
                if (!(t > tmin && t < h.t)) continue;
            }
        } else if (b.shape == OWE_SHAPE_CYLINDER) {
            if (!cylinderHitOut(b.a, b.z0, b.z1, lo3, ldv, tmin, h.t, own, t)) continue;
        } else if (b.shape == OWE_SHAPE_SAG) {
            f3 q;
            const float smin = own && b.sag.nA != 0 ? fmaxf(b.depTol, tmin) : tmin;
            if (!sagHitOut(b.sag, lo3, ldv, smin, h.t, own, t, q)) continue;
        } else {
            if (!waveHitOut(b.wave, sc.waveData, lo3, ldv, tmin, h.t, own, b.depTol, t)) continue;
        }
        h.t = t;
        analytic = i;
    }
    if (analytic != kNoBoundary) {
        h.boundary = analytic;
        h.face = 0;
        h.slot = kEmptyChild;
        if (!kGeometry) return true;
        const DBoundary& b = sc.boundaries[analytic];
        const f3 lo3 = xfPoint(b.toLocal, o), ldv = xfVector(b.toLocal, d);
        f3 pl, nl;
        analyticGeometry(sc, b, lo3, ldv, h.t, pl, nl);
        h.pLocal = pl;
        h.nLocal = nl;
        h.p = xfPoint(b.toWorld, pl);
        h.n = norm3(xfVector(b.toWorld, nl));
        return true;
    }
    if (h.boundary == kNoBoundary) return false;
    if (kGeometry) h.p = o + d * h.t;
    return true;
}

// The first region inconsistency of a renderer (TransportStats::first, scene/transport_stats.hpp:12-31;
// noteInconsistency, transport.slang:579-597): word 0 claims the record; 1 the boundary met, 2 the
// ray's region, 3 the boundary's region on the side met, 4 the boundary the ray left, 5–7 the
// camera-relative point (float bits), 8–10 the segment's origin, 11–13 its direction, 14 the
// primitive it left (not tracked: ~0), 15 the one it met (the mesh triangle's input index, ~0 for an
// analytic boundary).
constexpr int kFirstWords = 16;
__device__ __forceinline__ void noteInconsistency(const DScene& sc, uint32_t boundary, uint32_t rayRegion,
                                                  uint32_t boundaryRegion, uint32_t departed, f3 p, f3 o, f3 d,
                                                  uint32_t slot) {
    uint32_t* w = sc.firstInconsistency;
    if (!w || atomicCAS(w, 0u, 1u) != 0u) return;
    w[1] = boundary;
    w[2] = rayRegion;
    w[3] = boundaryRegion;
    w[4] = departed;
    const float v[9] = {p.x, p.y, p.z, o.x, o.y, o.z, d.x, d.y, d.z};
#pragma unroll
    for (int k = 0; k < 9; ++k) w[5 + k] = __float_as_uint(v[k]);
    w[14] = ~0u;
    w[15] = slot != kEmptyChild ? sc.tris[slot].prim : ~0u;
}

// Contact tolerance at a point (camera-relative coordinates): 2^-21 of its largest coordinate, two to
// four float ulps.
__device__ __forceinline__ float contactTol(f3 p) {
    return ldexpf(fmaxf(fmaxf(fabsf(p.x), fabsf(p.y)), fabsf(p.z)), -21);
}

// World::intersect with contact resolution, for scenes whose bodies touch (sc.contacts): a spawned
// ray meets other boundaries from t > −tol (tol the contact tolerance at its origin), and a crossing
// met from another region than the ray's, or a detector crossing, gives way to a crossing at the same
// place (within tol of its point; anywhere ahead for a contact at the ray's own origin) that is met
// from the ray's region and is not a detector — two rounds at most (one inlined query in a loop of
// at most three). Without contacts it is the plain intersectWorld.
// kContacts: compiled for scenes with contacts (the launch picks it from sc.contacts); without it
// this is the plain intersectWorld, with no contact code in the kernel.
template <bool kGeometry = true, bool kContacts = true>
__device__ __forceinline__ bool intersectContact(const DScene& sc, f3 o, f3 d, float tmax, uint32_t departed,
                                                 uint32_t departedFace, uint32_t exclude, uint32_t region,
                                                 uint32_t* stack, uint32_t stride, Hit& h) {
    if (!kContacts || !sc.contacts)
        return intersectWorld<kGeometry>(sc, o, d, tmax, departed, departedFace, exclude, stack, stride, h);
    const bool contacts = true;
    const float dep = departed != kNoBoundary ? contactTol(o) : 0.0f;
    float lo = -dep, hi = tmax;
    uint32_t skipB = kNoBoundary, skipF = ~0u;
    bool found = false, fits = false;
    for (int round = 0; round < 3; ++round) {
        Hit x;
        if (!intersectWorld<true, true>(sc, o, d, hi, departed, departedFace, exclude, stack, stride, x, lo, skipB, skipF))
            break;
        const DBoundary& xb = sc.boundaries[x.boundary];
        const bool xfits = (dot3(d, x.n) < 0.0f ? xb.front : xb.back) == region;
        const bool xdet = sc.optics[xb.optics].type == OWE_SURF_DETECTOR;
        // An alternative replaces the first crossing only if it fits and is not a detector where
        // the first one fitted.
        if (round > 0 && (!xfits || (xdet && fits))) break;
        h = x;
        found = true;
        fits = xfits;
        if (!contacts || region == kRegionNone || (xfits && !xdet)) break;
        const bool atOrigin = x.t <= dep;
        // The window about the crossing: the point's tolerance plus a few ulps of t itself.
        const float tol = contactTol(x.p) + ldexpf(x.t, -21);
        lo = atOrigin ? -dep : x.t - tol;
        hi = atOrigin ? tmax : fminf(tmax, x.t + tol);
        skipB = x.slot == kEmptyChild ? x.boundary : kNoBoundary;
        skipF = x.slot == kEmptyChild ? ~0u : x.face;
    }
    return found;
}

}  // namespace owehip
