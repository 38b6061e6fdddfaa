// TEST INFRASTRUCTURE ONLY — the double-precision CPU oracle of the hip transport backend.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this code, and only as
// the checker or the timed CPU baseline; the product never links it.
//
// Core arithmetic restated from the reference engine (IEEE-754 double):
//   vectors, frames and slab tests      src/owe/core/math.hpp:84-219, core/vector.hpp
//   PCG32 stream and 53-bit uniform     src/owe/core/sampling.hpp:11-31, core/shared/warps.hpp:12-29
//   fixed draw order of Draw2/Draw3     src/owe/core/sampling.hpp:37-51
//   warps, MIS, Henyey–Greenstein       src/owe/core/shared/warps.hpp:31-160
//   Fresnel, Snell, conductor, GGX      src/owe/core/shared/optics.hpp:14-256
//   CIE fit, λ sampling, Planck, n(λ)   src/owe/core/shared/spectra.hpp:9-61, core/spectrum.cpp:11-23
//   watertight triangle test            src/owe/core/shared/surfaces.hpp:265-290
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>

namespace oracle {

constexpr double kPi = 3.14159265358979323846;
constexpr double kInvPi = 0.31830988618379067154;
constexpr double kInf = std::numeric_limits<double>::infinity();
// Precision-dependent tolerances of the double reference (src/owe/core/shared/prelude.hpp:56-74).
constexpr double kTiny = 1e-300;
constexpr double kGrazing = 1e-12;
constexpr double kRelTol = 1e-9;

inline double sq(double x) { return x * x; }
inline double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }
inline double sqrt0(double x) { return std::sqrt(std::max(0.0, x)); }

struct V3 {
    double x = 0, y = 0, z = 0;
    V3() = default;
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(double s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
// Division by a scalar multiplies by its reciprocal, as the reference vectors do (core/vector.hpp:60).
inline V3 operator/(V3 a, double s) { return a * (1.0 / s); }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) { return a / length(a); }
inline V3 vmin(V3 a, V3 b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }
inline double maxAbs(V3 a) { return std::max(std::max(std::fabs(a.x), std::fabs(a.y)), std::fabs(a.z)); }
inline V3 reflect(V3 v, V3 n) { return v - n * (2 * dot(v, n)); }

// ---------------------------------------------------------------- hashing and PCG32

inline uint64_t mix64(uint64_t v) {
    v = (v ^ (v >> 31)) * 0x7fb5d329728ea185ULL;
    v = (v ^ (v >> 27)) * 0x81dadef4bc2dd44dULL;
    return v ^ (v >> 33);
}
inline uint64_t hashPair(uint64_t a, uint64_t b) { return mix64(a ^ (mix64(b) + 0x9e3779b97f4a7c15ULL)); }

inline uint32_t pcgStep(uint64_t& state, uint64_t inc) {
    const uint64_t s = state;
    state = s * 6364136223846793005ULL + inc;
    const uint32_t xs = uint32_t(((s >> 18) ^ s) >> 27);
    const uint32_t r = uint32_t(s >> 59);
    return (xs >> r) | (xs << ((32u - r) & 31u));
}

// A PCG32 stream keyed by (seed, stream id); pcg-c-basic's srandom order.
struct Pcg {
    uint64_t state = 0, inc = 1;
    Pcg() { seed(0, 0); }
    Pcg(uint64_t s, uint64_t stream) { seed(s, stream); }
    void seed(uint64_t s, uint64_t stream) {
        state = 0;
        inc = (stream << 1) | 1u;
        next();
        state += s;
        next();
    }
    uint32_t next() { return pcgStep(state, inc); }
    // 53-bit uniform in [0, 1): 27 bits of the first word, 26 of the second.
    double uniform() {
        const uint64_t hi = next() >> 5, lo = next() >> 6;
        return (double(hi) * 67108864.0 + double(lo)) * (1.0 / 9007199254740992.0);
    }
    // Two and three uniforms in the reference's fixed order: the last argument is drawn first.
    void draw2(double& u1, double& u2) {
        u2 = uniform();
        u1 = uniform();
    }
    void draw3(double& u1, double& u2, double& u3) {
        u3 = uniform();
        u2 = uniform();
        u1 = uniform();
    }
};

// ---------------------------------------------------------------- frames and warps

// Duff et al. (2017) branchless orthonormal basis around unit n.
inline void basis(V3 n, V3& b1, V3& b2) {
    const double s = std::copysign(1.0, n.z);
    const double a = -1 / (s + n.z);
    const double b = n.x * n.y * a;
    b1 = {1 + s * n.x * n.x * a, s * b, -s * n.x};
    b2 = {b, s + n.y * n.y * a, -n.y};
}
struct Frame {
    V3 s, t, n;
    explicit Frame(V3 normal) : n(normal) { basis(n, s, t); }
    V3 toLocal(V3 v) const { return {dot(v, s), dot(v, t), dot(v, n)}; }
    V3 toWorld(V3 v) const { return s * v.x + t * v.y + n * v.z; }
};

// Shirley–Chiu concentric square-to-disk map.
inline void concentricDisk(double u1, double u2, double& x, double& y) {
    const double a = 2 * u1 - 1, b = 2 * u2 - 1;
    x = y = 0;
    if (a == 0 && b == 0) return;
    double r, phi;
    if (std::fabs(a) > std::fabs(b)) {
        r = a;
        phi = (kPi / 4) * (b / a);
    } else {
        r = b;
        phi = kPi / 2 - (kPi / 4) * (a / b);
    }
    x = r * std::cos(phi);
    y = r * std::sin(phi);
}
inline V3 cosineHemisphere(double u1, double u2) {
    double x, y;
    concentricDisk(u1, u2, x, y);
    return {x, y, sqrt0(1 - x * x - y * y)};
}
inline V3 uniformSphere(double u1, double u2) {
    const double z = 1 - 2 * u1, r = sqrt0(1 - z * z), phi = 2 * kPi * u2;
    return {r * std::cos(phi), r * std::sin(phi), z};
}
inline V3 uniformCone(double u1, double u2, double cosMax) {
    const double c = (1 - u1) + u1 * cosMax, s = sqrt0(1 - c * c), phi = 2 * kPi * u2;
    return {std::cos(phi) * s, std::sin(phi) * s, c};
}
// Box–Muller pair.
inline void standardNormal2(double u1, double u2, double& g1, double& g2) {
    const double r = std::sqrt(-2 * std::log(1 - u1)), phi = 2 * kPi * u2;
    g1 = r * std::cos(phi);
    g2 = r * std::sin(phi);
}
// Solid-angle density at p of directions toward a uniform point of a disc (optics.hpp:248-256);
// generatingDisk: wi is known to come from the disc, so no rim test.
inline double diskAimPdf(V3 p, V3 wi, V3 center, V3 normal, double radius, bool generatingDisk) {
    if (!(radius > 0)) return 0;
    const double dn = dot(wi, normal);
    if (std::fabs(dn) < 1e-12) return 0;
    const double t = dot(center - p, normal) / dn;
    const V3 delta = p + wi * t - center;
    if (t <= 0 || (!generatingDisk && dot(delta, delta) > radius * radius)) return 0;
    return t * t / (std::fabs(dn) * kPi * radius * radius);
}
inline double powerHeuristic(double f, double g) {
    const double a = f * f, b = g * g;
    if (std::isinf(a)) return 1;
    if (a + b == 0) return 0;
    return a / (a + b);
}
// Henyey–Greenstein phase (wo and wi both pointing away from the scattering point).
inline double hgPhase(double cosTheta, double g) {
    const double d = 1 + g * g + 2 * g * cosTheta;
    return (1 - g * g) / (4 * kPi * d * std::sqrt(std::max(d, kTiny)));
}
inline double sampleHg(V3 wo, double g, double u1, double u2, V3& wi) {
    double c;
    if (std::fabs(g) < 1e-3) c = 1 - 2 * u1;
    else c = -1 / (2 * g) * (1 + g * g - sq((1 - g * g) / (1 + g - 2 * g * u1)));
    const double s = sqrt0(1 - c * c), phi = 2 * kPi * u2;
    V3 a, b;
    basis(wo, a, b);
    wi = a * (s * std::cos(phi)) + b * (s * std::sin(phi)) + wo * c;
    return hgPhase(dot(wo, wi), g);
}

// ---------------------------------------------------------------- interface optics

// Snell refraction of propagation direction d (normal n of either orientation), ni → nt.
inline bool refractDir(V3 d, V3 n, double ni, double nt, V3& out) {
    const V3 nn = dot(d, n) < 0 ? n : -n;
    const double ci = -dot(d, nn), eta = ni / nt;
    const double k = 1 - eta * eta * (1 - ci * ci);
    out = d;
    if (k < 0) return false;
    out = normalize(d * eta + nn * (eta * ci - std::sqrt(k)));
    return true;
}
// Unpolarised dielectric reflectance; eta = n_back / n_front, cosI signed w.r.t. the front normal.
inline double fresnelDielectric(double cosI, double eta) {
    cosI = clampd(cosI, -1, 1);
    if (cosI < 0) {
        eta = 1 / eta;
        cosI = -cosI;
    }
    const double s2t = (1 - cosI * cosI) / (eta * eta);
    if (s2t >= 1) return 1;
    const double ct = sqrt0(1 - s2t);
    const double rpa = (eta * cosI - ct) / (eta * cosI + ct);
    const double rpe = (cosI - eta * ct) / (cosI + eta * ct);
    return (rpa * rpa + rpe * rpe) / 2;
}
// Complex helpers and conductor reflectance with relative index (n + ik) / n_medium.
struct C2 {
    double re, im;
};
inline C2 cmul(C2 a, C2 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline C2 cdiv(C2 a, C2 b) {
    const double d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
inline C2 csqrt(C2 z) {
    const double r = std::sqrt(z.re * z.re + z.im * z.im);
    return {std::sqrt(std::max(0.0, (r + z.re) / 2)), std::copysign(std::sqrt(std::max(0.0, (r - z.re) / 2)), z.im)};
}
inline double fresnelConductor(double cosI, C2 eta) {
    cosI = clampd(cosI, 0, 1);
    const C2 s2t = cdiv({1 - cosI * cosI, 0}, cmul(eta, eta));
    const C2 ct = csqrt({1 - s2t.re, -s2t.im});
    const C2 ec{eta.re * cosI, eta.im * cosI}, ect = cmul(eta, ct);
    const C2 rpa = cdiv({ec.re - ct.re, ec.im - ct.im}, {ec.re + ct.re, ec.im + ct.im});
    const C2 rpe = cdiv({cosI - ect.re, -ect.im}, {cosI + ect.re, ect.im});
    return (rpa.re * rpa.re + rpa.im * rpa.im + rpe.re * rpe.re + rpe.im * rpe.im) / 2;
}
inline bool sameHemi(V3 a, V3 b) { return a.z * b.z > 0; }
// pbrt-v4 refraction of wi (pointing away) about n; eta = n_t / n_i relative to n's side.
inline bool refractLocal(V3 wi, V3 n, double eta, double& etap, V3& wt) {
    double ci = dot(n, wi);
    etap = 1;
    wt = wi;
    if (ci < 0) {
        eta = 1 / eta;
        ci = -ci;
        n = -n;
    }
    const double s2t = std::max(0.0, 1 - ci * ci) / (eta * eta);
    if (s2t >= 1) return false;
    const double ct = sqrt0(1 - s2t);
    wt = -wi / eta + n * (ci / eta - ct);
    etap = eta;
    return true;
}
inline double branchProbability(double R, double floor) { return floor > 0 && R < 1 ? clampd(R, floor, 1 - floor) : R; }

// Isotropic GGX with visible-normal sampling.
constexpr double kSmoothRoughness = 1e-3;
inline double ggxAlpha(double roughness) { return std::max(roughness, 1e-4); }
inline double ggxD(double a, V3 m) {
    const double c2 = m.z * m.z;
    if (c2 <= 0) return 0;
    const double c4 = c2 * c2;
    if (c4 < 1e-16) return 0;
    const double e = (1 - c2) / c2 / (a * a);
    return 1 / (kPi * a * a * c4 * sq(1 + e));
}
inline double ggxLambda(double a, V3 w) {
    const double c2 = w.z * w.z;
    if (c2 <= 0) return 0;
    return (std::sqrt(1 + a * a * ((1 - c2) / c2)) - 1) / 2;
}
inline double ggxG1(double a, V3 w) { return 1 / (1 + ggxLambda(a, w)); }
inline double ggxG(double a, V3 wo, V3 wi) { return 1 / (1 + ggxLambda(a, wo) + ggxLambda(a, wi)); }
inline double ggxPdf(double a, V3 w, V3 m) {
    if (w.z == 0) return 0;
    return ggxG1(a, w) / std::fabs(w.z) * ggxD(a, m) * std::fabs(dot(w, m));
}
inline V3 ggxSampleWm(double a, V3 w, double u1, double u2) {
    V3 wh = normalize(V3(a * w.x, a * w.y, w.z));
    if (wh.z < 0) wh = -wh;
    const V3 t1 = wh.z < 0.99999 ? normalize(cross(V3(0, 0, 1), wh)) : V3(1, 0, 0);
    const V3 t2 = cross(wh, t1);
    const double r = std::sqrt(u1), phi = 2 * kPi * u2;
    const double px = r * std::cos(phi);
    double py = r * std::sin(phi);
    const double h = std::sqrt(std::max(0.0, 1 - px * px)), s = (1 + wh.z) / 2;
    py = (1 - s) * h + s * py;
    const double pz = sqrt0(1 - px * px - py * py);
    const V3 nh = t1 * px + t2 * py + wh * pz;
    return normalize(V3(a * nh.x, a * nh.y, std::max(1e-6, nh.z)));
}
struct Lobe {
    V3 wi;
    double weight = 0, pdf = 0;
    bool transmitted = false;
};
// Walter et al. (2007) rough dielectric in pbrt-v4's form; eta = n_back / n_front.
inline bool sampleRoughDielectric(double a, V3 wo, double eta, double uc, double u1, double u2, bool radiance, Lobe& s) {
    const V3 wm = ggxSampleWm(a, wo, u1, u2);
    const double R = fresnelDielectric(dot(wo, wm), eta), T = 1 - R;
    s = Lobe{};
    s.wi = wo;
    if (uc < R / (R + T)) {
        s.wi = reflect(-wo, wm);
        if (!sameHemi(wo, s.wi)) return false;
        s.pdf = ggxPdf(a, wo, wm) / (4 * std::fabs(dot(wo, wm))) * R / (R + T);
        const double f = ggxD(a, wm) * ggxG(a, wo, s.wi) * R / (4 * std::fabs(s.wi.z) * std::fabs(wo.z));
        s.weight = f * std::fabs(s.wi.z) / s.pdf;
    } else {
        double etap;
        V3 wi;
        if (!refractLocal(wo, wm, eta, etap, wi)) return false;
        if (sameHemi(wo, wi) || wi.z == 0) return false;
        s.wi = wi;
        s.transmitted = true;
        const double den = sq(dot(wi, wm) + dot(wo, wm) / etap);
        s.pdf = ggxPdf(a, wo, wm) * (std::fabs(dot(wi, wm)) / den) * T / (R + T);
        double ft = T * ggxD(a, wm) * ggxG(a, wo, wi) * std::fabs(dot(wi, wm) * dot(wo, wm) / (wi.z * wo.z * den));
        if (radiance) ft /= sq(etap);
        s.weight = ft * std::fabs(wi.z) / s.pdf;
    }
    return std::isfinite(s.weight) && s.pdf > 0;
}
inline double evalRoughDielectric(double a, V3 wo, V3 wi, double eta, bool radiance, double& pdf) {
    pdf = 0;
    const double co = wo.z, ci = wi.z;
    if (co == 0 || ci == 0) return 0;
    const bool refl = co * ci > 0;
    double etap = 1;
    if (!refl) etap = co > 0 ? eta : 1 / eta;
    V3 wm = wi * etap + wo;
    if (dot(wm, wm) == 0) return 0;
    wm = normalize(wm);
    if (wm.z < 0) wm = -wm;
    if (dot(wm, wi) * ci < 0 || dot(wm, wo) * co < 0) return 0;
    const double R = fresnelDielectric(dot(wo, wm), eta), T = 1 - R;
    if (refl) {
        pdf = ggxPdf(a, wo, wm) / (4 * std::fabs(dot(wo, wm))) * R / (R + T);
        return ggxD(a, wm) * ggxG(a, wo, wi) * R / std::fabs(4 * ci * co);
    }
    const double den = sq(dot(wi, wm) + dot(wo, wm) / etap);
    pdf = ggxPdf(a, wo, wm) * std::fabs(dot(wi, wm)) / den * T / (R + T);
    double ft = ggxD(a, wm) * T * ggxG(a, wo, wi) * std::fabs(dot(wi, wm) * dot(wo, wm) / (den * ci * co));
    if (radiance) ft /= sq(etap);
    return ft;
}
// Rough conductor from the upper hemisphere, reflectance F·tint.
inline bool sampleRoughConductor(double a, V3 w, C2 eta, double tint, double u1, double u2, Lobe& s) {
    const V3 wm = ggxSampleWm(a, w, u1, u2);
    s = Lobe{};
    s.wi = reflect(-w, wm);
    const double F = fresnelConductor(std::fabs(dot(w, wm)), eta) * tint;
    if (s.wi.z <= 0) return false;
    s.pdf = ggxPdf(a, w, wm) / (4 * std::fabs(dot(w, wm)));
    if (!(s.pdf > 0)) return false;
    const double f = ggxD(a, wm) * F * ggxG(a, w, s.wi) / (4 * s.wi.z * w.z);
    s.weight = f * s.wi.z / s.pdf;
    return std::isfinite(s.weight);
}
inline double evalRoughConductor(double a, V3 wo, V3 wi, C2 eta, double tint, double& pdf) {
    pdf = 0;
    V3 wm = wo + wi;
    if (dot(wm, wm) == 0) return 0;
    wm = normalize(wm);
    const double F = fresnelConductor(std::fabs(dot(wo, wm)), eta) * tint;
    pdf = ggxPdf(a, wo, wm) / (4 * std::fabs(dot(wo, wm)));
    return ggxD(a, wm) * F * ggxG(a, wo, wi) / (4 * wi.z * wo.z);
}

// ---------------------------------------------------------------- spectra

constexpr double kLambdaMin = 360.0, kLambdaMax = 830.0;
// Wyman–Sloan–Shirley (2013) multi-lobe fit of the CIE 1931 2° observer.
inline double cieLobe(double x, double mu, double s1, double s2) {
    const double t = (x - mu) / (x < mu ? s1 : s2);
    return std::exp(-0.5 * t * t);
}
inline double cieX(double l) {
    return 1.056 * cieLobe(l, 599.8, 37.9, 31.0) + 0.362 * cieLobe(l, 442.0, 16.0, 26.7) - 0.065 * cieLobe(l, 501.1, 20.4, 26.2);
}
inline double cieY(double l) { return 0.821 * cieLobe(l, 568.8, 46.9, 40.5) + 0.286 * cieLobe(l, 530.9, 16.3, 31.1); }
inline double cieZ(double l) { return 1.217 * cieLobe(l, 437.0, 11.8, 36.0) + 0.681 * cieLobe(l, 459.0, 26.0, 13.8); }
// Visible-wavelength importance sampling, pdf ∝ sech²(0.0072 (λ − 538)) on [360, 830] nm.
inline double sampleVisible(double u) {
    return clampd(538 - 138.888889 * std::atanh(0.85691062 - 1.82750197 * u), kLambdaMin, kLambdaMax);
}
inline double visiblePdf(double l) {
    if (l < kLambdaMin || l > kLambdaMax) return 0;
    return 0.0039398042 / sq(std::cosh(0.0072 * (l - 538)));
}
// Composite Simpson over [360, 830] nm with 0.5 nm steps.
template <class F>
double integrateVisible(F f) {
    const int n = int((kLambdaMax - kLambdaMin) * 2);
    const double h = (kLambdaMax - kLambdaMin) / n;
    double s = f(kLambdaMin) + f(kLambdaMax);
    for (int i = 1; i < n; ++i) s += f(kLambdaMin + i * h) * ((i & 1) ? 4 : 2);
    return s * h / 3;
}
inline double cieYIntegral() {
    static const double v = integrateVisible([](double l) { return cieY(l); });
    return v;
}
// Planck radiance for λ in nm, evaluated in µm units (c1 = 2hc², c2 = hc/k).
inline double planck(double lnm, double kelvin) {
    const double c1 = 1.191042972e8, c2 = 1.438776877e4;
    const double l = lnm * 1e-3, l2 = l * l;
    return 1e6 * c1 / (l2 * l2 * l * (std::exp(c2 / (l * kelvin)) - 1));
}
inline double ciddorAir(double l2) {
    const double s2 = 1 / l2;
    return 1 + 0.05792105 / (238.0185 - s2) + 0.00167917 / (57.362 - s2);
}

// ---------------------------------------------------------------- triangles

// The cross product of a triangle's two shorter edges (robust for slivers).
inline V3 triNormal(V3 v0, V3 v1, V3 v2) {
    const V3 a = v1 - v0, b = v2 - v0, c = v2 - v1;
    const double la = dot(a, a), lb = dot(b, b), lc = dot(c, c);
    if (lc >= la && lc >= lb) return cross(a, b);
    return la >= lb ? cross(b, c) : cross(a, c);
}
// Woop–Benthin–Wald (JCGT 2013) watertight line–triangle crossing, edges included.
inline bool lineCrossesTri(V3 o, V3 d, V3 v0, V3 v1, V3 v2) {
    const double ax_ = std::fabs(d.x), ay_ = std::fabs(d.y), az_ = std::fabs(d.z);
    const int kz = ax_ > ay_ ? (ax_ > az_ ? 0 : 2) : (ay_ > az_ ? 1 : 2);
    const int kx = kz == 2 ? 0 : kz + 1, ky = kx == 2 ? 0 : kx + 1;
    const double sx = d[kx] / d[kz], sy = d[ky] / d[kz];
    const V3 a = v0 - o, b = v1 - o, c = v2 - o;
    const double px = a[kx] - sx * a[kz], py = a[ky] - sy * a[kz];
    const double qx = b[kx] - sx * b[kz], qy = b[ky] - sy * b[kz];
    const double rx = c[kx] - sx * c[kz], ry = c[ky] - sy * c[kz];
    const double u = rx * qy - ry * qx, v = px * ry - py * rx, w = qx * py - qy * px;
    if ((u < 0 || v < 0 || w < 0) && (u > 0 || v > 0 || w > 0)) return false;
    return u + v + w != 0;
}

}  // namespace oracle
