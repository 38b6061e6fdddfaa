// TEST INFRASTRUCTURE ONLY — the double-precision CPU oracle (see oracle_core.hpp).
//
// Sequential lens analysis restated in double from the reference (sag surfaces: oracle_shapes.hpp):
//   transmission through a smooth dielectric (weight T)       Tracer::walk PrimaryTransmission,
//                                                             src/owe/transport/transport.cpp:531-595,
//                                                             scattering.cpp:65-108, optics.hpp:14-54
//   ray acceptance (every surface refracts, then escape)      traceThrough, src/owe/analysis/analysis.cpp:49-70
//   paraxial system matrix, EFL, BFL, stop, entrance pupil    src/owe/scene/prescription.cpp:193-321
//   spot: best focus −cov/var, RMS radius, centroid           focusedSpot, src/owe/analysis/analysis.cpp:85-113
// The lens is traced surface after surface (each ray meets the next row's sag): for the rays
// traceThrough accepts — refracted once by every surface, then escaping — that is the order the
// world intersection finds them in.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../include/owe_hip_lens.h"
#include "oracle_scene.hpp"

using namespace oracle;

namespace {

using shapes::intersectSag;
using shapes::sag;

// A lens row: its sag (a mirror's annulus outside its hole), the vertex position along the fixed
// chart, and whether it reflects.
struct Surf : shapes::Sag {
    double z = 0;
    bool reflect = false;
};

Surf makeSurf(const owe_lens_surface& r, double z) {
    Surf s;
    static_cast<shapes::Sag&>(s) =
        shapes::makeSag(r.curvature, r.conic, r.asphere, r.semi_diameter, r.reflect ? r.hole_radius : 0.0);
    s.z = z;
    s.reflect = r.reflect != 0;
    return s;
}

struct Lens {
    std::vector<Surf> surf;
    std::vector<double> nAfter;  // per surface at the wavelength
    double nAmbient = 1;
};

// Whether the segment a → b (b's z may be infinite along d) crosses the body of a mirror other than
// surfaces `from` and `to`: its vertex-plane annulus hole ≤ r ≤ semi-diameter.
bool obstructed(const Lens& L, V3 a, V3 d, double zEnd, size_t from, size_t to) {
    for (size_t j = 0; j < L.surf.size(); ++j) {
        const Surf& m = L.surf[j];
        if (!m.reflect || j == from || j == to) continue;
        const double lo = std::min(a.z, zEnd), hi = std::max(a.z, zEnd);
        if (!(m.z > lo && m.z < hi) || d.z == 0) continue;
        const V3 x = a + d * ((m.z - a.z) / d.z);
        const double r2 = x.x * x.x + x.y * x.y;
        if (r2 >= m.rMin * m.rMin && r2 <= m.rMax * m.rMax) return true;
    }
    return false;
}

// One ray through the lens, surface after surface: true when every surface takes it — a
// refracting row refracts it (traceThrough's acceptance), a mirror row reflects it — no mirror body
// stops it on the way, and it leaves along +z; p, d the last surface's point and the outgoing
// direction, T the product of 1 − R over the refractions.
bool traceRay(const Lens& L, V3 o, V3 d, V3& p, V3& dOut, double& T) {
    T = 1;
    double nPrev = L.nAmbient;
    const size_t N = L.surf.size();
    for (size_t i = 0; i < N; ++i) {
        const Surf& s = L.surf[i];
        double t;
        V3 q, n;
        const V3 ol(o.x, o.y, o.z - s.z);
        if (!intersectSag(s, ol, d, 0, kInf, t, q, n)) return false;
        const V3 qw(q.x, q.y, q.z + s.z);
        if (obstructed(L, o, d, qw.z, i == 0 ? N : i - 1, i)) return false;
        V3 w;
        if (s.reflect) {
            w = d - n * (2 * dot(d, n));
        } else {
            const double nNext = L.nAfter[i];
            const double R = fresnelDielectric(std::fabs(dot(d, n)), nNext / nPrev);
            if (!refractDir(d, n, nPrev, nNext, w)) return false;
            T *= 1 - R;
            nPrev = nNext;
        }
        o = qw;
        d = w;
    }
    if (obstructed(L, o, d, d.z > 0 ? kInf : -kInf, N - 1, N)) return false;
    p = o;
    dOut = d;
    return d.z > 0;
}

// Pupil point i of n on the entrance pupil of radius rEP: with rings > 0 spotAt's hexapolar
// pattern (analysis.cpp:124-131), walked ring by ring; else a Fibonacci lattice over the disc of
// radius 0.999 rEP (the batch's sampling for arbitrary ray counts).
void pupilPoint(uint64_t i, uint64_t n, int rings, double rEP, double& x, double& y) {
    if (rings > 0) {
        uint64_t seen = 0;
        for (int k = 0; k <= rings; ++k) {
            const uint64_t m = k == 0 ? 1 : 6 * uint64_t(k);
            if (i < seen + m) {
                const uint64_t j = i - seen;
                const double rr = rEP * double(k) / rings * 0.999;
                const double a = 2 * kPi * double(j) / double(m);
                x = rr * std::cos(a);
                y = rr * std::sin(a);
                return;
            }
            seen += m;
        }
        x = y = std::nan("");
        return;
    }
    const double rr = 0.999 * rEP * std::sqrt((double(i) + 0.5) / double(n));
    double f = double(i) * 0.6180339887498949;
    f -= std::floor(f);
    const double a = 2 * kPi * f;
    x = rr * std::cos(a);
    y = rr * std::sin(a);
}

struct Paraxial {
    double efl = kInf, bfl = kInf, epz = 0, epr = 0, length = 0;
    double lastZ = 0;  // the last vertex along the fixed chart (signed gaps)
    bool afocal = false;
    double angularMagnification = 0;
};

// paraxialAnalysis with the object at infinity (prescription.cpp:193-321): a refracting row has
// power (n' − n) c, a mirror −2 dir n c with dir the travel along the fixed chart, which it reverses.
Paraxial paraxial(const std::vector<owe_lens_surface>& s, const std::vector<double>& nAfter, double n0) {
    const size_t N = s.size();
    std::vector<double> nBefore(N), phi(N), t(N);
    double nPrev = n0, dir = 1;
    for (size_t i = 0; i < N; ++i) {
        nBefore[i] = nPrev;
        phi[i] = s[i].reflect ? -2 * dir * nAfter[i] * s[i].curvature : (nAfter[i] - nPrev) * s[i].curvature;
        if (s[i].reflect) dir = -dir;
        t[i] = std::fabs(s[i].thickness);
        nPrev = nAfter[i];
    }
    double a = 1, b = 0, c = 0, d = 1;
    for (size_t i = 0; i < N; ++i) {
        // R = [1 0; −φ 1], then T = [1 t/n; 0 1]
        const double a2 = a, b2 = b, c2 = c - phi[i] * a, d2 = d - phi[i] * b;
        a = a2;
        b = b2;
        c = c2;
        d = d2;
        if (i + 1 < N) {
            const double k = t[i] / nAfter[i];
            a = a + k * c;
            b = b + k * d;
        }
    }
    Paraxial r;
    for (size_t i = 0; i + 1 < N; ++i) {
        r.length += t[i];
        r.lastZ += s[i].thickness;
    }
    const double nk = nAfter[N - 1];
    const double scale = 1e-12 * (1 + std::fabs(a) + std::fabs(d));
    r.afocal = std::fabs(c) * std::max(1.0, r.length) < scale * 1e3 || std::fabs(c) < 1e-9;
    if (!r.afocal) {
        r.efl = -nk / c;
        r.bfl = -nk * a / c;
    } else {
        r.angularMagnification = d * n0 / nk;
    }
    int stop = -1;
    for (size_t i = 0; i < N; ++i)
        if (s[i].stop) stop = int(i);
    if (stop < 0) {
        double best = -1, y = 1, nu = 0;
        for (size_t i = 0; i < N; ++i) {
            const double ratio = std::fabs(y) / std::max(s[i].semi_diameter, 1e-12);
            if (ratio > best) {
                best = ratio;
                stop = int(i);
            }
            nu -= phi[i] * y;
            if (i + 1 < N) y += t[i] * nu / nAfter[i];
        }
    }
    const double stopR = s[size_t(stop)].semi_diameter;
    double y = 0, nu = 1e-3 * nBefore[size_t(stop)], yM = stopR, nuM = 0;
    for (int i = stop - 1; i >= 0; --i) {
        y -= t[size_t(i)] * nu / nAfter[size_t(i)];
        yM -= t[size_t(i)] * nuM / nAfter[size_t(i)];
        nu += phi[size_t(i)] * y;
        nuM += phi[size_t(i)] * yM;
    }
    if (stop == 0) {
        r.epz = 0;
        r.epr = stopR;
    } else {
        const double u = nu / n0;
        if (u != 0) {
            const double dz = -y / u;
            r.epz = dz;
            r.epr = std::fabs(yM + dz * nuM / n0);
        }
    }
    return r;
}

IndexModel indexOf(const owe_medium& m) {
    IndexModel x;
    x.kind = m.index_kind;
    x.terms = m.terms;
    x.relAir = m.rel_air != 0;
    for (int k = 0; k < 4; ++k) {
        x.a[k] = m.coeff_a[k];
        x.b[k] = m.coeff_b[k];
    }
    return x;
}

struct Setup {
    Lens L;
    V3 dir;
    double zStart = 0, zEP = 0, rEP = 0;
    Paraxial px;
};

Setup setup(const owe_medium* media, const owe_lens_surface* surfaces, const owe_lens& lens, double fieldDeg,
            double lambda, double lambdaRef) {
    Setup u;
    std::vector<owe_lens_surface> rows(surfaces + lens.first_surface, surfaces + lens.first_surface + lens.surface_count);
    std::vector<double> nRef, nAt;
    double z = 0;
    for (const owe_lens_surface& r : rows) {
        u.L.surf.push_back(makeSurf(r, z));
        z += r.thickness;
        nRef.push_back(indexOf(media[r.medium]).n(lambdaRef));
        nAt.push_back(indexOf(media[r.medium]).n(lambda));
    }
    u.L.nAfter = nAt;
    u.L.nAmbient = indexOf(media[lens.ambient]).n(lambda);
    u.px = paraxial(rows, nRef, indexOf(media[lens.ambient]).n(lambdaRef));
    const double f = fieldDeg * kPi / 180;
    u.dir = V3(0, std::sin(f), std::cos(f));
    u.zStart = -0.05 - 0.1 * u.px.length;
    u.zEP = u.px.epz;
    u.rEP = u.px.epr;
    return u;
}

}  // namespace

extern "C" {

// Traces rays index[i] of `lens` at wavelength lambda: out n × 7 = p, d, T (T = 0 and NaN p, d when
// the ray does not arrive).
void oracle_lens_trace(const owe_medium* media, const owe_lens_surface* surfaces, const owe_lens* lens, double fieldDeg,
                       double lambda, double lambdaRef, const uint64_t* index, uint64_t n, double* out) {
    const Setup u = setup(media, surfaces, *lens, fieldDeg, lambda, lambdaRef);
    for (uint64_t k = 0; k < n; ++k) {
        double x, y;
        pupilPoint(index[k], lens->rays, lens->pupil_rings, u.rEP, x, y);
        const V3 P(x, y, u.zEP);
        const V3 o = P + u.dir * ((u.zStart - u.zEP) / u.dir.z);
        V3 p, d;
        double T;
        double* r = out + 7 * k;
        if (traceRay(u.L, o, u.dir, p, d, T)) {
            r[0] = p.x, r[1] = p.y, r[2] = p.z, r[3] = d.x, r[4] = d.y, r[5] = d.z, r[6] = T;
        } else {
            for (int q = 0; q < 6; ++q) r[q] = std::nan("");
            r[6] = 0;
        }
    }
}

// The spot sums of one lens at one wavelength over all its rays (OWE_LENS_MOMENTS doubles), and the
// paraxial quantities (efl, bfl, entrance pupil z and radius, length) at lambdaRef.
void oracle_lens_moments(const owe_medium* media, const owe_lens_surface* surfaces, const owe_lens* lens,
                         double fieldDeg, double lambda, double lambdaRef, int threads, double* moments,
                         double* paraxialOut) {
    const Setup u = setup(media, surfaces, *lens, fieldDeg, lambda, lambdaRef);
    if (threads <= 0) threads = int(std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
    std::vector<std::vector<double>> part(size_t(threads), std::vector<double>(OWE_LENS_MOMENTS, 0.0));
    std::vector<std::thread> pool;
    const uint64_t n = lens->rays;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            std::vector<double>& m = part[size_t(t)];
            for (uint64_t i = n * uint64_t(t) / uint64_t(threads); i < n * uint64_t(t + 1) / uint64_t(threads); ++i) {
                double x, y;
                pupilPoint(i, n, lens->pupil_rings, u.rEP, x, y);
                const V3 P(x, y, u.zEP);
                const V3 o = P + u.dir * ((u.zStart - u.zEP) / u.dir.z);
                V3 p, d;
                double T;
                if (!traceRay(u.L, o, u.dir, p, d, T)) continue;
                const double bx = d.x / d.z, by = d.y / d.z;
                const double ax = u.px.afocal ? std::atan(bx) : p.x - bx * p.z;
                const double ay = u.px.afocal ? std::atan(by) : p.y - by * p.z;
                m[0] += 1;
                m[1] += ax;
                m[2] += ay;
                m[3] += bx;
                m[4] += by;
                m[5] += ax * ax + ay * ay;
                m[6] += bx * bx + by * by;
                m[7] += ax * bx + ay * by;
                m[8] += T;
            }
        });
    for (auto& th : pool) th.join();
    for (int q = 0; q < OWE_LENS_MOMENTS; ++q) {
        moments[q] = 0;
        for (int t = 0; t < threads; ++t) moments[q] += part[size_t(t)][size_t(q)];
    }
    if (paraxialOut) {
        paraxialOut[0] = u.px.efl;
        paraxialOut[1] = u.px.bfl;
        paraxialOut[2] = u.px.epz;
        paraxialOut[3] = u.px.epr;
        paraxialOut[4] = u.px.length;
        paraxialOut[5] = u.px.afocal ? 1 : 0;
        paraxialOut[6] = u.px.angularMagnification;
        paraxialOut[7] = u.px.lastZ;
    }
}

// Rays index[i] of `lens` traced with its surfaces perturbed — delta[k] added to parameter k
// (surface params[2k], field params[2k + 1]: curvature, conic, gap after the surface, A4..A10) —
// from the launch of the unperturbed lens: out n × 7 as oracle_lens_trace.
void oracle_lens_trace_perturbed(const owe_medium* media, const owe_lens_surface* surfaces, const owe_lens* lens,
                                 double fieldDeg, double lambda, double lambdaRef, uint32_t nParams,
                                 const int32_t* params, const double* delta, const uint64_t* index, uint64_t n,
                                 double* out) {
    const Setup u = setup(media, surfaces, *lens, fieldDeg, lambda, lambdaRef);
    std::vector<owe_lens_surface> rows(surfaces + lens->first_surface, surfaces + lens->first_surface + lens->surface_count);
    for (uint32_t k = 0; k < nParams; ++k) {
        owe_lens_surface& r = rows[size_t(params[2 * k])];
        const int f = params[2 * k + 1];
        if (f == OWE_LENS_PARAM_CURVATURE) r.curvature += delta[k];
        else if (f == OWE_LENS_PARAM_CONIC) r.conic += delta[k];
        else if (f == OWE_LENS_PARAM_THICKNESS) r.thickness += delta[k];
        else r.asphere[f - OWE_LENS_PARAM_A4] += delta[k];
    }
    Lens L = u.L;
    double z = 0;
    for (size_t i = 0; i < rows.size(); ++i) {
        L.surf[i] = makeSurf(rows[i], z);
        z += rows[i].thickness;
    }
    for (uint64_t k = 0; k < n; ++k) {
        double x, y;
        pupilPoint(index[k], lens->rays, lens->pupil_rings, u.rEP, x, y);
        const V3 P(x, y, u.zEP);
        const V3 o = P + u.dir * ((u.zStart - u.zEP) / u.dir.z);
        V3 p, d;
        double T;
        double* r = out + 7 * k;
        if (traceRay(L, o, u.dir, p, d, T)) {
            r[0] = p.x, r[1] = p.y, r[2] = p.z, r[3] = d.x, r[4] = d.y, r[5] = d.z, r[6] = T;
        } else {
            for (int q = 0; q < 6; ++q) r[q] = std::nan("");
            r[6] = 0;
        }
    }
}

// SagSurface::intersect alone, in the surface's frame: t, point, normal (NaN t when no hit).
void oracle_sag_intersect(const owe_lens_surface* row, const double* o, const double* d, double tmin, double tmax,
                          double* out) {
    const Surf s = makeSurf(*row, 0);
    double t;
    V3 p, n;
    if (intersectSag(s, V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2]), tmin, tmax, t, p, n)) {
        out[0] = t, out[1] = p.x, out[2] = p.y, out[3] = p.z, out[4] = n.x, out[5] = n.y, out[6] = n.z;
    } else {
        for (int q = 0; q < 7; ++q) out[q] = std::nan("");
    }
}

double oracle_sag_value(const owe_lens_surface* row, double r) { return sag(makeSurf(*row, 0), r); }

// One ray (origin o, unit direction d, lens frame) through `lens` at wavelength lambda: out = p, d, T
// (T = 0 and NaN p, d when it does not arrive).
void oracle_lens_trace_ray(const owe_medium* media, const owe_lens_surface* surfaces, const owe_lens* lens, double lambda,
                           const double* o, const double* d, double* out) {
    const Setup u = setup(media, surfaces, *lens, 0, lambda, lambda);
    V3 p, w;
    double T;
    if (traceRay(u.L, V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2]), p, w, T)) {
        out[0] = p.x, out[1] = p.y, out[2] = p.z, out[3] = w.x, out[4] = w.y, out[5] = w.z, out[6] = T;
    } else {
        for (int q = 0; q < 6; ++q) out[q] = std::nan("");
        out[6] = 0;
    }
}

}  // extern "C"
