// TEST INFRASTRUCTURE ONLY — the double-precision CPU oracle (see oracle_core.hpp).
//
// Analytic shapes beyond planes and spheres, restated in double from the reference, each in its
// own frame:
//   quadric roots, refinement from the approximate hit,    src/owe/core/shared/surfaces.hpp:13-94,
//   first refined hit                                      src/owe/scene/geometry.cpp:34-48
//   conic sag, slope, normal, conic sheet                  src/owe/core/shared/surfaces.hpp:140-171
//   SagSurface: quadric path and numeric asphere (96       src/owe/scene/geometry.cpp:69-166
//   samples + bisection), annulus acceptance
//   CylinderShape (geometric-form roots, z trim)           surfaces.hpp:118-137, geometry.cpp:490-515
//   WaveSurface: plane-wave sum with smoothstep taper or    geometry.hpp:84-142, geometry.cpp:188-344,
//   sampled bicubic Hermite nodes; Lipschitz march +       surfaces.hpp:178-200
//   Newton, bisection inside a minimum step
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "oracle_core.hpp"

namespace oracle {
namespace shapes {

constexpr double kTinyD = 1e-300, kRefineD = 1e-6, kBisectD = 1e-15, kRelTolD = 1e-9;
constexpr int kBisectStepsD = 80;

inline int solveQuadratic(double a, double b, double c, double& t0, double& t1) {
    t0 = t1 = 0;
    if (std::fabs(a) < kTinyD || std::fabs(a) < 1e-14 * std::fabs(b)) {
        if (b == 0) return 0;
        t0 = t1 = -c / b;
        return 1;
    }
    const double disc = b * b - 4 * a * c;
    if (disc < 0) return 0;
    const double q = -(b + std::copysign(std::sqrt(disc), b)) / 2;
    t0 = q / a;
    t1 = q != 0 ? c / q : t0;
    if (t0 > t1) std::swap(t0, t1);
    return 2;
}

// Quadric coefficients A t² + B t + C of the sag c(x² + y²) + c(1 + k) z² − 2z = 0 (kind 0) or the
// cylinder x² + y² = R² (kind 1, a = R²).
struct Quadric {
    int kind;
    double a, b;
};
inline void quadricCoeffs(const Quadric& q, V3 o, V3 d, double& A, double& B, double& C) {
    if (q.kind == 0) {
        A = q.a * (d.x * d.x + d.y * d.y) + q.b * d.z * d.z;
        B = 2 * (q.a * (o.x * d.x + o.y * d.y) + q.b * o.z * d.z - d.z);
        C = q.a * (o.x * o.x + o.y * o.y) + q.b * o.z * o.z - 2 * o.z;
    } else {
        A = d.x * d.x + d.y * d.y;
        B = 2 * (o.x * d.x + o.y * d.y);
        C = o.x * o.x + o.y * o.y - q.a;
    }
}
inline double refineRoot(const Quadric& q, V3 o, V3 d, double t) {
    double A, B, C, u0, u1;
    quadricCoeffs(q, o + d * t, d, A, B, C);
    if (solveQuadratic(A, B, C, u0, u1) == 0) return t;
    const double du = std::fabs(u0) < std::fabs(u1) ? u0 : u1;
    return std::fabs(du) < kRefineD * (1 + std::fabs(t)) ? t + du : t;
}
inline bool mayLieIn(double t, double tmin, double tmax) {
    const double m = kRefineD * (1 + std::fabs(t));
    return !(t + m <= tmin || t - m >= tmax);
}

// ---------------------------------------------------------------- sag

struct Sag {
    double c = 0, k = 0, A[4] = {0, 0, 0, 0};
    int nA = 0;
    double rMax = 0, rMin = 0, zMin = 0, zMax = 0;
};

inline double sag(const Sag& s, double r) {
    const double r2 = r * r;
    double z = s.c * r2 / (1 + sqrt0(1 - (1 + s.k) * s.c * s.c * r2));
    double rp = r2 * r2;
    for (int i = 0; i < s.nA; ++i) {
        z += s.A[i] * rp;
        rp *= r2;
    }
    return z;
}
inline double dsag(const Sag& s, double r) {
    const double root = sqrt0(1 - (1 + s.k) * s.c * s.c * (r * r));
    double d = root > 0 ? s.c * r / root : (s.c != 0 ? kInf : 0.0);
    double m = 4, rp = r * r * r;
    for (int i = 0; i < s.nA; ++i) {
        d += s.A[i] * m * rp;
        rp *= r * r;
        m += 2;
    }
    return d;
}
inline V3 sagNormal(const Sag& s, V3 p) {
    const double r = std::sqrt(p.x * p.x + p.y * p.y);
    if (r < kTinyD) return V3(0, 0, 1);
    const double sl = dsag(s, r);
    return normalize(V3(-sl * p.x / r, -sl * p.y / r, 1));
}
inline bool sagAccept(const Sag& s, V3 p) {
    const double r2 = p.x * p.x + p.y * p.y;
    return r2 <= s.rMax * s.rMax && r2 >= s.rMin * s.rMin;
}
inline bool onConicSheet(const Sag& s, double z) { return s.c == 0 || (1 + s.k) * s.c * z <= 1 + kRelTolD; }

// SagSurface's constructor: trailing zero coefficients dropped, the slab [zMin, zMax] from 513
// samples padded by 1e-9 + 1e-3 of its height.
inline Sag makeSag(double c, double k, const double* A4, double rMax, double rMin) {
    Sag s;
    s.c = c;
    s.k = k;
    s.nA = 4;
    for (int i = 0; i < 4; ++i) s.A[i] = A4[i];
    while (s.nA > 0 && s.A[s.nA - 1] == 0) --s.nA;
    s.rMax = rMax;
    s.rMin = rMin;
    s.zMin = kInf;
    s.zMax = -kInf;
    for (int i = 0; i <= 512; ++i) {
        const double v = sag(s, rMin + (rMax - rMin) * i / 512);
        s.zMin = std::min(s.zMin, v);
        s.zMax = std::max(s.zMax, v);
    }
    const double pad = 1e-9 + 1e-3 * (s.zMax - s.zMin);
    s.zMin -= pad;
    s.zMax += pad;
    return s;
}

// Clips (ta, tb) to the cylinder of radius R about z; false when nothing remains.
inline bool clipCylinder(V3 o, V3 d, double R, double& ta, double& tb) {
    const double a = d.x * d.x + d.y * d.y;
    if (a < kTinyD) return o.x * o.x + o.y * o.y <= R * R && ta < tb;
    double q0, q1;
    if (solveQuadratic(a, 2 * (o.x * d.x + o.y * d.y), o.x * o.x + o.y * o.y - R * R, q0, q1) < 2) return false;
    ta = std::max(ta, q0);
    tb = std::min(tb, q1);
    return ta < tb;
}
inline bool clipSlab(double o, double d, double lo, double hi, double& ta, double& tb) {
    if (std::fabs(d) < kTinyD) return o >= lo && o <= hi && ta < tb;
    const double s0 = (lo - o) / d, s1 = (hi - o) / d;
    ta = std::max(ta, std::min(s0, s1));
    tb = std::min(tb, std::max(s0, s1));
    return ta < tb;
}

template <class F>
double bisect(F f, double lo, double hi, double flo) {
    for (int it = 0; it < kBisectStepsD && hi - lo > kBisectD * (1 + std::fabs(hi)); ++it) {
        const double mid = 0.5 * (lo + hi), fm = f(mid);
        if ((fm <= 0) == (flo <= 0)) {
            lo = mid;
            flo = fm;
        } else {
            hi = mid;
        }
    }
    return 0.5 * (lo + hi);
}

// SagSurface::intersect, t in (tmin, tmax).
inline bool intersectSag(const Sag& s, V3 o, V3 d, double tmin, double tmax, double& t, V3& p, V3& n) {
    if (s.nA == 0) {
        if (s.c == 0) {
            if (d.z == 0) return false;
            t = -o.z / d.z;
            if (t <= tmin || t >= tmax) return false;
            p = o + d * t;
            if (!sagAccept(s, p)) return false;
            n = sagNormal(s, p);
            return true;
        }
        const Quadric q{0, s.c, s.c * (1 + s.k)};
        double A, B, C, t0, t1;
        quadricCoeffs(q, o, d, A, B, C);
        if (solveQuadratic(A, B, C, t0, t1) == 0) return false;
        for (double tc : {t0, t1}) {
            if (!mayLieIn(tc, tmin, tmax)) continue;
            const double tr = refineRoot(q, o, d, tc);
            if (!(tr > tmin && tr < tmax)) continue;
            const V3 x = o + d * tr;
            if (!sagAccept(s, x) || !onConicSheet(s, x.z)) continue;
            t = tr;
            p = x;
            n = sagNormal(s, x);
            return true;
        }
        return false;
    }
    double ta = tmin, tb = tmax;
    if (!clipCylinder(o, d, s.rMax, ta, tb) || !clipSlab(o.z, d.z, s.zMin, s.zMax, ta, tb)) return false;
    auto f = [&](double tt) {
        const V3 x = o + d * tt;
        return sag(s, std::sqrt(x.x * x.x + x.y * x.y)) - x.z;
    };
    const int N = 96;
    double tPrev = ta, fPrev = f(ta);
    for (int i = 1; i <= N; ++i) {
        const double tCur = ta + (tb - ta) * i / N, fCur = f(tCur);
        if ((fPrev <= 0 && fCur >= 0) || (fPrev >= 0 && fCur <= 0)) {
            if (fPrev == 0 && tPrev <= tmin) {
                tPrev = tCur;
                fPrev = fCur;
                continue;
            }
            const double tr = bisect(f, tPrev, tCur, fPrev);
            const V3 x = o + d * tr;
            if (tr > tmin && tr < tmax && sagAccept(s, x)) {
                t = tr;
                p = x;
                n = sagNormal(s, x);
                return true;
            }
        }
        tPrev = tCur;
        fPrev = fCur;
    }
    return false;
}

// ---------------------------------------------------------------- cylinder

// Geometric-form roots of the infinite cylinder of radius R about z (cylinderRoots), ascending.
inline bool cylinderRoots(V3 o, V3 d, double R, double& t0, double& t1) {
    t0 = t1 = 0;
    const double A = d.x * d.x + d.y * d.y;
    if (A < kTinyD) return false;
    const double b = o.x * d.x + o.y * d.y;
    const double px = o.x - d.x * (b / A), py = o.y - d.y * (b / A);
    const double disc = A * (R * R - (px * px + py * py));
    if (disc < 0) return false;
    const double q = -(b + std::copysign(std::sqrt(disc), b));
    const double c = o.x * o.x + o.y * o.y - R * R;
    t0 = q / A;
    t1 = q != 0 ? c / q : t0;
    if (t0 > t1) std::swap(t0, t1);
    return true;
}

// CylinderShape::intersect: open cylinder x² + y² = R², z0 ≤ z ≤ z1, outward normal.
inline bool intersectCylinder(double R, double z0, double z1, V3 o, V3 d, double tmin, double tmax, double& t, V3& p,
                              V3& n) {
    double t0, t1;
    if (!cylinderRoots(o, d, R, t0, t1)) return false;
    const Quadric q{1, R * R, 0};
    for (double tc : {t0, t1}) {
        if (!mayLieIn(tc, tmin, tmax)) continue;
        const double tr = refineRoot(q, o, d, tc);
        if (!(tr > tmin && tr < tmax)) continue;
        const V3 x = o + d * tr;
        if (!(x.z >= z0 && x.z <= z1)) continue;
        t = tr;
        p = x;
        n = normalize(V3(x.x, x.y, 0));
        return true;
    }
    return false;
}

// ---------------------------------------------------------------- wave surface

struct PlaneWave {
    double amplitude, kx, ky, phase;
};

struct Wave {
    double hx = 0, hy = 0, margin = 0;
    double amp = 0, slope = 0;  // |h| ≤ amp, |∇h| ≤ slope
    std::vector<PlaneWave> waves;
    int nx = 0, ny = 0;         // sampled spline when nx > 0
    double dx = 0, dy = 0;
    std::vector<double> nodes;  // (h, ∂x, ∂y, ∂xy) per node, row by row from (−hx, −hy)
};

inline void waveTaper(double u, double half, double margin, double& w, double& dw) {
    w = 1;
    dw = 0;
    if (!(margin > 0)) return;
    const double e = half - std::fabs(u);
    if (e >= margin) return;
    if (e <= 0) {
        w = 0;
        return;
    }
    const double t = e / margin;
    w = t * t * (3 - 2 * t);
    dw = -(u < 0 ? -1.0 : 1.0) * 6 * t * (1 - t) / margin;
}

inline void hermiteBases(double u, double P[2], double Q[2], double dP[2], double dQ[2]) {
    P[0] = 2 * u * u * u - 3 * u * u + 1;
    P[1] = -2 * u * u * u + 3 * u * u;
    Q[0] = u * u * u - 2 * u * u + u;
    Q[1] = u * u * u - u * u;
    dP[0] = 6 * u * u - 6 * u;
    dP[1] = -6 * u * u + 6 * u;
    dQ[0] = 3 * u * u - 4 * u + 1;
    dQ[1] = 3 * u * u - 2 * u;
}

// WaveSurface::height: the tapered plane-wave sum, or the bicubic Hermite patch of the sampled
// nodes; zero outside the rectangle.
inline void waveHeight(const Wave& w, double x, double y, double& h, double& gx, double& gy) {
    h = gx = gy = 0;
    if (std::fabs(x) > w.hx || std::fabs(y) > w.hy) return;
    if (w.nx == 0) {
        double g = 0, sx = 0, sy = 0;
        for (const PlaneWave& p : w.waves) {
            const double arg = p.kx * x + p.ky * y + p.phase;
            const double s = std::sin(arg), c = std::cos(arg);
            g += p.amplitude * s;
            sx += p.amplitude * c * p.kx;
            sy += p.amplitude * c * p.ky;
        }
        double wx, dwx, wy, dwy;
        waveTaper(x, w.hx, w.margin, wx, dwx);
        waveTaper(y, w.hy, w.margin, wy, dwy);
        h = wx * wy * g;
        gx = dwx * wy * g + wx * wy * sx;
        gy = wx * dwy * g + wx * wy * sy;
        return;
    }
    const double fx = (x + w.hx) / w.dx, fy = (y + w.hy) / w.dy;
    const int i = std::clamp(int(fx), 0, w.nx - 2), j = std::clamp(int(fy), 0, w.ny - 2);
    double P[2], Q[2], dP[2], dQ[2], Pv[2], Qv[2], dPv[2], dQv[2];
    hermiteBases(fx - i, P, Q, dP, dQ);
    hermiteBases(fy - j, Pv, Qv, dPv, dQv);
    double f = 0, fu = 0, fv = 0;
    for (int b = 0; b < 2; ++b)
        for (int a = 0; a < 2; ++a) {
            const double* n = &w.nodes[4 * (size_t(j + b) * w.nx + (i + a))];
            const double c0 = n[0], cu = n[1] * w.dx, cv = n[2] * w.dy, cuv = n[3] * w.dx * w.dy;
            f += c0 * P[a] * Pv[b] + cu * Q[a] * Pv[b] + cv * P[a] * Qv[b] + cuv * Q[a] * Qv[b];
            fu += c0 * dP[a] * Pv[b] + cu * dQ[a] * Pv[b] + cv * dP[a] * Qv[b] + cuv * dQ[a] * Qv[b];
            fv += c0 * P[a] * dPv[b] + cu * Q[a] * dPv[b] + cv * P[a] * dQv[b] + cuv * Q[a] * dQv[b];
        }
    h = f;
    gx = fu / w.dx;
    gy = fv / w.dy;
}

inline V3 waveNormal(const Wave& w, double x, double y) {
    double h, gx, gy;
    waveHeight(w, x, y, h, gx, gy);
    return normalize(V3(-gx, -gy, 1));
}

// WaveSurface::intersect without the contact record: a ray leaving this surface (departed) ignores
// roots within 1e-9 of the surface's scale along the surface normal — the root it stands on.
inline bool intersectWave(const Wave& w, V3 o, V3 d, double tmin, double tmax, bool departed, double& t, V3& p, V3& n) {
    const double L = std::fabs(d.z) + w.slope * std::hypot(d.x, d.y);
    const double scale = 1 + std::max({w.hx, w.hy, w.amp});
    if (departed) tmin = std::max(tmin, 1e-9 * scale / std::max(L, 1e-300));
    double ta = tmin, tb = tmax;
    if (!clipSlab(o.x, d.x, -w.hx, w.hx, ta, tb) || !clipSlab(o.y, d.y, -w.hy, w.hy, ta, tb) ||
        !clipSlab(o.z, d.z, -w.amp, w.amp, ta, tb))
        return false;
    const double tol = 1e-12 * scale;
    auto f = [&](double tt, double& dfdt) {
        const V3 x = o + d * tt;
        double h, gx, gy;
        waveHeight(w, x.x, x.y, h, gx, gy);
        dfdt = d.z - gx * d.x - gy * d.y;
        return x.z - h;
    };
    double tc = ta, dfdt;
    double ft = f(tc, dfdt);
    if (tc <= tmin && std::fabs(ft) < tol) {
        tc = tmin + tol / std::max(L, 1e-300);
        ft = f(tc, dfdt);
    }
    const double minStep = 1e-7 * (tb - ta), nearSurface = 1e-6 * (w.amp + tol);
    auto accept = [&](double tr) {
        const V3 x = o + d * tr;
        if (!(tr > tmin && tr < tmax) || std::fabs(x.x) > w.hx || std::fabs(x.y) > w.hy) return false;
        t = tr;
        p = x;
        n = waveNormal(w, x.x, x.y);
        return true;
    };
    for (int i = 0; i < 4096 && tc < tb; ++i) {
        double step = std::max(std::fabs(ft) / L, minStep);
        if (std::fabs(ft) < nearSurface) {
            double tn = tc;
            for (int k = 0; k < 8; ++k) {
                const double fn = f(tn, dfdt);
                const V3 x = o + d * tn;
                const double scale = std::max({1e-12, w.amp, std::fabs(x.x), std::fabs(x.y), std::fabs(x.z)});
                if (std::fabs(fn) < 8 * std::numeric_limits<double>::epsilon() * scale || std::fabs(dfdt) < 1e-12) break;
                tn -= fn / dfdt;
            }
            double dd;
            if (tn >= tc - 1e3 * tol && std::fabs(f(tn, dd)) < 1e3 * tol && accept(tn)) return true;
            step = std::max(minStep, 2 * std::fabs(ft) / L);
        }
        const double tn = std::min(tc + step, tb);
        const double fn = f(tn, dfdt);
        if ((fn < 0) != (ft < 0)) {
            auto value = [&](double x) {
                double dx;
                return f(x, dx);
            };
            if (accept(bisect(value, tc, tn, ft))) return true;
        }
        tc = tn;
        ft = fn;
    }
    return false;
}

}  // namespace shapes
}  // namespace oracle
