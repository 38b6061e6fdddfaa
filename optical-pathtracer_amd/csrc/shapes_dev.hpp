// Analytic shapes beyond planes and spheres, in float32 with float64 where precision needs it,
// each in its own frame — shared by the world walk (transport_dev.hpp) and the lens batch (lens.hip):
//
//   sag surface z = sag(r), conic + even asphere,   SagSurface::intersectQuadric/Numeric,
//   r in [rMin, rMax]                                src/owe/scene/geometry.cpp:69-166,
//                                                    src/owe/core/shared/surfaces.hpp:13-94, 140-171
//   open cylinder x² + y² = R², z0 ≤ z ≤ z1          CylinderShape, geometry.cpp:490-515;
//                                                    cylinderRoots, surfaces.hpp:118-137
//   liquid surface z = h(x, y): tapered plane-wave   WaveSurface, geometry.hpp:84-142,
//   sum or sampled bicubic Hermite nodes             geometry.cpp:188-344; waveTaper/hermiteBases,
//                                                    surfaces.hpp:178-200
//
// A quadric root is found in float and polished from the approximate hit (refineRoot): in double on
// steep surfaces (DSag::polish64; every surface of the world walk), in float on the lens batch's
// shallow ones;
// a numeric asphere root is bracketed over 96 samples, bisected in float and polished by Newton
// steps in double. The wave surface is marched in steps bounded by its Lipschitz constant along
// the ray (no crossing skipped), then refined by Newton's method. A ray leaving the shape it
// stands on takes a quadric's other root; for the numeric shapes, roots within the boundary's
// departure tolerance of its origin (along the surface normal) are its own point.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "dmath.hpp"
#include "scene_dev.hpp"

namespace owehip {

// ---------------------------------------------------------------- quadratics

__device__ __forceinline__ int solveQuadraticF(float a, float b, float c, float& t0, float& t1) {
    t0 = t1 = 0.0f;
    if (fabsf(a) < 1e-30f || fabsf(a) < 1e-14f * fabsf(b)) {
        if (b == 0.0f) return 0;
        t0 = t1 = -c / b;
        return 1;
    }
    const float disc = b * b - 4.0f * a * c;
    if (disc < 0.0f) return 0;
    const float q = -(b + copysignf(sqrtf(disc), b)) * 0.5f;
    t0 = q / a;
    t1 = q != 0.0f ? c / q : t0;
    if (t0 > t1) {
        const float s = t0;
        t0 = t1;
        t1 = s;
    }
    return 2;
}
__device__ __forceinline__ int solveQuadraticD(double a, double b, double c, double& t0, double& t1) {
    t0 = t1 = 0.0;
    if (fabs(a) < 1e-300 || fabs(a) < 1e-14 * fabs(b)) {
        if (b == 0.0) return 0;
        t0 = t1 = -c / b;
        return 1;
    }
    const double disc = b * b - 4.0 * a * c;
    if (disc < 0.0) return 0;
    const double q = -(b + copysign(sqrt(disc), b)) * 0.5;
    t0 = q / a;
    t1 = q != 0.0 ? c / q : t0;
    if (t0 > t1) {
        const double s = t0;
        t0 = t1;
        t1 = s;
    }
    return 2;
}

// ---------------------------------------------------------------- sag

// refineRoot in double: the sag quadric re-solved from the approximate hit point o + t d.
__device__ __forceinline__ double refineSagD(const DSag& s, f3 of, f3 df, double t) {
    const double ox = double(of.x) + double(df.x) * t, oy = double(of.y) + double(df.y) * t,
                 oz = double(of.z) + double(df.z) * t;
    const double dx = df.x, dy = df.y, dz = df.z;
    const double A = s.c * (dx * dx + dy * dy) + s.ck * dz * dz;
    const double B = 2.0 * (s.c * (ox * dx + oy * dy) + s.ck * oz * dz - dz);
    const double C = s.c * (ox * ox + oy * oy) + s.ck * oz * oz - 2.0 * oz;
    double u0, u1;
    if (solveQuadraticD(A, B, C, u0, u1) == 0) return t;
    const double du = fabs(u0) < fabs(u1) ? u0 : u1;
    return fabs(du) < 1e-3 * (1.0 + fabs(t)) ? t + du : t;
}

// refineRoot in float, for surfaces flat enough that the re-solved quadric from a point near the
// surface loses no digits that matter (DSag::polish64 clear).
__device__ __forceinline__ float refineSagF(const DSag& s, f3 o, f3 d, float t) {
    const f3 q = o + d * t;
    const float A = s.cf * (d.x * d.x + d.y * d.y) + s.ckf * d.z * d.z;
    const float B = 2.0f * (s.cf * (q.x * d.x + q.y * d.y) + s.ckf * q.z * d.z - d.z);
    const float C = s.cf * (q.x * q.x + q.y * q.y) + s.ckf * q.z * q.z - 2.0f * q.z;
    float u0, u1;
    if (solveQuadraticF(A, B, C, u0, u1) == 0) return t;
    const float du = fabsf(u0) < fabsf(u1) ? u0 : u1;
    return fabsf(du) < 1e-3f * (1.0f + fabsf(t)) ? t + du : t;
}

__device__ __forceinline__ float sagF(const DSag& s, float r) {
    const float r2 = r * r;
    float z = s.cf * r2 / (1.0f + sqrtf(fmaxf(0.0f, 1.0f - (1.0f + s.kf) * s.cf * s.cf * r2)));
    float rp = r2 * r2;
    for (int i = 0; i < s.nA; ++i) {
        z += s.A[i] * rp;
        rp *= r2;
    }
    return z;
}
__device__ __forceinline__ double sagD(const DSag& s, double r) {
    const double r2 = r * r, k = double(s.kf);
    double z = s.c * r2 / (1.0 + sqrt(fmax(0.0, 1.0 - (1.0 + k) * s.c * s.c * r2)));
    double rp = r2 * r2;
    for (int i = 0; i < s.nA; ++i) {
        z += double(s.A[i]) * rp;
        rp *= r2;
    }
    return z;
}
__device__ __forceinline__ double dsagD(const DSag& s, double r) {
    const double k = double(s.kf), root = sqrt(fmax(0.0, 1.0 - (1.0 + k) * s.c * s.c * r * r));
    double d = root > 0.0 ? s.c * r / root : 0.0;
    double m = 4.0, rp = r * r * r;
    for (int i = 0; i < s.nA; ++i) {
        d += double(s.A[i]) * m * rp;
        rp *= r * r;
        m += 2.0;
    }
    return d;
}
__device__ __forceinline__ float dsagF(const DSag& s, float r) {
    const float root = sqrtf(fmaxf(0.0f, 1.0f - (1.0f + s.kf) * s.cf * s.cf * r * r));
    float d = root > 0.0f ? s.cf * r / root : 0.0f;
    float m = 4.0f, rp = r * r * r;
    for (int i = 0; i < s.nA; ++i) {
        d += s.A[i] * m * rp;
        rp *= r * r;
        m += 2.0f;
    }
    return d;
}
__device__ __forceinline__ f3 sagNormalF(const DSag& s, f3 p) {
    const float r = sqrtf(p.x * p.x + p.y * p.y);
    if (r < 1e-30f) return mk3(0.0f, 0.0f, 1.0f);
    const float sl = s.polish64 ? float(dsagD(s, double(r))) : dsagF(s, r);
    return norm3(mk3(-sl * p.x / r, -sl * p.y / r, 1.0f));
}
__device__ __forceinline__ bool sagAccept(const DSag& s, f3 q) {
    const float r2 = q.x * q.x + q.y * q.y;
    return r2 <= s.rMax2 && r2 >= s.rMin2;
}

// SagSurface::intersect for t in (tmin, tmax): t and the point (from the double-precision root).
// `departed`: the ray starts on this surface; the quadric's root at its origin is skipped by taking
// the other one (−B/A, polished), a numeric sag's by the caller's tmin.
__device__ __forceinline__ bool sagHit(const DSag& s, f3 o, f3 d, float tmin, float tmax, bool departed, float& tOut,
                                       f3& p) {
    if (s.nA == 0) {
        if (s.cf == 0.0f) {
            if (departed || d.z == 0.0f) return false;
            const float t = -o.z / d.z;
            if (!(t > tmin && t < tmax)) return false;
            p = o + d * t;
            p.z = 0.0f;
            if (!sagAccept(s, p)) return false;
            tOut = t;
            return true;
        }
        const float A = s.cf * (d.x * d.x + d.y * d.y) + s.ckf * d.z * d.z;
        const float B = 2.0f * (s.cf * (o.x * d.x + o.y * d.y) + s.ckf * o.z * d.z - d.z);
        const float C = s.cf * (o.x * o.x + o.y * o.y) + s.ckf * o.z * o.z - 2.0f * o.z;
        float t0, t1;
        if (departed) {
            if (fabsf(A) < 1e-30f) return false;
            t0 = t1 = -B / A;
        } else if (solveQuadraticF(A, B, C, t0, t1) == 0) {
            return false;
        }
        for (int k = 0; k < 2; ++k) {
            const float tc = k == 0 ? t0 : t1;
            const float m = 1e-3f * (1.0f + fabsf(tc));
            if (tc + m <= tmin || tc - m >= tmax) continue;  // cannot be refined into (tmin, tmax)
            f3 q;
            double tr;
            if (s.polish64) {
                tr = refineSagD(s, o, d, double(tc));
                if (!(tr > double(tmin) && tr < double(tmax))) continue;
                q = mk3(float(double(o.x) + double(d.x) * tr), float(double(o.y) + double(d.y) * tr),
                        float(double(o.z) + double(d.z) * tr));
            } else {
                const float tf = refineSagF(s, o, d, tc);
                if (!(tf > tmin && tf < tmax)) continue;
                tr = double(tf);
                q = o + d * tf;
            }
            if (!sagAccept(s, q)) continue;
            if (!((1.0f + s.kf) * s.cf * q.z <= 1.0f + 1e-6f)) continue;  // the conic's other sheet
            p = q;
            tOut = float(tr);
            return true;
        }
        return false;
    }
    // Numeric asphere: clip to the aperture cylinder and the sag's slab, 96 samples, bisection in
    // float, then Newton steps on f(t) = sag(r) − z in double.
    float ta = tmin, tb = tmax;
    const float a = d.x * d.x + d.y * d.y;
    if (a < 1e-30f) {
        if (o.x * o.x + o.y * o.y > s.rMax2) return false;
    } else {
        float q0, q1;
        if (solveQuadraticF(a, 2.0f * (o.x * d.x + o.y * d.y), o.x * o.x + o.y * o.y - s.rMax2, q0, q1) < 2) return false;
        ta = fmaxf(ta, q0);
        tb = fminf(tb, q1);
    }
    if (fabsf(d.z) < 1e-30f) {
        if (o.z < s.zMin || o.z > s.zMax) return false;
    } else {
        const float s0 = (s.zMin - o.z) / d.z, s1 = (s.zMax - o.z) / d.z;
        ta = fmaxf(ta, fminf(s0, s1));
        tb = fminf(tb, fmaxf(s0, s1));
    }
    if (!(ta < tb)) return false;
    auto f = [&](float tt) {
        const f3 q = o + d * tt;
        return sagF(s, sqrtf(q.x * q.x + q.y * q.y)) - q.z;
    };
    float tPrev = ta, fPrev = f(ta);
    for (int i = 1; i <= 96; ++i) {
        const float tCur = ta + (tb - ta) * float(i) / 96.0f, fCur = f(tCur);
        if ((fPrev <= 0.0f && fCur >= 0.0f) || (fPrev >= 0.0f && fCur <= 0.0f)) {
            if (!(fPrev == 0.0f && tPrev <= tmin)) {
                float lo = tPrev, hi = tCur, flo = fPrev;
                for (int it = 0; it < 40 && hi - lo > 2e-7f * (1.0f + fabsf(hi)); ++it) {
                    const float mid = 0.5f * (lo + hi), fm = f(mid);
                    if ((fm <= 0.0f) == (flo <= 0.0f)) {
                        lo = mid;
                        flo = fm;
                    } else {
                        hi = mid;
                    }
                }
                double t = 0.5 * (double(lo) + double(hi));
                for (int it = 0; it < 2; ++it) {
                    const double x = double(o.x) + double(d.x) * t, y = double(o.y) + double(d.y) * t,
                                 z = double(o.z) + double(d.z) * t, r = sqrt(x * x + y * y);
                    const double fv = sagD(s, r) - z;
                    const double fd = (r > 0.0 ? dsagD(s, r) * (x * d.x + y * d.y) / r : 0.0) - double(d.z);
                    if (fd == 0.0) break;
                    const double step = fv / fd;
                    if (!(fabs(step) <= double(hi - lo) + 1e-12)) break;
                    t -= step;
                }
                const f3 q = mk3(float(double(o.x) + double(d.x) * t), float(double(o.y) + double(d.y) * t),
                                 float(double(o.z) + double(d.z) * t));
                if (t > double(tmin) && t < double(tmax) && sagAccept(s, q)) {
                    p = q;
                    tOut = float(t);
                    return true;
                }
            }
        }
        tPrev = tCur;
        fPrev = fCur;
    }
    return false;
}

// ---------------------------------------------------------------- cylinder

// CylinderShape::intersect in float: geometric-form roots, the z trim; a ray leaving the wall
// takes the other root −2 (o·d)_xy / |d_xy|².
__device__ __forceinline__ bool cylinderHit(float R, float z0, float z1, f3 o, f3 d, float tmin, float tmax,
                                            bool departed, float& tOut) {
    const float A = d.x * d.x + d.y * d.y;
    if (A < 1e-30f) return false;
    const float b = o.x * d.x + o.y * d.y;
    float t0, t1;
    if (departed) {
        t0 = t1 = -2.0f * b / A;
    } else {
        const float px = o.x - d.x * (b / A), py = o.y - d.y * (b / A);
        const float disc = A * (R * R - (px * px + py * py));
        if (disc < 0.0f) return false;
        const float q = -(b + copysignf(sqrtf(disc), b));
        const float c = o.x * o.x + o.y * o.y - R * R;
        t0 = q / A;
        t1 = q != 0.0f ? c / q : t0;
        if (t0 > t1) {
            const float s = t0;
            t0 = t1;
            t1 = s;
        }
    }
    for (int k = 0; k < 2; ++k) {
        const float t = k == 0 ? t0 : t1;
        if (!(t > tmin && t < tmax)) continue;
        const float z = o.z + d.z * t;
        if (z < z0 || z > z1) continue;
        tOut = t;
        return true;
    }
    return false;
}

// ---------------------------------------------------------------- wave surface

__device__ __forceinline__ void waveTaperF(float u, float half, float margin, float& w, float& dw) {
    w = 1.0f;
    dw = 0.0f;
    if (!(margin > 0.0f)) return;
    const float e = half - fabsf(u);
    if (e >= margin) return;
    if (e <= 0.0f) {
        w = 0.0f;
        return;
    }
    const float t = e / margin;
    w = t * t * (3.0f - 2.0f * t);
    dw = -(u < 0.0f ? -1.0f : 1.0f) * 6.0f * t * (1.0f - t) / margin;
}

// WaveSurface::height: the tapered plane-wave sum (records (amplitude, kx, ky, phase)) or the
// bicubic Hermite patch of the cell (records (h, ∂x·dx, ∂y·dy, ∂xy·dx·dy) per node); zero outside.
__device__ __forceinline__ void waveHeightF(const DWave& w, const float4* data, float x, float y, float& h, float& gx,
                                            float& gy) {
    h = gx = gy = 0.0f;
    if (fabsf(x) > w.hx || fabsf(y) > w.hy) return;
    if (w.nx == 0) {
        float g = 0.0f, sx = 0.0f, sy = 0.0f;
        for (uint32_t i = 0; i < w.ny; ++i) {
            const float4 r = data[w.data + i];
            float sn, cs;
            sincosf(r.y * x + r.z * y + r.w, &sn, &cs);
            g += r.x * sn;
            sx += r.x * cs * r.y;
            sy += r.x * cs * r.z;
        }
        float wx, dwx, wy, dwy;
        waveTaperF(x, w.hx, w.margin, wx, dwx);
        waveTaperF(y, w.hy, w.margin, wy, dwy);
        h = wx * wy * g;
        gx = dwx * wy * g + wx * wy * sx;
        gy = wx * dwy * g + wx * wy * sy;
        return;
    }
    const float fx = (x + w.hx) / w.dx, fy = (y + w.hy) / w.dy;
    const int i = min(max(int(fx), 0), int(w.nx) - 2), j = min(max(int(fy), 0), int(w.ny) - 2);
    const float u = fx - float(i), v = fy - float(j);
    const float P[2] = {(2.0f * u - 3.0f) * u * u + 1.0f, (3.0f - 2.0f * u) * u * u};
    const float Q[2] = {((u - 2.0f) * u + 1.0f) * u, (u - 1.0f) * u * u};
    const float dP[2] = {6.0f * u * (u - 1.0f), 6.0f * u * (1.0f - u)};
    const float dQ[2] = {(3.0f * u - 4.0f) * u + 1.0f, (3.0f * u - 2.0f) * u};
    const float Pv[2] = {(2.0f * v - 3.0f) * v * v + 1.0f, (3.0f - 2.0f * v) * v * v};
    const float Qv[2] = {((v - 2.0f) * v + 1.0f) * v, (v - 1.0f) * v * v};
    const float dPv[2] = {6.0f * v * (v - 1.0f), 6.0f * v * (1.0f - v)};
    const float dQv[2] = {(3.0f * v - 4.0f) * v + 1.0f, (3.0f * v - 2.0f) * v};
    float f = 0.0f, fu = 0.0f, fv = 0.0f;
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int ai = 0; ai < 2; ++ai) {
            const float4 n = data[w.data + uint32_t(j + bj) * w.nx + uint32_t(i + ai)];
            f += n.x * P[ai] * Pv[bj] + n.y * Q[ai] * Pv[bj] + n.z * P[ai] * Qv[bj] + n.w * Q[ai] * Qv[bj];
            fu += n.x * dP[ai] * Pv[bj] + n.y * dQ[ai] * Pv[bj] + n.z * dP[ai] * Qv[bj] + n.w * dQ[ai] * Qv[bj];
            fv += n.x * P[ai] * dPv[bj] + n.y * Q[ai] * dPv[bj] + n.z * P[ai] * dQv[bj] + n.w * Q[ai] * dQv[bj];
        }
    h = f;
    gx = fu / w.dx;
    gy = fv / w.dy;
}

__device__ __forceinline__ f3 waveNormalF(const DWave& w, const float4* data, float x, float y) {
    float h, gx, gy;
    waveHeightF(w, data, x, y, h, gx, gy);
    return norm3(mk3(-gx, -gy, 1.0f));
}

// WaveSurface::intersect in float for t in (tlo, tmax): march f(t) = z − h in steps |f| / L (L the
// Lipschitz bound of f along the ray), Newton near the surface, bisection of a sign change inside a
// step. Every loop is bounded. A departing ray skips roots within `depTol` of its origin along the
// surface normal.
__device__ __forceinline__ bool waveHit(const DWave& w, const float4* data, f3 o, f3 d, float tlo, float tmax,
                                        bool departed, float depTol, float& tOut) {
    const float L = fabsf(d.z) + w.slope * sqrtf(d.x * d.x + d.y * d.y);
    if (!(L > 0.0f)) return false;
    const float tmin = departed ? fmaxf(depTol / L, tlo) : tlo;
    float ta = tmin, tb = tmax;
    auto clip = [&](float oc, float dc, float lo, float hi) {
        if (fabsf(dc) < 1e-30f) return oc >= lo && oc <= hi;
        const float s0 = (lo - oc) / dc, s1 = (hi - oc) / dc;
        ta = fmaxf(ta, fminf(s0, s1));
        tb = fminf(tb, fmaxf(s0, s1));
        return ta < tb;
    };
    if (!clip(o.x, d.x, -w.hx, w.hx) || !clip(o.y, d.y, -w.hy, w.hy) || !clip(o.z, d.z, -w.amp, w.amp)) return false;
    const float scale = fmaxf(fmaxf(w.hx, w.hy), w.amp);
    const float tol = 1e-6f * scale;
    auto f = [&](float tt, float& dfdt) {
        const f3 q = o + d * tt;
        float h, gx, gy;
        waveHeightF(w, data, q.x, q.y, h, gx, gy);
        dfdt = d.z - gx * d.x - gy * d.y;
        return q.z - h;
    };
    auto accept = [&](float tr) {
        if (!(tr > tmin && tr < tmax)) return false;
        const f3 q = o + d * tr;
        if (fabsf(q.x) > w.hx || fabsf(q.y) > w.hy) return false;
        if (departed) {
            float h, gx, gy;
            waveHeightF(w, data, q.x, q.y, h, gx, gy);
            const float dn = fabsf(d.z - gx * d.x - gy * d.y) * rsqrtf(1.0f + gx * gx + gy * gy);
            if (tr * dn <= depTol) return false;
        }
        tOut = tr;
        return true;
    };
    float t = ta, dfdt;
    float ft = f(t, dfdt);
    if (fabsf(ft) < tol) {
        t += tol / L;
        ft = f(t, dfdt);
    }
    const float minStep = 1e-6f * (tb - ta), nearSurface = 1e-4f * (w.amp + tol);
    for (int i = 0; i < 1024 && t < tb; ++i) {
        float step = fmaxf(fabsf(ft) / L, minStep);
        if (fabsf(ft) < nearSurface) {
            float tn = t, fn = ft, dn = dfdt;
            for (int k = 0; k < 8; ++k) {
                if (fabsf(fn) < 4e-7f * scale || fabsf(dn) < 1e-12f) break;
                tn -= fn / dn;
                fn = f(tn, dn);
            }
            if (tn >= t - 1e3f * tol && fabsf(fn) < 1e2f * tol && accept(tn)) return true;
            step = fmaxf(minStep, 2.0f * fabsf(ft) / L);
        }
        const float tn = fminf(t + step, tb);
        float dn;
        const float fn = f(tn, dn);
        if ((fn < 0.0f) != (ft < 0.0f)) {
            float lo = t, hi = tn, flo = ft, dd;
            for (int it = 0; it < 40 && hi - lo > 2e-7f * (1.0f + fabsf(hi)); ++it) {
                const float mid = 0.5f * (lo + hi), fm = f(mid, dd);
                if ((fm <= 0.0f) == (flo <= 0.0f)) {
                    lo = mid;
                    flo = fm;
                } else {
                    hi = mid;
                }
            }
            if (accept(0.5f * (lo + hi))) return true;
        }
        t = tn;
        ft = fn;
        dfdt = dn;
    }
    return false;
}

}  // namespace owehip
