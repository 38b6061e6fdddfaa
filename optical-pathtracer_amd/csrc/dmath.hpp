// Float32 vector arithmetic and the watertight triangle crossing for the gfx950 kernels.
//
// The crossing is the Woop–Benthin–Wald (JCGT 2013) test the reference uses in both precisions
// (src/owe/core/shared/surfaces.hpp:278-290; float kernels: shaders/geometry.slang:424-431): in a
// frame sheared so the ray runs along +z, each edge's side is a 2-D cross product of that edge's two
// vertices. Two triangles sharing an edge then compute exactly opposite values for it, which is what
// makes the test watertight — provided no product is fused into an FMA with the subtraction, so the
// edge functions below are compiled with floating-point contraction off.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace owehip {

struct f3 {
    float x, y, z;
};
__host__ __device__ inline f3 mk3(float x, float y, float z) { return {x, y, z}; }
__host__ __device__ inline f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__host__ __device__ inline f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline f3 operator*(float s, f3 a) { return {a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__host__ __device__ inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline f3 cross3(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline float len3(f3 a) { return sqrtf(dot3(a, a)); }
__host__ __device__ inline f3 norm3(f3 a) { return a / len3(a); }
__host__ __device__ inline float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// Per-ray constants of the sheared crossing test: the dominant axis kz, the two others kx, ky in
// cyclic order, and the shears sx = d[kx]/d[kz], sy = d[ky]/d[kz].
struct Shear {
    int kx, ky, kz;
    float sx, sy;
};
__device__ inline Shear makeShear(f3 d) {
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    Shear s;
    s.kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    s.kx = s.kz == 2 ? 0 : s.kz + 1;
    s.ky = s.kx == 2 ? 0 : s.kx + 1;
    s.sx = comp(d, s.kx) / comp(d, s.kz);
    s.sy = comp(d, s.ky) / comp(d, s.kz);
    return s;
}

// Whether the line through o crosses triangle (v0, v1, v2), edges included (edge-on: no).
__device__ inline bool lineCrosses(const Shear& s, f3 o, f3 v0, f3 v1, f3 v2) {
#pragma clang fp contract(off)
    const f3 a = v0 - o, b = v1 - o, c = v2 - o;
    const float ax = comp(a, s.kx) - s.sx * comp(a, s.kz), ay = comp(a, s.ky) - s.sy * comp(a, s.kz);
    const float bx = comp(b, s.kx) - s.sx * comp(b, s.kz), by = comp(b, s.ky) - s.sy * comp(b, s.kz);
    const float cx = comp(c, s.kx) - s.sx * comp(c, s.kz), cy = comp(c, s.ky) - s.sy * comp(c, s.kz);
    const float u = cx * by - cy * bx, v = ax * cy - ay * cx, w = bx * ay - by * ax;
    if ((u < 0.0f || v < 0.0f || w < 0.0f) && (u > 0.0f || v > 0.0f || w > 0.0f)) return false;
    return u + v + w != 0.0f;
}

// The cross product of the triangle's two shorter edges (robust for slivers).
__device__ inline f3 triNormal(f3 v0, f3 v1, f3 v2) {
    const f3 a = v1 - v0, b = v2 - v0, c = v2 - v1;
    const float la = dot3(a, a), lb = dot3(b, b), lc = dot3(c, c);
    if (lc >= la && lc >= lb) return cross3(a, b);
    return la >= lb ? cross3(b, c) : cross3(a, c);
}

// Crossing distance along d from o to the triangle's plane (t in units of |d|), after the crossing
// test; the plane form keeps slivers' hit points on their surface.
__device__ inline bool hitTriangle(const Shear& s, f3 o, f3 d, f3 v0, f3 v1, f3 v2, float tmin, float tmax, float& t) {
    if (!lineCrosses(s, o, v0, v1, v2)) return false;
    const f3 np = triNormal(v0, v1, v2);
    t = -dot3(o - v0, np) / dot3(d, np);
    return t > tmin && t < tmax;
}

}  // namespace owehip
