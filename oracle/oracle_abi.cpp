// TEST INFRASTRUCTURE ONLY — the double-precision CPU oracle (see oracle_core.hpp).
// C entry points the Python tests load with ctypes: known-answer probes of the core formulas, and
// the triangle-mesh ray queries (closest hit, any hit, brute force, traversal statistics) that check
// the hip backend's batch queries. Threads split a batch into contiguous ranges; every ray's answer
// is independent of the split.
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle_bvh.hpp"

using namespace oracle;

namespace {

template <class F>
void forRanges(size_t n, int threads, F&& f) {
    if (threads <= 0) threads = int(std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
    threads = int(std::min<size_t>(size_t(threads), std::max<size_t>(1, n / 1024 + 1)));
    if (threads <= 1) {
        f(size_t(0), n);
        return;
    }
    std::vector<std::thread> pool;
    for (int k = 0; k < threads; ++k) {
        const size_t b = n * size_t(k) / size_t(threads), e = n * size_t(k + 1) / size_t(threads);
        pool.emplace_back([&f, b, e] { f(b, e); });
    }
    for (auto& t : pool) t.join();
}

// A query ray: 8 floats (origin, tmin, direction, tmax), as the hip backend's batch format.
struct QueryRay {
    V3 o, d;
    double tmin, tmax;
};
QueryRay loadRay(const float* r) {
    return {{r[0], r[1], r[2]}, {r[4], r[5], r[6]}, double(r[3]), double(r[7])};
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- known-answer probes

void oracle_pcg_u32(uint64_t seed, uint64_t stream, int n, uint32_t* out) {
    Pcg r(seed, stream);
    for (int i = 0; i < n; ++i) out[i] = r.next();
}
void oracle_pcg_uniform(uint64_t seed, uint64_t stream, int n, double* out) {
    Pcg r(seed, stream);
    for (int i = 0; i < n; ++i) out[i] = r.uniform();
}
uint64_t oracle_mix64(uint64_t v) { return mix64(v); }
uint64_t oracle_hash_pair(uint64_t a, uint64_t b) { return hashPair(a, b); }

double oracle_fresnel_dielectric(double cosI, double eta) { return fresnelDielectric(cosI, eta); }
double oracle_fresnel_conductor(double cosI, double re, double im) { return fresnelConductor(cosI, {re, im}); }
double oracle_sample_visible(double u) { return sampleVisible(u); }
double oracle_visible_pdf(double l) { return visiblePdf(l); }
void oracle_cie(double l, double* xyz) {
    xyz[0] = cieX(l);
    xyz[1] = cieY(l);
    xyz[2] = cieZ(l);
}
double oracle_cie_y_integral() { return cieYIntegral(); }
double oracle_planck(double l, double kelvin) { return planck(l, kelvin); }
double oracle_ciddor_air(double lnm) { return ciddorAir(sq(lnm * 1e-3)); }
double oracle_hg_phase(double c, double g) { return hgPhase(c, g); }
double oracle_sample_hg(const double* wo, double g, double u1, double u2, double* wi) {
    V3 w;
    const double p = sampleHg({wo[0], wo[1], wo[2]}, g, u1, u2, w);
    wi[0] = w.x;
    wi[1] = w.y;
    wi[2] = w.z;
    return p;
}
double oracle_ggx_d(double a, const double* m) { return ggxD(a, {m[0], m[1], m[2]}); }
double oracle_ggx_pdf(double a, const double* w, const double* m) { return ggxPdf(a, {w[0], w[1], w[2]}, {m[0], m[1], m[2]}); }
void oracle_ggx_sample_wm(double a, const double* w, double u1, double u2, double* m) {
    const V3 r = ggxSampleWm(a, {w[0], w[1], w[2]}, u1, u2);
    m[0] = r.x;
    m[1] = r.y;
    m[2] = r.z;
}
// Rough dielectric sample: returns 1 on success; out = (wi.xyz, weight, pdf, transmitted).
int oracle_sample_rough_dielectric(double a, const double* wo, double eta, double uc, double u1, double u2, int radiance,
                                   double* out) {
    Lobe s;
    const bool ok = sampleRoughDielectric(a, {wo[0], wo[1], wo[2]}, eta, uc, u1, u2, radiance != 0, s);
    out[0] = s.wi.x;
    out[1] = s.wi.y;
    out[2] = s.wi.z;
    out[3] = s.weight;
    out[4] = s.pdf;
    out[5] = s.transmitted ? 1 : 0;
    return ok ? 1 : 0;
}
double oracle_eval_rough_dielectric(double a, const double* wo, const double* wi, double eta, int radiance, double* pdf) {
    return evalRoughDielectric(a, {wo[0], wo[1], wo[2]}, {wi[0], wi[1], wi[2]}, eta, radiance != 0, *pdf);
}
int oracle_sample_rough_conductor(double a, const double* w, double re, double im, double tint, double u1, double u2,
                                  double* out) {
    Lobe s;
    const bool ok = sampleRoughConductor(a, {w[0], w[1], w[2]}, {re, im}, tint, u1, u2, s);
    out[0] = s.wi.x;
    out[1] = s.wi.y;
    out[2] = s.wi.z;
    out[3] = s.weight;
    out[4] = s.pdf;
    return ok ? 1 : 0;
}
double oracle_eval_rough_conductor(double a, const double* wo, const double* wi, double re, double im, double tint,
                                   double* pdf) {
    return evalRoughConductor(a, {wo[0], wo[1], wo[2]}, {wi[0], wi[1], wi[2]}, {re, im}, tint, *pdf);
}
int oracle_line_crosses_triangle(const double* o, const double* d, const double* v) {
    return lineCrossesTri({o[0], o[1], o[2]}, {d[0], d[1], d[2]}, {v[0], v[1], v[2]}, {v[3], v[4], v[5]},
                          {v[6], v[7], v[8]})
               ? 1
               : 0;
}

// ---------------------------------------------------------------- triangle-mesh queries

void* oracle_mesh_create(const double* pos, uint32_t nVerts, const uint32_t* tris, uint32_t nTris) {
    std::vector<V3> p(nVerts);
    for (uint32_t i = 0; i < nVerts; ++i) p[i] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    std::vector<std::array<uint32_t, 3>> t(nTris);
    for (uint32_t i = 0; i < nTris; ++i) t[i] = {tris[3 * i], tris[3 * i + 1], tris[3 * i + 2]};
    return new Mesh(std::move(p), std::move(t));
}
void oracle_mesh_destroy(void* m) { delete static_cast<Mesh*>(m); }
uint64_t oracle_mesh_node_count(const void* m) { return static_cast<const Mesh*>(m)->bvh().nodes().size(); }
// The BVH's primitive order (leaf slots → triangle), nTris entries.
void oracle_mesh_order(const void* m, uint32_t* out) {
    const auto& o = static_cast<const Mesh*>(m)->bvh().order();
    std::memcpy(out, o.data(), o.size() * sizeof(uint32_t));
}

// Closest hits for n rays; prim = 0xFFFFFFFF and t = +inf on a miss.
void oracle_mesh_closest(const void* mh, const float* rays, uint64_t n, uint32_t* prim, double* t, int threads) {
    const Mesh& m = *static_cast<const Mesh*>(mh);
    forRanges(size_t(n), threads, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
            const QueryRay r = loadRay(rays + 8 * i);
            double tt;
            uint32_t p;
            if (m.closest(r.o, r.d, r.tmin, r.tmax, tt, p)) {
                prim[i] = p;
                t[i] = tt;
            } else {
                prim[i] = 0xFFFFFFFFu;
                t[i] = kInf;
            }
        }
    });
}
void oracle_mesh_any(const void* mh, const float* rays, uint64_t n, uint8_t* hit, int threads) {
    const Mesh& m = *static_cast<const Mesh*>(mh);
    forRanges(size_t(n), threads, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
            const QueryRay r = loadRay(rays + 8 * i);
            hit[i] = m.any(r.o, r.d, r.tmin, r.tmax) ? 1 : 0;
        }
    });
}
void oracle_mesh_brute(const void* mh, const float* rays, uint64_t n, uint32_t* prim, double* t, int threads) {
    const Mesh& m = *static_cast<const Mesh*>(mh);
    forRanges(size_t(n), threads, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
            const QueryRay r = loadRay(rays + 8 * i);
            double tt;
            uint32_t p;
            if (m.bruteClosest(r.o, r.d, r.tmin, r.tmax, tt, p)) {
                prim[i] = p;
                t[i] = tt;
            } else {
                prim[i] = 0xFFFFFFFFu;
                t[i] = kInf;
            }
        }
    });
}
// Mean work of the reference's closest-hit traversal over the batch: out[0] = slab tests per ray,
// out[1] = triangle tests per ray, out[2] = leaves per ray (each leaf ≈ one node fetch), out[3] =
// inner nodes expanded per ray.
void oracle_mesh_traversal_stats(const void* mh, const float* rays, uint64_t n, double* out, int threads) {
    const Mesh& m = *static_cast<const Mesh*>(mh);
    std::vector<uint64_t> boxes(64, 0), tris(64, 0);
    int used = threads <= 0 ? int(std::min(16u, std::max(1u, std::thread::hardware_concurrency()))) : threads;
    used = std::min(used, 64);
    std::vector<std::pair<size_t, size_t>> ranges;
    for (int k = 0; k < used; ++k) ranges.push_back({size_t(n) * k / used, size_t(n) * (k + 1) / used});
    std::vector<std::thread> pool;
    for (int k = 0; k < used; ++k)
        pool.emplace_back([&, k] {
            for (size_t i = ranges[k].first; i < ranges[k].second; ++i) {
                const QueryRay r = loadRay(rays + 8 * i);
                double tt;
                uint32_t p;
                m.closest(r.o, r.d, r.tmin, r.tmax, tt, p, ~0u, &boxes[k], &tris[k]);
            }
        });
    for (auto& t : pool) t.join();
    uint64_t B = 0, T = 0;
    for (int k = 0; k < used; ++k) {
        B += boxes[k];
        T += tris[k];
    }
    out[0] = double(B) / double(n);
    out[1] = double(T) / double(n);
    // Every expanded inner node performs two child slab tests; the root adds one.
    out[3] = (double(B) - double(n)) / 2.0 / double(n);
    out[2] = out[1] / 4.0;
}

}  // extern "C"
