// TEST INFRASTRUCTURE ONLY. A C-ABI harness over the reference engine's own sources, compiled by
// oracle/Makefile (target `ref`) from /root/reference/src into oracle/_ref/libowe_ref.so. It links the
// reference's bvh.cpp, mesh.cpp, geometry.cpp, spectrum.cpp and medium.cpp unchanged (they build with
// no external dependency); its entry points call reference functions directly so the tests can pin
// the repository's own restatement (oracle_*.cpp) against the real thing. World, Tracer and the
// scene loader need stb/cgltf (FetchContent, absent offline) and are not built.
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "owe/core/medium.hpp"
#include "owe/core/sampling.hpp"
#include "owe/core/shared/optics.hpp"
#include "owe/core/shared/spectra.hpp"
#include "owe/core/shared/surfaces.hpp"
#include "owe/core/shared/textures.hpp"
#include "owe/core/shared/warps.hpp"
#include "owe/core/spectrum.hpp"
#include "owe/core/wavelengths.hpp"
#include "owe/scene/geometry.hpp"

using namespace owe;

extern "C" {

void ref_pcg_u32(uint64_t seed, uint64_t stream, int n, uint32_t* out) {
    Rng r(seed, stream);
    for (int i = 0; i < n; ++i) out[i] = r.nextU32();
}
void ref_pcg_uniform(uint64_t seed, uint64_t stream, int n, double* out) {
    Rng r(seed, stream);
    for (int i = 0; i < n; ++i) out[i] = r.uniform();
}
uint64_t ref_mix64(uint64_t v) { return mixBits(v); }
uint64_t ref_hash_pair(uint64_t a, uint64_t b) { return hashCombine(a, b); }
double ref_fresnel_dielectric(double c, double eta) { return fresnelDielectric(c, eta); }
double ref_fresnel_conductor(double c, double re, double im) { return fresnelConductor(c, Vec2(re, im)); }
double ref_sample_visible(double u) { return sampleVisibleWavelength(u); }
double ref_visible_pdf(double l) { return visibleWavelengthPdf(l); }
void ref_cie(double l, double* xyz) {
    xyz[0] = cieX(l);
    xyz[1] = cieY(l);
    xyz[2] = cieZ(l);
}
double ref_cie_y_integral() { return cieYIntegral(); }
double ref_planck(double l, double T) { return planck(l, T); }
double ref_hg_phase(double c, double g) { return hgPhase(c, g); }
double ref_sample_hg(const double* wo, double g, double u1, double u2, double* wi) {
    Vec3 w;
    double p = sampleHg(Vec3(wo[0], wo[1], wo[2]), g, u1, u2, w);
    wi[0] = w.x;
    wi[1] = w.y;
    wi[2] = w.z;
    return p;
}
double ref_ggx_d(double a, const double* m) { return ggxD(a, Vec3(m[0], m[1], m[2])); }
double ref_ggx_pdf(double a, const double* w, const double* m) {
    return ggxPdf(a, Vec3(w[0], w[1], w[2]), Vec3(m[0], m[1], m[2]));
}
void ref_ggx_sample_wm(double a, const double* w, double u1, double u2, double* m) {
    Vec3 r = ggxSampleWm(a, Vec3(w[0], w[1], w[2]), u1, u2);
    m[0] = r.x;
    m[1] = r.y;
    m[2] = r.z;
}
int ref_sample_rough_dielectric(double a, const double* wo, double eta, double uc, double u1, double u2, int radiance,
                                double* out) {
    LobeSample s;
    bool ok = sampleRoughDielectric(a, Vec3(wo[0], wo[1], wo[2]), eta, uc, u1, u2, radiance != 0, s);
    out[0] = s.wi.x;
    out[1] = s.wi.y;
    out[2] = s.wi.z;
    out[3] = s.weight;
    out[4] = s.pdf;
    out[5] = s.transmitted ? 1 : 0;
    return ok ? 1 : 0;
}
double ref_eval_rough_dielectric(double a, const double* wo, const double* wi, double eta, int radiance, double* pdf) {
    return evalRoughDielectric(a, Vec3(wo[0], wo[1], wo[2]), Vec3(wi[0], wi[1], wi[2]), eta, radiance != 0, *pdf);
}
int ref_sample_rough_conductor(double a, const double* w, double re, double im, double tint, double u1, double u2,
                               double* out) {
    LobeSample s;
    bool ok = sampleRoughConductor(a, Vec3(w[0], w[1], w[2]), Vec2(re, im), tint, u1, u2, s);
    out[0] = s.wi.x;
    out[1] = s.wi.y;
    out[2] = s.wi.z;
    out[3] = s.weight;
    out[4] = s.pdf;
    return ok ? 1 : 0;
}
double ref_eval_rough_conductor(double a, const double* wo, const double* wi, double re, double im, double tint,
                                double* pdf) {
    return evalRoughConductor(a, Vec3(wo[0], wo[1], wo[2]), Vec3(wi[0], wi[1], wi[2]), Vec2(re, im), tint, *pdf);
}
int ref_line_crosses_triangle(const double* o, const double* d, const double* v) {
    return lineCrossesTriangle(Vec3(o[0], o[1], o[2]), Vec3(d[0], d[1], d[2]), Vec3(v[0], v[1], v[2]),
                               Vec3(v[3], v[4], v[5]), Vec3(v[6], v[7], v[8]))
               ? 1
               : 0;
}
double ref_value_noise3(const double* p) { return valueNoise3(Vec3(p[0], p[1], p[2])); }
// Catalog medium index n(λ) (absolute) by name; NaN for an unknown name.
double ref_catalog_index(const char* name, double lambdaNm) {
    Medium m;
    if (!catalogMedium(name, m)) return std::nan("");
    return m.n(lambdaNm);
}
double ref_sellmeier_index(const double* B, const double* C, int terms, int relAir, double lambdaNm) {
    IndexModel m = IndexModel::sellmeier(std::vector<double>(B, B + terms), std::vector<double>(C, C + terms));
    if (relAir) m = m.relativeToAir();
    return m.n(lambdaNm);
}
double ref_blackbody(double kelvin, double y, double lambdaNm) { return Spectrum::blackbody(kelvin, y).eval(lambdaNm); }
// Hero wavelengths for stratum u: out = 4 λ then 4 pdf.
void ref_wavelengths(double u, double* out) {
    Wavelengths w = Wavelengths::sample(u);
    for (int i = 0; i < NW; ++i) {
        out[i] = w.lambda[size_t(i)];
        out[NW + i] = w.pdf[size_t(i)];
    }
}

// ---------------------------------------------------------------- meshes (MeshShape + Bvh)

void* ref_mesh_create(const double* pos, uint32_t nVerts, const uint32_t* tris, uint32_t nTris) {
    std::vector<Vec3> p(nVerts);
    for (uint32_t i = 0; i < nVerts; ++i) p[i] = Vec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    std::vector<std::array<uint32_t, 3>> t(nTris);
    for (uint32_t i = 0; i < nTris; ++i) t[i] = {tris[3 * i], tris[3 * i + 1], tris[3 * i + 2]};
    return new MeshShape(std::move(p), std::move(t));
}
void ref_mesh_destroy(void* m) { delete static_cast<MeshShape*>(m); }
uint64_t ref_mesh_node_count(const void* m) { return static_cast<const MeshShape*>(m)->bvh().nodeCount(); }
void ref_mesh_order(const void* m, uint32_t* out) {
    const auto& o = static_cast<const MeshShape*>(m)->bvh().order();
    std::memcpy(out, o.data(), o.size() * sizeof(uint32_t));
}
void ref_mesh_closest(const void* mh, const float* rays, uint64_t n, uint32_t* prim, double* t, int threads) {
    const MeshShape& m = *static_cast<const MeshShape*>(mh);
    if (threads <= 0) threads = int(std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    for (int k = 0; k < threads; ++k)
        pool.emplace_back([&, k] {
            const uint64_t b = n * uint64_t(k) / uint64_t(threads), e = n * uint64_t(k + 1) / uint64_t(threads);
            for (uint64_t i = b; i < e; ++i) {
                const float* r = rays + 8 * i;
                Ray ray{Vec3(r[0], r[1], r[2]), Vec3(r[4], r[5], r[6])};
                LocalHit h;
                if (m.intersect(ray, double(r[3]), double(r[7]), h)) {
                    prim[i] = h.prim;
                    t[i] = h.t;
                } else {
                    prim[i] = 0xFFFFFFFFu;
                    t[i] = Inf;
                }
            }
        });
    for (auto& th : pool) th.join();
}

}  // extern "C"
