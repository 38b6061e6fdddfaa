// Batched lens analysis: owe_hip_lens_batch_* of include/owe_hip_lens.h.
//
// Work is a flat list of items — (lens, wavelength, block of `raysPerItem` rays) — drained by a
// persistent grid: each wave takes the next item with one atomic, its lanes trace the block's rays
// through the lens surface after surface, reduce the spot's sums across the wave in double and add
// them to the (lens, wavelength) totals with one atomic per sum; arrived rays also land on the lens's
// detector grid as 64-bit fixed-point weights, so the irradiance is independent of order. Lenses
// of very different ray counts (4k to 1M) share the launch without tails of idle blocks.
//
// Ray–surface roots follow SagSurface (src/owe/scene/geometry.cpp:115-166,
// src/owe/core/shared/surfaces.hpp:13-94) in float32 with the reference's float tolerances, and
// each root is then polished in float64 — a quadric root re-solved from the approximate hit point
// (refineRoot), a numeric asphere root by Newton steps — the only fp64 on the GPU path, where
// strongly curved surfaces would otherwise lose digits. Transmission is the walk's deterministic
// refraction with weight T = 1 − R (transport.cpp:531-595); TIR or a missed aperture ends a ray.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "owe_hip_lens.h"
#include "shapes_dev.hpp"
#include "transport_dev.hpp"

using namespace owehip;

namespace owehip {
namespace {

constexpr int kLensBlock = 256;
constexpr double kHistScale = 4294967296.0;  // 2^32 units of weight per unit

// A lens row: its sag surface (shapes_dev.hpp) and the vertex position along the axis.
struct DSurf : DSag {
    float z;
};

struct DLensInfo {
    uint32_t first, count;    // surfaces
    uint64_t rays;            // per wavelength
    uint64_t itemBase;        // first work item
    uint32_t itemsPerLambda;
    uint32_t rings;           // 0: Fibonacci lattice; r: hexapolar rings
    double rEP, zEP, zStart;  // entrance pupil (sampling radius 0.999 rEP) and launch plane
    float detX, detY, detZ, detHalf;  // detector grid centre, plane, half size
};

struct LensArgs {
    const DSurf* surf;
    const DLensInfo* lens;
    const float* nAfter;      // [surface × nλ]: index after each surface
    const float* nAmbient;    // [lens × nλ]
    uint32_t nLenses, nLambda, raysPerItem, pixels;
    uint64_t totalItems;
    double dirY, dirZ;        // collimated field direction (0, sin θ, cos θ)
    unsigned long long* next;
    double* moments;          // [lens × nλ × OWE_LENS_MOMENTS]
    unsigned long long* hist; // [lens × pixels²]
};

// Pupil point i of n on the entrance pupil of radius rEP, in double: a Fibonacci lattice over the
// disc of radius 0.999 rEP, or with rings > 0 spotAt's hexapolar pattern (analysis.cpp:124-131) —
// point 0 the centre, ring k holding the 6k points after the 1 + 3k(k − 1) before it.
__device__ __forceinline__ void pupilPoint(uint64_t i, uint64_t n, uint32_t rings, double rEP, double& x, double& y) {
    if (rings > 0) {
        if (i == 0) {
            x = y = 0.0;
            return;
        }
        uint64_t k = uint64_t((3.0 + sqrt(12.0 * double(i) - 3.0)) / 6.0);
        while (k > 1 && 1 + 3 * k * (k - 1) > i) --k;
        while (1 + 3 * k * (k + 1) <= i) ++k;
        const uint64_t j = i - (1 + 3 * k * (k - 1)), m = 6 * k;
        const double rr = rEP * double(k) / double(rings) * 0.999;
        const double a = 2.0 * 3.14159265358979323846 * double(j) / double(m);
        x = rr * cos(a);
        y = rr * sin(a);
        return;
    }
    const double rr = 0.999 * rEP * sqrt((double(i) + 0.5) / double(n));
    double f = double(i) * 0.6180339887498949;
    f -= floor(f);
    double sa, ca;
    sincospi(2.0 * f, &sa, &ca);
    x = rr * ca;
    y = rr * sa;
}

// One ray i of lens L at wavelength w: true when it crosses every surface; p, d after the last one.
__device__ __forceinline__ bool traceLensRay(const LensArgs& a, const DLensInfo& L, uint32_t li, uint32_t w, uint64_t i,
                                             f3& p, f3& d, float& T) {
    double x, y;
    pupilPoint(i, L.rays, L.rings, L.rEP, x, y);
    const double s = (L.zStart - L.zEP) / a.dirZ;
    f3 o = mk3(float(x), float(y + a.dirY * s), float(L.zEP + a.dirZ * s));
    d = mk3(0.0f, float(a.dirY), float(a.dirZ));
    T = 1.0f;
    float nPrev = a.nAmbient[li * a.nLambda + w];
    for (uint32_t k = 0; k < L.count; ++k) {
        const DSurf& sf = a.surf[L.first + k];
        f3 q, n;
        float th;
        if (!sagHit(sf, mk3(o.x, o.y, o.z - sf.z), d, 0.0f, __builtin_inff(), false, th, q)) return false;
        n = sagNormalF(sf, q);
        const float nNext = a.nAfter[(L.first + k) * a.nLambda + w];
        const float R = fresnelDielectric(fabsf(dot3(d, n)), nNext / nPrev);
        f3 wv;
        if (!refractDir(d, n, nPrev, nNext, wv)) return false;
        T *= 1.0f - R;
        o = mk3(q.x, q.y, q.z + sf.z);
        d = wv;
        nPrev = nNext;
    }
    p = o;
    return d.z > 0.0f;
}

__device__ __forceinline__ double waveSumD(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__global__ __launch_bounds__(kLensBlock) void lensKernel(LensArgs a) {
    const uint32_t lane = __lane_id();
    for (;;) {
        unsigned long long item = 0;
        if (lane == 0) item = atomicAdd(a.next, 1ull);
        item = __shfl(item, 0);
        if (item >= a.totalItems) break;
        // The lens of this item: the last whose first item is ≤ item.
        uint32_t lo = 0, len = a.nLenses;
        while (len > 0) {
            const uint32_t half = len / 2;
            if (a.lens[lo + half].itemBase <= item) {
                lo += half + 1;
                len -= half + 1;
            } else {
                len = half;
            }
        }
        const uint32_t li = lo - 1;
        const DLensInfo& L = a.lens[li];
        const uint64_t local = item - L.itemBase;
        const uint32_t w = uint32_t(local / L.itemsPerLambda);
        const uint64_t first = (local % L.itemsPerLambda) * a.raysPerItem;
        double m[OWE_LENS_MOMENTS] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (uint32_t k = lane; k < a.raysPerItem; k += 64) {
            const uint64_t i = first + k;
            if (i >= L.rays) break;
            f3 p, d;
            float T;
            if (!traceLensRay(a, L, li, w, i, p, d, T)) continue;
            const double bx = double(d.x) / double(d.z), by = double(d.y) / double(d.z);
            const double ax = double(p.x) - bx * double(p.z), ay = double(p.y) - by * double(p.z);
            m[0] += 1.0;
            m[1] += ax;
            m[2] += ay;
            m[3] += bx;
            m[4] += by;
            m[5] += ax * ax + ay * ay;
            m[6] += bx * bx + by * by;
            m[7] += ax * bx + ay * by;
            m[8] += double(T);
            if (a.pixels > 0) {
                const float s = (L.detZ - p.z) / d.z;
                const float gx = (p.x + d.x * s - L.detX) / L.detHalf, gy = (p.y + d.y * s - L.detY) / L.detHalf;
                const float fx = (gx + 1.0f) * 0.5f * float(a.pixels), fy = (1.0f - gy) * 0.5f * float(a.pixels);
                if (fx >= 0.0f && fy >= 0.0f && fx < float(a.pixels) && fy < float(a.pixels)) {
                    const uint32_t px = min(uint32_t(fx), a.pixels - 1), py = min(uint32_t(fy), a.pixels - 1);
                    atomicAdd(&a.hist[(size_t(li) * a.pixels + py) * a.pixels + px],
                              (unsigned long long)llrint(double(T) * kHistScale));
                }
            }
        }
        double* dst = a.moments + (size_t(li) * a.nLambda + w) * OWE_LENS_MOMENTS;
#pragma unroll
        for (int q = 0; q < OWE_LENS_MOMENTS; ++q) {
            const double v = waveSumD(m[q]);
            if (lane == 0 && v != 0.0) atomicAdd(dst + q, v);
        }
    }
}

__global__ void lensBegin(LensArgs a, size_t nMoments, size_t nHist) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x, stride = size_t(gridDim.x) * blockDim.x;
    if (i == 0) *a.next = 0;
    for (size_t k = i; k < nMoments; k += stride) a.moments[k] = 0.0;
    for (size_t k = i; k < nHist; k += stride) a.hist[k] = 0;
}

__global__ void lensProbe(LensArgs a, uint32_t li, uint32_t w, const uint64_t* index, uint64_t n, double* out) {
    const uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    f3 p, d;
    float T;
    double* r = out + 7 * k;
    if (traceLensRay(a, a.lens[li], li, w, index[k], p, d, T)) {
        r[0] = p.x, r[1] = p.y, r[2] = p.z, r[3] = d.x, r[4] = d.y, r[5] = d.z, r[6] = T;
    } else {
        for (int q = 0; q < 6; ++q) r[q] = __builtin_nan("");
        r[6] = 0.0;
    }
}

// ---------------------------------------------------------------- host: media and paraxial optics

double ciddor(double lnm) {
    const double s2 = (1e3 / lnm) * (1e3 / lnm);
    return 1 + 0.05792105 / (238.0185 - s2) + 0.00167917 / (57.362 - s2);
}
// IndexModel::n (src/owe/core/medium.cpp:46-66): absolute index, ×n_air when relative to air.
double mediumIndexHost(const owe_medium& m, double lnm) {
    const double l = lnm * 1e-3, l2 = l * l;
    double n = 1;
    switch (m.index_kind) {
        case OWE_INDEX_CONSTANT: n = m.coeff_a[0]; break;
        case OWE_INDEX_CAUCHY: n = m.coeff_a[0] + m.coeff_a[1] / l2 + m.coeff_a[2] / (l2 * l2); break;
        case OWE_INDEX_SELLMEIER: {
            double s = 1;
            for (int k = 0; k < m.terms; ++k) s += m.coeff_a[k] * l2 / (l2 - m.coeff_b[k]);
            n = std::sqrt(s);
            break;
        }
        case OWE_INDEX_CIDDOR: n = ciddor(lnm); break;
        default: fail(OWE_HIP_ERR_INVALID, "unknown index model");
    }
    return m.rel_air ? n * ciddor(lnm) : n;
}

struct ParaxialOut {
    double efl, bfl, epz, epr, length;
};

// paraxialAnalysis for refracting rows, object at infinity (src/owe/scene/prescription.cpp:193-321):
// the (y, n·u) system matrix, EFL −n'/C, BFL −n'A/C, the stop (marked, else the surface that limits
// a marginal ray) and its image through the rows before it, the entrance pupil.
ParaxialOut paraxialHost(const owe_lens_surface* s, size_t N, const std::vector<double>& nAfter, double n0) {
    std::vector<double> nBefore(N), phi(N), t(N);
    double nPrev = n0;
    for (size_t i = 0; i < N; ++i) {
        nBefore[i] = nPrev;
        phi[i] = (nAfter[i] - nPrev) * s[i].curvature;
        t[i] = std::fabs(s[i].thickness);
        nPrev = nAfter[i];
    }
    double a = 1, b = 0, c = 0, d = 1;
    for (size_t i = 0; i < N; ++i) {
        c -= phi[i] * a;
        d -= phi[i] * b;
        if (i + 1 < N) {
            const double k = t[i] / nAfter[i];
            a += k * c;
            b += k * d;
        }
    }
    ParaxialOut r{};
    for (size_t i = 0; i + 1 < N; ++i) r.length += t[i];
    require(std::fabs(c) > 1e-9, "lens is afocal (no finite focus for the spot analysis)");
    const double nk = nAfter[N - 1];
    r.efl = -nk / c;
    r.bfl = -nk * a / c;
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
        require(u != 0, "entrance pupil at infinity");
        const double dz = -y / u;
        r.epz = dz;
        r.epr = std::fabs(yM + dz * nuM / n0);
    }
    return r;
}

}  // namespace
}  // namespace owehip

struct owe_hip_lens_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocations;
    std::vector<owe_lens> lenses;
    std::vector<owe_lens_result> paraxial;  // launched/efl/bfl/pupil/last vertex per lens
    std::vector<DLensInfo> info;
    owe_lens_batch_settings settings{};
    LensArgs args{};
    size_t nMoments = 0, nHist = 0;
    int computeUnits = 256;
    int blocksPerCU = 4;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double lastMs = 0;
    bool ran = false;
};

namespace owehip {
namespace {

template <class T>
T* lensAlloc(owe_hip_lens_batch* b, size_t n) {
    void* p = nullptr;
    OWE_HIP_CHECK(hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T)));
    b->allocations.push_back(p);
    return static_cast<T*>(p);
}

// The work items of each lens for the current raysPerItem.
void layoutItems(owe_hip_lens_batch* b, uint32_t raysPerItem) {
    uint64_t base = 0;
    for (DLensInfo& L : b->info) {
        L.itemBase = base;
        L.itemsPerLambda = uint32_t((L.rays + raysPerItem - 1) / raysPerItem);
        base += uint64_t(L.itemsPerLambda) * b->args.nLambda;
    }
    OWE_HIP_CHECK(hipMemcpy(const_cast<DLensInfo*>(b->args.lens), b->info.data(), b->info.size() * sizeof(DLensInfo),
                            hipMemcpyHostToDevice));
    b->args.raysPerItem = raysPerItem;
    b->args.totalItems = base;
}

double runOnce(owe_hip_lens_batch* b) {
    hipLaunchKernelGGL(lensBegin, dim3{1024}, dim3{256}, 0, b->stream, b->args, b->nMoments, b->nHist);
    OWE_HIP_CHECK(hipEventRecord(b->ev[0], b->stream));
    hipLaunchKernelGGL(lensKernel, dim3{unsigned(std::max(1, b->computeUnits * b->blocksPerCU))}, dim3{kLensBlock}, 0,
                       b->stream, b->args);
    OWE_HIP_CHECK(hipGetLastError());
    OWE_HIP_CHECK(hipEventRecord(b->ev[1], b->stream));
    OWE_HIP_CHECK(hipEventSynchronize(b->ev[1]));
    float ms = 0;
    OWE_HIP_CHECK(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    b->lastMs = ms;
    b->ran = true;
    return ms;
}

}  // namespace
}  // namespace owehip

extern "C" {

int owe_hip_lens_batch_create(int device, const owe_medium* media, uint32_t n_media, const owe_lens_surface* surfaces,
                              uint32_t n_surfaces, const owe_lens* lenses, uint32_t n_lenses,
                              const owe_lens_batch_settings* settings, owe_hip_lens_batch** out) {
    owe_hip_lens_batch* b = nullptr;
    const int st = guard([&] {
        require(out && media && surfaces && lenses && settings, "null argument");
        *out = nullptr;
        require(n_media > 0 && n_surfaces > 0 && n_lenses > 0, "empty batch");
        const owe_lens_batch_settings& s = *settings;
        require(s.n_wavelengths >= 1 && s.n_wavelengths <= OWE_LENS_MAX_WAVELENGTHS, "1 to 16 wavelengths");
        for (int w = 0; w < s.n_wavelengths; ++w)
            require(s.wavelengths[w] >= 300 && s.wavelengths[w] <= 2000, "wavelengths must lie in [300, 2000] nm");
        require(s.reference_wavelength >= 300 && s.reference_wavelength <= 2000, "reference wavelength out of range");
        require(std::fabs(s.field_deg) < 80, "field angle must lie in (-80, 80) degrees");
        require(s.detector_pixels >= 0 && s.detector_pixels <= 4096, "detector pixels must lie in [0, 4096]");
        require(s.detector_half_size >= 0, "detector half size must be non-negative");
        for (uint32_t i = 0; i < n_surfaces; ++i) {
            const owe_lens_surface& r = surfaces[i];
            const std::string name = "lens surface " + std::to_string(i);
            require(r.medium >= 0 && uint32_t(r.medium) < n_media, name + ": medium out of range");
            require(r.semi_diameter > 0 && std::isfinite(r.semi_diameter), name + ": semi-diameter must be positive");
            require(std::isfinite(r.curvature) && std::isfinite(r.conic) && std::isfinite(r.thickness),
                    name + ": values must be finite");
            require(1 - (1 + r.conic) * r.curvature * r.curvature * r.semi_diameter * r.semi_diameter >= 0,
                    name + ": aperture exceeds the domain of the conic");
        }
        selectDevice(device);
        b = new owe_hip_lens_batch;
        b->device = device;
        b->settings = s;
        hipDeviceProp_t prop;
        OWE_HIP_CHECK(hipGetDeviceProperties(&prop, device));
        b->computeUnits = prop.multiProcessorCount;
        OWE_HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        OWE_HIP_CHECK(hipEventCreate(&b->ev[0]));
        OWE_HIP_CHECK(hipEventCreate(&b->ev[1]));
        const uint32_t nw = uint32_t(s.n_wavelengths);
        std::vector<DSurf> ds;
        std::vector<float> nAfter(size_t(n_surfaces) * nw, 1.0f), nAmb(size_t(n_lenses) * nw, 1.0f);
        std::vector<uint8_t> used(n_surfaces, 0);
        const double f = s.field_deg * 3.14159265358979323846 / 180;
        for (uint32_t li = 0; li < n_lenses; ++li) {
            const owe_lens& L = lenses[li];
            const std::string name = "lens " + std::to_string(li);
            require(L.surface_count >= 1 && uint64_t(L.first_surface) + L.surface_count <= n_surfaces,
                    name + ": surfaces out of range");
            require(L.rays >= 1 && L.rays <= (uint64_t(1) << 32), name + ": rays must lie in [1, 2^32]");
            require(L.ambient >= 0 && uint32_t(L.ambient) < n_media, name + ": ambient medium out of range");
            require(L.pupil_rings >= 0 && L.pupil_rings <= 30000, name + ": pupil rings must lie in [0, 30000]");
            require(L.pupil_rings == 0 || L.rays == 1 + 3 * uint64_t(L.pupil_rings) * uint64_t(L.pupil_rings + 1),
                    name + ": with pupil rings r the lens traces 1 + 3 r (r + 1) rays");
            const owe_lens_surface* rows = surfaces + L.first_surface;
            std::vector<double> nRef(L.surface_count);
            for (uint32_t k = 0; k < L.surface_count; ++k) {
                require(!used[L.first_surface + k], name + ": surfaces are shared with another lens");
                used[L.first_surface + k] = 1;
                nRef[k] = mediumIndexHost(media[rows[k].medium], s.reference_wavelength);
                for (uint32_t w = 0; w < nw; ++w)
                    nAfter[size_t(L.first_surface + k) * nw + w] = float(mediumIndexHost(media[rows[k].medium], s.wavelengths[w]));
            }
            for (uint32_t w = 0; w < nw; ++w) nAmb[size_t(li) * nw + w] = float(mediumIndexHost(media[L.ambient], s.wavelengths[w]));
            const ParaxialOut px = paraxialHost(rows, L.surface_count, nRef, mediumIndexHost(media[L.ambient], s.reference_wavelength));
            owe_lens_result r{};
            r.launched = L.rays * nw;
            r.efl = px.efl;
            r.bfl = px.bfl;
            r.entrance_pupil_z = px.epz;
            r.entrance_pupil_radius = px.epr;
            r.last_vertex_z = px.length;
            b->paraxial.push_back(r);
            DLensInfo I{};
            I.first = L.first_surface;
            I.count = L.surface_count;
            I.rays = L.rays;
            I.rings = uint32_t(L.pupil_rings);
            I.rEP = px.epr;
            I.zEP = px.epz;
            I.zStart = -0.05 - 0.1 * px.length;
            // The detector: the paraxial image plane at the reference wavelength, centred on the
            // chief ray's paraxial image height EFL·tan θ.
            I.detZ = float(px.length + px.bfl);
            I.detX = 0.0f;
            I.detY = float(px.efl * std::tan(f));
            I.detHalf = float(s.detector_half_size > 0 ? s.detector_half_size : 5e-3 * std::fabs(px.efl));
            b->info.push_back(I);
        }
        ds.resize(n_surfaces);
        for (uint32_t li = 0; li < n_lenses; ++li) {
            const owe_lens& L = lenses[li];
            double z = 0;
            for (uint32_t k = 0; k < L.surface_count; ++k) {
                const owe_lens_surface& r = surfaces[L.first_surface + k];
                DSurf d{};
                static_cast<DSag&>(d) = makeSag(r.curvature, r.conic, r.asphere, r.semi_diameter, 0.0);
                d.z = float(z);
                ds[L.first_surface + k] = d;
                z += r.thickness;
            }
        }
        b->lenses.assign(lenses, lenses + n_lenses);
        LensArgs& a = b->args;
        DSurf* dsurf = lensAlloc<DSurf>(b, ds.size());
        OWE_HIP_CHECK(hipMemcpy(dsurf, ds.data(), ds.size() * sizeof(DSurf), hipMemcpyHostToDevice));
        a.surf = dsurf;
        a.lens = lensAlloc<DLensInfo>(b, n_lenses);
        float* na = lensAlloc<float>(b, nAfter.size());
        OWE_HIP_CHECK(hipMemcpy(na, nAfter.data(), nAfter.size() * sizeof(float), hipMemcpyHostToDevice));
        a.nAfter = na;
        float* nb = lensAlloc<float>(b, nAmb.size());
        OWE_HIP_CHECK(hipMemcpy(nb, nAmb.data(), nAmb.size() * sizeof(float), hipMemcpyHostToDevice));
        a.nAmbient = nb;
        a.nLenses = n_lenses;
        a.nLambda = nw;
        a.pixels = uint32_t(s.detector_pixels);
        a.dirY = std::sin(f);
        a.dirZ = std::cos(f);
        a.next = lensAlloc<unsigned long long>(b, 1);
        b->nMoments = size_t(n_lenses) * nw * OWE_LENS_MOMENTS;
        a.moments = lensAlloc<double>(b, b->nMoments);
        b->nHist = size_t(n_lenses) * a.pixels * a.pixels;
        a.hist = lensAlloc<unsigned long long>(b, b->nHist);
        layoutItems(b, 1024);
        *out = b;
    });
    if (st != OWE_HIP_OK && b) {
        for (void* p : b->allocations) (void)hipFree(p);
        for (hipEvent_t e : b->ev)
            if (e) (void)hipEventDestroy(e);
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
    }
    return st;
}

int owe_hip_lens_batch_destroy(owe_hip_lens_batch* b) {
    return guard([&] {
        if (!b) return;
        (void)hipSetDevice(b->device);
        if (b->stream) (void)hipStreamSynchronize(b->stream);
        for (void* p : b->allocations) OWE_HIP_CHECK(hipFree(p));
        for (hipEvent_t e : b->ev)
            if (e) OWE_HIP_CHECK(hipEventDestroy(e));
        if (b->stream) OWE_HIP_CHECK(hipStreamDestroy(b->stream));
        delete b;
    });
}

int owe_hip_lens_batch_run(owe_hip_lens_batch* b) {
    return guard([&] {
        require(b != nullptr, "batch is null");
        OWE_HIP_CHECK(hipSetDevice(b->device));
        runOnce(b);
    });
}

int owe_hip_lens_batch_last_ms(const owe_hip_lens_batch* b, double* ms) {
    return guard([&] {
        require(b != nullptr && ms != nullptr, "null argument");
        *ms = b->lastMs;
    });
}

int owe_hip_lens_batch_set_shape(owe_hip_lens_batch* b, int blocks_per_cu, int rays_per_item) {
    return guard([&] {
        require(b != nullptr, "batch is null");
        require(blocks_per_cu >= 0 && blocks_per_cu <= 64, "blocks per compute unit must lie in [0, 64]");
        require(rays_per_item >= 0 && rays_per_item % 64 == 0 && rays_per_item <= (1 << 20),
                "rays per item must be a multiple of 64 up to 2^20");
        OWE_HIP_CHECK(hipSetDevice(b->device));
        if (blocks_per_cu > 0) b->blocksPerCU = blocks_per_cu;
        if (rays_per_item > 0) layoutItems(b, uint32_t(rays_per_item));
    });
}

int owe_hip_lens_batch_autotune(owe_hip_lens_batch* b, double* best_ms) {
    return guard([&] {
        require(b != nullptr, "batch is null");
        OWE_HIP_CHECK(hipSetDevice(b->device));
        double best = INFINITY;
        int bestB = b->blocksPerCU;
        uint32_t bestR = b->args.raysPerItem;
        for (int blocks : {1, 2, 4, 8})
            for (uint32_t rays : {256u, 1024u, 4096u, 16384u}) {
                b->blocksPerCU = blocks;
                layoutItems(b, rays);
                const double ms = runOnce(b);
                if (ms < best) {
                    best = ms;
                    bestB = blocks;
                    bestR = rays;
                }
            }
        b->blocksPerCU = bestB;
        layoutItems(b, bestR);
        b->ran = false;
        if (best_ms) *best_ms = best;
    });
}

int owe_hip_lens_batch_results(owe_hip_lens_batch* b, owe_lens_result* lenses, double* moments, float* irradiance) {
    return guard([&] {
        require(b != nullptr, "batch is null");
        require(b->ran, "no run yet");
        OWE_HIP_CHECK(hipSetDevice(b->device));
        const uint32_t nl = uint32_t(b->lenses.size()), nw = b->args.nLambda, px = b->args.pixels;
        std::vector<double> m(b->nMoments);
        OWE_HIP_CHECK(hipMemcpy(m.data(), b->args.moments, m.size() * sizeof(double), hipMemcpyDeviceToHost));
        if (moments) std::copy(m.begin(), m.end(), moments);
        if (lenses) {
            for (uint32_t li = 0; li < nl; ++li) {
                owe_lens_result r = b->paraxial[li];
                double s[OWE_LENS_MOMENTS] = {0};
                for (uint32_t w = 0; w < nw; ++w)
                    for (int q = 0; q < OWE_LENS_MOMENTS; ++q) s[q] += m[(size_t(li) * nw + w) * OWE_LENS_MOMENTS + q];
                r.arrived = uint64_t(s[0]);
                r.transmitted = s[8] / double(r.launched);
                // focusedSpot (analysis.cpp:85-113) from the sums.
                const double zImg = r.last_vertex_z + r.bfl;
                if (s[0] >= 3) {
                    const double n = s[0], max_ = s[1] / n, may = s[2] / n, mbx = s[3] / n, mby = s[4] / n;
                    const double varA = s[5] / n - (max_ * max_ + may * may), varB = s[6] / n - (mbx * mbx + mby * mby);
                    const double cov = s[7] / n - (max_ * mbx + may * mby);
                    const double zBest = varB > 0 ? -cov / varB : zImg;
                    auto rms = [&](double z) { return std::sqrt(std::max(0.0, varA + 2 * z * cov + z * z * varB)); };
                    r.best_focus_z = zBest - r.last_vertex_z;
                    r.rms_best = rms(zBest);
                    r.rms_paraxial = rms(zImg);
                    r.centroid_y = may + mby * zImg;
                } else {
                    r.best_focus_z = r.rms_best = r.rms_paraxial = r.centroid_y = NAN;
                }
                lenses[li] = r;
            }
        }
        if (irradiance && px > 0) {
            std::vector<unsigned long long> h(b->nHist);
            OWE_HIP_CHECK(hipMemcpy(h.data(), b->args.hist, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
            const double cosT = b->args.dirZ;
            for (uint32_t li = 0; li < nl; ++li) {
                // Unit irradiance across the beam: each ray carries π r² cos θ / (rays · nλ) of flux.
                const DLensInfo& I = b->info[li];
                const double pix = 2.0 * I.detHalf / px;
                const double k = 3.14159265358979323846 * I.rEP * I.rEP * 0.999 * 0.999 * cosT /
                                 (double(I.rays) * nw * kHistScale * pix * pix);
                for (size_t q = 0; q < size_t(px) * px; ++q)
                    irradiance[size_t(li) * px * px + q] = float(double(h[size_t(li) * px * px + q]) * k);
            }
        }
    });
}

int owe_hip_lens_batch_trace(owe_hip_lens_batch* b, uint32_t lens, int wavelength, const uint64_t* index, uint64_t n,
                             double* out) {
    return guard([&] {
        require(b != nullptr, "batch is null");
        require(lens < b->lenses.size(), "lens out of range");
        require(wavelength >= 0 && uint32_t(wavelength) < b->args.nLambda, "wavelength out of range");
        if (n == 0) return;
        require(index != nullptr && out != nullptr, "null buffer");
        for (uint64_t k = 0; k < n; ++k) require(index[k] < b->lenses[lens].rays, "ray index out of range");
        OWE_HIP_CHECK(hipSetDevice(b->device));
        uint64_t* di = nullptr;
        double* dout = nullptr;
        OWE_HIP_CHECK(hipMalloc(&di, n * sizeof(uint64_t)));
        OWE_HIP_CHECK(hipMalloc(&dout, 7 * n * sizeof(double)));
        OWE_HIP_CHECK(hipMemcpy(di, index, n * sizeof(uint64_t), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(lensProbe, dim3{unsigned((n + 255) / 256)}, dim3{256}, 0, b->stream, b->args, lens,
                           uint32_t(wavelength), di, n, dout);
        const hipError_t e = hipGetLastError();
        if (e == hipSuccess) {
            (void)hipStreamSynchronize(b->stream);
            (void)hipMemcpy(out, dout, 7 * n * sizeof(double), hipMemcpyDeviceToHost);
        }
        (void)hipFree(di);
        (void)hipFree(dout);
        OWE_HIP_CHECK(e);
    });
}

}  // extern "C"
