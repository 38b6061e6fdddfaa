// TEST INFRASTRUCTURE ONLY — the double-precision CPU oracle (see oracle_scene.hpp).
// C entry points over the world and the camera-path tracer: whole passes into an XYZ film
// (CpuRenderer::pass, cpu_backend.cpp:52-94, in 16×16 tiles on host threads) and single paths by
// (pixel, sample) for per-path comparisons with the hip backend.
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle_scene.hpp"

using namespace oracle;

namespace {
Settings toSettings(const owe_render_settings& r) {
    Settings s;
    s.seed = r.seed;
    s.maxDepth = r.max_depth;
    s.rrDepth = r.rr_depth;
    s.spectral = r.spectral;
    s.lambdaMin = r.lambda_min;
    s.lambdaMax = r.lambda_max;
    s.fresnelFloor = r.fresnel_floor;
    return s;
}
// The stream of sample s of a pixel in the per-sample layout.
Pcg sampleStream(uint64_t seed, uint64_t pixel, uint64_t s, uint64_t pass) {
    return Pcg(hashPair(hashPair(seed, pixel), s), pass);
}
}  // namespace

extern "C" {

void* oracle_world_create(const owe_scene* s) { return new World(*s); }
void oracle_world_destroy(void* w) { delete static_cast<World*>(w); }
// Whether rays of this world resolve contacts (two distinct boundaries' boxes meet).
int oracle_world_contacts(const void* wh) { return static_cast<const World*>(wh)->contacts ? 1 : 0; }

// Light-selection pdfs in table order (boundaries, then the sun); returns the count.
int oracle_light_table(const void* wh, double* pdf, uint32_t* boundary, int cap) {
    const World& w = *static_cast<const World*>(wh);
    for (size_t i = 0; i < w.lights.size() && int(i) < cap; ++i) {
        pdf[i] = w.lightPdf[i];
        boundary[i] = w.lights[i].boundary;
    }
    return int(w.lights.size());
}

// Adds one pass of spp samples per pixel into xyz (W·H·3) and samples (W·H); stats += paths,
// segments, inconsistencies, leaks. layout: 0 one stream per pixel (the reference's), 1 one per sample.
void oracle_render(const void* wh, const owe_render_settings* rs, int spp, int pass, int layout, int threads,
                   double* xyz, double* samples, uint64_t* stats) {
    const World& w = *static_cast<const World*>(wh);
    const Settings set = toSettings(*rs);
    const Tracer tr(w, set);
    const int W = w.observer.width, H = w.observer.height, tile = 16;
    const int tilesX = (W + tile - 1) / tile, tilesY = (H + tile - 1) / tile;
    if (threads <= 0) threads = int(std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
    std::atomic<int> next{0};
    std::vector<Stats> st{size_t(threads)};
    auto worker = [&](int tid) {
        for (;;) {
            const int t = next.fetch_add(1);
            if (t >= tilesX * tilesY) break;
            const int tx = t % tilesX, ty = t / tilesX;
            for (int y = ty * tile; y < std::min(H, (ty + 1) * tile); ++y)
                for (int x = tx * tile; x < std::min(W, (tx + 1) * tile); ++x) {
                    const uint64_t pixel = uint64_t(y) * W + x;
                    Pcg rng(hashPair(set.seed, pixel), uint64_t(pass));
                    const double rot = rng.uniform();
                    double acc[3] = {0, 0, 0};
                    for (int s = 0; s < spp; ++s) {
                        double c[3];
                        bool ok;
                        if (layout == kLayoutPixel) {
                            ok = tr.cameraSample(x, y, s, spp, rot, rng, st[size_t(tid)], c);
                        } else {
                            Pcg own = sampleStream(set.seed, pixel, uint64_t(s), uint64_t(pass));
                            ok = tr.cameraSample(x, y, s, spp, rot, own, st[size_t(tid)], c);
                        }
                        if (!ok) continue;
                        for (int k = 0; k < 3; ++k) acc[k] += c[k];
                    }
                    for (int k = 0; k < 3; ++k) xyz[3 * pixel + k] += acc[k];
                    samples[pixel] += spp;
                }
        }
    };
    std::vector<std::thread> pool;
    for (int k = 0; k < threads; ++k) pool.emplace_back(worker, k);
    for (auto& th : pool) th.join();
    Stats total;
    for (const Stats& s : st) total.add(s);
    stats[0] += total.paths;
    stats[1] += total.segments;
    stats[2] += total.inconsistencies;
    stats[3] += total.leaks;
}

// Per-path XYZ of n camera samples (pixel[i], sample[i]) in the per-sample stream layout; segments
// per path in segs.
void oracle_camera_paths(const void* wh, const owe_render_settings* rs, int spp, int pass, uint64_t n,
                         const uint32_t* pixel, const uint32_t* sample, double* xyz, uint32_t* segs, int threads) {
    const World& w = *static_cast<const World*>(wh);
    const Settings set = toSettings(*rs);
    const Tracer tr(w, set);
    const int W = w.observer.width;
    if (threads <= 0) threads = int(std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    for (int k = 0; k < threads; ++k)
        pool.emplace_back([&, k] {
            for (uint64_t i = n * uint64_t(k) / uint64_t(threads); i < n * uint64_t(k + 1) / uint64_t(threads); ++i) {
                const uint64_t px = pixel[i];
                Pcg pr(hashPair(set.seed, px), uint64_t(pass));
                const double rot = pr.uniform();
                Pcg own = sampleStream(set.seed, px, sample[i], uint64_t(pass));
                Stats st;
                double c[3];
                if (!tr.cameraSample(int(px % uint64_t(W)), int(px / uint64_t(W)), int(sample[i]), spp, rot, own, st, c))
                    c[0] = c[1] = c[2] = 0;
                for (int q = 0; q < 3; ++q) xyz[3 * i + q] = c[q];
                segs[i] = uint32_t(st.segments);
            }
        });
    for (auto& th : pool) th.join();
}

// Adds one light-tracing pass into xyz (W·H·3): particles [0, W·H·ppp), particle i on the stream
// (hashCombine(seed ^ 0x5bd1e995, i), pass) as particle_kernel.slang:15-16, each thread into its own
// film, summed in thread order. The caller divides by the particle count (gpu_renderer_film.cpp:17-20).
void oracle_render_particles(const void* wh, const owe_render_settings* rs, int ppp, int pass, int threads,
                             double* xyz, uint64_t* stats) {
    const World& w = *static_cast<const World*>(wh);
    const Settings set = toSettings(*rs);
    const Tracer tr(w, set);
    const int W = w.observer.width, H = w.observer.height;
    const uint64_t n = uint64_t(W) * uint64_t(H) * uint64_t(ppp);
    if (threads <= 0) threads = int(std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
    struct Film final : Tracer::Splat {
        int width = 0;
        std::vector<double> v;
        void add(int px, int py, const double c[3]) override {
            if (!std::isfinite(c[0] + c[1] + c[2])) return;
            for (int k = 0; k < 3; ++k) v[3 * (size_t(py) * size_t(width) + size_t(px)) + size_t(k)] += c[k];
        }
    };
    std::vector<Film> films(static_cast<size_t>(threads));
    std::vector<Stats> st(static_cast<size_t>(threads));
    std::vector<std::thread> pool;
    for (int k = 0; k < threads; ++k)
        pool.emplace_back([&, k] {
            Film& f = films[size_t(k)];
            f.width = W;
            f.v.assign(size_t(W) * size_t(H) * 3, 0.0);
            for (uint64_t i = n * uint64_t(k) / uint64_t(threads); i < n * uint64_t(k + 1) / uint64_t(threads); ++i) {
                Pcg rng(hashPair(set.seed ^ 0x5bd1e995ULL, i), uint64_t(pass));
                tr.particle(rng, st[size_t(k)], f);
            }
        });
    for (auto& th : pool) th.join();
    Stats total;
    for (int k = 0; k < threads; ++k) {
        for (size_t j = 0; j < films[size_t(k)].v.size(); ++j) xyz[j] += films[size_t(k)].v[j];
        total.add(st[size_t(k)]);
    }
    stats[0] += total.paths;
    stats[1] += total.segments;
    stats[2] += total.inconsistencies;
    stats[3] += total.leaks;
}

// The XYZ each of n particles (by index, of pass `pass`) deposits, summed over its splats; segments
// per particle in segs.
void oracle_particles(const void* wh, const owe_render_settings* rs, int pass, uint64_t n, const uint64_t* index,
                      double* xyz, uint32_t* segs, int threads) {
    const World& w = *static_cast<const World*>(wh);
    const Settings set = toSettings(*rs);
    const Tracer tr(w, set);
    if (threads <= 0) threads = int(std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
    struct Sum final : Tracer::Splat {
        double c[3] = {0, 0, 0};
        void add(int, int, const double v[3]) override {
            if (!std::isfinite(v[0] + v[1] + v[2])) return;
            for (int k = 0; k < 3; ++k) c[k] += v[k];
        }
    };
    std::vector<std::thread> pool;
    for (int k = 0; k < threads; ++k)
        pool.emplace_back([&, k] {
            for (uint64_t i = n * uint64_t(k) / uint64_t(threads); i < n * uint64_t(k + 1) / uint64_t(threads); ++i) {
                Pcg rng(hashPair(set.seed ^ 0x5bd1e995ULL, index[i]), uint64_t(pass));
                Stats st;
                Sum sum;
                tr.particle(rng, st, sum);
                for (int q = 0; q < 3; ++q) xyz[3 * i + q] = sum.c[q];
                segs[i] = uint32_t(st.segments);
            }
        });
    for (auto& th : pool) th.join();
}

// The particle emission set-up: the sky's share of emitted power, the scene radius and centre.
void oracle_emission_info(const void* wh, double* out) {
    const World& w = *static_cast<const World*>(wh);
    const V3 c = w.sceneCenter();
    out[0] = w.skyShare;
    out[1] = w.sceneRadius();
    out[2] = c.x;
    out[3] = c.y;
    out[4] = c.z;
    out[5] = w.environmentPower;
    out[6] = w.totalLightPower;
}

// World::intersect for n rays (origins and directions in world coordinates, the boundary each
// departs or ~0u): out n × 8 = t, point, normal, boundary (t = inf and boundary ~0u on a miss).
void oracle_world_intersect(const void* wh, uint64_t n, const double* o, const double* d, const uint32_t* departed,
                            double* out) {
    const World& w = *static_cast<const World*>(wh);
    for (uint64_t i = 0; i < n; ++i) {
        Ray r;
        r.o = V3(o[3 * i], o[3 * i + 1], o[3 * i + 2]);
        r.d = V3(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        r.departed = departed ? departed[i] : kNone;
        Hit h;
        double* q = out + 8 * i;
        if (w.intersect(r, kInf, h)) {
            q[0] = h.t, q[1] = h.p.x, q[2] = h.p.y, q[3] = h.p.z, q[4] = h.n.x, q[5] = h.n.y, q[6] = h.n.z;
            q[7] = double(h.boundary);
        } else {
            q[0] = kInf;
            for (int k = 1; k < 7; ++k) q[k] = std::nan("");
            q[7] = double(kNone);
        }
    }
}

double oracle_spectrum_luminance_blackbody(double kelvin) {
    Spectrum s;
    s.kind = OWE_SPEC_BLACKBODY;
    s.kelvin = kelvin;
    return s.luminance();
}

double oracle_medium_index(const owe_medium* m, double lnm) {
    IndexModel x;
    x.kind = m->index_kind;
    x.terms = m->terms;
    x.relAir = m->rel_air != 0;
    for (int k = 0; k < 4; ++k) {
        x.a[k] = m->coeff_a[k];
        x.b[k] = m->coeff_b[k];
    }
    return x.n(lnm);
}

}  // extern "C"
