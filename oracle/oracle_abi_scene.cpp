// TEST INFRASTRUCTURE ONLY — the double-precision CPU oracle (see oracle_scene.hpp).
// C entry points over the world and the camera-path tracer: whole passes into an XYZ film
// (CpuRenderer::pass, cpu_backend.cpp:52-94, in 16×16 tiles on host threads) and single paths by
// (pixel, sample) for per-path comparisons with the hip backend.
#include <algorithm>
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
    s.photonRadius = r.photon_radius > 0 ? r.photon_radius : 2.0;
    s.photonAlpha = r.photon_alpha > 0 ? r.photon_alpha : 2.0 / 3.0;
    s.photonsPerPixel = r.photons_per_pixel > 0 ? r.photons_per_pixel : 1.0;
    s.photonGroups = r.photon_groups > 0 ? r.photon_groups : 4;
    s.partition = r.integrator == OWE_INTEGRATOR_HYBRID;
    // Camera paths rendered under vcm or bdpt settings are the medium-only pass that accompanies
    // those integrators (light that reached the eye through a medium scattering vertex).
    s.mediumOnly = r.integrator == OWE_INTEGRATOR_VCM || r.integrator == OWE_INTEGRATOR_BDPT;
    return s;
}
// The stream of sample s of a pixel in the per-sample layout.
Pcg sampleStream(uint64_t seed, uint64_t pixel, uint64_t s, uint64_t pass) {
    return Pcg(hashPair(hashPair(seed, pixel), s), pass);
}

// SPPM streams and wavelengths (the hip backend's layout): an iteration's base rotation (a golden-ratio
// sequence from a seed offset, as gpu_renderer_passes.cpp:52-56), group g's rotation by g/(4 G), the
// camera path of a pixel in a group chosen by a hash of (pixel, iteration), photon i in group i mod G.
constexpr uint64_t kSppmCamera = 0x6a09e667f3bcc908ULL, kSppmGroup = 0xbb67ae8584caa73bULL,
                   kSppmPhoton = 0xa54ff53a5f1d36f1ULL;
double sppmRotation(uint64_t seed, uint64_t iteration) {
    const double offset = double(uint32_t(mix64(seed) >> 32)) / 4294967296.0;
    const double u = offset + 0.6180339887498949 * double(iteration + 1);
    return u - std::floor(u);
}
Wavelengths sppmWavelengths(const Settings& s, double rotation, uint32_t group) {
    double u = rotation + double(group) / double(4 * s.photonGroups);
    u -= std::floor(u);
    return s.spectral == OWE_SPECTRAL_HERO4 ? Wavelengths::hero(u) : Wavelengths::uniform1(u, s.lambdaMin, s.lambdaMax);
}
}  // namespace

extern "C" {

// The initial gather radius per metre of unfolded path and its floor (GpuRenderer::resetSppm,
// gpu_renderer_passes.cpp:238-257): photon_radius pixel footprints, a pixel's angle for an observer
// or for a sensor seen from its aim disc's centre; a sensor without one takes a fixed radius.
void oracle_sppm_radius(const void* wh, const owe_render_settings* rs, double out[2]) {
    const World& w = *static_cast<const World*>(wh);
    const Settings set = toSettings(*rs);
    const Observer& ob = w.observer;
    double pixelAngle = 0, floor = 0;
    if (!ob.sensor()) {
        pixelAngle = 2 * ob.tanY / ob.height;
    } else {
        const double pixel = 2 * ob.halfY / ob.height;
        const double d = ob.hasAim ? length(ob.aimCenter - ob.toWorld.point(V3{0, 0, 0})) : 0;
        if (d > 0) pixelAngle = pixel / d;
        else floor = set.photonRadius * pixel;
    }
    out[0] = set.photonRadius * pixelAngle;
    out[1] = std::max(floor, 1e-9);
}

// Iterations [start, start + n) of SPPM (sppm.slang; GpuRenderer::passSppm, gpu_renderer_passes.cpp:259-290)
// on caller-held state: direct (W·H·3) sums the camera passes' own XYZ; state (W·H·5) is τ (XYZ), N
// and the gather radius r per pixel; guideCounts (PhotonGuide::kCells) the photon guide's counts. A photon reaches every visible point within r: the reference's
// multi-level hash finds exactly these (a point lives in one cell of size ≥ 2r, and the 2×2×2 cells
// about the photon cover its reach), so the oracle searches them directly. stats += paths (camera
// paths and photons), segments, inconsistencies, leaks.
void oracle_sppm(const void* wh, const owe_render_settings* rs, int n, int start, int threads, double* direct,
                 double* state, uint64_t* guideCounts, uint64_t* stats) {
    const World& w = *static_cast<const World*>(wh);
    const Settings set = toSettings(*rs);
    const Tracer tr(w, set);
    const int W = w.observer.width, H = w.observer.height;
    const size_t pixels = size_t(W) * size_t(H);
    const uint64_t photons = std::max<uint64_t>(1, uint64_t(std::llround(double(pixels) * set.photonsPerPixel)));
    const uint32_t G = uint32_t(set.photonGroups);
    if (threads <= 0) threads = int(std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
    double r0[2];
    oracle_sppm_radius(wh, rs, r0);
    std::vector<Stats> st(static_cast<size_t>(threads));
    std::vector<VisiblePoint> vps(pixels);
    PhotonGuide guide;
    std::copy(guideCounts, guideCounts + PhotonGuide::kCells, guide.counts.begin());
    auto parallel = [&](uint64_t count, auto&& body) {
        std::vector<std::thread> pool;
        for (int k = 0; k < threads; ++k)
            pool.emplace_back([&, k] {
                for (uint64_t i = count * uint64_t(k) / uint64_t(threads); i < count * uint64_t(k + 1) / uint64_t(threads); ++i)
                    body(k, i);
            });
        for (auto& th : pool) th.join();
    };
    for (int it = start; it < start + n; ++it) {
        const double rot = sppmRotation(set.seed, uint64_t(it));
        // Camera pass: one path per pixel.
        parallel(pixels, [&](int k, uint64_t pixel) {
            const uint32_t g = uint32_t(hashPair(hashPair(set.seed ^ kSppmGroup, pixel), uint64_t(it)) % G);
            Pcg rng(hashPair(set.seed ^ kSppmCamera, pixel), uint64_t(it));
            double c[3];
            VisiblePoint& vp = vps[pixel];
            tr.sppmCamera(int(pixel % uint64_t(W)), int(pixel / uint64_t(W)), sppmWavelengths(set, rot, g), rng,
                          st[size_t(k)], c, vp);
            vp.group = g;
            for (int j = 0; j < 3; ++j) direct[3 * pixel + j] += c[j];
            if (vp.valid && state[5 * pixel + 4] == 0)
                state[5 * pixel + 4] = std::max(r0[0] * vp.pathLength, r0[1]);
        });
        // The points within reach of a photon: a sweep over x.
        std::vector<uint32_t> order;
        double rmax = 0;
        for (size_t i = 0; i < pixels; ++i)
            if (vps[i].valid) {
                order.push_back(uint32_t(i));
                rmax = std::max(rmax, state[5 * i + 4]);
            }
        std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return vps[a].p.x < vps[b].p.x; });
        std::vector<double> xs(order.size());
        for (size_t i = 0; i < order.size(); ++i) xs[i] = vps[order[i]].p.x;
        struct Gather final : Tracer::PhotonSink {
            const Tracer* tr;
            const std::vector<VisiblePoint>* vps;
            const std::vector<uint32_t>* order;
            const std::vector<double>* xs;
            const double* state;
            double rmax;
            uint32_t group, groups;
            std::vector<double> flux;  // per pixel: XYZ, M
            std::vector<uint32_t> taught;  // guide cells of this thread's useful photons
            void gather(const Hit& hit, V3 d, const Spec4& beta, const Wavelengths& wl) override {
                const auto lo = std::lower_bound(xs->begin(), xs->end(), hit.p.x - rmax);
                const auto hi = std::upper_bound(xs->begin(), xs->end(), hit.p.x + rmax);
                for (auto q = lo; q != hi; ++q) {
                    const uint32_t i = (*order)[size_t(q - xs->begin())];
                    const VisiblePoint& v = (*vps)[i];
                    const double r = state[5 * size_t(i) + 4];
                    const V3 dp = hit.p - v.p;
                    if (dot(dp, dp) > r * r) continue;
                    // A disc in the point's tangent plane, on a surface facing the same way.
                    if (std::fabs(dot(dp, v.n)) > 0.25 * r || dot(v.n, hit.n) < 0.5) continue;
                    if (!v.free && v.group != group) continue;  // wavelengths must agree
                    double c[3];
                    if (!tr->sppmValue(v, int(groups), d, beta, wl, c)) continue;
                    for (int j = 0; j < 3; ++j) flux[4 * size_t(i) + size_t(j)] += c[j];
                    flux[4 * size_t(i) + 3] += 1;
                    useful = true;
                }
            }
        };
        std::vector<Gather> sinks(static_cast<size_t>(threads));
        for (auto& g : sinks) {
            g.tr = &tr;
            g.vps = &vps;
            g.order = &order;
            g.xs = &xs;
            g.state = state;
            g.rmax = rmax;
            g.groups = G;
            g.flux.assign(4 * pixels, 0.0);
        }
        // The guide learnt from the iterations before this one (sppm_guide_kernel).
        guide.build();
        parallel(photons, [&](int k, uint64_t i) {
            Gather& g = sinks[size_t(k)];
            g.group = uint32_t(i % G);
            Pcg rng(hashPair(set.seed ^ kSppmPhoton, i), uint64_t(it));
            const uint32_t cell = tr.sppmPhoton(sppmWavelengths(set, rot, g.group), rng, st[size_t(k)], g, &guide);
            if (cell != kNone) g.taught.push_back(cell);
        });
        for (const Gather& g : sinks)
            for (uint32_t c : g.taught) guide.counts[c]++;
        // The progressive update (sppm_update_kernel): N' = N + αM, r' = r·√(N'/(N+M)), τ' = (τ+Φ)·N'/(N+M).
        for (size_t i = 0; i < pixels; ++i) {
            double phi[3] = {0, 0, 0}, M = 0;
            for (const Gather& g : sinks) {
                for (int j = 0; j < 3; ++j) phi[j] += g.flux[4 * i + size_t(j)];
                M += g.flux[4 * i + 3];
            }
            if (M == 0) continue;
            double* s = state + 5 * i;
            const double N = s[3], Nn = N + set.photonAlpha * M, ratio = Nn / (N + M);
            for (int j = 0; j < 3; ++j) s[j] = (s[j] + phi[j]) * ratio;
            s[3] = Nn;
            s[4] *= std::sqrt(ratio);
        }
    }
    std::copy(guide.counts.begin(), guide.counts.end(), guideCounts);
    Stats total;
    for (const Stats& s : st) total.add(s);
    stats[0] += total.paths;
    stats[1] += total.segments;
    stats[2] += total.inconsistencies;
    stats[3] += total.leaks;
}

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

// Iterations [start, start + n) of vertex connection and merging (vcm.slang; merging off: bdpt) in the
// hip backend's layout: N = ⌈round(W·H·photons_per_pixel)/G⌉·G light subpaths per iteration, subpath
// j of group g = j mod G on the stream (hashCombine(seed ^ kVcmLight, j), iteration) at the group's
// wavelengths; the cache is the subpaths' non-specular vertices in the order (g, j / G, vertex); the
// camera path of a pixel runs on (hashCombine(seed ^ kVcmCamera, pixel), iteration) in the group a
// hash of (iteration, pixel) picks, and merges with every cached vertex of its group within the
// radius (the multi-level hash finds the same set; here a sweep over the group's vertices by x). camera (W·H·3) += the camera paths' XYZ (the
// caller divides by the iterations); light (W·H·3) += the light subpaths' splats (the caller divides
// by N times the iterations); stats += paths (light and camera), segments, inconsistencies, leaks;
// cached[it - start] = the iteration's cache size when cached is given.
void oracle_vcm(const void* wh, const owe_render_settings* rs, int merging, int n, int start, int threads,
                double* camera, double* light, uint64_t* stats, uint64_t* cached) {
    constexpr uint64_t kVcmLight = 0x510e527fade682d1ULL, kVcmCamera = 0x9b05688c2b3e6c1fULL,
                       kVcmGroup = 0x1f83d9abfb41bd6bULL;
    const World& w = *static_cast<const World*>(wh);
    Settings set = toSettings(*rs);
    const uint32_t G = uint32_t(rs->photon_groups > 0 ? rs->photon_groups : 1);
    set.photonGroups = int(G);
    const Tracer tr(w, set);
    const int W = w.observer.width, H = w.observer.height;
    const size_t pixels = size_t(W) * size_t(H);
    const uint64_t want = std::max<uint64_t>(1, uint64_t(std::llround(double(pixels) * set.photonsPerPixel)));
    const uint64_t N = (want + G - 1) / G * G, perGroup = N / G;
    if (threads <= 0) threads = int(std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
    struct Film final : Tracer::Splat {
        int width = 0;
        std::vector<double> v;
        void add(int px, int py, const double c[3]) override {
            if (!std::isfinite(c[0] + c[1] + c[2])) return;
            for (int k = 0; k < 3; ++k) v[3 * (size_t(py) * size_t(width) + size_t(px)) + size_t(k)] += c[k];
        }
    };
    auto parallel = [&](uint64_t count, auto&& body) {
        std::vector<std::thread> pool;
        for (int k = 0; k < threads; ++k)
            pool.emplace_back([&, k] {
                for (uint64_t i = count * uint64_t(k) / uint64_t(threads); i < count * uint64_t(k + 1) / uint64_t(threads); ++i)
                    body(k, i);
            });
        for (auto& th : pool) th.join();
    };
    std::vector<Stats> st(static_cast<size_t>(threads));
    std::vector<Film> films(static_cast<size_t>(threads));
    for (Film& f : films) {
        f.width = W;
        f.v.assign(3 * pixels, 0.0);
    }
    const double perMetre = merging ? double(float(set.photonRadius * 2.0 * double(float(w.observer.tanY)) / double(H))) : 0.0;
    for (int it = start; it < start + n; ++it) {
        Tracer::VcmParams p;
        p.lightPaths = double(perGroup);
        p.groups = G;
        p.iteration = uint32_t(it);
        p.mergePerMetre = merging ? perMetre * std::pow(double(it + 1), (set.photonAlpha - 1.0) / 2.0) : 0.0;
        const double rot = sppmRotation(set.seed, uint64_t(it));
        std::vector<std::vector<Tracer::VcmVertex>> paths(N);
        parallel(N, [&](int k, uint64_t j) {
            const uint32_t g = uint32_t(j % G);
            Pcg rng(hashPair(set.seed ^ kVcmLight, j), uint64_t(it));
            tr.vcmLight(sppmWavelengths(set, rot, g), rng, st[size_t(k)], films[size_t(k)], p, g,
                        paths[size_t(g) * perGroup + j / G]);
        });
        std::vector<Tracer::VcmVertex> cache;
        std::vector<uint64_t> first(G + 1, 0);
        for (uint32_t g = 0; g < G; ++g) {
            first[g] = cache.size();
            for (uint64_t i = 0; i < perGroup; ++i)
                for (const auto& y : paths[size_t(g) * perGroup + i]) cache.push_back(y);
        }
        first[G] = cache.size();
        if (cached) cached[it - start] = cache.size();
        paths.clear();
        // Each group's vertices by x for merging's sweep.
        std::vector<uint32_t> index(cache.size());
        std::vector<double> xs(cache.size());
        for (uint32_t g = 0; g < G; ++g) {
            uint32_t* ix = index.data() + first[g];
            const uint32_t m = uint32_t(first[g + 1] - first[g]);
            for (uint32_t i = 0; i < m; ++i) ix[i] = i;
            const Tracer::VcmVertex* c = cache.data() + first[g];
            std::sort(ix, ix + m, [&](uint32_t a, uint32_t b) { return c[a].p.x < c[b].p.x; });
            for (uint32_t i = 0; i < m; ++i) xs[first[g] + i] = c[ix[i]].p.x;
        }
        parallel(pixels, [&](int k, uint64_t pixel) {
            const uint32_t g = uint32_t(mix64((uint64_t(uint32_t(it)) << 32 | pixel) ^ kVcmGroup) % uint64_t(G));
            Pcg rng(hashPair(set.seed ^ kVcmCamera, pixel), uint64_t(it));
            double c[3];
            if (tr.vcmCamera(int(pixel % uint64_t(W)), int(pixel / uint64_t(W)), sppmWavelengths(set, rot, g), rng,
                             st[size_t(k)], p, cache.data() + first[g], uint32_t(first[g + 1] - first[g]), c,
                             Tracer::VcmSweep{xs.data() + first[g], index.data() + first[g]}))
                for (int q = 0; q < 3; ++q) camera[3 * pixel + size_t(q)] += c[q];
        });
    }
    Stats total;
    for (int k = 0; k < threads; ++k) {
        for (size_t j = 0; j < 3 * pixels; ++j) light[j] += films[size_t(k)].v[j];
        total.add(st[size_t(k)]);
    }
    stats[0] += total.paths;
    stats[1] += total.segments;
    stats[2] += total.inconsistencies;
    stats[3] += total.leaks;
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
