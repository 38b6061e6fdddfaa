// TEST INFRASTRUCTURE ONLY — the double-precision CPU oracle (see oracle_core.hpp).
//
// The caustic slab of include/owe_hip_caustic.h in double, for checking the hip forward pass and its
// path-replay gradient:
//   bicubic Hermite patch over the node grid, node slopes by   WaveSurface::height, src/owe/scene/geometry.cpp:251-270,
//   central differences of the heights (clamped at the edges)  hermiteBases, src/owe/core/shared/surfaces.hpp:190-197
//   collimated particles, stratified over tiles of the entry   emitDistant, particle.slang:234-265
//   face (uniform in each), a
//   luminous spectral weight B(λ) ȳ(λ) / (pdf(λ) ∫ B ȳ)
//   transmission with weight 1 − R at both faces, Snell out   optics.hpp:14-54; transport.cpp:531-595
//   of the exit surface, the wall point, a quadratic B-spline splat over 3 × 3 pixels
// The particles' random numbers are the float kernels' (24 bits of the first PCG word of each draw),
// so both precisions trace the same particles. Finite differences of L = mean((E − target)²)
// re-trace only the particles whose cell a perturbed height reaches.
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "../include/owe_hip_caustic.h"
#include "oracle_scene.hpp"

using namespace oracle;

namespace {

struct Part {
    double x, y, weight, n1, n2;
    int ci, cj;  // the cell
};

struct Slab {
    const owe_caustic_settings& s;
    double cell, pix, invY, tileArea;
    uint32_t tiles;  // tiles per side of the entry face
    IndexModel slab, air;
    explicit Slab(const owe_caustic_settings& st) : s(st) {
        cell = 2 * s.slab_half / (s.nodes - 1);
        pix = 2 * s.wall_half / s.pixels;
        invY = 1 / integrateVisible([&](double l) { return planck(l, s.sun_kelvin) * cieY(l); });
        tiles = 1;
        while (tiles < 64 && uint64_t(4) * tiles * tiles * 64 <= s.particles) tiles *= 2;
        const double side = 2 * s.slab_half / tiles;
        tileArea = side * side / (pix * pix);
        slab.kind = s.medium.index_kind;
        slab.terms = s.medium.terms;
        slab.relAir = s.medium.rel_air != 0;
        for (int k = 0; k < 4; ++k) {
            slab.a[k] = s.medium.coeff_a[k];
            slab.b[k] = s.medium.coeff_b[k];
        }
        air.kind = OWE_INDEX_CIDDOR;
    }
    static double u24(Pcg& r) {
        const uint32_t w = r.next();
        r.next();
        return double(w >> 8) * (1.0 / 16777216.0);
    }
    Part particle(uint64_t i, uint64_t pass) const {
        Pcg r(hashPair(s.seed, i), pass);
        const double ul = u24(r);
        const double lambda = sampleVisible(ul), pdf = visiblePdf(lambda);
        const double u2 = u24(r), u1 = u24(r);
        // Stratified over tiles × tiles tiles in Morton order: tile t holds particles
        // [⌊t·N/T⌋, ⌊(t+1)·N/T⌋), uniform over the tile, weighted by the tile's area over its count.
        const uint64_t T = uint64_t(tiles) * tiles, N = s.particles;
        auto start = [&](uint64_t t) { return t * N / T; };
        uint64_t t = i * T / N;
        while (t + 1 < T && start(t + 1) <= i) ++t;
        while (t > 0 && start(t) > i) --t;
        uint32_t tx = 0, ty = 0;
        for (int b = 0; b < 16; ++b) {
            tx |= uint32_t((t >> (2 * b)) & 1) << b;
            ty |= uint32_t((t >> (2 * b + 1)) & 1) << b;
        }
        const double side = 2 * s.slab_half / tiles;
        Part p;
        p.x = -s.slab_half + (tx + u1) * side;
        p.y = -s.slab_half + (ty + u2) * side;
        p.n1 = slab.n(lambda);
        p.n2 = air.n(lambda);
        p.weight = planck(lambda, s.sun_kelvin) * cieY(lambda) / pdf * invY * (1 - fresnelDielectric(1, p.n1 / p.n2)) *
                   tileArea / double(start(t + 1) - start(t));
        const double fx = (p.x + s.slab_half) / cell, fy = (p.y + s.slab_half) / cell;
        p.ci = std::min(std::max(int(fx), 0), int(s.nodes) - 2);
        p.cj = std::min(std::max(int(fy), 0), int(s.nodes) - 2);
        return p;
    }
    double H(const double* h, int i, int j) const {
        const int N = int(s.nodes);
        i = std::min(std::max(i, 0), N - 1);
        j = std::min(std::max(j, 0), N - 1);
        return h[size_t(j) * s.nodes + size_t(i)];
    }
    // Height and gradient of the exit surface at (x, y) in cell (i, j).
    void surface(const double* h, const Part& p, double& z, double& gx, double& gy) const {
        const int i = p.ci, j = p.cj;
        const double u = (p.x + s.slab_half) / cell - i, v = (p.y + s.slab_half) / cell - j;
        auto bases = [](double t, double P[2], double Q[2], double dP[2], double dQ[2]) {
            P[0] = 2 * t * t * t - 3 * t * t + 1;
            P[1] = -2 * t * t * t + 3 * t * t;
            Q[0] = t * t * t - 2 * t * t + t;
            Q[1] = t * t * t - t * t;
            dP[0] = 6 * t * t - 6 * t;
            dP[1] = -6 * t * t + 6 * t;
            dQ[0] = 3 * t * t - 4 * t + 1;
            dQ[1] = 3 * t * t - 2 * t;
        };
        double P[2], Q[2], dP[2], dQ[2], Pv[2], Qv[2], dPv[2], dQv[2];
        bases(u, P, Q, dP, dQ);
        bases(v, Pv, Qv, dPv, dQv);
        double f = 0, fu = 0, fv = 0;
        for (int b = 0; b < 2; ++b)
            for (int a = 0; a < 2; ++a) {
                const int x = i + a, y = j + b;
                const double c0 = H(h, x, y), cu = (H(h, x + 1, y) - H(h, x - 1, y)) / 2,
                             cv = (H(h, x, y + 1) - H(h, x, y - 1)) / 2,
                             cuv = (H(h, x + 1, y + 1) - H(h, x - 1, y + 1) - H(h, x + 1, y - 1) + H(h, x - 1, y - 1)) / 4;
                f += c0 * P[a] * Pv[b] + cu * Q[a] * Pv[b] + cv * P[a] * Qv[b] + cuv * Q[a] * Qv[b];
                fu += c0 * dP[a] * Pv[b] + cu * dQ[a] * Pv[b] + cv * dP[a] * Qv[b] + cuv * dQ[a] * Qv[b];
                fv += c0 * P[a] * dPv[b] + cu * Q[a] * dPv[b] + cv * P[a] * dQv[b] + cuv * Q[a] * dQv[b];
            }
        z = f;
        gx = fu / cell;
        gy = fv / cell;
    }
    // The particle's splat: nine (pixel, value) pairs of a quadratic B-spline about the nearest pixel
    // centre (pixel −1 where off the grid); false when it does not reach the wall.
    bool splat(const double* h, const Part& p, int pixOut[9], double valOut[9]) const {
        double z, ax, ay;
        surface(h, p, z, ax, ay);
        const double inv = 1 / std::sqrt(ax * ax + ay * ay + 1);
        const double nx = -ax * inv, ny = -ay * inv, nz = inv;
        const double eta = p.n1 / p.n2;
        const double k = 1 - eta * eta * (1 - nz * nz);
        if (!(k >= 0)) return false;
        const double m = eta * nz - std::sqrt(k);
        const double dx = -m * nx, dy = -m * ny, dz = eta - m * nz;
        const double T2 = 1 - fresnelDielectric(nz, p.n2 / p.n1);
        const double L = s.wall_z - (s.thickness + z);
        const double xw = p.x + dx / dz * L, yw = p.y + dy / dz * L;
        const double w = p.weight * T2;
        const double gu = (xw + s.wall_half) / pix - 0.5, gv = (yw + s.wall_half) / pix - 0.5;
        const double fi = std::floor(gu + 0.5), fj = std::floor(gv + 0.5);
        const double tu = gu - fi, tv = gv - fj;
        const double wu[3] = {0.5 * (0.5 - tu) * (0.5 - tu), 0.75 - tu * tu, 0.5 * (0.5 + tu) * (0.5 + tu)};
        const double wv[3] = {0.5 * (0.5 - tv) * (0.5 - tv), 0.75 - tv * tv, 0.5 * (0.5 + tv) * (0.5 + tv)};
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a) {
                const double pi = fi + a - 1, pj = fj + b - 1;
                const bool in = pi >= 0 && pj >= 0 && pi < s.pixels && pj < s.pixels;
                pixOut[3 * b + a] = in ? int(pj) * int(s.pixels) + int(pi) : -1;
                valOut[3 * b + a] = w * wu[a] * wv[b];
            }
        return true;
    }
};

}  // namespace

extern "C" {

// The irradiance of pass `pass` (pixels² doubles, row = y) through heights (nodes² doubles).
void oracle_caustic_forward(const owe_caustic_settings* s, const double* heights, uint64_t pass, double* E,
                            int threads) {
    const Slab slab(*s);
    if (threads <= 0) threads = int(std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
    const size_t np = size_t(s->pixels) * s->pixels;
    std::vector<std::vector<double>> part(size_t(threads), std::vector<double>(np, 0.0));
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            std::vector<double>& e = part[size_t(t)];
            const uint64_t n = s->particles;
            for (uint64_t i = n * uint64_t(t) / uint64_t(threads); i < n * uint64_t(t + 1) / uint64_t(threads); ++i) {
                int px[9];
                double v[9];
                if (!slab.splat(heights, slab.particle(i, pass), px, v)) continue;
                for (int q = 0; q < 9; ++q)
                    if (px[q] >= 0) e[size_t(px[q])] += v[q];
            }
        });
    for (auto& th : pool) th.join();
    for (size_t k = 0; k < np; ++k) {
        E[k] = 0;
        for (int t = 0; t < threads; ++t) E[k] += part[size_t(t)][k];
    }
}

// Central finite differences (step h, metres) of L = mean((E − target)²) with respect to the heights
// nodes[k] (flat indices), on the particles of pass `pass`: grad[k].
void oracle_caustic_fd(const owe_caustic_settings* s, const double* heights, const double* target, uint64_t pass,
                       const uint32_t* nodes, int n, double h, double* grad) {
    const Slab slab(*s);
    const size_t np = size_t(s->pixels) * s->pixels;
    const int C = int(s->nodes) - 1;
    std::vector<Part> parts(s->particles);
    std::vector<std::vector<uint32_t>> byCell(size_t(C) * size_t(C));
    std::vector<double> E(np, 0.0);
    for (uint64_t i = 0; i < s->particles; ++i) {
        parts[i] = slab.particle(i, pass);
        byCell[size_t(parts[i].cj) * size_t(C) + size_t(parts[i].ci)].push_back(uint32_t(i));
        int px[9];
        double v[9];
        if (!slab.splat(heights, parts[i], px, v)) continue;
        for (int q = 0; q < 9; ++q)
            if (px[q] >= 0) E[size_t(px[q])] += v[q];
    }
    std::vector<double> hp(heights, heights + size_t(s->nodes) * s->nodes);
    std::vector<double> delta(np, 0.0);
    std::vector<char> mark(np, 0);
    std::vector<int> touched;
    for (int k = 0; k < n; ++k) {
        const int ni = int(nodes[k] % s->nodes), nj = int(nodes[k] / s->nodes);
        double L[2];
        for (int side = 0; side < 2; ++side) {
            const double sign = side == 0 ? 1 : -1;
            hp[nodes[k]] = heights[nodes[k]] + sign * h;
            touched.clear();
            for (int cj = std::max(0, nj - 2); cj <= std::min(C - 1, nj + 1); ++cj)
                for (int ci = std::max(0, ni - 2); ci <= std::min(C - 1, ni + 1); ++ci)
                    for (uint32_t i : byCell[size_t(cj) * size_t(C) + size_t(ci)]) {
                        int px[9];
                        double v[9];
                        for (int pass2 = 0; pass2 < 2; ++pass2) {
                            const double* hh = pass2 == 0 ? heights : hp.data();
                            if (!slab.splat(hh, parts[i], px, v)) continue;
                            for (int q = 0; q < 9; ++q) {
                                if (px[q] < 0) continue;
                                if (!mark[size_t(px[q])]) {
                                    mark[size_t(px[q])] = 1;
                                    touched.push_back(px[q]);
                                }
                                delta[size_t(px[q])] += pass2 == 0 ? -v[q] : v[q];
                            }
                        }
                    }
            double dL = 0;
            for (int p : touched) {
                const double a = E[size_t(p)] - target[size_t(p)], b = a + delta[size_t(p)];
                dL += b * b - a * a;
                delta[size_t(p)] = 0;
                mark[size_t(p)] = 0;
            }
            L[side] = dL / double(np);
        }
        hp[nodes[k]] = heights[nodes[k]];
        grad[k] = (L[0] - L[1]) / (2 * h);
    }
}

}  // extern "C"
