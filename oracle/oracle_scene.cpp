// TEST INFRASTRUCTURE ONLY — the double-precision CPU oracle (see oracle_scene.hpp for citations).
#include "oracle_scene.hpp"

#include <map>
#include <numeric>
#include <stdexcept>
#include <tuple>

namespace oracle {

// ---------------------------------------------------------------- spectra and media

double Spectrum::eval(double l) const {
    switch (kind) {
        case OWE_SPEC_CONSTANT: return scale * value;
        case OWE_SPEC_BLACKBODY: return scale * planck(l, kelvin);
        case OWE_SPEC_TABULATED: {
            if (lam.empty()) return 0;
            if (l <= lam.front()) return scale * val.front();
            if (l >= lam.back()) return scale * val.back();
            const size_t i = size_t(std::upper_bound(lam.begin(), lam.end(), l) - lam.begin());
            const double t = (l - lam[i - 1]) / (lam[i] - lam[i - 1]);
            return scale * (val[i - 1] * (1 - t) + val[i] * t);
        }
    }
    return 0;
}

double Spectrum::luminance() const {
    return integrateVisible([&](double l) { return eval(l) * cieY(l); }) / cieYIntegral();
}

Wavelengths Wavelengths::hero(double u) {
    Wavelengths w;
    for (int i = 0; i < NW; ++i) {
        double up = u + double(i) / NW;
        if (up >= 1) up -= 1;
        w.lambda[i] = sampleVisible(up);
        w.pdf[i] = visiblePdf(w.lambda[i]);
    }
    return w;
}

Wavelengths Wavelengths::uniform1(double u, double lo, double hi) {
    Wavelengths w;
    const double l = lo + (hi - lo) * u;
    for (int i = 0; i < NW; ++i) w.lambda[i] = l;
    w.pdf[0] = 1.0 / (hi - lo) / NW;  // one wavelength standing for the four strata
    return w;
}

void Wavelengths::toXYZ(const Spec4& L, double xyz[3]) const {
    xyz[0] = xyz[1] = xyz[2] = 0;
    const double inv = 1.0 / (NW * cieYIntegral());
    for (int i = 0; i < NW; ++i) {
        if (pdf[i] == 0 || L[i] == 0) continue;
        const double w = L[i] / pdf[i] * inv;
        xyz[0] += w * cieX(lambda[i]);
        xyz[1] += w * cieY(lambda[i]);
        xyz[2] += w * cieZ(lambda[i]);
    }
}

double IndexModel::n(double lnm) const {
    const double l = lnm * 1e-3, l2 = l * l;
    double v = 1;
    switch (kind) {
        case OWE_INDEX_CONSTANT: v = a[0]; break;
        case OWE_INDEX_CAUCHY: v = a[0] + a[1] / l2 + a[2] / (l2 * l2); break;
        case OWE_INDEX_SELLMEIER: {
            double s = 1;
            for (int i = 0; i < terms; ++i) s += a[i] * l2 / (l2 - b[i]);
            v = std::sqrt(s);
            break;
        }
        case OWE_INDEX_CIDDOR: v = ciddorAir(l2); break;
    }
    return relAir ? v * ciddorAir(sq(lnm * 1e-3)) : v;
}

Xf Xf::inverse() const {
    Xf r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.R[i][j] = R[j][i];
    r.t = -(r.vector(t));
    return r;
}

double Boundary::area() const {
    switch (shape) {
        case OWE_SHAPE_PLANE:
            if (aperture == OWE_APERTURE_DISK) return kPi * (a * a - b * b);
            if (aperture == OWE_APERTURE_RECT) return 4 * a * b;
            return kPi * a * b;
        case OWE_SHAPE_SPHERE: return 4 * kPi * a * a;
        case OWE_SHAPE_MESH: return mesh ? mesh->area() : 0;
    }
    return 0;
}

namespace {

Spectrum makeSpectrum(const owe_spectrum& s, const double* table) {
    Spectrum r;
    r.kind = s.kind;
    r.value = s.value;
    r.kelvin = s.kelvin;
    r.scale = s.scale;
    if (s.kind == OWE_SPEC_TABULATED)
        for (int i = 0; i < s.count; ++i) {
            r.lam.push_back(table[s.offset + 2 * i]);
            r.val.push_back(table[s.offset + 2 * i + 1]);
        }
    return r;
}

bool inAperture(int kind, double a, double b, V3 p) {
    if (kind == OWE_APERTURE_DISK) {
        const double r2 = p.x * p.x + p.y * p.y;
        return r2 <= a * a && r2 >= b * b;
    }
    if (kind == OWE_APERTURE_RECT) return std::fabs(p.x) <= a && std::fabs(p.y) <= b;
    return sq(p.x / a) + sq(p.y / b) <= 1;
}

// Geometric-form sphere roots (surfaces.hpp:98-117), ascending.
bool sphereRoots(V3 o, V3 d, double R, double& t0, double& t1) {
    const double b = dot(o, d);
    const V3 perp = o - d * b;
    const double disc = R * R - dot(perp, perp);
    t0 = t1 = 0;
    if (disc < 0) return false;
    const double c = dot(o, o) - R * R;
    const double q = -b - std::copysign(std::sqrt(disc), b);
    t0 = q;
    t1 = q != 0 ? c / q : q;
    if (t0 > t1) std::swap(t0, t1);
    return true;
}

bool areaToSolidAngle(V3 ref, V3 p, V3 n, double area, double& pdfW) {
    pdfW = 0;
    const V3 w = p - ref;
    const double d2 = dot(w, w);
    if (!(d2 > 0)) return false;
    const double cosL = std::fabs(dot(n, w)) / std::sqrt(d2);
    if (cosL <= kGrazing) return false;
    pdfW = d2 / (cosL * area);
    return true;
}

double sphereConeSize(V3 ref, double R) {
    const double dc2 = dot(ref, ref);
    if (dc2 <= R * R * (1 + kRelTol)) return 0;
    const double s2 = R * R / dc2;
    return s2 / (1 + sqrt0(1 - s2));
}

V3 sampleSphereCone(V3 ref, double omc, double u1, double u2) {
    const V3 axis = -ref / std::sqrt(dot(ref, ref));
    V3 s, t;
    basis(axis, s, t);
    const double omcT = u1 * omc, cosT = 1 - omcT, sinT = sqrt0(omcT * (2 - omcT)), phi = 2 * kPi * u2;
    return s * (sinT * std::cos(phi)) + t * (sinT * std::sin(phi)) + axis * cosT;
}

// Coplanar adjacent triangles (MeshShape::sameFace, mesh.cpp:50-71) joined into face groups.
std::vector<uint32_t> faceGroups(const Mesh& m) {
    const auto& tris = m.triangles();
    const auto& pos = m.positions();
    std::vector<uint32_t> parent(tris.size());
    std::iota(parent.begin(), parent.end(), 0u);
    auto find = [&](uint32_t x) {
        while (parent[x] != x) x = parent[x] = parent[parent[x]];
        return x;
    };
    std::map<std::tuple<double, double, double>, uint32_t> weld;
    auto key = [&](uint32_t v) {
        auto k = std::make_tuple(pos[v].x, pos[v].y, pos[v].z);
        return weld.emplace(k, uint32_t(weld.size())).first->second;
    };
    std::map<std::pair<uint32_t, uint32_t>, std::vector<uint32_t>> edges;
    for (uint32_t i = 0; i < tris.size(); ++i)
        for (int e = 0; e < 3; ++e) {
            uint32_t a = key(tris[i][e]), b = key(tris[i][(e + 1) % 3]);
            if (a > b) std::swap(a, b);
            edges[{a, b}].push_back(i);
        }
    for (auto& [k, list] : edges)
        for (size_t x = 0; x < list.size(); ++x)
            for (size_t y = x + 1; y < list.size(); ++y)
                if (m.sameFace(list[x], list[y])) parent[find(list[x])] = find(list[y]);
    std::vector<uint32_t> g(tris.size());
    for (uint32_t i = 0; i < tris.size(); ++i) g[i] = find(i);
    return g;
}

// WaveSurface's constructor (geometry.cpp:188-225): amplitude and slope bounds, and with a node
// spacing (given, or automatic beyond six waves) the sampled bicubic Hermite nodes and the slope
// bound of the spline.
shapes::Wave makeWave(const owe_boundary& ob, const owe_plane_wave* records) {
    shapes::Wave w;
    w.hx = ob.a;
    w.hy = ob.b;
    w.margin = std::max(0.0, std::min({ob.params[0], w.hx, w.hy}));
    for (uint32_t i = 0; i < ob.wave_count; ++i) {
        const owe_plane_wave& r = records[ob.first_wave + i];
        w.waves.push_back({r.amplitude, r.kx, r.ky, r.phase});
    }
    double slope = 0, curvature = 0, kMax = 0;
    for (const auto& p : w.waves) {
        const double k = std::hypot(p.kx, p.ky);
        w.amp += std::fabs(p.amplitude);
        slope += std::fabs(p.amplitude) * k;
        curvature += std::fabs(p.amplitude) * k * k;
        kMax = std::max(kMax, k);
    }
    const double taper = w.margin > 0 ? 1.5 * std::sqrt(2.0) / w.margin : 0;
    w.slope = slope + w.amp * taper;
    double spacing = ob.params[1];
    if (spacing < 0) spacing = w.waves.size() <= 6 || kMax == 0 ? 0 : (2 * kPi / kMax) / 12;
    if (!(spacing > 0)) return w;
    w.nx = std::clamp(int(std::ceil(2 * w.hx / spacing)) + 1, 2, 4096);
    w.ny = std::clamp(int(std::ceil(2 * w.hy / spacing)) + 1, 2, 4096);
    w.dx = 2 * w.hx / (w.nx - 1);
    w.dy = 2 * w.hy / (w.ny - 1);
    w.nodes.resize(4 * size_t(w.nx) * w.ny);
    double nodeSlope = 0;
    for (int j = 0; j < w.ny; ++j)
        for (int i = 0; i < w.nx; ++i) {
            double* n = &w.nodes[4 * (size_t(j) * w.nx + i)];
            const double x = -w.hx + i * w.dx, y = -w.hy + j * w.dy;
            double g = 0, gx = 0, gy = 0, gxy = 0;
            for (const auto& p : w.waves) {
                const double arg = p.kx * x + p.ky * y + p.phase, sn = std::sin(arg), cs = std::cos(arg);
                g += p.amplitude * sn;
                gx += p.amplitude * cs * p.kx;
                gy += p.amplitude * cs * p.ky;
                gxy -= p.amplitude * sn * p.kx * p.ky;
            }
            double wx, dwx, wy, dwy;
            shapes::waveTaper(x, w.hx, w.margin, wx, dwx);
            shapes::waveTaper(y, w.hy, w.margin, wy, dwy);
            if (std::fabs(x) > w.hx || std::fabs(y) > w.hy) wx = wy = dwx = dwy = 0;
            n[0] = wx * wy * g;
            n[1] = dwx * wy * g + wx * wy * gx;
            n[2] = wx * dwy * g + wx * wy * gy;
            n[3] = dwx * dwy * g + dwx * wy * gy + wx * dwy * gx + wx * wy * gxy;
            nodeSlope = std::max(nodeSlope, std::hypot(n[1], n[2]));
        }
    const double hessian = curvature + 2 * taper * slope + (w.margin > 0 ? 6 * w.amp / (w.margin * w.margin) : 0);
    w.slope = std::min(w.slope, 1.05 * (nodeSlope + hessian * 0.5 * std::hypot(w.dx, w.dy)));
    return w;
}

}  // namespace

// ---------------------------------------------------------------- world

World::World(const owe_scene& s) {
    for (uint32_t i = 0; i < s.n_media; ++i) {
        const owe_medium& m = s.media[i];
        Medium r;
        r.index.kind = m.index_kind;
        r.index.terms = m.terms;
        r.index.relAir = m.rel_air != 0;
        for (int k = 0; k < 4; ++k) {
            r.index.a[k] = m.coeff_a[k];
            r.index.b[k] = m.coeff_b[k];
        }
        r.absorption = makeSpectrum(m.absorption, s.table);
        r.scattering = makeSpectrum(m.scattering, s.table);
        r.g = m.g;
        r.opaque = m.opaque != 0;
        media.push_back(r);
    }
    regionMedium.assign(s.region_medium, s.region_medium + s.n_regions);
    for (uint32_t i = 0; i < s.n_optics; ++i) {
        const owe_optics& o = s.optics[i];
        Optics r;
        r.type = o.type;
        r.backAbsorbs = o.back_absorbs != 0;
        r.roughness = o.roughness;
        r.reflectance = makeSpectrum(o.reflectance, s.table);
        r.condN = makeSpectrum(o.conductor_n, s.table);
        r.condK = makeSpectrum(o.conductor_k, s.table);
        optics.push_back(r);
    }
    for (uint32_t i = 0; i < s.n_emissions; ++i) {
        Emission e;
        e.radiance = makeSpectrum(s.emissions[i].radiance, s.table);
        e.front = s.emissions[i].front != 0;
        e.back = s.emissions[i].back != 0;
        e.nee = s.emissions[i].nee != 0;
        emissions.push_back(e);
    }
    for (uint32_t i = 0; i < s.n_boundaries; ++i) {
        const owe_boundary& ob = s.boundaries[i];
        Boundary b;
        b.shape = ob.shape;
        b.aperture = ob.aperture;
        b.front = ob.front;
        b.back = ob.back;
        b.optics = ob.optics;
        b.emission = ob.emission;
        b.a = ob.a;
        b.b = ob.b;
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) b.toWorld.R[r][c] = ob.to_world[4 * r + c];
            b.toWorld.t[r] = ob.to_world[4 * r + 3];
        }
        b.toLocal = b.toWorld.inverse();
        Box local;
        if (b.shape == OWE_SHAPE_MESH) {
            std::vector<V3> pos;
            std::vector<std::array<uint32_t, 3>> tris;
            std::map<uint32_t, uint32_t> remap;
            for (uint32_t t = ob.first_triangle; t < ob.first_triangle + ob.triangle_count; ++t) {
                std::array<uint32_t, 3> tri{};
                for (int k = 0; k < 3; ++k) {
                    const uint32_t v = s.triangles[3 * size_t(t) + k];
                    auto it = remap.find(v);
                    if (it == remap.end()) {
                        it = remap.emplace(v, uint32_t(pos.size())).first;
                        pos.push_back({s.vertices[3 * size_t(v)], s.vertices[3 * size_t(v) + 1], s.vertices[3 * size_t(v) + 2]});
                    }
                    tri[k] = it->second;
                }
                tris.push_back(tri);
            }
            b.mesh = std::make_unique<Mesh>(std::move(pos), std::move(tris));
            b.faceGroup = faceGroups(*b.mesh);
            local = b.mesh->bounds();
        } else if (b.shape == OWE_SHAPE_PLANE) {
            const double ex = b.a, ey = b.aperture == OWE_APERTURE_DISK ? b.a : b.b;
            const double pad = 1e-9 * (1 + std::max(ex, ey));
            local.grow(V3(-ex, -ey, -pad));
            local.grow(V3(ex, ey, pad));
        } else if (b.shape == OWE_SHAPE_CYLINDER) {
            b.z0 = std::min(ob.params[0], ob.params[1]);
            b.z1 = std::max(ob.params[0], ob.params[1]);
            local.grow(V3(-b.a, -b.a, b.z0));
            local.grow(V3(b.a, b.a, b.z1));
        } else if (b.shape == OWE_SHAPE_SAG) {
            b.sag = shapes::makeSag(ob.params[0], ob.params[1], ob.params + 2, b.a, b.b);
            local.grow(V3(-b.a, -b.a, b.sag.zMin));
            local.grow(V3(b.a, b.a, b.sag.zMax));
        } else if (b.shape == OWE_SHAPE_WAVE) {
            b.wave = std::make_shared<shapes::Wave>(makeWave(ob, s.waves));
            local.grow(V3(-b.a, -b.b, -b.wave->amp));
            local.grow(V3(b.a, b.b, b.wave->amp));
        } else {
            local.grow(V3(-b.a, -b.a, -b.a));
            local.grow(V3(b.a, b.a, b.a));
        }
        for (int c = 0; c < 8; ++c)
            b.worldBox.grow(b.toWorld.point(V3((c & 1) ? local.hi.x : local.lo.x, (c & 2) ? local.hi.y : local.lo.y,
                                               (c & 4) ? local.hi.z : local.lo.z)));
        bounds.grow(b.worldBox);
        boundaries.push_back(std::move(b));
    }
    env.hasSun = s.env.has_sun != 0;
    env.sky = s.env.sky;
    env.sunDir = normalize(V3(s.env.sun_dir[0], s.env.sun_dir[1], s.env.sun_dir[2]));
    env.sunAngularRadius = s.env.sun_angular_radius;
    env.sunRadiance = makeSpectrum(s.env.sun_radiance, s.table);
    env.skyRadiance = makeSpectrum(s.env.sky_radiance, s.table);
    env.sunNeeShare = s.env.sun_nee_share;

    // Light table (World::buildLightTable, world.cpp:285-350): boundaries by power, then the sun.
    const double radius = sceneRadius();
    lightOfBoundary.assign(boundaries.size(), -1);
    for (size_t i = 0; i < boundaries.size(); ++i) {
        const Boundary& b = boundaries[i];
        // Next-event estimation samples planes and spheres; an emitting mesh is met by BSDF sampling
        // only (no entry: its MIS weight is 1), on the CPU and the GPU alike.
        if (b.emission < 0 || !b.canSample() || b.shape == OWE_SHAPE_MESH || !emissions[size_t(b.emission)].nee) continue;
        const Emission& e = emissions[size_t(b.emission)];
        const double sides = (e.front ? 1 : 0) + (e.back ? 1 : 0);
        lightOfBoundary[i] = int(lights.size());
        lights.push_back({0, uint32_t(i), std::max(1e-12, e.radiance.luminance()) * b.area() * kPi * sides});
    }
    if (env.hasSun)
        lights.push_back({1, kNone, std::max(1e-12, env.sunRadiance.luminance()) * env.sunSolidAngle() * kPi * radius * radius});
    double total = 0, boundaryPower = 0;
    for (const Light& l : lights) {
        total += l.power;
        if (l.kind == 0) boundaryPower += l.power;
    }
    const double share = env.hasSun && boundaryPower > 0 ? clampd(env.sunNeeShare, 0, 0.999) : 0;
    double acc = 0;
    for (const Light& l : lights) {
        double p = total > 0 ? l.power / total : 0;
        if (share > 0) p = l.kind == 1 ? share : (1 - share) * l.power / boundaryPower;
        lightPdf.push_back(p);
        acc += p;
        lightCdf.push_back(acc);
    }
    if (!lightCdf.empty()) lightCdf.back() = 1.0;
    totalLightPower = total;
    // computeEnvironmentPower (world.cpp:352-371): a uniform sky's radiance averaged over 400..700 nm
    // in 50 nm steps (the same over every direction), times π R² · 4π.
    if (env.sky == OWE_SKY_UNIFORM) {
        double y = 0;
        for (double l = 400; l <= 700; l += 50) y += env.skyRadiance.eval(l);
        environmentPower = std::max(0.0, y / 7) * kPi * radius * radius * 4 * kPi;
    }
    const double emitted = environmentPower + (lights.empty() ? 0.0 : totalLightPower);
    skyShare = emitted > 0 ? environmentPower / emitted : 0;

    const owe_observer& ob = s.observer;
    observer.kind = ob.kind;
    observer.width = ob.width;
    observer.height = ob.height;
    observer.sigma = ob.pixel_sigma;
    if (ob.kind == OWE_DETECTOR_SENSOR) {
        // SurfaceSensor::prepare (detector.cpp:100-108).
        const Boundary& b = boundaries.at(ob.sensor_boundary);
        observer.boundary = ob.sensor_boundary;
        observer.toWorld = b.toWorld;
        observer.normal = normalize(b.toWorld.vector(V3(0, 0, 1)));
        observer.region = b.front;
        observer.halfX = ob.half_x;
        observer.halfY = ob.half_y;
        observer.flipX = ob.flip_x != 0;
        observer.flipY = ob.flip_y != 0;
        observer.hasAim = ob.has_aim != 0;
        observer.aimCenter = V3(ob.aim_center[0], ob.aim_center[1], ob.aim_center[2]);
        observer.aimNormal = V3(ob.aim_normal[0], ob.aim_normal[1], ob.aim_normal[2]);
        observer.aimRadius = ob.aim_radius;
        observer.aimShare = ob.aim_share;
        rebase = b.toWorld.point(V3(0, 0, 0));
        findContacts();
        return;
    }
    // Ideal observer (IdealObserver::prepare, detector.cpp:36-45).
    observer.region = ob.region;
    observer.position = V3(ob.position[0], ob.position[1], ob.position[2]);
    rebase = observer.position;
    findContacts();
    observer.fwd = normalize(V3(ob.look_at[0], ob.look_at[1], ob.look_at[2]) - observer.position);
    V3 u(ob.up[0], ob.up[1], ob.up[2]);
    if (std::fabs(dot(normalize(u), observer.fwd)) > 0.9999)
        u = std::fabs(observer.fwd.z) < 0.9 ? V3(0, 0, 1) : V3(0, 1, 0);
    observer.right = normalize(cross(observer.fwd, u));
    observer.up = cross(observer.right, observer.fwd);
    observer.tanY = std::tan(0.5 * ob.fov_y);
    observer.tanX = observer.tanY * double(ob.width) / double(ob.height);
    observer.pupil = ob.pupil_radius;
    observer.focus = ob.focus_distance;
}

double World::sceneRadius() const {
    if (!bounds.valid()) return 1.0;
    return std::max(1e-3, 0.5 * length(bounds.hi - bounds.lo));
}

size_t World::pickLight(double u) const {
    const size_t i = size_t(std::upper_bound(lightCdf.begin(), lightCdf.end(), u) - lightCdf.begin());
    return std::min(i, lights.size() - 1);
}

// The root of a quadric other than the one a departing ray stands on (C ≈ 0 at its origin):
// −B/A, polished from the approximate point.
double otherRoot(const shapes::Quadric& q, V3 o, V3 d) {
    double A, B, C;
    shapes::quadricCoeffs(q, o, d, A, B, C);
    if (std::fabs(A) < 1e-300) return -1;
    return shapes::refineRoot(q, o, d, -B / A);
}

void World::findContacts() {
    contacts = false;
    for (size_t i = 0; i < boundaries.size() && !contacts; ++i)
        for (size_t j = i + 1; j < boundaries.size() && !contacts; ++j) {
            const Box& a = boundaries[i].worldBox;
            const Box& c = boundaries[j].worldBox;
            const V3 al = a.lo - rebase, ah = a.hi - rebase, cl = c.lo - rebase, ch = c.hi - rebase;
            double m = 0;
            for (const V3& v : {al, ah, cl, ch}) m = std::max({m, std::fabs(v.x), std::fabs(v.y), std::fabs(v.z)});
            const double tol = std::ldexp(m, -21);
            contacts = al.x - tol <= ch.x && cl.x - tol <= ah.x && al.y - tol <= ch.y && cl.y - tol <= ah.y &&
                       al.z - tol <= ch.z && cl.z - tol <= ah.z;
        }
}

double World::contactTol(V3 p) const {
    const V3 q = p - rebase;
    return std::ldexp(std::max({std::fabs(q.x), std::fabs(q.y), std::fabs(q.z)}), -21);
}

bool World::intersectRaw(const Ray& ray, double lo, double hi, Hit& hit, uint32_t skipB, uint32_t skipPrim) const {
    bool found = false;
    double best = hi;
    const Slabs rs(ray.o, ray.d);
    for (uint32_t i = 0; i < boundaries.size(); ++i) {
        const Boundary& b = boundaries[i];
        if (i == skipB && b.shape != OWE_SHAPE_MESH) continue;
        double tEnter;
        if (!rs.hit(b.worldBox, std::min(lo, 0.0), best, tEnter)) continue;
        const bool departed = i == ray.departed;
        // Other boundaries count from lo (below zero at a departure: a contact at the origin); the
        // departed one by its own rules, ahead of the origin.
        const double tmin = departed ? std::max(lo, 0.0) : lo;
        const V3 lo3 = b.toLocal.point(ray.o), ld = b.toLocal.vector(ray.d);
        double t;
        V3 pl, nl;
        uint32_t prim = 0;
        if (b.shape == OWE_SHAPE_PLANE) {
            if (departed || ld.z == 0) continue;
            t = -lo3.z / ld.z;
            if (!(t > tmin && t < best)) continue;
            pl = lo3 + ld * t;
            pl.z = 0;
            if (!inAperture(b.aperture, b.a, b.b, pl)) continue;
            nl = V3(0, 0, 1);
        } else if (b.shape == OWE_SHAPE_SPHERE) {
            if (departed) {
                t = -2 * dot(lo3, ld) / dot(ld, ld);  // the root other than the origin's own
                if (!(t > tmin && t < best)) continue;
            } else {
                double t0, t1;
                if (!sphereRoots(lo3, ld, b.a, t0, t1)) continue;
                t = t0 > tmin ? t0 : t1;
                if (!(t > tmin && t < best)) continue;
            }
            pl = lo3 + ld * t;
            nl = normalize(pl / b.a);
        } else if (b.shape == OWE_SHAPE_CYLINDER) {
            if (departed) {
                t = otherRoot(shapes::Quadric{1, b.a * b.a, 0}, lo3, ld);
                if (!(t > tmin && t < best)) continue;
                pl = lo3 + ld * t;
                if (!(pl.z >= b.z0 && pl.z <= b.z1)) continue;
                nl = normalize(V3(pl.x, pl.y, 0));
            } else if (!shapes::intersectCylinder(b.a, b.z0, b.z1, lo3, ld, tmin, best, t, pl, nl)) {
                continue;
            }
        } else if (b.shape == OWE_SHAPE_SAG) {
            const shapes::Sag& sg = b.sag;
            if (departed && sg.nA == 0) {
                if (sg.c == 0) continue;  // a plane: nothing beyond its own point
                t = otherRoot(shapes::Quadric{0, sg.c, sg.c * (1 + sg.k)}, lo3, ld);
                if (!(t > tmin && t < best)) continue;
                pl = lo3 + ld * t;
                if (!shapes::sagAccept(sg, pl) || !shapes::onConicSheet(sg, pl.z)) continue;
                nl = shapes::sagNormal(sg, pl);
            } else {
                // A numeric sag left by this ray: roots within 1e-9 of its scale are its own point.
                const double own = departed ? 1e-9 * (sg.rMax + std::max(std::fabs(sg.zMin), std::fabs(sg.zMax))) : tmin;
                if (!shapes::intersectSag(sg, lo3, ld, std::max(own, tmin), best, t, pl, nl)) continue;
            }
        } else if (b.shape == OWE_SHAPE_WAVE) {
            if (!shapes::intersectWave(*b.wave, lo3, ld, tmin, best, departed, t, pl, nl)) continue;
        } else {
            const uint32_t skip = departed && ray.departedPrim < b.faceGroup.size() ? ray.departedPrim : ~0u;
            const uint32_t skip2 = i == skipB && skipPrim < b.faceGroup.size() ? skipPrim : ~0u;
            if (!b.mesh->closest(lo3, ld, tmin, best, t, prim, skip, nullptr, nullptr,
                                 b.faceGroup.empty() ? nullptr : b.faceGroup.data(), skip2))
                continue;
            pl = lo3 + ld * t;
            nl = b.mesh->normalOf(prim);
        }
        best = t;
        found = true;
        hit.t = t;
        hit.boundary = i;
        hit.prim = prim;
        hit.pLocal = pl;
        hit.nLocal = nl;
    }
    if (!found) return false;
    const Boundary& b = boundaries[hit.boundary];
    hit.p = b.toWorld.point(hit.pLocal);
    hit.n = normalize(b.toWorld.vector(hit.nLocal));
    return true;
}

bool World::intersect(const Ray& ray, double tmax, Hit& hit, uint32_t region) const {
    const double dep = contacts && ray.departed != kNone ? contactTol(ray.o) : 0.0;
    if (!intersectRaw(ray, -dep, tmax, hit, kNone, kNone)) return false;
    if (region == kNone || !contacts) return true;
    auto incident = [&](const Hit& h) {
        const Boundary& b = boundaries[h.boundary];
        return dot(ray.d, h.n) < 0 ? b.front : b.back;
    };
    auto detector = [&](const Hit& h) { return optics[size_t(boundaries[h.boundary].optics)].type == OWE_SURF_DETECTOR; };
    // Two rounds at most: a crossing that does not fit the ray gives way to one that does at the same
    // place — anywhere ahead when the misfit is a contact at the ray's own origin.
    for (int round = 0; round < 2; ++round) {
        const bool fits = incident(hit) == region;
        if (fits && !detector(hit)) return true;
        const bool atOrigin = hit.t <= dep;
        const double tol = contactTol(hit.p) + std::ldexp(hit.t, -21);
        const double lo = atOrigin ? -dep : hit.t - tol, hi = atOrigin ? tmax : std::min(tmax, hit.t + tol);
        Hit alt;
        if (!intersectRaw(ray, lo, hi, alt, hit.boundary, hit.prim)) return true;
        if (incident(alt) != region || (detector(alt) && fits)) return true;
        hit = alt;
    }
    return true;
}

bool World::sampleFrom(const Boundary& b, V3 ref, double u1, double u2, double u3, V3& p, V3& n, double& pdfW) const {
    if (b.shape == OWE_SHAPE_SPHERE) {
        const double omc = sphereConeSize(ref, b.a);
        if (omc > 0) {
            const V3 w = sampleSphereCone(ref, omc, u1, u2);
            double t0, t1;
            if (sphereRoots(ref, w, b.a, t0, t1) && (t0 > 0 || t1 > 0)) p = ref + w * (t0 > 0 ? t0 : t1);
            else p = normalize(ref + w * dot(-ref, w)) * b.a;
            n = normalize(p);
            pdfW = 1 / (2 * kPi * omc);
            return true;
        }
    }
    uint32_t prim;
    if (!sampleArea(b, u1, u2, u3, p, n, prim)) return false;
    return areaToSolidAngle(ref, p, n, b.area(), pdfW);
}

bool World::sampleArea(const Boundary& b, double u1, double u2, double u3, V3& p, V3& n, uint32_t& prim) const {
    prim = 0;
    if (b.shape == OWE_SHAPE_SPHERE) {
        n = uniformSphere(u1, u2);
        p = n * b.a;
        return true;
    }
    if (!b.canSample()) return false;
    if (b.shape == OWE_SHAPE_PLANE) {
        n = V3(0, 0, 1);
        if (b.aperture == OWE_APERTURE_DISK) {
            const double r = std::sqrt(b.b * b.b + u1 * (b.a * b.a - b.b * b.b)), phi = 2 * kPi * u2;
            p = V3(r * std::cos(phi), r * std::sin(phi), 0);
        } else if (b.aperture == OWE_APERTURE_RECT) {
            p = V3((2 * u1 - 1) * b.a, (2 * u2 - 1) * b.b, 0);
        } else {
            const double r = std::sqrt(u1), phi = 2 * kPi * u2;
            p = V3(b.a * r * std::cos(phi), b.b * r * std::sin(phi), 0);
        }
    } else {
        const auto& cdf = b.mesh->areaCdf();
        size_t i = size_t(std::lower_bound(cdf.begin(), cdf.end(), u3 * b.mesh->area()) - cdf.begin());
        if (i >= cdf.size()) i = cdf.size() - 1;
        const auto& t = b.mesh->triangles()[i];
        const auto& P = b.mesh->positions();
        const double su = std::sqrt(u1), b0 = 1 - su, b1 = u2 * su;
        p = P[t[0]] * b0 + P[t[1]] * b1 + P[t[2]] * (1 - b0 - b1);
        n = b.mesh->normalOf(uint32_t(i));
        prim = uint32_t(i);
    }
    return true;
}

double World::pdfFrom(const Boundary& b, V3 ref, V3 p, V3 n) const {
    if (b.shape == OWE_SHAPE_SPHERE) {
        const double omc = sphereConeSize(ref, b.a);
        if (omc > 0) return 1 / (2 * kPi * omc);
    }
    double pdfW;
    return b.canSample() && areaToSolidAngle(ref, p, n, b.area(), pdfW) ? pdfW : 0;
}

// ---------------------------------------------------------------- observer

void Observer::filterOffset(Pcg& rng, double& fx, double& fy) const {
    fx = fy = 0;
    if (!(sigma > 0)) return;
    double u1, u2;
    rng.draw2(u1, u2);
    standardNormal2(u1, u2, fx, fy);
    fx *= sigma;
    fy *= sigma;
}

bool Observer::generate(double px, double py, Pcg& rng, Ray& ray, double& weight) const {
    if (sensor()) {
        // SurfaceSensor::generate (detector.cpp:110-149), without the focus disk.
        weight = 0;
        if (px < 0 || py < 0 || px >= width || py >= height) return false;
        if (flipX) px = width - px;
        if (flipY) py = height - py;
        const V3 p = toWorld.point(V3((2 * px / width - 1) * halfX, (1 - 2 * py / height) * halfY, 0));
        double u1, u2;
        V3 dir;
        if (hasAim) {
            const Frame aim(aimNormal);
            const bool useAim = aimShare >= 1 || rng.uniform() < aimShare;
            rng.draw2(u1, u2);
            double dx, dy;
            concentricDisk(u1, u2, dx, dy);
            const V3 w = aimCenter + (aim.s * dx + aim.t * dy) * aimRadius - p;
            dir = w / length(w);
            if (!useAim) dir = Frame(normal).toWorld(cosineHemisphere(u1, u2));
            const double cosS = dot(dir, normal);
            if (cosS <= 0) return false;
            const double pdfW = aimShare * diskAimPdf(p, dir, aimCenter, aimNormal, aimRadius, useAim) +
                                (1 - aimShare) * cosS * kInvPi;
            if (!(pdfW > 0)) return false;
            weight = cosS / pdfW;
        } else {
            rng.draw2(u1, u2);
            dir = Frame(normal).toWorld(cosineHemisphere(u1, u2));
            weight = kPi;
        }
        ray = Ray{p, dir, boundary, kNone};
        return true;
    }
    const double sx = (2 * px / width - 1) * tanX, sy = (1 - 2 * py / height) * tanY;
    const V3 dirPinhole = normalize(right * sx + up * sy + fwd);
    V3 q = position;
    weight = 1;
    if (pupil > 0) {
        double u1, u2, dx, dy;
        rng.draw2(u1, u2);
        concentricDisk(u1, u2, dx, dy);
        q = position + (right * dx + up * dy) * pupil;
    }
    V3 dir = dirPinhole;
    if (pupil > 0 && std::isfinite(focus)) dir = normalize(position + (right * sx + up * sy + fwd) * focus - q);
    ray = Ray{q, dir, kNone, kNone};
    return true;
}

}  // namespace oracle
