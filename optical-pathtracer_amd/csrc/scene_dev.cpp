// Host flattening of an owe_scene into the device layout (scene_dev.hpp): validation, the
// camera-relative rebase in double, the light table and the face groups, then float32.
//
// Light table: World::buildLightTable (src/owe/scene/world.cpp:285-350) — emitting planes and spheres
// by power max(1e-12, Y)·area·π·sides, then the sun with Y·Ω·π·R² (R the scene radius), the sun's
// share fixed by sun_nee_share when boundaries also emit. An emitting mesh is not in the table (it is
// met by BSDF sampling with MIS weight 1), matching the oracle.
// Face groups: adjacent, exactly coplanar, equally oriented triangles of one mesh
// (MeshShape::sameFace, src/owe/scene/mesh.cpp:50-71), joined transitively.
#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <numeric>
#include <string>
#include <tuple>

#include "common.hpp"
#include "scene_dev.hpp"

namespace owehip {

namespace {

constexpr double kPi = 3.14159265358979323846;

struct V {
    double x, y, z;
};
V sub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V crs(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double dt(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V nrm(V a) {
    const double l = std::sqrt(dt(a, a));
    return {a.x / l, a.y / l, a.z / l};
}

double cieY(double l) {
    auto lobe = [](double x, double mu, double s1, double s2) {
        const double t = (x - mu) / (x < mu ? s1 : s2);
        return std::exp(-0.5 * t * t);
    };
    return 0.821 * lobe(l, 568.8, 46.9, 40.5) + 0.286 * lobe(l, 530.9, 16.3, 31.1);
}
double planck(double lnm, double kelvin) {
    const double c1 = 1.191042972e8, c2 = 1.438776877e4, l = lnm * 1e-3, l2 = l * l;
    return 1e6 * c1 / (l2 * l2 * l * (std::exp(c2 / (l * kelvin)) - 1));
}
double evalHost(const owe_spectrum& s, const double* table, double l) {
    switch (s.kind) {
        case OWE_SPEC_CONSTANT: return s.scale * s.value;
        case OWE_SPEC_BLACKBODY: return s.scale * planck(l, s.kelvin);
        default: {
            const double* t = table + s.offset;
            if (l <= t[0]) return s.scale * t[1];
            if (l >= t[2 * (s.count - 1)]) return s.scale * t[2 * (s.count - 1) + 1];
            int i = 1;
            while (t[2 * i] <= l) ++i;
            const double u = (l - t[2 * (i - 1)]) / (t[2 * i] - t[2 * (i - 1)]);
            return s.scale * (t[2 * (i - 1) + 1] * (1 - u) + t[2 * i + 1] * u);
        }
    }
}
template <class F>
double simpson(F f) {
    const int n = int((830.0 - 360.0) * 2);
    const double h = (830.0 - 360.0) / n;
    double s = f(360.0) + f(830.0);
    for (int i = 1; i < n; ++i) s += f(360.0 + i * h) * ((i & 1) ? 4 : 2);
    return s * h / 3;
}
double luminance(const owe_spectrum& s, const double* table) {
    static const double ybar = simpson(cieY);
    return simpson([&](double l) { return evalHost(s, table, l) * cieY(l); }) / ybar;
}
bool spectrumZero(const owe_spectrum& s) { return s.kind == OWE_SPEC_CONSTANT && (s.value == 0 || s.scale == 0); }

void checkSpectrum(const owe_spectrum& s, uint32_t nTable, const char* what) {
    if (s.kind < 0 || s.kind > OWE_SPEC_TABULATED) fail(OWE_HIP_ERR_INVALID, std::string(what) + ": unknown spectrum kind");
    if (s.kind == OWE_SPEC_TABULATED) {
        if (s.count < 1 || s.offset < 0 || uint64_t(s.offset) + 2 * uint64_t(s.count) > nTable)
            fail(OWE_HIP_ERR_INVALID, std::string(what) + ": tabulated spectrum outside the table");
    }
    if (s.kind == OWE_SPEC_BLACKBODY && !(s.kelvin > 0))
        fail(OWE_HIP_ERR_INVALID, std::string(what) + ": blackbody temperature must be positive");
}

DSpec toDev(const owe_spectrum& s, const double* table, std::vector<float>& devTable) {
    DSpec d{};
    d.kind = s.kind;
    d.value = float(s.value);
    d.kelvin = float(s.kelvin);
    d.scale = float(s.scale);
    if (s.kind == OWE_SPEC_TABULATED) {
        d.count = s.count;
        d.offset = int32_t(devTable.size());
        for (int i = 0; i < 2 * s.count; ++i) devTable.push_back(float(table[s.offset + i]));
    }
    return d;
}

struct Xf {
    double R[3][3], t[3];
    V point(V p) const {
        return {R[0][0] * p.x + R[0][1] * p.y + R[0][2] * p.z + t[0], R[1][0] * p.x + R[1][1] * p.y + R[1][2] * p.z + t[1],
                R[2][0] * p.x + R[2][1] * p.y + R[2][2] * p.z + t[2]};
    }
    Xf inverse() const {
        Xf r{};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) r.R[i][j] = R[j][i];
        for (int i = 0; i < 3; ++i) r.t[i] = -(r.R[i][0] * t[0] + r.R[i][1] * t[1] + r.R[i][2] * t[2]);
        return r;
    }
    void store(float* out) const {
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) out[4 * r + c] = float(R[r][c]);
            out[4 * r + 3] = float(t[r]);
        }
    }
};

float down(double v) {
    const float f = float(v);
    return double(f) > v ? std::nextafter(f, -INFINITY) : f;
}
float upf(double v) {
    const float f = float(v);
    return double(f) < v ? std::nextafter(f, INFINITY) : f;
}

}  // namespace

namespace {

// WaveSurface's construction (src/owe/scene/geometry.cpp:188-225): amplitude and slope bounds; with
// a node spacing (given, or automatic beyond six waves: a twelfth of the shortest wavelength) the
// bicubic Hermite nodes (h, ∂x, ∂y, ∂xy) of the tapered sum and the spline's slope bound. Records
// are appended to `data` as float quadruples — waves (amplitude, kx, ky, phase), or nodes with
// their derivatives pre-multiplied by the spacings.
DWave buildWave(const owe_boundary& b, const owe_plane_wave* waves, std::vector<float>& data) {
    DWave w{};
    const double hx = b.a, hy = b.b, margin = std::max(0.0, std::min({b.params[0], hx, hy}));
    double amp = 0, slope = 0, curvature = 0, kMax = 0;
    for (uint32_t i = 0; i < b.wave_count; ++i) {
        const owe_plane_wave& p = waves[b.first_wave + i];
        const double k = std::hypot(p.kx, p.ky);
        amp += std::fabs(p.amplitude);
        slope += std::fabs(p.amplitude) * k;
        curvature += std::fabs(p.amplitude) * k * k;
        kMax = std::max(kMax, k);
    }
    const double taper = margin > 0 ? 1.5 * std::sqrt(2.0) / margin : 0;
    double slopeBound = slope + amp * taper;
    double spacing = b.params[1];
    if (spacing < 0) spacing = b.wave_count <= 6 || kMax == 0 ? 0 : (2 * kPi / kMax) / 12;
    w.hx = float(hx);
    w.hy = float(hy);
    w.margin = float(margin);
    w.amp = float(amp);
    require(data.size() / 4 < 0xFFFFFFFFull, "wave data too large");
    w.data = uint32_t(data.size() / 4);
    if (!(spacing > 0)) {
        w.nx = 0;
        w.ny = b.wave_count;
        for (uint32_t i = 0; i < b.wave_count; ++i) {
            const owe_plane_wave& p = waves[b.first_wave + i];
            data.insert(data.end(), {float(p.amplitude), float(p.kx), float(p.ky), float(p.phase)});
        }
        w.slope = float(slopeBound);
        return w;
    }
    const int nx = std::clamp(int(std::ceil(2 * hx / spacing)) + 1, 2, 4096);
    const int ny = std::clamp(int(std::ceil(2 * hy / spacing)) + 1, 2, 4096);
    const double dx = 2 * hx / (nx - 1), dy = 2 * hy / (ny - 1);
    auto taperOf = [&](double u, double half, double& wv, double& dw) {
        wv = 1;
        dw = 0;
        if (!(margin > 0)) return;
        const double e = half - std::fabs(u);
        if (e >= margin) return;
        if (e <= 0) {
            wv = 0;
            return;
        }
        const double t = e / margin;
        wv = t * t * (3 - 2 * t);
        dw = -(u < 0 ? -1.0 : 1.0) * 6 * t * (1 - t) / margin;
    };
    double nodeSlope = 0;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const double x = -hx + i * dx, y = -hy + j * dy;
            double g = 0, gx = 0, gy = 0, gxy = 0;
            for (uint32_t q = 0; q < b.wave_count; ++q) {
                const owe_plane_wave& p = waves[b.first_wave + q];
                const double arg = p.kx * x + p.ky * y + p.phase, sn = std::sin(arg), cs = std::cos(arg);
                g += p.amplitude * sn;
                gx += p.amplitude * cs * p.kx;
                gy += p.amplitude * cs * p.ky;
                gxy -= p.amplitude * sn * p.kx * p.ky;
            }
            double wx, dwx, wy, dwy;
            taperOf(x, hx, wx, dwx);
            taperOf(y, hy, wy, dwy);
            if (std::fabs(x) > hx || std::fabs(y) > hy) wx = wy = dwx = dwy = 0;
            const double h = wx * wy * g, hX = dwx * wy * g + wx * wy * gx, hY = wx * dwy * g + wx * wy * gy;
            const double hXY = dwx * dwy * g + dwx * wy * gy + wx * dwy * gx + wx * wy * gxy;
            nodeSlope = std::max(nodeSlope, std::hypot(hX, hY));
            data.insert(data.end(), {float(h), float(hX * dx), float(hY * dy), float(hXY * dx * dy)});
        }
    const double hessian = curvature + 2 * taper * slope + (margin > 0 ? 6 * amp / (margin * margin) : 0);
    slopeBound = std::min(slopeBound, 1.05 * (nodeSlope + hessian * 0.5 * std::hypot(dx, dy)));
    w.nx = uint32_t(nx);
    w.ny = uint32_t(ny);
    w.dx = float(dx);
    w.dy = float(dy);
    w.slope = float(slopeBound);
    return w;
}

}  // namespace

DSag makeSag(double c, double k, const double asphere[4], double rMax, double rMin) {
    DSag d{};
    d.c = c;
    d.ck = c * (1 + k);
    d.cf = float(c);
    d.kf = float(k);
    d.ckf = float(d.ck);
    d.rMax = float(rMax);
    d.rMax2 = float(rMax * rMax);
    d.rMin2 = float(rMin * rMin);
    d.polish64 = 1;
    d.nA = 4;
    for (int q = 0; q < 4; ++q) d.A[q] = float(asphere[q]);
    while (d.nA > 0 && asphere[d.nA - 1] == 0) --d.nA;
    auto sag = [&](double r) {
        const double r2 = r * r;
        double z = c * r2 / (1 + std::sqrt(std::max(0.0, 1 - (1 + k) * c * c * r2)));
        double rp = r2 * r2;
        for (int q = 0; q < 4; ++q) {
            z += asphere[q] * rp;
            rp *= r2;
        }
        return z;
    };
    double zmin = INFINITY, zmax = -INFINITY;
    for (int q = 0; q <= 512; ++q) {
        const double v = sag(rMin + (rMax - rMin) * q / 512);
        zmin = std::min(zmin, v);
        zmax = std::max(zmax, v);
    }
    const double pad = 1e-9 + 1e-3 * (zmax - zmin);
    d.zMin = float(zmin - pad);
    d.zMax = float(zmax + pad);
    return d;
}

HostScene flattenScene(const owe_scene& s) {
    HostScene h;
    require(s.n_regions >= 1, "a scene needs at least the ambient region 0");
    require(s.n_media >= 1, "a scene needs at least one medium");
    require(s.media && s.region_medium && (s.optics || s.n_optics == 0) && (s.boundaries || s.n_boundaries == 0),
            "null scene arrays");
    require(s.n_table == 0 || s.table, "null spectrum table");
    require(s.n_triangles == 0 || (s.triangles && s.vertices), "null mesh arrays");
    // Detector and rebase origin: the eye, or the sensor's centre.
    const owe_observer& ob = s.observer;
    require(ob.width > 0 && ob.height > 0, "detector resolution must be positive");
    require(ob.kind == OWE_DETECTOR_OBSERVER || ob.kind == OWE_DETECTOR_SENSOR, "unknown detector kind");
    require(ob.pixel_sigma >= 0, "pixel filter width must be non-negative");
    if (ob.kind == OWE_DETECTOR_OBSERVER) {
        require(ob.region < s.n_regions, "observer region out of range");
        require(ob.fov_y > 0 && ob.fov_y < kPi, "observer field of view must lie in (0, pi)");
        for (int k = 0; k < 3; ++k) h.origin[k] = ob.position[k];
    } else {
        // SurfaceSensor::prepare (detector.cpp:100-108): a detector surface, or a diffuse screen.
        require(ob.sensor_boundary < s.n_boundaries, "sensor boundary out of range");
        const owe_boundary& sb = s.boundaries[ob.sensor_boundary];
        require(sb.shape == OWE_SHAPE_PLANE, "a sensor is a plane boundary");
        require(sb.optics >= 0 && uint32_t(sb.optics) < s.n_optics &&
                    (s.optics[sb.optics].type == OWE_SURF_DETECTOR || s.optics[sb.optics].type == OWE_SURF_DIFFUSE),
                "sensor boundary is not a detector surface");
        require(ob.half_x > 0 && ob.half_y > 0, "sensor half sizes must be positive");
        require(!ob.has_aim || (ob.aim_radius > 0 && ob.aim_share > 0 && ob.aim_share <= 1),
                "sensor aim needs a positive radius and a share in (0, 1]");
        for (int k = 0; k < 3; ++k) h.origin[k] = sb.to_world[4 * k + 3];
    }
    const V origin{h.origin[0], h.origin[1], h.origin[2]};

    for (uint32_t i = 0; i < s.n_media; ++i) {
        const owe_medium& m = s.media[i];
        require(m.index_kind >= 0 && m.index_kind <= OWE_INDEX_CIDDOR, "medium " + std::to_string(i) + ": unknown index model");
        require(m.index_kind != OWE_INDEX_SELLMEIER || (m.terms >= 1 && m.terms <= 4), "medium: Sellmeier needs 1-4 terms");
        checkSpectrum(m.absorption, s.n_table, "medium absorption");
        checkSpectrum(m.scattering, s.n_table, "medium scattering");
        DMedium d{};
        d.indexKind = m.index_kind;
        d.terms = m.terms;
        d.relAir = m.rel_air;
        d.flags = (spectrumZero(m.absorption) ? 0 : 1) | (spectrumZero(m.scattering) ? 0 : 2);
        for (int k = 0; k < 4; ++k) {
            d.a[k] = float(m.coeff_a[k]);
            d.b[k] = float(m.coeff_b[k]);
        }
        d.absorption = toDev(m.absorption, s.table, h.table);
        d.scattering = toDev(m.scattering, s.table, h.table);
        d.g = float(m.g);
        h.media.push_back(d);
    }
    for (uint32_t r = 0; r < s.n_regions; ++r) {
        require(s.region_medium[r] < s.n_media, "region " + std::to_string(r) + ": medium out of range");
        h.regionMedium.push_back(s.region_medium[r]);
    }
    for (uint32_t i = 0; i < s.n_optics; ++i) {
        const owe_optics& o = s.optics[i];
        require(o.type >= OWE_SURF_NULL && o.type <= OWE_SURF_DETECTOR, "optics " + std::to_string(i) + ": unknown type");
        checkSpectrum(o.reflectance, s.n_table, "reflectance");
        checkSpectrum(o.conductor_n, s.n_table, "conductor n");
        checkSpectrum(o.conductor_k, s.n_table, "conductor k");
        DOptics d{};
        d.type = o.type;
        d.backAbsorbs = o.back_absorbs;
        d.roughness = float(o.roughness);
        d.delta = (o.type == OWE_SURF_DIELECTRIC || o.type == OWE_SURF_CONDUCTOR) ? (o.roughness < 1e-3) : (o.type == OWE_SURF_MIRROR);
        d.reflectance = toDev(o.reflectance, s.table, h.table);
        d.condN = toDev(o.conductor_n, s.table, h.table);
        d.condK = toDev(o.conductor_k, s.table, h.table);
        h.optics.push_back(d);
    }
    for (uint32_t i = 0; i < s.n_emissions; ++i) {
        checkSpectrum(s.emissions[i].radiance, s.n_table, "emission");
        DEmission d{};
        d.radiance = toDev(s.emissions[i].radiance, s.table, h.table);
        d.front = s.emissions[i].front;
        d.back = s.emissions[i].back;
        h.emissions.push_back(d);
    }

    // Boundaries; mesh triangles to world space (camera-relative) with owner and face group.
    std::vector<float> triPos;
    std::vector<uint32_t> triIdx, triOwner, triFace;
    uint32_t faceBase = 0;
    double blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    std::vector<std::array<double, 6>> boxes;  // per boundary, camera-relative: lo, hi
    for (uint32_t i = 0; i < s.n_boundaries; ++i) {
        const owe_boundary& b = s.boundaries[i];
        const std::string name = "boundary " + std::to_string(i);
        require(b.front < s.n_regions && b.back < s.n_regions, name + ": region out of range");
        require(b.optics >= 0 && uint32_t(b.optics) < s.n_optics, name + ": optics out of range");
        require(b.emission < 0 || uint32_t(b.emission) < s.n_emissions, name + ": emission out of range");
        const int ot = s.optics[b.optics].type;
        if (ot == OWE_SURF_DIELECTRIC || ot == OWE_SURF_NULL)
            require(!s.media[s.region_medium[b.front]].opaque && !s.media[s.region_medium[b.back]].opaque,
                    name + " is transmissive but borders an opaque medium");
        Xf xf{};
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) xf.R[r][c] = b.to_world[4 * r + c];
            xf.t[r] = b.to_world[4 * r + 3] - h.origin[r];
        }
        DBoundary d{};
        d.shape = b.shape;
        d.aperture = b.aperture;
        d.front = b.front;
        d.back = b.back;
        d.optics = b.optics;
        d.emission = b.emission;
        d.a = float(b.a);
        d.b = float(b.b);
        d.light = -1;
        xf.inverse().store(d.toLocal);
        xf.store(d.toWorld);
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        auto grow = [&](V p) {
            const double c[3] = {p.x, p.y, p.z};
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], c[k]);
                hi[k] = std::max(hi[k], c[k]);
            }
        };
        double area = 0;
        if (b.shape == OWE_SHAPE_PLANE) {
            require(b.aperture >= OWE_APERTURE_DISK && b.aperture <= OWE_APERTURE_ELLIPSE, name + ": unknown aperture");
            require(b.a > 0 && b.b >= 0 && (b.aperture == OWE_APERTURE_DISK ? b.b < b.a : b.b > 0), name + ": bad aperture size");
            const double ex = b.a, ey = b.aperture == OWE_APERTURE_DISK ? b.a : b.b;
            const double pad = 1e-9 * (1 + std::max(ex, ey));
            for (int c = 0; c < 8; ++c) grow(xf.point({(c & 1) ? ex : -ex, (c & 2) ? ey : -ey, (c & 4) ? pad : -pad}));
            area = b.aperture == OWE_APERTURE_DISK ? kPi * (b.a * b.a - b.b * b.b)
                   : b.aperture == OWE_APERTURE_RECT ? 4 * b.a * b.b : kPi * b.a * b.b;
        } else if (b.shape == OWE_SHAPE_SPHERE) {
            require(b.a > 0, name + ": sphere radius must be positive");
            for (int c = 0; c < 8; ++c) grow(xf.point({(c & 1) ? b.a : -b.a, (c & 2) ? b.a : -b.a, (c & 4) ? b.a : -b.a}));
            area = 4 * kPi * b.a * b.a;
        } else if (b.shape == OWE_SHAPE_CYLINDER) {
            const double z0 = std::min(b.params[0], b.params[1]), z1 = std::max(b.params[0], b.params[1]);
            require(b.a > 0 && z1 > z0, name + ": a cylinder needs a positive radius and z0 != z1");
            require(b.emission < 0, name + ": cylinders do not emit on this backend");
            for (int c = 0; c < 8; ++c) grow(xf.point({(c & 1) ? b.a : -b.a, (c & 2) ? b.a : -b.a, (c & 4) ? z1 : z0}));
            d.z0 = float(z0);
            d.z1 = float(z1);
            area = 2 * kPi * b.a * (z1 - z0);
        } else if (b.shape == OWE_SHAPE_SAG) {
            const double c = b.params[0], k = b.params[1];
            require(b.a > 0 && b.b >= 0 && b.b < b.a, name + ": a sag surface needs 0 <= r_min < r_max");
            require(1 - (1 + k) * c * c * b.a * b.a >= 0, name + ": aperture exceeds the domain of the conic");
            require(b.emission < 0, name + ": sag surfaces do not emit on this backend");
            d.sag = makeSag(c, k, b.params + 2, b.a, b.b);
            for (int q = 0; q < 8; ++q)
                grow(xf.point({(q & 1) ? b.a : -b.a, (q & 2) ? b.a : -b.a, double((q & 4) ? d.sag.zMax : d.sag.zMin)}));
            d.depTol = float(1e-5 * (b.a + std::max(std::fabs(double(d.sag.zMin)), std::fabs(double(d.sag.zMax)))));
            area = kPi * (b.a * b.a - b.b * b.b);
        } else if (b.shape == OWE_SHAPE_WAVE) {
            require(b.a > 0 && b.b > 0, name + ": a wave surface needs a positive extent");
            require(b.wave_count >= 1 && s.waves && uint64_t(b.first_wave) + b.wave_count <= s.n_waves,
                    name + ": waves out of range");
            require(b.emission < 0, name + ": wave surfaces do not emit on this backend");
            d.wave = buildWave(b, s.waves, h.waveData);
            const double amp = d.wave.amp;
            for (int q = 0; q < 8; ++q) grow(xf.point({(q & 1) ? b.a : -b.a, (q & 2) ? b.b : -b.b, (q & 4) ? amp : -amp}));
            d.depTol = float(1e-5 * (std::max(b.a, b.b) + amp));
            area = 4 * b.a * b.b;
        } else if (b.shape == OWE_SHAPE_MESH) {
            require(uint64_t(b.first_triangle) + b.triangle_count <= s.n_triangles, name + ": triangles out of range");
            // Face groups in the mesh's own frame (double, exact coplanarity).
            const uint32_t nt = b.triangle_count;
            auto P = [&](uint32_t v) {
                return V{s.vertices[3 * size_t(v)], s.vertices[3 * size_t(v) + 1], s.vertices[3 * size_t(v) + 2]};
            };
            auto T = [&](uint32_t t, int k) {
                const uint32_t v = s.triangles[3 * size_t(b.first_triangle + t) + k];
                require(v < s.n_vertices, name + ": vertex index out of range");
                return v;
            };
            std::vector<uint32_t> parent(nt);
            std::iota(parent.begin(), parent.end(), 0u);
            auto find = [&](uint32_t x) {
                while (parent[x] != x) x = parent[x] = parent[parent[x]];
                return x;
            };
            std::map<std::tuple<double, double, double>, uint32_t> weld;
            auto key = [&](uint32_t v) {
                const V p = P(v);
                return weld.emplace(std::make_tuple(p.x, p.y, p.z), uint32_t(weld.size())).first->second;
            };
            auto sameFace = [&](uint32_t a, uint32_t c) {
                auto onA = [&](uint32_t v) {
                    for (int k = 0; k < 3; ++k) {
                        const V q = P(T(a, k)), p = P(v);
                        if (p.x == q.x && p.y == q.y && p.z == q.z) return true;
                    }
                    return false;
                };
                int shared = 0;
                for (int k = 0; k < 3; ++k) shared += onA(T(c, k)) ? 1 : 0;
                if (shared < 2) return false;
                const V na = crs(sub(P(T(a, 1)), P(T(a, 0))), sub(P(T(a, 2)), P(T(a, 0))));
                if (dt(na, crs(sub(P(T(c, 1)), P(T(c, 0))), sub(P(T(c, 2)), P(T(c, 0))))) <= 0) return false;
                for (int k = 0; k < 3; ++k)
                    if (!onA(T(c, k)) && dt(na, sub(P(T(c, k)), P(T(a, 0)))) != 0) return false;
                return true;
            };
            std::map<std::pair<uint32_t, uint32_t>, std::vector<uint32_t>> edges;
            for (uint32_t t = 0; t < nt; ++t)
                for (int e = 0; e < 3; ++e) {
                    uint32_t a = key(T(t, e)), c = key(T(t, (e + 1) % 3));
                    if (a > c) std::swap(a, c);
                    edges[{a, c}].push_back(t);
                }
            for (auto& [k, list] : edges)
                for (size_t x = 0; x < list.size(); ++x)
                    for (size_t y = x + 1; y < list.size(); ++y)
                        if (sameFace(list[x], list[y])) parent[find(list[x])] = find(list[y]);
            std::map<uint32_t, uint32_t> groupId;
            for (uint32_t t = 0; t < nt; ++t) {
                const uint32_t g = find(t);
                auto it = groupId.emplace(g, faceBase + uint32_t(groupId.size())).first;
                triFace.push_back(it->second);
                triOwner.push_back(i);
                for (int k = 0; k < 3; ++k) {
                    const V w = xf.point(P(T(t, k)));
                    grow(w);
                    triIdx.push_back(uint32_t(triPos.size() / 3));
                    triPos.push_back(float(w.x));
                    triPos.push_back(float(w.y));
                    triPos.push_back(float(w.z));
                }
                const V e1 = sub(P(T(t, 1)), P(T(t, 0))), e2 = sub(P(T(t, 2)), P(T(t, 0)));
                const V c = crs(e1, e2);
                area += 0.5 * std::sqrt(dt(c, c));
            }
            faceBase += uint32_t(groupId.size());
        } else {
            fail(OWE_HIP_ERR_UNSUPPORTED, name + ": unknown shape");
        }
        for (int k = 0; k < 3; ++k) {
            d.lo[k] = down(lo[k]);
            d.hi[k] = upf(hi[k]);
            blo[k] = std::min(blo[k], lo[k]);
            bhi[k] = std::max(bhi[k], hi[k]);
        }
        boxes.push_back({lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]});
        d.area = float(area);
        h.boundaries.push_back(d);
    }
    h.hasTriangles = !triOwner.empty();
    // Contacts are possible only where two distinct boundaries' boxes meet, within the contact
    // tolerance (2^-21 of the largest coordinate of either box).
    for (size_t i = 0; i < boxes.size() && !h.contacts; ++i)
        for (size_t j = i + 1; j < boxes.size() && !h.contacts; ++j) {
            double m = 0;
            for (int k = 0; k < 6; ++k) m = std::max({m, std::fabs(boxes[i][k]), std::fabs(boxes[j][k])});
            const double tol = std::ldexp(m, -21);
            bool meet = true;
            for (int k = 0; k < 3; ++k)
                meet = meet && boxes[i][k] - tol <= boxes[j][k + 3] && boxes[j][k] - tol <= boxes[i][k + 3];
            h.contacts = meet;
        }
    h.bvh = buildBvh4(triPos.data(), uint32_t(triPos.size() / 3), triIdx.data(), uint32_t(triOwner.size()),
                      triOwner.data(), triFace.data());

    // Environment and lights.
    const owe_environment& e = s.env;
    checkSpectrum(e.sun_radiance, s.n_table, "sun radiance");
    checkSpectrum(e.sky_radiance, s.n_table, "sky radiance");
    h.env.hasSun = e.has_sun;
    h.env.sky = e.sky;
    const V sd = nrm({e.sun_dir[0], e.sun_dir[1], e.sun_dir[2]});
    h.env.sunDir[0] = float(sd.x);
    h.env.sunDir[1] = float(sd.y);
    h.env.sunDir[2] = float(sd.z);
    const double omc = 2 * std::sin(0.5 * e.sun_angular_radius) * std::sin(0.5 * e.sun_angular_radius);
    h.env.sunOneMinusCos = float(omc);
    h.env.sunSolidAngle = float(2 * kPi * omc);
    h.env.sunRadiance = toDev(e.sun_radiance, s.table, h.table);
    h.env.skyRadiance = toDev(e.sky_radiance, s.table, h.table);
    const bool boundsValid = blo[0] <= bhi[0];
    const double ext[3] = {bhi[0] - blo[0], bhi[1] - blo[1], bhi[2] - blo[2]};
    const double radius = boundsValid ? std::max(1e-3, 0.5 * std::sqrt(ext[0] * ext[0] + ext[1] * ext[1] + ext[2] * ext[2])) : 1.0;
    struct L {
        int kind;
        uint32_t b;
        double power;
    };
    std::vector<L> lights;
    for (uint32_t i = 0; i < s.n_boundaries; ++i) {
        const owe_boundary& b = s.boundaries[i];
        if (b.emission < 0 || b.shape == OWE_SHAPE_MESH || !s.emissions[b.emission].nee || !(h.boundaries[i].area > 0))
            continue;
        const owe_emission& em = s.emissions[b.emission];
        const double sides = (em.front ? 1 : 0) + (em.back ? 1 : 0);
        h.boundaries[i].light = int32_t(lights.size());
        lights.push_back({0, i, std::max(1e-12, luminance(em.radiance, s.table)) * double(h.boundaries[i].area) * kPi * sides});
    }
    if (e.has_sun)
        lights.push_back({1, kNoBoundary, std::max(1e-12, luminance(e.sun_radiance, s.table)) * (2 * kPi * omc) * kPi * radius * radius});
    double total = 0, boundaryPower = 0;
    for (const L& l : lights) {
        total += l.power;
        if (l.kind == 0) boundaryPower += l.power;
    }
    const double share = e.has_sun && boundaryPower > 0 ? std::min(0.999, std::max(0.0, e.sun_nee_share)) : 0;
    double acc = 0;
    for (size_t i = 0; i < lights.size(); ++i) {
        double p = total > 0 ? lights[i].power / total : 0;
        if (share > 0) p = lights[i].kind == 1 ? share : (1 - share) * lights[i].power / boundaryPower;
        acc += p;
        h.lights.push_back({lights[i].kind, lights[i].b, float(p), float(i + 1 == lights.size() ? 1.0 : acc)});
    }
    h.env.nLights = int32_t(h.lights.size());
    h.env.sunPdf = e.has_sun ? float(h.lights.back().pdf / (2 * kPi * omc)) : 0.0f;
    // Particles: computeEnvironmentPower (world.cpp:352-371) for a uniform sky, against the lights'
    // total power (gpu_scene.cpp:753-758).
    double skyPower = 0;
    if (e.sky == OWE_SKY_UNIFORM) {
        double y = 0;
        for (double l = 400; l <= 700; l += 50) y += evalHost(e.sky_radiance, s.table, l);
        skyPower = std::max(0.0, y / 7) * kPi * radius * radius * 4 * kPi;
    }
    h.emittedPower = skyPower + (lights.empty() ? 0.0 : total);
    h.env.skyShare = h.emittedPower > 0 ? float(skyPower / h.emittedPower) : 0.0f;
    h.env.emitRadius = float(radius * 1.001);
    for (int k = 0; k < 3; ++k) h.env.center[k] = boundsValid ? float(0.5 * (blo[k] + bhi[k])) : 0.0f;

    h.obs.width = ob.width;
    h.obs.height = ob.height;
    h.obs.kind = ob.kind;
    h.obs.sigma = float(ob.pixel_sigma);
    if (ob.kind == OWE_DETECTOR_SENSOR) {
        const DBoundary& sb = h.boundaries[ob.sensor_boundary];
        h.obs.sensor = ob.sensor_boundary;
        h.obs.region = sb.front;
        h.obs.halfX = float(ob.half_x);
        h.obs.halfY = float(ob.half_y);
        h.obs.flipX = ob.flip_x != 0;
        h.obs.flipY = ob.flip_y != 0;
        h.obs.hasAim = ob.has_aim != 0;
        h.obs.aimShare = float(ob.aim_share);
        h.obs.aimRadius = float(ob.aim_radius);
        const owe_boundary& b = s.boundaries[ob.sensor_boundary];
        const V nrmW = nrm({b.to_world[2], b.to_world[6], b.to_world[10]});
        const V an = nrm({ob.aim_normal[0], ob.aim_normal[1], ob.aim_normal[2]});
        for (int k = 0; k < 3; ++k) {
            h.obs.normal[k] = float(k == 0 ? nrmW.x : (k == 1 ? nrmW.y : nrmW.z));
            h.obs.aimCenter[k] = float(ob.aim_center[k] - h.origin[k]);
            h.obs.aimNormal[k] = float(k == 0 ? an.x : (k == 1 ? an.y : an.z));
        }
        std::copy(sb.toWorld, sb.toWorld + 12, h.obs.toWorld);
        return h;
    }
    // Observer (IdealObserver::prepare, src/owe/scene/detector.cpp:36-45), in the rebased frame.
    const V pos{ob.position[0], ob.position[1], ob.position[2]};
    const V fwd = nrm(sub({ob.look_at[0], ob.look_at[1], ob.look_at[2]}, pos));
    V u{ob.up[0], ob.up[1], ob.up[2]};
    if (std::fabs(dt(nrm(u), fwd)) > 0.9999) u = std::fabs(fwd.z) < 0.9 ? V{0, 0, 1} : V{0, 1, 0};
    const V right = nrm(crs(fwd, u)), up = crs(right, fwd);
    h.obs.region = ob.region;
    const double tanY = std::tan(0.5 * ob.fov_y);
    h.obs.tanY = float(tanY);
    h.obs.tanX = float(tanY * double(ob.width) / double(ob.height));
    const V vs[3] = {fwd, right, up};
    float* dst[3] = {h.obs.fwd, h.obs.right, h.obs.up};
    for (int k = 0; k < 3; ++k) {
        dst[k][0] = float(vs[k].x);
        dst[k][1] = float(vs[k].y);
        dst[k][2] = float(vs[k].z);
    }
    h.obs.pupil = float(ob.pupil_radius);
    h.obs.focus = float(ob.focus_distance);
    (void)origin;
    return h;
}

}  // namespace owehip
