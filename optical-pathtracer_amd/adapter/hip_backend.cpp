// The hip backend's adapter: owe::Backend / owe::Renderer (src/owe/render/backend.hpp:31-52,
// src/owe/render/renderer.hpp:19-79) implemented over the C ABI of include/owe_hip.h, and the
// flattening of a built owe::Scene into include/owe_hip_scene.h. Status codes come back from the ABI
// and are rethrown as std::runtime_error, the interface's error contract (renderer.hpp:22-36,
// backend.hpp:44-46). Maintainer-side source, compiled inside the engine as
// src/owe/backends/hip/hip_backend.cpp.
#include "hip_backend.hpp"

#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "owe/scene/detector.hpp"
#include "owe/scene/geometry.hpp"
#include "owe/scene/scene.hpp"
#include "owe/scene/world.hpp"
#include "owe_hip.h"

namespace owe {
namespace {

[[noreturn]] void unsupported(const std::string& what) {
    throw std::runtime_error("hip backend: " + what + " is not supported");
}

void check(int status) {
    if (status != OWE_HIP_OK) throw std::runtime_error(std::string("hip backend: ") + owe_hip_last_error());
}

// Spectrum → owe_spectrum. Constant, blackbody and tabulated spectra cross exactly; the RGB bases
// are tabulated every nanometre over 360–830 nm (their smooth partition is linear within 1 nm).
owe_spectrum spectrum(const Spectrum& s, std::vector<double>& table) {
    owe_spectrum o{};
    o.scale = s.paramScale();
    switch (s.kind()) {
        case Spectrum::Kind::Constant:
            o.kind = OWE_SPEC_CONSTANT;
            o.value = s.paramA();
            return o;
        case Spectrum::Kind::Blackbody:
            o.kind = OWE_SPEC_BLACKBODY;
            o.kelvin = s.paramA();
            return o;
        case Spectrum::Kind::Tabulated: {
            const std::vector<double>& l = *s.tableLambdas();
            const std::vector<double>& v = *s.tableValues();
            o.kind = OWE_SPEC_TABULATED;
            o.count = int32_t(l.size());
            o.offset = int32_t(table.size() / 2);
            for (size_t i = 0; i < l.size(); ++i) {
                table.push_back(l[i]);
                table.push_back(v[i]);
            }
            return o;
        }
        default: {
            o.kind = OWE_SPEC_TABULATED;
            o.scale = 1.0;
            o.offset = int32_t(table.size() / 2);
            o.count = 0;
            for (int nm = 360; nm <= 830; ++nm, ++o.count) {
                table.push_back(double(nm));
                table.push_back(s.eval(double(nm)));
            }
            return o;
        }
    }
}

owe_medium medium(const Medium& m, std::vector<double>& table) {
    owe_medium o{};
    const IndexModel& ix = m.index;
    o.rel_air = ix.isRelativeToAir() ? 1 : 0;
    o.opaque = m.opaque ? 1 : 0;
    switch (ix.kind()) {
        case IndexModel::Kind::Constant:
            o.index_kind = OWE_INDEX_CONSTANT;
            o.coeff_a[0] = ix.coeffA().at(0);
            break;
        case IndexModel::Kind::Cauchy:
            o.index_kind = OWE_INDEX_CAUCHY;
            for (size_t k = 0; k < ix.coeffA().size() && k < 3; ++k) o.coeff_a[k] = ix.coeffA()[k];
            break;
        case IndexModel::Kind::Sellmeier:
            if (ix.coeffA().size() > 4) unsupported("a Sellmeier index of more than 4 terms (medium '" + m.name + "')");
            o.index_kind = OWE_INDEX_SELLMEIER;
            o.terms = int32_t(ix.coeffA().size());
            for (size_t k = 0; k < ix.coeffA().size(); ++k) {
                o.coeff_a[k] = ix.coeffA()[k];
                o.coeff_b[k] = ix.coeffB().at(k);
            }
            break;
        case IndexModel::Kind::Ciddor: o.index_kind = OWE_INDEX_CIDDOR; break;
        case IndexModel::Kind::Tabulated: unsupported("a tabulated refractive index (medium '" + m.name + "')");
    }
    o.absorption = spectrum(m.absorption, table);
    o.scattering = spectrum(m.scattering, table);
    o.g = m.g;
    return o;
}

owe_optics optics(const SurfaceOptics& s, std::vector<double>& table) {
    owe_optics o{};
    switch (s.type) {
        case SurfaceType::Null: o.type = OWE_SURF_NULL; break;
        case SurfaceType::Dielectric: o.type = OWE_SURF_DIELECTRIC; break;
        case SurfaceType::Diffuse: o.type = OWE_SURF_DIFFUSE; break;
        case SurfaceType::Conductor: o.type = OWE_SURF_CONDUCTOR; break;
        case SurfaceType::Mirror: o.type = OWE_SURF_MIRROR; break;
        case SurfaceType::Absorber: o.type = OWE_SURF_ABSORBER; break;
        case SurfaceType::Detector: o.type = OWE_SURF_DETECTOR; break;
        case SurfaceType::StainedGlass: unsupported("stained glass (material '" + s.name + "')");
    }
    if (s.texture.kind != Texture::Kind::None) unsupported("a textured material ('" + s.name + "')");
    if (!s.roughnessMap.empty() || s.reliefDepth != 0 || s.normalImage.image || s.roughnessImage.image)
        unsupported("a material map, relief or image map ('" + s.name + "')");
    if (s.sampleAimShare > 0 && s.sampleAimRadius > 0) unsupported("a diffuse aim guide ('" + s.name + "')");
    o.roughness = s.roughness;
    o.back_absorbs = s.backAbsorbs ? 1 : 0;
    o.reflectance = spectrum(s.reflectance, table);
    o.conductor_n = spectrum(s.conductor.n, table);
    o.conductor_k = spectrum(s.conductor.k, table);
    return o;
}

void rowMajor(const Transform& xf, double out[12]) {
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) out[4 * i + j] = xf.R.m[i][j];
        out[4 * i + 3] = xf.t[i];
    }
}

// One boundary's shape, in its own frame, with its triangles or plane waves appended.
void shape(const Boundary& b, owe_boundary& o, HipFlatScene& f) {
    const Shape* s = b.shape.get();
    if (auto m = dynamic_cast<const MeshShape*>(s)) {
        o.shape = OWE_SHAPE_MESH;
        const uint32_t base = uint32_t(f.vertices.size() / 3);
        for (const Vec3& p : m->positions()) {
            f.vertices.push_back(p.x);
            f.vertices.push_back(p.y);
            f.vertices.push_back(p.z);
        }
        o.first_triangle = uint32_t(f.triangles.size() / 3);
        o.triangle_count = uint32_t(m->triangles().size());
        for (const auto& t : m->triangles())
            for (uint32_t v : t) f.triangles.push_back(base + v);
    } else if (auto p = dynamic_cast<const PlaneShape*>(s)) {
        o.shape = OWE_SHAPE_PLANE;
        switch (p->aperture()) {
            case PlaneShape::Aperture::Disk: o.aperture = OWE_APERTURE_DISK; break;
            case PlaneShape::Aperture::Rect: o.aperture = OWE_APERTURE_RECT; break;
            case PlaneShape::Aperture::Ellipse: o.aperture = OWE_APERTURE_ELLIPSE; break;
            case PlaneShape::Aperture::DiskHole: unsupported("a disk with an off-centre hole ('" + b.name + "')");
        }
        o.a = p->a();
        o.b = p->b();
    } else if (auto sp = dynamic_cast<const SphereShape*>(s)) {
        o.shape = OWE_SHAPE_SPHERE;
        o.a = sp->radius();
    } else if (auto c = dynamic_cast<const CylinderShape*>(s)) {
        o.shape = OWE_SHAPE_CYLINDER;
        o.a = c->radius();
        o.params[0] = c->z0();
        o.params[1] = c->z1();
    } else if (auto g = dynamic_cast<const SagSurface*>(s)) {
        if (g->asphere().size() > 4) unsupported("an asphere beyond A10 ('" + b.name + "')");
        o.shape = OWE_SHAPE_SAG;
        o.a = g->rMax();
        o.b = g->rMin();
        o.params[0] = g->curvature();
        o.params[1] = g->conic();
        for (size_t k = 0; k < g->asphere().size(); ++k) o.params[2 + k] = g->asphere()[k];
    } else if (auto w = dynamic_cast<const WaveSurface*>(s)) {
        o.shape = OWE_SHAPE_WAVE;
        o.a = w->halfX();
        o.b = w->halfY();
        o.params[0] = w->margin();
        // The built surface keeps its realised node spacing; a hair above it reproduces the same
        // node count (⌈2·halfX / spacing⌉ + 1) and spacing on the device. 0: the analytic sum.
        o.params[1] = w->sampled() ? w->spacingX() * (1 + 1e-9) : 0.0;
        o.first_wave = uint32_t(f.waves.size());
        o.wave_count = uint32_t(w->waves().size());
        for (const PlaneWave& pw : w->waves()) f.waves.push_back({pw.amplitude, pw.kx, pw.ky, pw.phase});
    } else {
        unsupported("the shape of boundary '" + b.name + "' (" + s->describe() + ")");
    }
}

}  // namespace

void flattenForHip(const Scene& scene, const Detector& detector, HipFlatScene& f) {
    const World& w = scene.world;
    f = HipFlatScene{};
    for (const Medium& m : w.media()) f.media.push_back(medium(m, f.table));
    for (const Region& r : w.regions()) f.regionMedium.push_back(r.medium);
    for (const SurfaceOptics& o : w.optics()) f.optics.push_back(optics(o, f.table));
    for (const Emission& e : w.emissions()) {
        owe_emission o{};
        o.radiance = spectrum(e.radiance, f.table);
        o.front = e.front ? 1 : 0;
        o.back = e.back ? 1 : 0;
        o.nee = e.nee ? 1 : 0;
        f.emissions.push_back(o);
    }
    for (const Boundary& b : w.boundaries()) {
        owe_boundary o{};
        o.front = b.front;
        o.back = b.back;
        o.optics = int32_t(b.optics);
        o.emission = b.emission;
        rowMajor(b.toWorld, o.to_world);
        shape(b, o, f);
        f.boundaries.push_back(o);
    }
    owe_scene& s = f.scene;
    const Environment& env = w.env;
    if (env.skyModel == Environment::Sky::Gradient || env.skyModel == Environment::Sky::Map)
        unsupported("a gradient or image sky");
    s.env.sky = env.skyModel == Environment::Sky::Uniform ? OWE_SKY_UNIFORM : OWE_SKY_NONE;
    s.env.sky_radiance = spectrum(env.zenith, f.table);
    s.env.has_sun = env.hasSun ? 1 : 0;
    for (int k = 0; k < 3; ++k) s.env.sun_dir[k] = env.sunDir[k];
    s.env.sun_angular_radius = env.sunAngularRadius;
    s.env.sun_radiance = spectrum(env.sunRadiance, f.table);
    s.env.sun_nee_share = env.sunNeeShare;
    owe_observer& ob = s.observer;
    ob.width = detector.width;
    ob.height = detector.height;
    ob.region = detector.region;
    ob.pixel_sigma = detector.pixelSigma;
    if (auto eye = dynamic_cast<const IdealObserver*>(&detector)) {
        ob.kind = OWE_DETECTOR_OBSERVER;
        for (int k = 0; k < 3; ++k) {
            ob.position[k] = eye->position[k];
            ob.look_at[k] = eye->lookAt[k];
            ob.up[k] = eye->up[k];
        }
        ob.fov_y = eye->fovY;
        ob.pupil_radius = eye->pupilRadius;
        ob.focus_distance = eye->focusDistance;
        // An exit-pupil guide only redistributes pupil samples (IdealObserver::PupilGuide); the
        // backend samples the pupil uniformly, which estimates the same image.
    } else if (auto sensor = dynamic_cast<const SurfaceSensor*>(&detector)) {
        if (sensor->focusShare > 0 && sensor->focusRadius > 0) unsupported("a sensor focus guide");
        ob.kind = OWE_DETECTOR_SENSOR;
        ob.sensor_boundary = sensor->boundaryIndex;
        ob.half_x = sensor->halfX;
        ob.half_y = sensor->halfY;
        ob.flip_x = sensor->flipX ? 1 : 0;
        ob.flip_y = sensor->flipY ? 1 : 0;
        ob.has_aim = sensor->hasAim ? 1 : 0;
        for (int k = 0; k < 3; ++k) {
            ob.aim_center[k] = sensor->aimCenter[k];
            ob.aim_normal[k] = sensor->aimNormal[k];
        }
        ob.aim_radius = sensor->aimRadius;
        ob.aim_share = sensor->aimShare;
    } else {
        unsupported("detector '" + detector.name + "' (" + detector.describe() + ")");
    }
    s.n_media = uint32_t(f.media.size());
    s.n_regions = uint32_t(f.regionMedium.size());
    s.n_optics = uint32_t(f.optics.size());
    s.n_emissions = uint32_t(f.emissions.size());
    s.n_boundaries = uint32_t(f.boundaries.size());
    s.n_vertices = uint32_t(f.vertices.size() / 3);
    s.n_triangles = uint32_t(f.triangles.size() / 3);
    s.n_table = uint32_t(f.table.size() / 2);
    s.n_waves = uint32_t(f.waves.size());
    s.media = f.media.data();
    s.region_medium = f.regionMedium.data();
    s.optics = f.optics.data();
    s.emissions = f.emissions.data();
    s.boundaries = f.boundaries.data();
    s.vertices = f.vertices.data();
    s.triangles = f.triangles.data();
    s.table = f.table.data();
    s.waves = f.waves.data();
}

namespace {

std::vector<std::string> integratorList() {
    std::vector<std::string> out;
    std::string all = owe_hip_backend_integrators(), item;
    for (char ch : all + ",") {
        if (ch == ',') {
            if (!item.empty()) out.push_back(item);
            item.clear();
        } else {
            item += ch;
        }
    }
    return out;
}

class HipRenderer final : public Renderer {
public:
    HipRenderer(const Scene& scene, int detectorIndex, const RenderSettings& rs)
        : Renderer(scene, detectorIndex, rs, integratorList()) {
        flattenForHip(scene, *det_, flat_);
        owe_render_settings s{};
        s.seed = rs.seed;
        s.max_depth = rs.maxDepth;
        s.rr_depth = rs.rrDepth;
        s.integrator = rs.integrator == "light"    ? OWE_INTEGRATOR_LIGHT
                       : rs.integrator == "sppm"   ? OWE_INTEGRATOR_SPPM
                       : rs.integrator == "vcm"    ? OWE_INTEGRATOR_VCM
                       : rs.integrator == "bdpt"   ? OWE_INTEGRATOR_BDPT
                       : rs.integrator == "hybrid" ? OWE_INTEGRATOR_HYBRID
                                                   : OWE_INTEGRATOR_PATH;
        s.spectral = OWE_SPECTRAL_HERO4;
        s.lambda_min = 380.0;
        s.lambda_max = 780.0;
        s.fresnel_floor = rs.fresnelFloor;
        s.photon_radius = rs.photonRadius;
        s.photon_alpha = rs.photonAlpha;
        s.photons_per_pixel = rs.photonsPerPixel;
        s.photon_groups = rs.integrator == "vcm" || rs.integrator == "bdpt" ? rs.vcmWavelengthGroups
                                                                            : rs.photonWavelengthGroups;
        settings_c_ = s;
        check(owe_hip_renderer_create(rs.device < 0 ? 0 : rs.device, &flat_.scene, &s, &r_));
    }
    ~HipRenderer() override {
        if (r_) owe_hip_renderer_destroy(r_);
    }
    HipRenderer(const HipRenderer&) = delete;
    HipRenderer& operator=(const HipRenderer&) = delete;

    Image resolve() const override {
        Image img;
        img.width = det_->width;
        img.height = det_->height;
        img.xyz.assign(size_t(img.width) * size_t(img.height) * 3, 0.0);
        check(owe_hip_renderer_resolve(r_, img.xyz.data()));
        return img;
    }
    std::string backend() const override { return std::string("hip: ") + owe_hip_backend_summary(); }
    bool resetObserver() override {
        // The device scene is rebased on the eye (camera-relative floats), so a new pose or
        // resolution is a new flattened scene and device renderer; the samples are discarded, as
        // the interface says.
        HipFlatScene next;
        flattenForHip(scene_, *det_, next);
        owe_hip_renderer* fresh = nullptr;
        check(owe_hip_renderer_create(settings_.device < 0 ? 0 : settings_.device, &next.scene, &settings_c_, &fresh));
        owe_hip_renderer_destroy(r_);
        r_ = fresh;
        flat_ = std::move(next);
        resetProgress();
        return true;
    }

private:
    void pass(int spp) override {
        check(owe_hip_renderer_pass(r_, spp));
        owe_transport_stats t{};
        check(owe_hip_renderer_stats(r_, &t));
        stats_ = {};
        stats_.paths = t.paths;
        stats_.segments = t.segments;
        stats_.inconsistencies = t.inconsistencies;
        stats_.leaks = t.leaks;
        owe_inconsistency f{};
        int found = 0;
        check(owe_hip_renderer_first_inconsistency(r_, &f, &found));
        if (found) {
            auto none = [](uint32_t v) { return v == 0xFFFFFFFFu ? kNone : v; };
            stats_.first = {f.boundary, none(f.ray_region), none(f.boundary_region), none(f.previous),
                            Vec3(f.p[0], f.p[1], f.p[2]), Vec3(f.origin[0], f.origin[1], f.origin[2]),
                            Vec3(f.direction[0], f.direction[1], f.direction[2]), none(f.previous_primitive),
                            none(f.primitive)};
        }
    }

    HipFlatScene flat_;
    owe_render_settings settings_c_{};
    owe_hip_renderer* r_ = nullptr;
};

class HipBackend final : public Backend {
public:
    std::string name() const override { return owe_hip_backend_name(); }
    std::string summary() const override { return owe_hip_backend_summary(); }
    bool available(std::string* why) const override {
        char buf[256] = {0};
        const bool ok = owe_hip_available(buf, sizeof buf) != 0;
        if (!ok && why) *why = buf;
        return ok;
    }
    std::vector<DeviceInfo> devices() const override {
        int n = 0;
        if (owe_hip_device_count(&n) != OWE_HIP_OK) return {};
        std::vector<DeviceInfo> out;
        for (int i = 0; i < n; ++i) {
            owe_hip_device_info d{};
            if (owe_hip_device_info_get(i, &d) != OWE_HIP_OK) continue;
            DeviceInfo info;
            info.index = i;
            info.name = d.name;
            info.type = d.type;
            info.driver = std::string(d.arch) + ", " + std::to_string(d.compute_units) + " CUs";
            out.push_back(info);
        }
        return out;
    }
    std::vector<std::string> integrators() const override { return integratorList(); }
    std::unique_ptr<Renderer> createRenderer(const Scene& scene, int detectorIndex,
                                             const RenderSettings& settings) const override {
        std::string why;
        if (!available(&why)) throw std::runtime_error("hip backend unavailable: " + why);
        return std::make_unique<HipRenderer>(scene, detectorIndex, settings);
    }
};

}  // namespace

const Backend& hipBackend() {
    static const HipBackend backend;
    return backend;
}

}  // namespace owe
