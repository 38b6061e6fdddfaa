// TEST INFRASTRUCTURE ONLY — the double-precision CPU oracle (see oracle_scene.hpp for citations).
// Camera-path transport restated from src/owe/transport/transport.cpp:35-529 and
// src/owe/transport/scattering.cpp:65-222, over the simplified World of oracle_scene.cpp; particle
// transport (light tracing) restated in double from the reference's only implementation, the float
// kernel src/owe/backends/gpu/shaders/particle.slang:23-427 with particle_kernel.slang:11-19.
#include "oracle_scene.hpp"

namespace oracle {

namespace {

constexpr bool kRadiance = true;

// Hit-local optics state at one wavelength (scattering.hpp:31-44).
struct Interface {
    const Optics* o;
    V3 n, p;
    double nFront, nBack, lambda, floor;
};

struct Sample {
    V3 wi;
    double weight = 0, pdf = 0;
    bool delta = false, transmitted = false;
};

C2 conductorEta(const Optics& o, double lambda, double nMed) {
    const double inv = 1.0 / nMed;  // a Vec2 divided by a scalar multiplies by its reciprocal
    return {o.condN.eval(lambda) * inv, o.condK.eval(lambda) * inv};
}

bool specularReflection(const Frame& f, V3 wo, double weight, Sample& s) {
    s.wi = f.toWorld(V3(-wo.x, -wo.y, wo.z));
    s.weight = weight;
    s.delta = true;
    return weight > 0;
}

void fromLobe(const Frame& f, const Lobe& l, Sample& s) {
    s.wi = f.toWorld(l.wi);
    s.weight = l.weight;
    s.pdf = l.pdf;
    s.delta = false;
    s.transmitted = l.transmitted;
}

// sampleScatter (scattering.cpp:65-157): uc picks a branch, (u1, u2) the direction.
bool sampleScatter(const Interface& it, V3 d, double uc, double u1, double u2, bool radiance, Sample& s) {
    const Optics& o = *it.o;
    const Frame frame(it.n);
    const V3 wo = frame.toLocal(-d);
    const bool fromFront = wo.z > 0;
    s = Sample{};
    switch (o.type) {
        case OWE_SURF_DIELECTRIC: {
            const double eta = it.nBack / it.nFront;
            if (o.roughness >= kSmoothRoughness && eta != 1) {
                Lobe l;
                if (!sampleRoughDielectric(ggxAlpha(o.roughness), wo, eta, uc, u1, u2, radiance, l)) return false;
                fromLobe(frame, l, s);
                return true;
            }
            const double ni = fromFront ? it.nFront : it.nBack, nt = fromFront ? it.nBack : it.nFront;
            const double R = fresnelDielectric(wo.z, eta), pR = branchProbability(R, it.floor);
            s.delta = true;
            if (uc < pR) {
                specularReflection(frame, wo, R / pR, s);
                return true;
            }
            if (!refractDir(d, it.n, ni, nt, s.wi)) return false;
            s.weight = (1 - R) / (1 - pR) * (radiance ? sq(ni / nt) : 1.0);
            s.transmitted = true;
            return true;
        }
        case OWE_SURF_DIFFUSE: {
            if (!fromFront && o.backAbsorbs) return false;
            const double albedo = o.reflectance.eval(it.lambda);
            if (albedo <= 0) return false;
            V3 w = cosineHemisphere(u1, u2);
            if (!fromFront) w.z = -w.z;
            s.wi = frame.toWorld(w);
            s.pdf = std::fabs(w.z) * kInvPi;
            s.weight = albedo;
            return s.pdf > 0;
        }
        case OWE_SURF_MIRROR:
            if (!fromFront && o.backAbsorbs) return false;
            return specularReflection(frame, wo, o.reflectance.eval(it.lambda), s);
        case OWE_SURF_CONDUCTOR: {
            if (!fromFront && o.backAbsorbs) return false;
            const C2 eta = conductorEta(o, it.lambda, fromFront ? it.nFront : it.nBack);
            const double tint = o.reflectance.eval(it.lambda);
            const V3 w = fromFront ? wo : V3(wo.x, wo.y, -wo.z);
            if (o.roughness < kSmoothRoughness) return specularReflection(frame, wo, fresnelConductor(w.z, eta) * tint, s);
            Lobe l;
            if (!sampleRoughConductor(ggxAlpha(o.roughness), w, eta, tint, u1, u2, l)) return false;
            if (!fromFront) l.wi.z = -l.wi.z;
            fromLobe(frame, l, s);
            return true;
        }
        default: return false;
    }
}

// evalScatter (scattering.cpp:159-188): f without the cosine, and the pdf of sampling wi.
double evalScatter(const Interface& it, V3 d, V3 wiWorld, bool radiance, double& pdf) {
    pdf = 0;
    const Optics& o = *it.o;
    if (o.isDelta()) return 0;
    const Frame frame(it.n);
    V3 wo = frame.toLocal(-d), wi = frame.toLocal(wiWorld);
    const bool fromFront = wo.z > 0;
    switch (o.type) {
        case OWE_SURF_DIFFUSE:
            if (!fromFront && o.backAbsorbs) return 0;
            if (!sameHemi(wo, wi)) return 0;
            pdf = std::fabs(wi.z) * kInvPi;
            return o.reflectance.eval(it.lambda) * kInvPi;
        case OWE_SURF_CONDUCTOR: {
            if (!fromFront && o.backAbsorbs) return 0;
            if (!sameHemi(wo, wi)) return 0;
            if (!fromFront) {
                wo.z = -wo.z;
                wi.z = -wi.z;
            }
            const C2 eta = conductorEta(o, it.lambda, fromFront ? it.nFront : it.nBack);
            return evalRoughConductor(ggxAlpha(o.roughness), wo, wi, eta, o.reflectance.eval(it.lambda), pdf);
        }
        case OWE_SURF_DIELECTRIC:
            return evalRoughDielectric(ggxAlpha(o.roughness), wo, wi, it.nBack / it.nFront, radiance, pdf);
        default: return 0;
    }
}

bool isDispersive(const Interface& hero, const Interface& other) {
    const double a = hero.nBack / hero.nFront, b = other.nBack / other.nFront;
    return std::fabs(a - b) > 1e-12 * std::fabs(a);
}

double secondaryWeight(const Interface& it, V3 d, const Sample& s, bool radiance) {
    const Optics& o = *it.o;
    if (!s.delta) {
        double pdf;
        const double f = evalScatter(it, d, s.wi, radiance, pdf);
        return s.pdf > 0 ? f * std::fabs(dot(s.wi, it.n)) / s.pdf : 0.0;
    }
    const double cosI = std::fabs(dot(d, it.n));
    switch (o.type) {
        case OWE_SURF_DIELECTRIC: return s.weight;
        case OWE_SURF_MIRROR: return o.reflectance.eval(it.lambda);
        case OWE_SURF_CONDUCTOR: {
            const double nMed = dot(d, it.n) < 0 ? it.nFront : it.nBack;
            return fresnelConductor(cosI, conductorEta(o, it.lambda, nMed)) * o.reflectance.eval(it.lambda);
        }
        default: return 0;
    }
}

Spec4 sigmaT(const Medium& m, const Wavelengths& wl) {
    Spec4 r;
    for (int i = 0; i < NW; ++i) r[i] = m.absorption.eval(wl.lambda[i]) + m.scattering.eval(wl.lambda[i]);
    return r;
}

void maskTerminated(const Wavelengths& wl, Spec4& beta) {
    for (int i = 1; i < NW; ++i)
        if (wl.pdf[i] == 0) beta[i] = 0;
}

enum class MediumEvent { Pass, Scatter, End };

// sampleMedium (transport.cpp:95-113).
MediumEvent sampleMedium(const Medium& m, const Wavelengths& wl, double seg, Pcg& rng, Spec4& beta, double& t) {
    if (m.scatters()) {
        const Spec4 sT = sigmaT(m, wl);
        t = -std::log(1 - rng.uniform()) / sT[0];
        if (t < seg) {
            const double pdf0 = sT[0] * std::exp(-sT[0] * t);
            for (int i = 0; i < NW; ++i) beta[i] *= m.scattering.eval(wl.lambda[i]) * std::exp(-sT[i] * t) / pdf0;
            return MediumEvent::Scatter;
        }
        if (std::isinf(seg)) return MediumEvent::End;
        const double pass0 = std::exp(-sT[0] * seg);
        for (int i = 0; i < NW; ++i) beta[i] *= std::exp(-sT[i] * seg) / pass0;
    } else if (m.absorbs()) {
        if (std::isinf(seg)) return MediumEvent::End;
        for (int i = 0; i < NW; ++i) beta[i] *= std::exp(-m.absorption.eval(wl.lambda[i]) * seg);
    }
    return MediumEvent::Pass;
}

}  // namespace

struct Tracer::Path {
    Ray ray;
    uint32_t region;
    Spec4 L, beta = Spec4(1.0);
    bool specular = true;
    double prevPdf = 0;
    V3 prevP;
    // SPPM camera paths (transport.slang:749-762): where the visible point goes, and what decides it.
    VisiblePoint* vp = nullptr;
    int scatterEvents = 0, nonDeltaEvents = 0;
    // The hybrid partition (transport.slang:745-747): delta events since the last non-delta one,
    // whether the first event was non-delta; and whether a medium scattering vertex was met.
    int deltaSinceNonDelta = 0;
    bool firstIsNonDelta = false, mediumScattered = false;
    bool counts = true;  // this step's light counts (medium-only paths: after a medium scattering vertex)
    bool chromatic = false;  // an absorbing, non-scattering medium on the way
    double pathLength = 0;
    bool vpDone() const { return vp && vp->valid; }
    // noteScatter (transport.slang:783-791).
    void noteScatter(bool delta) {
        if (delta) {
            ++deltaSinceNonDelta;
        } else {
            if (scatterEvents == 0) firstIsNonDelta = true;
            ++nonDeltaEvents;
            deltaSinceNonDelta = 0;
        }
        ++scatterEvents;
    }
    // Russian roulette against the throughput (transport.cpp:369-374).
    bool survives(Pcg& rng) {
        const double q = std::min(1.0, beta.maxValue());
        if (rng.uniform() >= q) return false;
        beta *= 1 / q;
        return true;
    }
};

// Tracer::transmittance (transport.cpp:237-285) along dir from a start point: through null
// boundaries with Beer–Lambert on the way; zero at matter. `exclude` is the destination boundary
// (an emitter's own surface), which the segment ends on.
Spec4 Tracer::transmittance(V3 p, const Ray& start, uint32_t region, V3 dir, double dist, const Wavelengths& wl,
                            uint32_t* finalRegion, uint32_t exclude) const {
    Spec4 Tr(1.0);
    Ray ray = start;
    ray.o = p;
    ray.d = dir;
    double remaining = dist;
    for (int guard = 0; guard < 256; ++guard) {
        Hit h;
        bool found = w_.intersect(ray, remaining, h, region);
        while (found && h.boundary == exclude) found = false;  // the segment's own end
        const double seg = found ? std::min(h.t, remaining) : remaining;
        const Medium& m = w_.mediumOf(region);
        if (m.absorbs() || m.scatters()) {
            if (std::isinf(seg)) return Spec4(0.0);
            const Spec4 sT = sigmaT(m, wl);
            for (int i = 0; i < NW; ++i) Tr[i] *= std::exp(-sT[i] * seg);
        }
        if (!found) {
            if (std::isinf(dist) && region != 0) return Spec4(0.0);
            if (finalRegion) *finalRegion = region;
            return Tr;
        }
        const Boundary& b = w_.boundaries[h.boundary];
        const uint32_t incident = dot(dir, h.n) < 0 ? b.front : b.back;
        if (incident != region) return Spec4(0.0);
        if (w_.optics[size_t(b.optics)].type != OWE_SURF_NULL) return Spec4(0.0);
        region = dot(dir, h.n) < 0 ? b.back : b.front;
        ray = w_.spawn(h, dir);
        if (!std::isinf(remaining)) remaining -= h.t;
    }
    return Spec4(0.0);
}

// Tracer::directLighting with sampleLightFrom (transport.cpp:35-70, 323-348).
Spec4 Tracer::directLighting(const Hit& hit, bool surface, const Optics* o, double nFront, double nBack, V3 d,
                             const Wavelengths& wl, Pcg& rng, double g, uint32_t region) const {
    if (w_.lights.empty()) return Spec4(0.0);
    const size_t li = w_.pickLight(rng.uniform());
    const double sel = w_.lightPdf[li];
    const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
    const Light& L = w_.lights[li];
    V3 wi;
    double dist = kInf, lpdf = 0;
    Spec4 Le;
    uint32_t lightRegion = kNone, lightBoundary = kNone;
    if (L.kind == 1) {
        wi = Frame(w_.env.sunDir).toWorld(uniformCone(u1, u2, std::cos(w_.env.sunAngularRadius)));
        Le = wl.eval(w_.env.sunRadiance);
        lpdf = sel / w_.env.sunSolidAngle();
    } else {
        const Boundary& b = w_.boundaries[L.boundary];
        V3 pL, nL;
        double pdfW;
        if (!w_.sampleFrom(b, b.toLocal.point(hit.p), u1, u2, u3, pL, nL, pdfW)) return Spec4(0.0);
        const V3 ep = b.toWorld.point(pL), en = normalize(b.toWorld.vector(nL));
        const V3 dv = ep - hit.p;
        dist = length(dv);
        if (!(dist > 0)) return Spec4(0.0);
        wi = dv / dist;
        const Emission& e = w_.emissions[size_t(b.emission)];
        const double cosL = dot(en, -wi);
        Le = ((cosL > 0 && e.front) || (cosL < 0 && e.back)) ? wl.eval(e.radiance) : Spec4(0.0);
        lpdf = sel * pdfW;
        lightRegion = cosL > 0 ? b.front : b.back;
        lightBoundary = L.boundary;
    }
    if (Le.isZero() || lpdf <= 0) return Spec4(0.0);
    Spec4 f;
    double pdfScatter = 0, cosFactor = 1;
    uint32_t startRegion = region;
    Ray start;
    if (!surface) {
        pdfScatter = hgPhase(dot(-d, wi), g);
        f = Spec4(pdfScatter);
    } else {
        const Interface base{o, hit.n, hit.p, nFront, nBack, wl.lambda[0], s_.fresnelFloor};
        const Boundary& b = w_.boundaries[hit.boundary];
        for (int i = 0; i < NW; ++i) {
            if (wl.pdf[i] == 0) continue;
            Interface it = base;
            if (i > 0) {
                it.nFront = w_.indexOf(b.front, wl.lambda[i]);
                it.nBack = w_.indexOf(b.back, wl.lambda[i]);
                it.lambda = wl.lambda[i];
            }
            double pdf;
            f[i] = evalScatter(it, d, wi, kRadiance, pdf);
            if (i == 0) pdfScatter = pdf;
        }
        if (f.isZero()) return Spec4(0.0);
        cosFactor = std::fabs(dot(wi, hit.n));
        startRegion = dot(wi, hit.n) > 0 ? b.front : b.back;
        start.departed = hit.boundary;
        start.departedPrim = hit.prim;
    }
    Spec4 Tr;
    if (std::isfinite(dist)) {
        uint32_t arrival = kNone;
        Tr = transmittance(hit.p, start, startRegion, wi, dist, wl, &arrival, lightBoundary);
        if (arrival != lightRegion) Tr = Spec4(0.0);
    } else {
        Tr = transmittance(hit.p, start, startRegion, wi, kInf, wl, nullptr, kNone);
    }
    if (Tr.isZero()) return Spec4(0.0);
    const double w = powerHeuristic(lpdf, pdfScatter);
    return f * Tr * Le * (cosFactor * w / lpdf);
}

// Tracer::scatterAtSurface (transport.cpp:437-471).
bool Tracer::scatterAtSurface(const Hit& hit, const Optics& o, uint32_t other, int depth, Wavelengths& wl, Pcg& rng,
                              Path& path) const {
    const Boundary& b = w_.boundaries[hit.boundary];
    const double hero = wl.lambda[0];
    Interface it{&o, hit.n, hit.p, w_.indexOf(b.front, hero), w_.indexOf(b.back, hero), hero, s_.fresnelFloor};
    auto at = [&](int i) {
        Interface r = it;
        r.nFront = w_.indexOf(b.front, wl.lambda[i]);
        r.nBack = w_.indexOf(b.back, wl.lambda[i]);
        r.lambda = wl.lambda[i];
        return r;
    };
    // At a dispersive dielectric the secondary wavelengths end; the hero carries the path.
    if (o.type == OWE_SURF_DIELECTRIC && !wl.secondaryTerminated())
        for (int i = 1; i < NW; ++i)
            if (wl.pdf[i] != 0 && isDispersive(it, at(i))) {
                wl.terminateSecondary();
                maskTerminated(wl, path.beta);
                break;
            }
    if (path.vp && !o.isDelta() && path.nonDeltaEvents == 0) {
        // recordVisiblePoint (transport.slang:846-858).
        VisiblePoint& v = *path.vp;
        v.valid = true;
        v.p = hit.p;
        v.n = hit.n;
        v.boundary = hit.boundary;
        v.d = path.ray.d;
        v.beta = path.beta;
        v.terminated = wl.secondaryTerminated();
        v.free = path.scatterEvents == 0 && !path.chromatic;
        v.pathLength = path.pathLength;
    }
    if (!o.isDelta() && path.counts) {
        const Spec4 dl = directLighting(hit, true, &o, it.nFront, it.nBack, path.ray.d, wl, rng, 0, path.region);
        if (!dl.isZero()) path.L += path.beta * dl;
    }
    double u1, u2, u3;
    rng.draw3(u1, u2, u3);
    Sample s;
    const bool ok = sampleScatter(it, path.ray.d, u1, u2, u3, kRadiance, s);
    if (!ok || s.weight <= 0) return false;
    path.beta[0] *= s.weight;
    for (int i = 1; i < NW; ++i)
        if (wl.pdf[i] != 0) path.beta[i] *= secondaryWeight(at(i), path.ray.d, s, kRadiance);
    if (s.transmitted) path.region = other;
    path.specular = s.delta;
    path.prevPdf = s.pdf;
    path.prevP = hit.p;
    path.noteScatter(s.delta);
    path.ray = w_.spawn(hit, s.wi);
    if (depth >= s_.rrDepth && !path.survives(rng)) return false;
    return true;
}

Spec4 Tracer::radiance(Ray ray, uint32_t region, Wavelengths& wl, Pcg& rng, Stats& st, VisiblePoint* vp) const {
    Path path{ray, region, Spec4(), Spec4(1.0), true, 0.0, V3()};
    path.prevP = ray.o;
    path.vp = vp;
    maskTerminated(wl, path.beta);
    st.paths++;
    for (int depth = 0;; ++depth) {
        Hit hit;
        const bool found = w_.intersect(path.ray, kInf, hit, path.region);
        st.segments++;
        const Medium& med = w_.mediumOf(path.region);
        double t;
        const MediumEvent ev = sampleMedium(med, wl, found ? hit.t : kInf, rng, path.beta, t);
        if (ev == MediumEvent::End) break;
        if (ev == MediumEvent::Scatter) {
            // scatterInMedium (transport.cpp:379-397). Light an SPPM visible point's scatter segment
            // meets in a medium is the photons'.
            const V3 p = path.ray.o + path.ray.d * t;
            if (path.vpDone()) break;
            if (depth >= s_.maxDepth) break;
            path.mediumScattered = true;
            Hit mh;
            mh.p = p;
            const Spec4 dl = directLighting(mh, false, nullptr, 1, 1, path.ray.d, wl, rng, med.g, path.region);
            if (!dl.isZero()) path.L += path.beta * dl;
            double u1, u2;
            rng.draw2(u1, u2);
            V3 wi;
            path.prevPdf = sampleHg(-path.ray.d, med.g, u1, u2, wi);
            path.specular = false;
            path.prevP = p;
            path.noteScatter(false);
            path.ray = Ray{p, wi, kNone, kNone};
            if (!(depth < s_.rrDepth || path.survives(rng))) break;
            continue;
        }
        if (path.vp && med.absorbs() && !med.scatters()) path.chromatic = true;
        // Light this path may count (stepPath, transport.slang:868-891): under the hybrid partition
        // not the paths light tracing owns; medium-only, only after a medium scattering vertex.
        const bool owned = s_.partition && path.nonDeltaEvents == 1 && path.firstIsNonDelta && path.deltaSinceNonDelta >= 1;
        const bool counts = !s_.mediumOnly || path.mediumScattered;
        path.counts = counts;
        if (!found) {
            if (path.region == 0) {
                if (owned || !counts) break;
                // skyMet (transport.cpp:401-417): the uniform sky, and the sun weighted against NEE.
                Spec4 sky, sun;
                for (int i = 0; i < NW; ++i) {
                    sky[i] = w_.env.skyAt(wl.lambda[i]);
                    sun[i] = w_.env.sunAt(path.ray.d, wl.lambda[i]);
                }
                double w = 1;
                if (!path.specular && !sun.isZero()) {
                    const double sunPdf = w_.env.hasSun ? w_.lightPdf.back() / w_.env.sunSolidAngle() : 0;
                    w = powerHeuristic(path.prevPdf, sunPdf);
                }
                path.L += path.beta * (sky + sun * w);
            } else {
                st.leaks++;
            }
            break;
        }
        const Boundary& b = w_.boundaries[hit.boundary];
        const bool fromFront = dot(path.ray.d, hit.n) < 0;
        const uint32_t inc = fromFront ? b.front : b.back, other = fromFront ? b.back : b.front;
        if (inc != path.region) {
            st.inconsistencies++;
            path.region = inc;
        }
        if (path.vp && !path.vpDone()) path.pathLength += hit.t;
        const Optics& o = w_.optics[size_t(b.optics)];
        // emissionMet (transport.cpp:420-433).
        if (b.emission >= 0 && !owned && counts) {
            const Emission& e = w_.emissions[size_t(b.emission)];
            if (fromFront ? e.front : e.back) {
                double w = 1;
                if (!path.specular) {
                    const int li = w_.lightOfBoundary[hit.boundary];
                    const double pl = li < 0 ? 0
                                             : w_.lightPdf[size_t(li)] *
                                                   w_.pdfFrom(b, b.toLocal.point(path.prevP), hit.pLocal, hit.nLocal);
                    if (pl > 0) w = powerHeuristic(path.prevPdf, pl);
                }
                path.L += path.beta * wl.eval(e.radiance) * w;
            }
        }
        if (o.type == OWE_SURF_NULL) {
            path.region = other;
            path.ray = w_.spawn(hit, path.ray.d);
            if (depth >= 4 * s_.maxDepth) break;
            continue;
        }
        if (path.vpDone()) break;  // the visible point's scatter segment has met matter
        if (o.type == OWE_SURF_ABSORBER || o.type == OWE_SURF_DETECTOR) break;
        if (depth >= s_.maxDepth) break;
        if (!scatterAtSurface(hit, o, other, depth, wl, rng, path)) break;
    }
    return path.L;
}

// One sample of cpu_backend.cpp:69-84: wavelengths, position in the pixel, filter offset, ray.
bool Tracer::cameraSample(int x, int y, int s, int spp, double rot, Pcg& rng, Stats& st, double xyz[3]) const {
    xyz[0] = xyz[1] = xyz[2] = 0;
    double u = (s + rng.uniform()) / spp + rot;
    u -= std::floor(u);
    Wavelengths wl = s_.spectral == OWE_SPECTRAL_HERO4 ? Wavelengths::hero(u)
                                                        : Wavelengths::uniform1(u, s_.lambdaMin, s_.lambdaMax);
    const double px = x + rng.uniform(), py = y + rng.uniform();
    double fx = 0, fy = 0;
    const Observer& ob = w_.observer;
    if (ob.sigma > 0) {
        double u1, u2;
        rng.draw2(u1, u2);
        standardNormal2(u1, u2, fx, fy);
        fx *= ob.sigma;
        fy *= ob.sigma;
    }
    Ray ray;
    double weight;
    if (!ob.generate(px + fx, py + fy, rng, ray, weight) || weight <= 0) return false;
    const Spec4 L = radiance(ray, ob.region, wl, rng, st);
    if (L.isZero()) return false;
    wl.toXYZ(L * weight, xyz);
    return std::isfinite(xyz[0] + xyz[1] + xyz[2]);
}

// ---------------------------------------------------------------- particles (light tracing)

// IdealObserver::connect as the particle kernel does it (particle.slang:59-95): a pupil point q,
// the pixel, and the factor W with which the pixel receives β·f·|cos|·W·Tr / N.
struct Tracer::Link {
    V3 q, dir;
    int px = 0, py = 0;
    double factor = 0;
};

// A particle leaving an emitter (particle.slang:184-196).
struct Tracer::Emitted {
    Ray ray;
    uint32_t region = 0;
    Spec4 beta;
    bool distant = false;
    uint32_t guideCell = kNone;  // SPPM: the guide cell a sun photon of the uniform share teaches
    // VCM (Emitted, particle.slang:192-202): the pdf of the emitted position and direction with the
    // light's choice, next-event estimation's pdf of a distant direction, |cos| at an area emitter
    // (1 for distant light), the area emitter's light index, point and normal in its own frame.
    double emissionPdfW = 0, directPdfW = 0, usedCos = 1;
    size_t light = size_t(-1);
    V3 pL, nL;
};

void PhotonGuide::build() {
    total = 0;
    for (uint64_t c : counts) total += double(c);
    double run = 0;
    for (int i = 0; i < kCells; ++i) {
        run += double(counts[size_t(i)]);
        cdf[size_t(i)] = total > 0 ? float(run / total) : 0.0f;
    }
}

uint32_t PhotonGuide::sample(float u) const {
    uint32_t lo = 0, len = kCells;
    while (len > 0) {
        const uint32_t half = len / 2;
        if (cdf[lo + half] <= u) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return std::min<uint32_t>(lo, kCells - 1);
}

double PhotonGuide::pdf(double qx, double qy, uint32_t& cell) const {
    const double size = 2.0 / kRes;
    const int cx = int(std::floor((qx + 1) / size)), cy = int(std::floor((qy + 1) / size));
    const bool in = cx >= 0 && cy >= 0 && cx < kRes && cy < kRes;
    cell = in ? uint32_t(cy * kRes + cx) : kNone;
    const double uniform = qx * qx + qy * qy <= 1 ? 1 / kPi : 0;
    if (!(total > 0)) return uniform;
    const double p = in ? double(cdf[cell]) - (cell > 0 ? double(cdf[cell - 1]) : 0.0) : 0.0;
    return (in ? kShare * p / (size * size) : 0.0) + (1 - kShare) * uniform;
}

bool Tracer::linkObserver(V3 x, Pcg& rng, Link& k) const {
    const Observer& ob = w_.observer;
    k = Link{};
    k.q = ob.position;
    if (ob.pupil > 0) {
        double u1, u2, dx, dy;
        rng.draw2(u1, u2);
        concentricDisk(u1, u2, dx, dy);
        k.q = ob.position + (ob.right * dx + ob.up * dy) * ob.pupil;
    }
    const V3 w = x - k.q;
    const double dist2 = dot(w, w);
    if (dist2 <= 0) return false;
    const V3 om = w / std::sqrt(dist2);
    const double c = dot(om, ob.fwd);
    if (c <= 1e-9) return false;
    double sx, sy;
    if (ob.pupil > 0 && std::isfinite(ob.focus)) {
        const V3 F = (k.q - ob.position) + om * (ob.focus / c);
        sx = dot(F, ob.right) / ob.focus;
        sy = dot(F, ob.up) / ob.focus;
    } else {
        sx = dot(om, ob.right) / c;
        sy = dot(om, ob.up) / c;
    }
    double ox, oy;
    ob.filterOffset(rng, ox, oy);
    const double W = ob.width, H = ob.height;
    const double fx = (sx / ob.tanX + 1) * 0.5 * W + ox, fy = (1 - sy / ob.tanY) * 0.5 * H + oy;
    if (!(fx >= 0 && fy >= 0 && fx < W && fy < H)) return false;
    k.px = std::min(int(fx), ob.width - 1);
    k.py = std::min(int(fy), ob.height - 1);
    const double pixelArea = (2 * ob.tanX / W) * (2 * ob.tanY / H);
    k.factor = 1 / (dist2 * pixelArea * c * c * c);
    k.dir = normalize(k.q - x);
    return true;
}

void Tracer::deposit(Splat& out, int px, int py, const Spec4& v, const Wavelengths& wl) const {
    double xyz[3];
    wl.toXYZ(v, xyz);
    out.add(px, py, xyz);
}

// connectVertex (particle.slang:118-138): a surface vertex (its optics o at the hit) or a medium
// point (phase anisotropy g) joined to the observer; the connection's walk runs from the vertex.
void Tracer::connectVertex(const Hit& hit, bool surface, const Optics* o, V3 d, double g, uint32_t region,
                           const Spec4& beta, const Wavelengths& wl, Pcg& rng, Splat& out) const {
    Link k;
    if (!linkObserver(hit.p, rng, k)) return;
    Spec4 f;
    double cosX = 1;
    uint32_t start = region;
    Ray from;
    if (!surface) {
        f = Spec4(hgPhase(dot(-d, k.dir), g));
    } else {
        const Boundary& b = w_.boundaries[hit.boundary];
        for (int i = 0; i < NW; ++i) {
            if (wl.pdf[i] == 0) continue;
            const Interface it{o, hit.n, hit.p, w_.indexOf(b.front, wl.lambda[i]), w_.indexOf(b.back, wl.lambda[i]),
                               wl.lambda[i], s_.fresnelFloor};
            double pdf;
            f[i] = evalScatter(it, d, k.dir, !kRadiance, pdf);
        }
        if (f.isZero()) return;
        cosX = std::fabs(dot(k.dir, hit.n));
        start = dot(k.dir, hit.n) > 0 ? b.front : b.back;
        from.departed = hit.boundary;
        from.departedPrim = hit.prim;
    }
    uint32_t arrival = kNone;
    Spec4 Tr = transmittance(hit.p, from, start, k.dir, length(k.q - hit.p), wl, &arrival, kNone);
    if (arrival != w_.observer.region || Tr.isZero()) return;
    deposit(out, k.px, k.py, beta * f * Tr * (cosX * k.factor), wl);
}

// emitParticle with emitDistant / emitFromArea (particle.slang:234-344), without photon guiding.
bool Tracer::emitParticle(Wavelengths& wl, Pcg& rng, bool connect, Emitted& e, Splat& out,
                          const PhotonGuide* guide) const {
    e = Emitted{};
    const double pSky = w_.skyShare;
    const bool fromSky = rng.uniform() < pSky;
    size_t li = 0;
    if (!fromSky) {
        if (w_.lights.empty()) return false;
        li = w_.pickLight(rng.uniform());
    }
    if (fromSky || w_.lights[li].kind == 1) {
        // A direction, then a point of a disc of radius 1.001·R across the scene, facing it.
        const double sel = fromSky ? pSky : (1 - pSky) * w_.lightPdf[li];
        const double R = w_.sceneRadius() * 1.001;
        double u1, u2;
        rng.draw2(u1, u2);
        V3 toward;
        double pdfDir;
        Spec4 Le;
        if (fromSky) {
            toward = uniformSphere(u1, u2);
            pdfDir = 1 / (4 * kPi);
            for (int i = 0; i < NW; ++i) Le[i] = w_.env.skyAt(wl.lambda[i]);
        } else {
            toward = Frame(w_.env.sunDir).toWorld(uniformCone(u1, u2, std::cos(w_.env.sunAngularRadius)));
            pdfDir = 1 / w_.env.sunSolidAngle();
            Le = wl.eval(w_.env.sunRadiance);
        }
        const V3 dir = -toward;
        const Frame disc(dir);
        double v1, v2, qx, qy, pdfUnit = 1 / kPi;  // the point's density per unit-disc area
        if (guide && !fromSky && guide->total > 0) {
            // sampleEmissionDisc (particle.slang:205-232): the map's share, or uniform over the disc.
            const bool guided = rng.uniform() < PhotonGuide::kShare;
            rng.draw2(v1, v2);
            if (guided) {
                const uint32_t c = guide->sample(float(v1));
                qx = (double(c % PhotonGuide::kRes) + v2) * (2.0 / PhotonGuide::kRes) - 1;
                qy = (double(c / PhotonGuide::kRes) + rng.uniform()) * (2.0 / PhotonGuide::kRes) - 1;
            } else {
                concentricDisk(v1, v2, qx, qy);
            }
            uint32_t cell;
            pdfUnit = guide->pdf(qx, qy, cell);
            if (!(pdfUnit > 0)) return false;
            if (!guided) e.guideCell = cell;  // the uniform share teaches the map
        } else {
            rng.draw2(v1, v2);
            concentricDisk(v1, v2, qx, qy);
            if (guide && !fromSky) guide->pdf(qx, qy, e.guideCell);  // before the map has learnt: every sun photon
        }
        e.ray = Ray{w_.sceneCenter() - dir * (1.01 * R) + disc.s * (qx * R) + disc.t * (qy * R), dir, kNone, kNone};
        e.region = 0;
        e.emissionPdfW = sel * pdfDir * (pdfUnit / (R * R));
        e.beta = Le * (1 / e.emissionPdfW);
        e.distant = true;
        e.directPdfW = fromSky ? 0.0 : w_.lightPdf[li] / w_.env.sunSolidAngle();
    } else {
        const double sel = (1 - pSky) * w_.lightPdf[li];
        const uint32_t bi = w_.lights[li].boundary;
        const Boundary& b = w_.boundaries[bi];
        const Emission& em = w_.emissions[size_t(b.emission)];
        const Spec4 Le = wl.eval(em.radiance);
        double u1, u2, u3;
        rng.draw3(u1, u2, u3);
        V3 pL, nL;
        uint32_t prim;
        if (!w_.sampleArea(b, u1, u2, u3, pL, nL, prim)) return false;
        const V3 p = b.toWorld.point(pL), n = normalize(b.toWorld.vector(nL));
        const double pdfA = 1 / b.area();
        if (connect) {
            // areaEmitterToObserver (particle.slang:267-279): the emitter seen directly.
            Link k;
            if (linkObserver(p, rng, k)) {
                const double cosL = dot(n, k.dir);
                if ((cosL > 0 && em.front) || (cosL < 0 && em.back)) {
                    uint32_t arrival = kNone;
                    const Ray from{p, k.dir, bi, prim};
                    const Spec4 Tr = transmittance(p, from, cosL > 0 ? b.front : b.back, k.dir, length(k.q - p), wl,
                                                   &arrival, kNone);
                    if (arrival == w_.observer.region && !Tr.isZero())
                        deposit(out, k.px, k.py, Le * Tr * (std::fabs(cosL) * k.factor / (sel * pdfA)), wl);
                }
            }
        }
        double sidePdf = 1;
        bool front = em.front;
        if (em.front && em.back) {
            front = rng.uniform() < 0.5;
            sidePdf = 0.5;
        }
        double c1, c2;
        rng.draw2(c1, c2);
        const V3 dir = Frame(front ? n : -n).toWorld(cosineHemisphere(c1, c2));
        e.ray = Ray{p, dir, bi, prim};
        e.region = front ? b.front : b.back;
        e.beta = Le * (kPi / (sel * pdfA * sidePdf));
        e.usedCos = std::fabs(dot(n, dir));
        e.emissionPdfW = sel * pdfA * sidePdf * e.usedCos / kPi;
        e.light = li;
        e.pL = pL;
        e.nL = nL;
    }
    maskTerminated(wl, e.beta);
    return !e.beta.isZero();
}

// A light-tracing particle (its wavelengths from the first uniform of rng), or with `sink` an SPPM
// photon at the fixed wavelengths: no observer connections or sensor splats; a photon gathers at
// every non-specular surface after its first scattering event (sppm.slang:150-201).
uint32_t Tracer::particleWalk(const Wavelengths* fixed, Pcg& rng, Stats& st, Splat& out, PhotonSink* sink,
                              const PhotonGuide* guide) const {
    Wavelengths wl;
    if (fixed) {
        wl = *fixed;
    } else {
        const double u = rng.uniform();
        wl = s_.spectral == OWE_SPECTRAL_HERO4 ? Wavelengths::hero(u) : Wavelengths::uniform1(u, s_.lambdaMin, s_.lambdaMax);
    }
    const Observer& ob = w_.observer;
    const bool observer = !ob.sensor() && !sink;
    const bool sensor = ob.sensor() && !sink;
    // The hybrid partition (particle.slang:358-417): a particle connects to the eye only at its
    // first non-specular vertex and only after at least one specular event, and ends there.
    const bool partition = s_.partition && !sink;
    int events = 0, nonDelta = 0, deltas = 0;
    st.paths++;
    Emitted em;
    if (!emitParticle(wl, rng, observer && !partition, em, out, guide)) return kNone;
    Ray ray = em.ray;
    uint32_t region = em.region;
    Spec4 beta = em.beta;
    for (int depth = 0;; ++depth) {
        Hit hit;
        const bool found = w_.intersect(ray, kInf, hit, region);
        st.segments++;
        const Medium& med = w_.mediumOf(region);
        double t;
        const MediumEvent ev = sampleMedium(med, wl, found ? hit.t : kInf, rng, beta, t);
        if (ev == MediumEvent::End) break;
        const bool mayConnect = !partition || (nonDelta == 0 && deltas >= 1);
        if (ev == MediumEvent::Scatter) {
            Hit mh;
            mh.p = ray.o + ray.d * t;
            if (observer && mayConnect) connectVertex(mh, false, nullptr, ray.d, med.g, region, beta, wl, rng, out);
            ++nonDelta;
            if (partition) break;
            if (depth >= s_.maxDepth) break;
            double u1, u2;
            rng.draw2(u1, u2);
            V3 wi;
            sampleHg(-ray.d, med.g, u1, u2, wi);
            ray = Ray{mh.p, wi, kNone, kNone};
            ++events;
            if (depth >= s_.rrDepth) {
                if (rng.uniform() >= 0.95) break;
                beta *= 1 / 0.95;
            }
            continue;
        }
        if (!found) {
            if (region != 0) st.leaks++;
            break;
        }
        const Boundary& b = w_.boundaries[hit.boundary];
        const bool fromFront = dot(ray.d, hit.n) < 0;
        const uint32_t inc = fromFront ? b.front : b.back, other = fromFront ? b.back : b.front;
        if (inc != region) {
            st.inconsistencies++;
            region = inc;
        }
        const Optics& o = w_.optics[size_t(b.optics)];
        if (o.type == OWE_SURF_NULL) {
            region = other;
            ray = w_.spawn(hit, ray.d);
            if (depth >= 4 * s_.maxDepth) break;
            continue;
        }
        // The sensor records the flux reaching its front (sensorPixel, particle.slang:98-114).
        if (sensor && fromFront && hit.boundary == ob.boundary) {
            double fx = (hit.pLocal.x / ob.halfX + 1) * 0.5 * ob.width;
            double fy = (1 - hit.pLocal.y / ob.halfY) * 0.5 * ob.height;
            if (ob.flipX) fx = ob.width - fx;
            if (ob.flipY) fy = ob.height - fy;
            double ox, oy;
            ob.filterOffset(rng, ox, oy);
            fx += ox;
            fy += oy;
            const double area = (2 * ob.halfX / ob.width) * (2 * ob.halfY / ob.height);
            if (fx >= 0 && fy >= 0 && fx < ob.width && fy < ob.height)
                deposit(out, std::min(int(fx), ob.width - 1), std::min(int(fy), ob.height - 1), beta * (1 / area), wl);
        }
        if (o.type == OWE_SURF_DETECTOR || o.type == OWE_SURF_ABSORBER) break;
        if (!o.isDelta() && observer && mayConnect) connectVertex(hit, true, &o, ray.d, 0, region, beta, wl, rng, out);
        if (!o.isDelta() && sink && events >= 1) sink->gather(hit, ray.d, beta, wl);
        if (!o.isDelta()) {
            ++nonDelta;
            if (partition) break;
        }
        if (depth >= s_.maxDepth) break;
        // terminateIfDispersive, then the importance-mode scatter (transport.slang:639-684).
        Interface it{&o, hit.n, hit.p, w_.indexOf(b.front, wl.lambda[0]), w_.indexOf(b.back, wl.lambda[0]),
                     wl.lambda[0], s_.fresnelFloor};
        auto at = [&](int i) {
            Interface r = it;
            r.nFront = w_.indexOf(b.front, wl.lambda[i]);
            r.nBack = w_.indexOf(b.back, wl.lambda[i]);
            r.lambda = wl.lambda[i];
            return r;
        };
        if (o.type == OWE_SURF_DIELECTRIC && !wl.secondaryTerminated())
            for (int i = 1; i < NW; ++i)
                if (wl.pdf[i] != 0 && isDispersive(it, at(i))) {
                    wl.terminateSecondary();
                    maskTerminated(wl, beta);
                    break;
                }
        double u1, u2, u3;
        rng.draw3(u1, u2, u3);
        Sample s;
        if (!sampleScatter(it, ray.d, u1, u2, u3, !kRadiance, s) || s.weight <= 0) break;
        beta[0] *= s.weight;
        for (int i = 1; i < NW; ++i)
            if (wl.pdf[i] != 0) beta[i] *= secondaryWeight(at(i), ray.d, s, !kRadiance);
        if (s.transmitted) region = other;
        ray = w_.spawn(hit, s.wi);
        ++events;
        if (s.delta) ++deltas;
        // Particles carry flux: roulette against the event's own throughput.
        if (depth >= s_.rrDepth && s.weight < 1) {
            if (rng.uniform() >= s.weight) break;
            beta *= 1 / s.weight;
        }
    }
    return em.guideCell;
}

// ---------------------------------------------------------------- SPPM (sppm.slang)

bool Tracer::sppmCamera(int x, int y, Wavelengths wl, Pcg& rng, Stats& st, double xyz[3], VisiblePoint& vp) const {
    xyz[0] = xyz[1] = xyz[2] = 0;
    vp = VisiblePoint{};
    const double px = x + rng.uniform(), py = y + rng.uniform();
    double fx = 0, fy = 0;
    const Observer& ob = w_.observer;
    if (ob.sigma > 0) {
        double u1, u2;
        rng.draw2(u1, u2);
        standardNormal2(u1, u2, fx, fy);
        fx *= ob.sigma;
        fy *= ob.sigma;
    }
    Ray ray;
    double weight;
    if (!ob.generate(px + fx, py + fy, rng, ray, weight) || weight <= 0) return false;
    const Spec4 L = radiance(ray, ob.region, wl, rng, st, &vp);
    if (vp.valid) vp.beta *= weight;
    if (L.isZero()) return true;
    wl.toXYZ(L * weight, xyz);
    if (!std::isfinite(xyz[0] + xyz[1] + xyz[2])) xyz[0] = xyz[1] = xyz[2] = 0;
    return true;
}

uint32_t Tracer::sppmPhoton(Wavelengths wl, Pcg& rng, Stats& st, PhotonSink& sink, const PhotonGuide* guide) const {
    struct None final : Splat {
        void add(int, int, const double*) override {}
    } none;
    sink.useful = false;
    const uint32_t cell = particleWalk(&wl, rng, st, none, &sink, guide);
    return sink.useful ? cell : kNone;
}

bool Tracer::sppmValue(const VisiblePoint& v, int groups, V3 d, const Spec4& beta, const Wavelengths& wl,
                       double xyz[3]) const {
    Wavelengths w = wl;  // a free point is evaluated at the photon's wavelengths
    if (!v.free && v.terminated) w.terminateSecondary();
    const Boundary& b = w_.boundaries[v.boundary];
    const Optics& o = w_.optics[size_t(b.optics)];
    Spec4 value;
    for (int i = 0; i < NW; ++i) {
        if (w.pdf[i] == 0) continue;
        const Interface it{&o, v.n, v.p, w_.indexOf(b.front, w.lambda[i]), w_.indexOf(b.back, w.lambda[i]), w.lambda[i],
                           s_.fresnelFloor};
        double pdf;
        const double f = evalScatter(it, v.d, -d, kRadiance, pdf);
        value[i] = (v.free ? v.beta[0] : v.beta[i] * groups) * f * beta[i];
    }
    maskTerminated(w, value);
    if (value.isZero()) return false;
    w.toXYZ(value, xyz);
    return std::isfinite(xyz[0] + xyz[1] + xyz[2]);
}

// ---------------------------------------------------------------- vertex connection and merging (vcm.slang)

namespace {

// Georgiev's recursive MIS quantities in the balance heuristic (vcm.slang:150-178).
struct VcmMis {
    double dVCM = 0, dVC = 0, dVM = 0;
    void arrive(double dist, double cosIn) {
        dVCM *= dist * dist;
        dVCM /= cosIn;
        dVC /= cosIn;
        dVM /= cosIn;
    }
    void scatter(V3 n, const Sample& s, double pdfRev, double eta) {
        const double cosOut = std::fabs(dot(n, s.wi));
        if (s.delta) {
            dVCM = 0;
            dVC *= cosOut;
            dVM *= cosOut;
            return;
        }
        dVC = cosOut / s.pdf * (dVC * pdfRev + dVCM + eta);
        dVM = cosOut / s.pdf * (dVM * pdfRev + dVCM * (eta > 0 ? 1 / eta : 0.0) + 1);
        dVCM = 1 / s.pdf;
    }
};

// f per wavelength at the interface (hero `it`, boundary b) for arrival d and departure wi, with the
// hero's pdf.
Spec4 vcmEval(const World& w, const Interface& it, const Boundary& b, const Wavelengths& wl, V3 d, V3 wi, bool radiance,
              double& pdf) {
    Spec4 f(0.0);
    pdf = 0;
    for (int i = 0; i < NW; ++i) {
        if (wl.pdf[i] == 0) continue;
        Interface a = it;
        if (i > 0) {
            a.nFront = w.indexOf(b.front, wl.lambda[i]);
            a.nBack = w.indexOf(b.back, wl.lambda[i]);
            a.lambda = wl.lambda[i];
        }
        double q;
        f[i] = evalScatter(a, d, wi, radiance, q);
        if (i == 0) pdf = q;
    }
    return f;
}

double reversePdf(const Interface& it, V3 arriving, V3 leaving) {
    double pdf;
    evalScatter(it, arriving, leaving, kRadiance, pdf);
    return pdf;
}

// XYZ of a contribution at the camera path's wavelengths, with only the hero when the light side left
// only its hero (joinedXYZ, vcm.slang:319-324).
void joinedXYZ(Wavelengths wl, bool lightTerminated, const Spec4& v, double xyz[3]) {
    if (lightTerminated) wl.terminateSecondary();
    Spec4 m = v;
    maskTerminated(wl, m);
    wl.toXYZ(m, xyz);
}

uint32_t vcmPixelGroup(uint32_t pixel, uint32_t iteration, uint32_t groups);

}  // namespace

double Tracer::vcmEta(const VcmParams& p, V3 x) const {
    if (!(p.mergePerMetre > 0)) return 0;
    const double r = std::max(p.mergePerMetre * length(x - w_.observer.position), 1e-9);
    return kPi * r * r * p.lightPaths;
}

void Tracer::vcmLight(Wavelengths wl, Pcg& rng, Stats& st, Splat& out, const VcmParams& p, uint32_t group,
                      std::vector<VcmVertex>& cache) const {
    st.paths++;
    Emitted em;
    if (!emitParticle(wl, rng, false, em, out, nullptr) || !(em.emissionPdfW > 0)) return;
    Ray ray = em.ray;
    uint32_t region = em.region;
    Spec4 beta = em.beta;
    VcmMis mis;
    double segment = 0;
    bool first = true;
    const int W = w_.observer.width;
    for (int depth = 0;; ++depth) {
        Hit hit;
        const bool found = w_.intersect(ray, kInf, hit, region);
        st.segments++;
        const Medium& med = w_.mediumOf(region);
        if (med.absorbs() || med.scatters()) {  // extinction only: scattered light is the medium-only pass's
            if (!found) return;
            const Spec4 sT = sigmaT(med, wl);
            for (int i = 0; i < NW; ++i) beta[i] *= std::exp(-sT[i] * hit.t);
        }
        if (!found) {
            if (region != 0) st.leaks++;
            return;
        }
        const Boundary& b = w_.boundaries[hit.boundary];
        const bool fromFront = dot(ray.d, hit.n) < 0;
        const uint32_t inc = fromFront ? b.front : b.back, other = fromFront ? b.back : b.front;
        if (inc != region) {
            st.inconsistencies++;
            region = inc;
        }
        segment += hit.t;
        const Optics& o = w_.optics[size_t(b.optics)];
        if (o.type == OWE_SURF_NULL) {
            region = other;
            ray = w_.spawn(hit, ray.d);
            if (depth >= 4 * s_.maxDepth) return;
            continue;
        }
        if (o.type == OWE_SURF_DETECTOR || o.type == OWE_SURF_ABSORBER) return;
        const V3 d = ray.d;
        if (first) {
            // firstLightVertexMis (vcm.slang:185-197).
            mis.dVC = em.usedCos / em.emissionPdfW;
            if (em.distant) {
                mis.dVCM = em.directPdfW / em.emissionPdfW;
            } else {
                const Boundary& lb = w_.boundaries[w_.lights[em.light].boundary];
                const double directPdfW = w_.lightPdf[em.light] * w_.pdfFrom(lb, lb.toLocal.point(hit.p), em.pL, em.nL);
                mis.dVCM = directPdfW * em.usedCos / em.emissionPdfW;
            }
            const double eta = vcmEta(p, hit.p);
            mis.dVM = mis.dVC * (eta > 0 ? 1 / eta : 0.0);
        }
        mis.arrive(first ? 1 : segment, std::fabs(dot(hit.n, d)));
        first = false;
        segment = 0;
        const Interface it{&o, hit.n, hit.p, w_.indexOf(b.front, wl.lambda[0]), w_.indexOf(b.back, wl.lambda[0]),
                           wl.lambda[0], s_.fresnelFloor};
        if (!o.isDelta()) {
            VcmVertex y;
            y.p = hit.p;
            y.n = hit.n;
            y.d = d;
            y.beta = beta;
            y.dVCM = mis.dVCM;
            y.dVC = mis.dVC;
            y.dVM = mis.dVM;
            y.boundary = hit.boundary;
            y.terminated = wl.secondaryTerminated();
            cache.push_back(y);
            // lightVertexToEye (vcm.slang:229-244): only to the pixels of this group, scaled by G.
            Link k;
            if (linkObserver(hit.p, rng, k) &&
                vcmPixelGroup(uint32_t(k.py) * uint32_t(W) + uint32_t(k.px), p.iteration, p.groups) == group) {
                double pdfFwd;
                const Spec4 f = vcmEval(w_, it, b, wl, d, k.dir, !kRadiance, pdfFwd);
                if (!f.isZero()) {
                    const double pdfRev = reversePdf(it, -k.dir, -d);
                    const double cosX = std::fabs(dot(hit.n, k.dir));
                    const double wLight = k.factor * cosX / p.lightPaths * (vcmEta(p, hit.p) + mis.dVCM + mis.dVC * pdfRev);
                    const double wgt = 1 / (wLight + 1);
                    uint32_t arrival = kNone;
                    const Ray from{hit.p, k.dir, hit.boundary, hit.prim};
                    const Spec4 Tr = transmittance(hit.p, from, dot(k.dir, hit.n) > 0 ? b.front : b.back, k.dir,
                                                   length(k.q - hit.p), wl, &arrival, kNone);
                    if (arrival == w_.observer.region && !Tr.isZero())
                        deposit(out, k.px, k.py, beta * f * Tr * (cosX * k.factor * wgt * double(p.groups)), wl);
                }
            }
        }
        if (depth >= s_.maxDepth) return;
        auto at = [&](int i) {
            Interface r = it;
            r.nFront = w_.indexOf(b.front, wl.lambda[i]);
            r.nBack = w_.indexOf(b.back, wl.lambda[i]);
            r.lambda = wl.lambda[i];
            return r;
        };
        if (o.type == OWE_SURF_DIELECTRIC && !wl.secondaryTerminated())
            for (int i = 1; i < NW; ++i)
                if (wl.pdf[i] != 0 && isDispersive(it, at(i))) {
                    wl.terminateSecondary();
                    maskTerminated(wl, beta);
                    break;
                }
        double u1, u2, u3;
        rng.draw3(u1, u2, u3);
        Sample s;
        if (!sampleScatter(it, d, u1, u2, u3, !kRadiance, s) || s.weight <= 0) return;
        beta[0] *= s.weight;
        for (int i = 1; i < NW; ++i)
            if (wl.pdf[i] != 0) beta[i] *= secondaryWeight(at(i), d, s, !kRadiance);
        const double pdfRev = s.delta ? 0.0 : reversePdf(it, -s.wi, -d);
        mis.scatter(hit.n, s, pdfRev, vcmEta(p, hit.p));
        if (s.transmitted) region = other;
        ray = w_.spawn(hit, s.wi);
        if (depth >= s_.rrDepth && s.weight < 1) {
            if (rng.uniform() >= s.weight) return;
            beta *= 1 / s.weight;
        }
    }
}

bool Tracer::vcmCamera(int x, int y, Wavelengths wl, Pcg& rng, Stats& st, const VcmParams& p, const VcmVertex* cache,
                       uint32_t M, double xyz[3], const VcmSweep& sweep) const {
    xyz[0] = xyz[1] = xyz[2] = 0;
    const Observer& ob = w_.observer;
    const double px = x + rng.uniform(), py = y + rng.uniform();
    double fx = 0, fy = 0;
    if (ob.sigma > 0) {
        double u1, u2;
        rng.draw2(u1, u2);
        standardNormal2(u1, u2, fx, fy);
        fx *= ob.sigma;
        fy *= ob.sigma;
    }
    Ray ray;
    double weight;
    if (!ob.generate(px + fx, py + fy, rng, ray, weight) || weight <= 0) return false;
    const double cosCam = dot(ray.d, ob.fwd);
    if (!(cosCam > 0)) return false;
    st.paths++;
    const double pixelArea = (2 * ob.tanX / ob.width) * (2 * ob.tanY / ob.height);
    const double cameraPdfW = 1 / (pixelArea * cosCam * cosCam * cosCam);
    VcmMis mis;
    mis.dVCM = p.lightPaths / cameraPdfW;
    Spec4 beta(weight);
    maskTerminated(wl, beta);
    uint32_t region = ob.region;
    double segment = 0;
    V3 prevP = ray.o;
    int pathLength = 0;
    double L[3] = {0, 0, 0};
    auto add = [&](bool lightTerminated, const Spec4& v) {
        double c[3];
        joinedXYZ(wl, lightTerminated, v, c);
        for (int k = 0; k < 3; ++k) L[k] += c[k];
    };
    const double R = w_.sceneRadius() * 1.001;
    double sunLightPdf = 0;
    for (size_t li = 0; li < w_.lights.size(); ++li)
        if (w_.lights[li].kind == 1) sunLightPdf = w_.lightPdf[li];
    for (int depth = 0;; ++depth) {
        Hit hit;
        const bool found = w_.intersect(ray, kInf, hit, region);
        st.segments++;
        const Medium& med = w_.mediumOf(region);
        if (med.absorbs() || med.scatters()) {
            if (!found) break;
            const Spec4 sT = sigmaT(med, wl);
            for (int i = 0; i < NW; ++i) beta[i] *= std::exp(-sT[i] * hit.t);
        }
        if (!found) {
            if (region == 0) {
                // cameraSkyMet (vcm.slang:360-376).
                double wSky = 1, wSun = 1;
                const bool inSun = w_.env.hasSun && dot(ray.d, w_.env.sunDir) >= std::cos(w_.env.sunAngularRadius);
                if (pathLength > 0) {
                    wSky = 1 / (1 + w_.skyShare / (4 * kPi) / (kPi * R * R) * mis.dVC);
                    if (inSun && w_.env.sunSolidAngle() > 0)
                        wSun = 1 / (1 + sunLightPdf / w_.env.sunSolidAngle() * mis.dVCM +
                                    (1 - w_.skyShare) * sunLightPdf / w_.env.sunSolidAngle() / (kPi * R * R) * mis.dVC);
                }
                Spec4 v;
                for (int i = 0; i < NW; ++i) {
                    if (wl.pdf[i] == 0) continue;
                    const double sky = w_.env.skyAt(wl.lambda[i]);
                    const double sun = inSun ? w_.env.sunRadiance.eval(wl.lambda[i]) : 0.0;
                    v[i] = beta[i] * (sky * wSky + sun * wSun);
                }
                add(false, v);
            } else {
                st.leaks++;
            }
            break;
        }
        const Boundary& b = w_.boundaries[hit.boundary];
        const bool fromFront = dot(ray.d, hit.n) < 0;
        const uint32_t inc = fromFront ? b.front : b.back, other = fromFront ? b.back : b.front;
        if (inc != region) {
            st.inconsistencies++;
            region = inc;
        }
        segment += hit.t;
        const Optics& o = w_.optics[size_t(b.optics)];
        if (o.type == OWE_SURF_NULL) {
            region = other;
            ray = w_.spawn(hit, ray.d);
            if (depth >= 4 * s_.maxDepth) break;
            continue;
        }
        const V3 d = ray.d;
        const double cosIn = std::fabs(dot(hit.n, d));
        const double dist = segment;
        mis.arrive(dist, cosIn);
        segment = 0;
        ++pathLength;
        // cameraEmissionMet (vcm.slang:380-399).
        if (b.emission >= 0) {
            const Emission& e = w_.emissions[size_t(b.emission)];
            if (fromFront ? e.front : e.back) {
                double wgt = 1;
                const int li = w_.lightOfBoundary[hit.boundary];
                if (pathLength > 1 && li >= 0) {
                    const double directPdfA =
                        w_.lightPdf[size_t(li)] * w_.pdfFrom(b, b.toLocal.point(prevP), hit.pLocal, hit.nLocal) * cosIn / (dist * dist);
                    const double side = e.front && e.back ? 0.5 : 1.0;
                    const double emissionPdfW = (1 - w_.skyShare) * w_.lightPdf[size_t(li)] / b.area() * side * cosIn / kPi;
                    wgt = 1 / (1 + directPdfA * mis.dVCM + emissionPdfW * mis.dVC);
                }
                add(false, beta * wl.eval(e.radiance) * wgt);
            }
        }
        if (o.type == OWE_SURF_ABSORBER || o.type == OWE_SURF_DETECTOR) break;
        if (depth >= s_.maxDepth) break;
        const Interface it{&o, hit.n, hit.p, w_.indexOf(b.front, wl.lambda[0]), w_.indexOf(b.back, wl.lambda[0]),
                           wl.lambda[0], s_.fresnelFloor};
        auto at = [&](int i) {
            Interface r = it;
            r.nFront = w_.indexOf(b.front, wl.lambda[i]);
            r.nBack = w_.indexOf(b.back, wl.lambda[i]);
            r.lambda = wl.lambda[i];
            return r;
        };
        if (o.type == OWE_SURF_DIELECTRIC && !wl.secondaryTerminated())
            for (int i = 1; i < NW; ++i)
                if (wl.pdf[i] != 0 && isDispersive(it, at(i))) {
                    wl.terminateSecondary();
                    maskTerminated(wl, beta);
                    break;
                }
        const double eta = vcmEta(p, hit.p);
        if (!o.isDelta()) {
            // Next-event estimation (cameraNextEvent, vcm.slang:402-425).
            if (!w_.lights.empty()) {
                const size_t li = w_.pickLight(rng.uniform());
                const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
                const Light& Lt = w_.lights[li];
                V3 wi;
                double ldist = kInf, lpdf = 0, cosAtLight = 1, emissionPdf = 0;
                Spec4 Le;
                uint32_t arrival = kNone, exclude = kNone;
                bool ok = true;
                if (Lt.kind == 1) {
                    wi = Frame(w_.env.sunDir).toWorld(uniformCone(u1, u2, std::cos(w_.env.sunAngularRadius)));
                    Le = wl.eval(w_.env.sunRadiance);
                    lpdf = w_.lightPdf[li] / w_.env.sunSolidAngle();
                    emissionPdf = (1 - w_.skyShare) * w_.lightPdf[li] / w_.env.sunSolidAngle() / (kPi * R * R);
                } else {
                    const Boundary& lb = w_.boundaries[Lt.boundary];
                    V3 pL, nL;
                    double pdfW;
                    ok = w_.sampleFrom(lb, lb.toLocal.point(hit.p), u1, u2, u3, pL, nL, pdfW);
                    if (ok) {
                        const V3 ep = lb.toWorld.point(pL), en = normalize(lb.toWorld.vector(nL));
                        const V3 dv = ep - hit.p;
                        ldist = length(dv);
                        ok = ldist > 0;
                        if (ok) {
                            wi = dv / ldist;
                            const Emission& e = w_.emissions[size_t(lb.emission)];
                            const double cosL = dot(en, -wi);
                            Le = ((cosL > 0 && e.front) || (cosL < 0 && e.back)) ? wl.eval(e.radiance) : Spec4(0.0);
                            lpdf = w_.lightPdf[li] * pdfW;
                            cosAtLight = std::fabs(cosL);
                            const double side = e.front && e.back ? 0.5 : 1.0;
                            emissionPdf = (1 - w_.skyShare) * w_.lightPdf[li] / lb.area() * side * std::fabs(cosL) / kPi;
                            arrival = cosL > 0 ? lb.front : lb.back;
                            exclude = Lt.boundary;
                        }
                    }
                }
                maskTerminated(wl, Le);
                if (ok && lpdf > 0 && cosAtLight > 0 && !Le.isZero()) {
                    double pdfFwd;
                    const Spec4 f = vcmEval(w_, it, b, wl, d, wi, kRadiance, pdfFwd);
                    if (!f.isZero()) {
                        const double pdfRev = reversePdf(it, -wi, -d);
                        const double cosToLight = std::fabs(dot(hit.n, wi));
                        const double wLight = pdfFwd / lpdf;
                        const double wCamera = emissionPdf * cosToLight / (lpdf * cosAtLight) * (eta + mis.dVCM + mis.dVC * pdfRev);
                        const double wgt = 1 / (wLight + 1 + wCamera);
                        const Ray from{hit.p, wi, hit.boundary, hit.prim};
                        const uint32_t start = dot(wi, hit.n) > 0 ? b.front : b.back;
                        uint32_t reached = kNone;
                        Spec4 Tr = transmittance(hit.p, from, start, wi, ldist, wl, std::isfinite(ldist) ? &reached : nullptr,
                                                 exclude);
                        if (std::isfinite(ldist) && reached != arrival) Tr = Spec4(0.0);
                        if (!Tr.isZero()) add(false, beta * f * Le * Tr * (cosToLight * wgt / lpdf));
                    }
                }
            }
            // A connection to one cached vertex drawn uniformly (cameraConnect, vcm.slang:428-462).
            if (M > 0) {
                // The pick as the float path forms it: the uniform's leading 24 bits times M, rounded in
                // single precision, so that both draw the same vertex.
                const float u24 = float(std::floor(rng.uniform() * 16777216.0) / 16777216.0);
                const uint32_t pick = std::min(uint32_t(u24 * float(M)), M - 1);
                const VcmVertex& yv = cache[pick];
                const V3 dv = yv.p - hit.p;
                const double dist2 = dot(dv, dv);
                if (dist2 > 0) {
                    const double dl = std::sqrt(dist2);
                    const V3 dir = dv / dl;
                    Wavelengths wj = wl;
                    if (yv.terminated) wj.terminateSecondary();
                    double pdfCamFwd, pdfLightFwd;
                    const Spec4 fc = vcmEval(w_, it, b, wj, d, dir, kRadiance, pdfCamFwd);
                    const Boundary& lb = w_.boundaries[yv.boundary];
                    const Optics& lo = w_.optics[size_t(lb.optics)];
                    const Interface hl{&lo, yv.n, yv.p, w_.indexOf(lb.front, wj.lambda[0]), w_.indexOf(lb.back, wj.lambda[0]),
                                       wj.lambda[0], s_.fresnelFloor};
                    const Spec4 fl = vcmEval(w_, hl, lb, wj, yv.d, -dir, !kRadiance, pdfLightFwd);
                    if (!(fc * fl).isZero()) {
                        const double pdfCamRev = reversePdf(it, -dir, -d), pdfLightRev = reversePdf(hl, dir, -yv.d);
                        const double cosX = std::fabs(dot(hit.n, dir)), cosY = std::fabs(dot(yv.n, dir));
                        const double camPdfA = pdfCamFwd * cosY / dist2, lightPdfA = pdfLightFwd * cosX / dist2;
                        const double wLight = camPdfA * (vcmEta(p, yv.p) + yv.dVCM + yv.dVC * pdfLightRev);
                        const double wCamera = lightPdfA * (eta + mis.dVCM + mis.dVC * pdfCamRev);
                        const double wgt = 1 / (wLight + 1 + wCamera);
                        const Ray from{hit.p, dir, hit.boundary, hit.prim};
                        uint32_t reached = kNone;
                        Spec4 Tr = transmittance(hit.p, from, dot(dir, hit.n) > 0 ? b.front : b.back, dir, dl * (1 - 1e-5), wj,
                                                 &reached, kNone);
                        if (reached != (dot(-dir, yv.n) > 0 ? lb.front : lb.back)) Tr = Spec4(0.0);
                        if (!Tr.isZero())
                            add(yv.terminated, beta * fc * yv.beta * fl * Tr * (cosX * cosY / dist2 * wgt * double(M) / p.lightPaths));
                    }
                }
            }
            // Merging with the group's vertices within the merging radius (cameraMerge, vcm.slang:465-520).
            if (p.mergePerMetre > 0) {
                const double r = std::max(p.mergePerMetre * length(hit.p - ob.position), 1e-9), r2 = r * r;
                const double invEta = 1 / eta;
                uint32_t lo = 0, hi = M;
                if (sweep.x) {
                    lo = uint32_t(std::lower_bound(sweep.x, sweep.x + M, hit.p.x - r) - sweep.x);
                    hi = uint32_t(std::upper_bound(sweep.x, sweep.x + M, hit.p.x + r) - sweep.x);
                }
                for (uint32_t m = lo; m < hi; ++m) {
                    const VcmVertex& yv = cache[sweep.x ? sweep.index[m] : m];
                    const V3 dp = yv.p - hit.p;
                    if (dot(dp, dp) > r2) continue;
                    if (std::fabs(dot(dp, hit.n)) > 0.25 * r || dot(yv.n, hit.n) < 0.5) continue;
                    const V3 wi = -yv.d;
                    Wavelengths wj = wl;
                    if (yv.terminated) wj.terminateSecondary();
                    double pdfFwd;
                    const Spec4 fc = vcmEval(w_, it, b, wj, d, wi, kRadiance, pdfFwd);
                    if (fc.isZero()) continue;
                    const double pdfRev = reversePdf(it, -wi, -d);
                    const double wLight = yv.dVCM * invEta + yv.dVM * pdfFwd;
                    const double wCamera = mis.dVCM * invEta + mis.dVM * pdfRev;
                    const double wgt = 1 / (wLight + 1 + wCamera);
                    add(yv.terminated, beta * fc * yv.beta * (wgt * invEta));
                }
            }
        }
        double u1, u2, u3;
        rng.draw3(u1, u2, u3);
        Sample s;
        if (!sampleScatter(it, d, u1, u2, u3, kRadiance, s) || s.weight <= 0) break;
        beta[0] *= s.weight;
        for (int i = 1; i < NW; ++i)
            if (wl.pdf[i] != 0) beta[i] *= secondaryWeight(at(i), d, s, kRadiance);
        const double pdfRev = s.delta ? 0.0 : reversePdf(it, -s.wi, -d);
        mis.scatter(hit.n, s, pdfRev, eta);
        if (s.transmitted) region = other;
        prevP = hit.p;
        ray = w_.spawn(hit, s.wi);
        if (depth >= s_.rrDepth) {
            double q = 0;
            for (int i = 0; i < NW; ++i) q = std::max(q, beta[i]);
            q = std::min(1.0, q);
            if (rng.uniform() >= q) break;
            beta *= 1 / q;
        }
    }
    if (std::isfinite(L[0] + L[1] + L[2]))
        for (int k = 0; k < 3; ++k) xyz[k] = L[k];
    return true;
}

namespace {
uint32_t vcmPixelGroup(uint32_t pixel, uint32_t iteration, uint32_t groups) {
    return uint32_t(mix64((uint64_t(iteration) << 32 | pixel) ^ 0x1f83d9abfb41bd6bULL) % uint64_t(groups));
}
}  // namespace

}  // namespace oracle
