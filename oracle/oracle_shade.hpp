// Oracle shading: FP64 restatement of the reference's scattering, material,
// texture and light code. Test infrastructure only (see oracle_core.hpp).
#pragma once

#include "oracle_core.hpp"

#include <complex>
#include <optional>

namespace oracle {

struct Rgb {
    double r{0}, g{0}, b{0};
};
inline Rgb operator+(Rgb a, Rgb b) { return {a.r + b.r, a.g + b.g, a.b + b.b}; }
inline Rgb operator*(Rgb a, Rgb b) { return {a.r * b.r, a.g * b.g, a.b * b.b}; }
inline Rgb operator*(Rgb a, double s) { return {a.r * s, a.g * s, a.b * s}; }
inline Rgb operator/(Rgb a, double s) { return {a.r / s, a.g / s, a.b / s}; }
inline bool isBlack(Rgb c) { return c.r == 0 && c.g == 0 && c.b == 0; }
inline Rgb gray(double v) { return {v, v, v}; }

// --------------------------------------------------------------- Fresnel
// reference scattering/fresnel.cpp:17-58 (std::complex<double>, as there).
inline double reflectance(double cos_i, std::complex<double> eta)
{
    using C = std::complex<double>;
    const double sin2 = 1 - cos_i * cos_i;
    const C cos_t = std::sqrt(C(1) - sin2 / (eta * eta));
    const C rpar = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    const C rper = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    return (std::norm(rpar) + std::norm(rper)) / 2;
}
inline double fresnelDielectric(double cos_i, double eta)
{
    if (cos_i < 0) {
        cos_i = -cos_i;
        eta = 1 / eta;
    }
    cos_i = std::min(cos_i, 1.0);
    if (eta == 1) return 0;
    const double sin2t = (1 - cos_i * cos_i) / (eta * eta);
    if (sin2t >= 1) return 1;
    return reflectance(cos_i, std::complex<double>(eta, 0));
}
inline Rgb fresnelConductor(double cos_i, Rgb eta, Rgb k)
{
    cos_i = std::clamp(cos_i, 0.0, 1.0);
    return {reflectance(cos_i, {eta.r, k.r}), reflectance(cos_i, {eta.g, k.g}),
            reflectance(cos_i, {eta.b, k.b})};
}

// ---------------------------------------------------------------- specular
// reference scattering/specular.hpp:13-16, specular.cpp:9-22
inline V3d reflect(V3d w, V3d n) { return n * (2 * dot(w, n)) - w; }
inline std::optional<V3d> refract(V3d w, V3d n, double eta)
{
    const double ci = dot(w, n);
    const double s2t = std::max(0.0, 1 - ci * ci) / (eta * eta);
    if (s2t >= 1) return std::nullopt;
    const double ct = std::sqrt(1 - s2t);
    return (-w) / eta + n * (ci / eta - ct);
}

// ------------------------------------------------------------------- GGX
// reference scattering/ggx_distribution.cpp:28-111
struct Ggx {
    double alpha;
    double density(V3d m) const
    {
        if (m.z <= 0) return 0;
        const double a2 = alpha * alpha;
        const double den = m.z * m.z * (a2 - 1) + 1;
        return a2 / (kPi * den * den);
    }
    double lambda(V3d w) const
    {
        const double c2 = w.z * w.z;
        if (c2 == 0) return std::numeric_limits<double>::infinity();
        const double t2 = (w.x * w.x + w.y * w.y) / c2;
        return (std::sqrt(1 + alpha * alpha * t2) - 1) / 2;
    }
    double masking(V3d w) const { return 1 / (1 + lambda(w)); }
    double maskingShadowing(V3d wo, V3d wi) const { return 1 / (1 + lambda(wo) + lambda(wi)); }
    V3d sampleVisibleNormal(V3d wo, double ux, double uy) const
    {
        const V3d view = normalized(V3d{alpha * wo.x, alpha * wo.y, wo.z});
        const double l2 = view.x * view.x + view.y * view.y;
        const V3d t1a = l2 > 0 ? V3d{-view.y, view.x, 0} / std::sqrt(l2) : V3d{1, 0, 0};
        const V3d t2a = cross(view, t1a);
        const double r = std::sqrt(ux);
        const double phi = 2 * kPi * uy;
        const double t1 = r * std::cos(phi);
        const double blend = (1 + view.z) / 2;
        const double t2 = (1 - blend) * std::sqrt(std::max(0.0, 1 - t1 * t1)) + blend * r * std::sin(phi);
        const double lift = std::sqrt(std::max(0.0, 1 - t1 * t1 - t2 * t2));
        const V3d h = t1a * t1 + t2a * t2 + view * lift;
        return normalized(V3d{alpha * h.x, alpha * h.y, std::max(0.0, h.z)});
    }
    double visibleNormalPdf(V3d wo, V3d m) const
    {
        const double f = dot(wo, m);
        if (wo.z <= 0 || f <= 0) return 0;
        return masking(wo) * f * density(m) / wo.z;
    }
};

// ------------------------------------------------------------------ lobes
struct ScatterSample {
    V3d direction;
    Rgb weight;
    double pdf{0};
    bool specular{false};
};

enum class LobeKind { Lambert, Conductor, Dielectric };
struct Lobe {
    LobeKind kind{LobeKind::Lambert};
    Rgb albedo;         // Lambert
    Rgb eta, k;         // conductor
    double alpha{0};    // conductor, dielectric
    double eta_d{1.5};  // dielectric

    static constexpr double kSmooth = 1e-4;

    // Lambert: reference scattering/lambert_lobe.cpp:17-41
    // Conductor: reference scattering/conductor_lobe.cpp:18-73
    // Dielectric: reference scattering/dielectric_lobe.cpp:17-186
    bool dielectricSmooth() const { return alpha < kSmooth || eta_d == 1; }

    static std::optional<V3d> refractionHalf(V3d wo, V3d wi, double eta)
    {
        const V3d s = wo + wi * eta;
        if (lengthSquared(s) == 0) return std::nullopt;
        V3d m = normalized(s);
        if (m.z < 0) m = -m;
        if (dot(wo, m) <= 0 || dot(wi, m) >= 0) return std::nullopt;
        return m;
    }
    static double refrDen(V3d wo, V3d wi, V3d m, double eta)
    {
        const double d = eta * dot(wi, m) + dot(wo, m);
        return d * d;
    }

    Rgb evaluate(V3d wo, V3d wi) const
    {
        switch (kind) {
        case LobeKind::Lambert:
            return (wo.z > 0 && wi.z > 0) ? albedo * kInvPi : Rgb{};
        case LobeKind::Conductor: {
            if (alpha < kSmooth || wo.z <= 0 || wi.z <= 0) return {};
            const V3d m = normalized(wo + wi);
            const Ggx g{alpha};
            return fresnelConductor(dot(wo, m), eta, k) *
                   (g.density(m) * g.maskingShadowing(wo, wi) / (4 * wo.z * wi.z));
        }
        case LobeKind::Dielectric: {
            if (dielectricSmooth() || wo.z <= 0 || wi.z == 0) return {};
            const Ggx g{alpha};
            if (wi.z > 0) {
                const V3d m = normalized(wo + wi);
                const double F = fresnelDielectric(dot(wo, m), eta_d);
                return gray(F * g.density(m) * g.maskingShadowing(wo, wi) / (4 * wo.z * wi.z));
            }
            const auto m = refractionHalf(wo, wi, eta_d);
            if (!m) return {};
            const double T = 1 - fresnelDielectric(dot(wo, *m), eta_d);
            const double G = g.masking(wo) * g.masking(wi);
            return gray(std::abs(dot(wi, *m)) * dot(wo, *m) * T * g.density(*m) * G /
                        (-wi.z * wo.z * refrDen(wo, wi, *m, eta_d)));
        }
        }
        return {};
    }

    double pdf(V3d wo, V3d wi) const
    {
        switch (kind) {
        case LobeKind::Lambert:
            return wo.z > 0 ? cosineHemispherePdf(wi.z) : 0.0;
        case LobeKind::Conductor: {
            if (alpha < kSmooth || wo.z <= 0 || wi.z <= 0) return 0;
            const V3d m = normalized(wo + wi);
            const double f = dot(wo, m);
            if (f <= 0) return 0;
            return Ggx{alpha}.visibleNormalPdf(wo, m) / (4 * f);
        }
        case LobeKind::Dielectric: {
            if (dielectricSmooth() || wo.z <= 0 || wi.z == 0) return 0;
            const Ggx g{alpha};
            if (wi.z > 0) {
                const V3d m = normalized(wo + wi);
                const double f = dot(wo, m);
                return g.visibleNormalPdf(wo, m) * fresnelDielectric(f, eta_d) / (4 * f);
            }
            const auto m = refractionHalf(wo, wi, eta_d);
            if (!m) return 0;
            const double jac = eta_d * eta_d * std::abs(dot(wi, *m)) / refrDen(wo, wi, *m, eta_d);
            return g.visibleNormalPdf(wo, *m) * (1 - fresnelDielectric(dot(wo, *m), eta_d)) * jac;
        }
        }
        return 0;
    }

    std::optional<ScatterSample> sample(V3d wo, double uc, double ux, double uy) const
    {
        switch (kind) {
        case LobeKind::Lambert: {
            if (wo.z <= 0) return std::nullopt;
            const V3d wi = cosineHemisphere(ux, uy);
            if (wi.z <= 0) return std::nullopt;
            return ScatterSample{wi, albedo, cosineHemispherePdf(wi.z), false};
        }
        case LobeKind::Conductor: {
            if (wo.z <= 0) return std::nullopt;
            if (alpha < kSmooth)
                return ScatterSample{reflect(wo, {0, 0, 1}), fresnelConductor(wo.z, eta, k), 1, true};
            const Ggx g{alpha};
            const V3d m = g.sampleVisibleNormal(wo, ux, uy);
            const V3d wi = reflect(wo, m);
            const double p = pdf(wo, wi);
            if (wi.z <= 0 || p <= 0) return std::nullopt;
            return ScatterSample{wi,
                                 fresnelConductor(dot(wo, m), eta, k) *
                                     (g.maskingShadowing(wo, wi) / g.masking(wo)),
                                 p, false};
        }
        case LobeKind::Dielectric: {
            if (wo.z <= 0) return std::nullopt;
            if (dielectricSmooth()) {
                const double R = fresnelDielectric(wo.z, eta_d);
                if (uc < R) return ScatterSample{reflect(wo, {0, 0, 1}), gray(1), R, true};
                const auto wi = refract(wo, {0, 0, 1}, eta_d);
                if (!wi) return std::nullopt;
                return ScatterSample{*wi, gray(1 / (eta_d * eta_d)), 1 - R, true};
            }
            const Ggx g{alpha};
            const V3d m = g.sampleVisibleNormal(wo, ux, uy);
            const double f = dot(wo, m);
            if (f <= 0) return std::nullopt;
            std::optional<ScatterSample> res;
            if (uc < fresnelDielectric(f, eta_d)) {
                const V3d wi = reflect(wo, m);
                if (wi.z > 0)
                    res = ScatterSample{wi, gray(g.maskingShadowing(wo, wi) / g.masking(wo)), 0, false};
            } else {
                const auto wi = refract(wo, m, eta_d);
                if (wi && wi->z < 0)
                    res = ScatterSample{*wi, gray(g.masking(*wi) / (eta_d * eta_d)), 0, false};
            }
            if (!res) return std::nullopt;
            res->pdf = pdf(wo, res->direction);
            if (res->pdf <= 0) return std::nullopt;
            return res;
        }
        }
        return std::nullopt;
    }
};

// Duff et al. 2017 basis; reference core/orthonormal_basis.cpp:17-35.
struct Onb {
    V3d t, b, n;
    static Onb from(V3d n)
    {
        const double s = std::copysign(1.0, n.z);
        const double a = -1 / (s + n.z);
        const double bb = n.x * n.y * a;
        return {{1 + s * n.x * n.x * a, s * bb, -s * n.x}, {bb, s + n.y * n.y * a, -n.y}, n};
    }
    V3d toLocal(V3d v) const { return {dot(v, t), dot(v, b), dot(v, n)}; }
    V3d toWorld(V3d v) const { return t * v.x + b * v.y + n * v.z; }
};

// World-frame scattering; reference scattering/scattering.cpp:5-54.
struct Scattering {
    Lobe lobe;
    Onb basis;
    V3d geometric;
    bool crosses(V3d d) const { return (dot(d, basis.n) < 0) != (dot(d, geometric) < 0); }
    Rgb evaluate(V3d wo, V3d wi) const
    {
        if (crosses(wo) || crosses(wi)) return {};
        return lobe.evaluate(basis.toLocal(wo), basis.toLocal(wi));
    }
    double pdf(V3d wo, V3d wi) const
    {
        if (crosses(wo) || crosses(wi)) return 0;
        return lobe.pdf(basis.toLocal(wo), basis.toLocal(wi));
    }
    std::optional<ScatterSample> sample(V3d wo, double uc, double ux, double uy) const
    {
        if (crosses(wo)) return std::nullopt;
        auto r = lobe.sample(basis.toLocal(wo), uc, ux, uy);
        if (!r) return r;
        r->direction = basis.toWorld(r->direction);
        if (crosses(r->direction)) return std::nullopt;
        return r;
    }
};

// ---------------------------------------------------------------- Perlin
// Ken Perlin's published permutation ("Improving Noise", SIGGRAPH 2002); the
// reference uses the same table (texture/perlin_noise.cpp:18-32) so patterns
// match the published ImprovedNoise.
inline const uint8_t* perlinPermutation()
{
    static const uint8_t p[256] = {
        151, 160, 137, 91, 90, 15, 131, 13, 201, 95, 96, 53, 194, 233, 7, 225, 140, 36, 103, 30, 69,
        142, 8, 99, 37, 240, 21, 10, 23, 190, 6, 148, 247, 120, 234, 75, 0, 26, 197, 62, 94, 252,
        219, 203, 117, 35, 11, 32, 57, 177, 33, 88, 237, 149, 56, 87, 174, 20, 125, 136, 171, 168,
        68, 175, 74, 165, 71, 134, 139, 48, 27, 166, 77, 146, 158, 231, 83, 111, 229, 122, 60, 211,
        133, 230, 220, 105, 92, 41, 55, 46, 245, 40, 244, 102, 143, 54, 65, 25, 63, 161, 1, 216, 80,
        73, 209, 76, 132, 187, 208, 89, 18, 169, 200, 196, 135, 130, 116, 188, 159, 86, 164, 100,
        109, 198, 173, 186, 3, 64, 52, 217, 226, 250, 124, 123, 5, 202, 38, 147, 118, 126, 255, 82,
        85, 212, 207, 206, 59, 227, 47, 16, 58, 17, 182, 189, 28, 42, 223, 183, 170, 213, 119, 248,
        152, 2, 44, 154, 163, 70, 221, 153, 101, 155, 167, 43, 172, 9, 129, 22, 39, 253, 19, 98,
        108, 110, 79, 113, 224, 232, 178, 185, 112, 104, 218, 246, 97, 228, 251, 34, 242, 193, 238,
        210, 144, 12, 191, 179, 162, 241, 81, 51, 145, 235, 249, 14, 239, 107, 49, 192, 214, 31,
        181, 199, 106, 157, 184, 84, 204, 176, 115, 121, 50, 45, 127, 4, 150, 254, 138, 236, 205,
        93, 222, 114, 67, 29, 24, 72, 243, 141, 128, 195, 78, 66, 215, 61, 156, 180};
    return p;
}

// reference texture/perlin_noise.cpp:70-145
inline double perlinNoise(V3d p)
{
    if (!isFiniteV(p)) return 0;
    const uint8_t* perm = perlinPermutation();
    auto hashed = [&](size_t i) -> size_t { return perm[i % 256]; };
    auto cell = [](double c, size_t& idx, double& off) {
        const double f = std::floor(c);
        double w = std::fmod(f, 256.0);
        if (w < 0) w += 256.0;
        idx = size_t(w);
        off = c - f;
    };
    size_t xi, yi, zi;
    double x, y, z;
    cell(p.x, xi, x);
    cell(p.y, yi, y);
    cell(p.z, zi, z);
    auto fade = [](double t) { return t * t * t * (t * (t * 6 - 15) + 10); };
    auto lerp = [](double t, double a, double b) { return a + t * (b - a); };
    auto grad = [](size_t hash, double gx, double gy, double gz) {
        const size_t h = hash & 15u;
        const double u = h < 8 ? gx : gy;
        double v = gz;
        if (h < 4) v = gy;
        else if (h == 12 || h == 14) v = gx;
        return ((h & 1u) == 0 ? u : -u) + ((h & 2u) == 0 ? v : -v);
    };
    const double u = fade(x), v = fade(y), w = fade(z);
    const size_t a = hashed(xi) + yi, aa = hashed(a) + zi, ab = hashed(a + 1) + zi;
    const size_t b = hashed(xi + 1) + yi, ba = hashed(b) + zi, bb = hashed(b + 1) + zi;
    return lerp(w,
                lerp(v, lerp(u, grad(hashed(aa), x, y, z), grad(hashed(ba), x - 1, y, z)),
                     lerp(u, grad(hashed(ab), x, y - 1, z), grad(hashed(bb), x - 1, y - 1, z))),
                lerp(v, lerp(u, grad(hashed(aa + 1), x, y, z - 1), grad(hashed(ba + 1), x - 1, y, z - 1)),
                     lerp(u, grad(hashed(ab + 1), x, y - 1, z - 1),
                          grad(hashed(bb + 1), x - 1, y - 1, z - 1))));
}
inline double turbulence(V3d p)
{
    double sum = 0, weight = 1;
    for (int o = 0; o < 7; ++o) {
        sum += weight * std::abs(perlinNoise(p));
        weight *= 0.5;
        p = {p.x * 2, p.y * 2, p.z * 2};
    }
    return sum;
}
// reference texture/perlin_pattern.cpp:11-23
inline double patternWeight(V3d p, double scale, double veins)
{
    const V3d s{p.x * scale, p.y * scale, p.z * scale};
    const double swirl = turbulence(s);
    if (veins == 0) return std::min(swirl, 1.0);
    const double w = (1 + std::sin(s.x + veins * swirl)) / 2;
    return std::isnan(w) ? 0.5 : w;
}

// ---------------------------------------------------------------- lights
struct LightSample {
    V3d direction;
    Rgb radiance;
    double distance{0};
    double pdf{0};
    bool delta{false};
};

struct Cone {
    V3d axis;
    double distance, tangent2, cos_max, one_minus_cos_max;
};
// reference light/area_light.cpp:31-51
inline std::optional<Cone> coneTowards(V3d center, double radius, V3d p)
{
    const V3d tc = center - p;
    const double d2 = lengthSquared(tc);
    const double r2 = radius * radius;
    if (d2 <= r2) return std::nullopt;
    const double d = std::sqrt(d2);
    const double s2 = r2 / d2;
    const double cm = std::sqrt(1 - s2);
    const double omc = s2 / (1 + cm);
    if (!std::isnormal(omc)) return std::nullopt;
    return Cone{tc / d, d, d2 - r2, cm, omc};
}

} // namespace oracle
