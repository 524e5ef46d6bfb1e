// Device-side shading in FP32: sampling warps, complex-index Fresnel, GGX with
// visible-normal sampling, the Lambert / conductor / dielectric lobes, textures
// (constant, checker, Perlin), lights, and the camera. Each function names the
// reference code whose behaviour it restates.
#pragma once

#include <hip/hip_runtime.h>

#include "layout.hpp"

namespace ab {

constexpr float kPiF = 3.14159265358979323846f;
constexpr float kInvPiF = 0.31830988618379067154f;

struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 operator-(F3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ F3 operator*(F3 a, F3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ F3 operator/(F3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 cross(F3 a, F3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len2(F3 a) { return dot(a, a); }
__device__ __forceinline__ F3 normalize(F3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
__device__ __forceinline__ bool isBlack(F3 c) { return c.x == 0.0f && c.y == 0.0f && c.z == 0.0f; }
__device__ __forceinline__ float maxc(F3 c) { return fmaxf(c.x, fmaxf(c.y, c.z)); }
__device__ __forceinline__ bool isNormalF(float x) { return x >= 1.17549435e-38f && x <= 3.40282347e38f; }

// ------------------------------------------------------------------ random

// PCG32 step and output (reference sampling/pcg32.hpp:43-51).
struct Rng {
    uint64_t state;
    uint64_t inc;
    __device__ __forceinline__ uint32_t next()
    {
        const uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    // A draw in [0, 1) that FP32 holds exactly: (u32 >> 8) * 2^-24.
    __device__ __forceinline__ float uniform() { return float(next() >> 8) * 5.9604644775390625e-08f; }
};

__host__ __device__ __forceinline__ uint64_t mixBits(uint64_t v)
{
    v = (v ^ (v >> 30u)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27u)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31u);
}

// The generator of sample `sample` of pixel `pixel_index`: the pixel's stream
// (reference SampleGenerator::forPixel, sample_generator.hpp:35-45) advanced by
// sample * kSampleStride steps (Brown's arbitrary-stride jump).
constexpr uint64_t kSampleStride = 1024;
__device__ __forceinline__ Rng sampleRng(uint64_t seed, uint64_t pixel_index, uint64_t sample)
{
    Rng r;
    r.inc = (mixBits(pixel_index) << 1u) | 1u;
    r.state = 0;
    (void)r.next();
    r.state += seed;
    (void)r.next();
    uint64_t delta = sample * kSampleStride;
    uint64_t am = 1, ap = 0, cm = 6364136223846793005ULL, cp = r.inc;
    while (delta) {
        if (delta & 1u) {
            am *= cm;
            ap = ap * cm + cp;
        }
        cp = (cm + 1) * cp;
        cm *= cm;
        delta >>= 1u;
    }
    r.state = am * r.state + ap;
    return r;
}

// Warps, reference sampling/warps.cpp:8-45.
__device__ __forceinline__ F3 cosineHemisphere(float ux, float uy)
{
    const float ox = 2.0f * ux - 1.0f, oy = 2.0f * uy - 1.0f;
    float dx = 0.0f, dy = 0.0f;
    if (ox != 0.0f || oy != 0.0f) {
        float r, a;
        if (fabsf(ox) > fabsf(oy)) {
            r = ox;
            a = (kPiF / 4) * (oy / ox);
        } else {
            r = oy;
            a = (kPiF / 2) - (kPiF / 4) * (ox / oy);
        }
        float s, c;
        sincosf(a, &s, &c);
        dx = r * c;
        dy = r * s;
    }
    return {dx, dy, sqrtf(fmaxf(0.0f, 1.0f - dx * dx - dy * dy))};
}
__device__ __forceinline__ float cosineHemispherePdf(float c) { return c > 0.0f ? c * kInvPiF : 0.0f; }
__device__ __forceinline__ F3 uniformSphere(float ux, float uy)
{
    const float z = 1.0f - 2.0f * ux;
    const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float s, c;
    sincosf(2.0f * kPiF * uy, &s, &c);
    return {r * c, r * s, z};
}
constexpr float kUniformSpherePdf = kInvPiF / 4;

// ----------------------------------------------------------------- Fresnel

// Principal square root of a + bi (non-negative real part).
__device__ __forceinline__ void csqrt(float a, float b, float& re, float& im)
{
    const float r = hypotf(a, b);
    if (r == 0.0f) {
        re = im = 0.0f;
        return;
    }
    if (a >= 0.0f) {
        const float t = sqrtf(0.5f * (r + a));
        re = t;
        im = b / (2.0f * t);
    } else {
        const float t = sqrtf(0.5f * (r - a));
        re = fabsf(b) / (2.0f * t);
        im = copysignf(t, b);
    }
}

// Unpolarised reflectance for a complex relative index n + ik (reference
// scattering/fresnel.cpp:17-25).
__device__ __forceinline__ float reflectanceC(float ci, float n, float k)
{
    const float s2 = 1.0f - ci * ci;
    // eta^2 = (n^2 - k^2) + 2nk i; s2 / eta^2
    const float er = n * n - k * k, ei = 2.0f * n * k;
    const float den = er * er + ei * ei;
    const float qr = s2 * er / den, qi = -s2 * ei / den;
    float ctr, cti;
    csqrt(1.0f - qr, -qi, ctr, cti);
    // eta * ci
    const float ecr = n * ci, eci = k * ci;
    // r_par = (eta ci - cos_t) / (eta ci + cos_t)
    const float pnr = ecr - ctr, pni = eci - cti, pdr = ecr + ctr, pdi = eci + cti;
    const float rpar = (pnr * pnr + pni * pni) / (pdr * pdr + pdi * pdi);
    // eta * cos_t
    const float ectr = n * ctr - k * cti, ecti = n * cti + k * ctr;
    const float snr = ci - ectr, sni = -ecti, sdr = ci + ectr, sdi = ecti;
    const float rper = (snr * snr + sni * sni) / (sdr * sdr + sdi * sdi);
    return 0.5f * (rpar + rper);
}

// Real index: reference fresnel.cpp:29-47.
__device__ __forceinline__ float fresnelDielectric(float ci, float eta)
{
    if (ci < 0.0f) {
        ci = -ci;
        eta = 1.0f / eta;
    }
    ci = fminf(ci, 1.0f);
    if (eta == 1.0f) return 0.0f;
    const float s2t = (1.0f - ci * ci) / (eta * eta);
    if (s2t >= 1.0f) return 1.0f;
    const float ct = sqrtf(1.0f - s2t);
    const float rpar = (eta * ci - ct) / (eta * ci + ct);
    const float rper = (ci - eta * ct) / (ci + eta * ct);
    return 0.5f * (rpar * rpar + rper * rper);
}

__device__ __forceinline__ F3 fresnelConductor(float ci, F3 eta, F3 k)
{
    ci = fminf(fmaxf(ci, 0.0f), 1.0f);
    return {reflectanceC(ci, eta.x, k.x), reflectanceC(ci, eta.y, k.y), reflectanceC(ci, eta.z, k.z)};
}

// ---------------------------------------------------------------- specular

__device__ __forceinline__ F3 reflectV(F3 w, F3 n) { return n * (2.0f * dot(w, n)) - w; }

__device__ __forceinline__ bool refractV(F3 w, F3 n, float eta, F3& out)
{
    const float ci = dot(w, n);
    const float s2t = fmaxf(0.0f, 1.0f - ci * ci) / (eta * eta);
    if (s2t >= 1.0f) return false;
    const float ct = sqrtf(1.0f - s2t);
    out = (-w) / eta + n * (ci / eta - ct);
    return true;
}

// --------------------------------------------------------------------- GGX
// reference scattering/ggx_distribution.cpp:28-111

__device__ __forceinline__ float ggxD(float alpha, F3 m)
{
    if (m.z <= 0.0f) return 0.0f;
    const float a2 = alpha * alpha;
    const float d = m.z * m.z * (a2 - 1.0f) + 1.0f;
    return a2 / (kPiF * d * d);
}
__device__ __forceinline__ float ggxLambda(float alpha, F3 w)
{
    const float c2 = w.z * w.z;
    if (c2 == 0.0f) return __builtin_huge_valf();
    const float t2 = (w.x * w.x + w.y * w.y) / c2;
    return 0.5f * (sqrtf(1.0f + alpha * alpha * t2) - 1.0f);
}
__device__ __forceinline__ float ggxG1(float alpha, F3 w) { return 1.0f / (1.0f + ggxLambda(alpha, w)); }
__device__ __forceinline__ float ggxG2(float alpha, F3 wo, F3 wi)
{
    return 1.0f / (1.0f + ggxLambda(alpha, wo) + ggxLambda(alpha, wi));
}
__device__ __forceinline__ F3 ggxSampleVisible(float alpha, F3 wo, float ux, float uy)
{
    const F3 v = normalize(f3(alpha * wo.x, alpha * wo.y, wo.z));
    const float l2 = v.x * v.x + v.y * v.y;
    const F3 t1a = l2 > 0.0f ? f3(-v.y, v.x, 0.0f) * (1.0f / sqrtf(l2)) : f3(1.0f, 0.0f, 0.0f);
    const F3 t2a = cross(v, t1a);
    const float r = sqrtf(ux);
    float s, c;
    sincosf(2.0f * kPiF * uy, &s, &c);
    const float t1 = r * c;
    const float blend = 0.5f * (1.0f + v.z);
    const float t2 = (1.0f - blend) * sqrtf(fmaxf(0.0f, 1.0f - t1 * t1)) + blend * r * s;
    const float lift = sqrtf(fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2));
    const F3 h = t1a * t1 + t2a * t2 + v * lift;
    return normalize(f3(alpha * h.x, alpha * h.y, fmaxf(0.0f, h.z)));
}
__device__ __forceinline__ float ggxVisiblePdf(float alpha, F3 wo, F3 m)
{
    const float f = dot(wo, m);
    if (wo.z <= 0.0f || f <= 0.0f) return 0.0f;
    return ggxG1(alpha, wo) * f * ggxD(alpha, m) / wo.z;
}

// ------------------------------------------------------------------- lobes

enum LobeKind : uint32_t { kLambert = 0, kConductor = 1, kDielectric = 2 };

struct Lobe {
    uint32_t kind;
    F3 albedo; // Lambert albedo / conductor eta
    F3 k;      // conductor k
    float alpha;
    float eta; // dielectric relative index
};

struct Sample {
    F3 dir;
    F3 weight;
    float pdf;
    bool specular;
};

constexpr float kSmoothAlpha = 1e-4f;

__device__ __forceinline__ bool refractionHalf(F3 wo, F3 wi, float eta, F3& m)
{
    const F3 s = wo + wi * eta;
    if (len2(s) == 0.0f) return false;
    m = normalize(s);
    if (m.z < 0.0f) m = -m;
    return !(dot(wo, m) <= 0.0f || dot(wi, m) >= 0.0f);
}

// f(wo, wi), local frame (reference lambert_lobe.cpp:28-31, conductor_lobe.cpp:25-37,
// dielectric_lobe.cpp:109-136).
__device__ __forceinline__ F3 lobeEval(const Lobe& l, F3 wo, F3 wi)
{
    if (l.kind == kLambert) return (wo.z > 0.0f && wi.z > 0.0f) ? l.albedo * kInvPiF : f3(0, 0, 0);
    if (l.kind == kConductor) {
        if (l.alpha < kSmoothAlpha || wo.z <= 0.0f || wi.z <= 0.0f) return f3(0, 0, 0);
        const F3 m = normalize(wo + wi);
        return fresnelConductor(dot(wo, m), l.albedo, l.k) *
               (ggxD(l.alpha, m) * ggxG2(l.alpha, wo, wi) / (4.0f * wo.z * wi.z));
    }
    if (l.alpha < kSmoothAlpha || l.eta == 1.0f || wo.z <= 0.0f || wi.z == 0.0f) return f3(0, 0, 0);
    if (wi.z > 0.0f) {
        const F3 m = normalize(wo + wi);
        const float F = fresnelDielectric(dot(wo, m), l.eta);
        const float v = F * ggxD(l.alpha, m) * ggxG2(l.alpha, wo, wi) / (4.0f * wo.z * wi.z);
        return f3(v, v, v);
    }
    F3 m;
    if (!refractionHalf(wo, wi, l.eta, m)) return f3(0, 0, 0);
    const float T = 1.0f - fresnelDielectric(dot(wo, m), l.eta);
    const float G = ggxG1(l.alpha, wo) * ggxG1(l.alpha, wi);
    const float dn = l.eta * dot(wi, m) + dot(wo, m);
    const float v = fabsf(dot(wi, m)) * dot(wo, m) * T * ggxD(l.alpha, m) * G / (-wi.z * wo.z * dn * dn);
    return f3(v, v, v);
}

__device__ __forceinline__ float lobePdf(const Lobe& l, F3 wo, F3 wi)
{
    if (l.kind == kLambert) return wo.z > 0.0f ? cosineHemispherePdf(wi.z) : 0.0f;
    if (l.kind == kConductor) {
        if (l.alpha < kSmoothAlpha || wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
        const F3 m = normalize(wo + wi);
        const float f = dot(wo, m);
        if (f <= 0.0f) return 0.0f;
        return ggxVisiblePdf(l.alpha, wo, m) / (4.0f * f);
    }
    if (l.alpha < kSmoothAlpha || l.eta == 1.0f || wo.z <= 0.0f || wi.z == 0.0f) return 0.0f;
    if (wi.z > 0.0f) {
        const F3 m = normalize(wo + wi);
        const float f = dot(wo, m);
        return ggxVisiblePdf(l.alpha, wo, m) * fresnelDielectric(f, l.eta) / (4.0f * f);
    }
    F3 m;
    if (!refractionHalf(wo, wi, l.eta, m)) return 0.0f;
    const float dn = l.eta * dot(wi, m) + dot(wo, m);
    const float jac = l.eta * l.eta * fabsf(dot(wi, m)) / (dn * dn);
    return ggxVisiblePdf(l.alpha, wo, m) * (1.0f - fresnelDielectric(dot(wo, m), l.eta)) * jac;
}

// reference lambert_lobe.cpp:33-43, conductor_lobe.cpp:39-61, dielectric_lobe.cpp:138-172
__device__ __forceinline__ bool lobeSample(const Lobe& l, F3 wo, float uc, float ux, float uy, Sample& s)
{
    if (wo.z <= 0.0f) return false;
    if (l.kind == kLambert) {
        s.dir = cosineHemisphere(ux, uy);
        if (s.dir.z <= 0.0f) return false;
        s.weight = l.albedo;
        s.pdf = cosineHemispherePdf(s.dir.z);
        s.specular = false;
        return true;
    }
    if (l.kind == kConductor) {
        if (l.alpha < kSmoothAlpha) {
            s.dir = f3(-wo.x, -wo.y, wo.z);
            s.weight = fresnelConductor(wo.z, l.albedo, l.k);
            s.pdf = 1.0f;
            s.specular = true;
            return true;
        }
        const F3 m = ggxSampleVisible(l.alpha, wo, ux, uy);
        const F3 wi = reflectV(wo, m);
        const float p = lobePdf(l, wo, wi);
        if (wi.z <= 0.0f || p <= 0.0f) return false;
        s.dir = wi;
        s.weight = fresnelConductor(dot(wo, m), l.albedo, l.k) * (ggxG2(l.alpha, wo, wi) / ggxG1(l.alpha, wo));
        s.pdf = p;
        s.specular = false;
        return true;
    }
    if (l.alpha < kSmoothAlpha || l.eta == 1.0f) {
        const float R = fresnelDielectric(wo.z, l.eta);
        if (uc < R) {
            s.dir = f3(-wo.x, -wo.y, wo.z);
            s.weight = f3(1, 1, 1);
            s.pdf = R;
            s.specular = true;
            return true;
        }
        F3 wi;
        if (!refractV(wo, f3(0, 0, 1), l.eta, wi)) return false;
        const float w = 1.0f / (l.eta * l.eta);
        s.dir = wi;
        s.weight = f3(w, w, w);
        s.pdf = 1.0f - R;
        s.specular = true;
        return true;
    }
    const F3 m = ggxSampleVisible(l.alpha, wo, ux, uy);
    const float f = dot(wo, m);
    if (f <= 0.0f) return false;
    if (uc < fresnelDielectric(f, l.eta)) {
        const F3 wi = reflectV(wo, m);
        if (wi.z <= 0.0f) return false;
        const float w = ggxG2(l.alpha, wo, wi) / ggxG1(l.alpha, wo);
        s.dir = wi;
        s.weight = f3(w, w, w);
    } else {
        F3 wi;
        if (!refractV(wo, m, l.eta, wi) || wi.z >= 0.0f) return false;
        const float w = ggxG1(l.alpha, wi) / (l.eta * l.eta);
        s.dir = wi;
        s.weight = f3(w, w, w);
    }
    s.pdf = lobePdf(l, wo, s.dir);
    s.specular = false;
    return s.pdf > 0.0f;
}

// Duff et al. 2017 basis (reference core/orthonormal_basis.cpp:17-35).
struct Frame {
    F3 t, b, n;
};
__device__ __forceinline__ Frame frameFrom(F3 n)
{
    const float s = copysignf(1.0f, n.z);
    const float a = -1.0f / (s + n.z);
    const float bb = n.x * n.y * a;
    return {f3(1.0f + s * n.x * n.x * a, s * bb, -s * n.x), f3(bb, s + n.y * n.y * a, -n.y), n};
}
__device__ __forceinline__ F3 toLocal(const Frame& f, F3 v) { return f3(dot(v, f.t), dot(v, f.b), dot(v, f.n)); }
__device__ __forceinline__ F3 toWorld(const Frame& f, F3 v) { return f.t * v.x + f.b * v.y + f.n * v.z; }

// World-space scattering with the shading / geometric normal guard
// (reference scattering/scattering.cpp:5-54).
struct Scatter {
    Lobe lobe;
    Frame frame;
    F3 geo;
    __device__ __forceinline__ bool crosses(F3 d) const
    {
        return (dot(d, frame.n) < 0.0f) != (dot(d, geo) < 0.0f);
    }
    __device__ __forceinline__ F3 eval(F3 wo, F3 wi) const
    {
        if (crosses(wo) || crosses(wi)) return f3(0, 0, 0);
        return lobeEval(lobe, toLocal(frame, wo), toLocal(frame, wi));
    }
    __device__ __forceinline__ float pdf(F3 wo, F3 wi) const
    {
        if (crosses(wo) || crosses(wi)) return 0.0f;
        return lobePdf(lobe, toLocal(frame, wo), toLocal(frame, wi));
    }
    __device__ __forceinline__ bool sample(F3 wo, float uc, float ux, float uy, Sample& s) const
    {
        if (crosses(wo)) return false;
        if (!lobeSample(lobe, toLocal(frame, wo), uc, ux, uy, s)) return false;
        s.dir = toWorld(frame, s.dir);
        return !crosses(s.dir);
    }
};

// ------------------------------------------------------------------ Perlin
// Ken Perlin's published 2002 permutation (reference texture/perlin_noise.cpp:18-32).
static __constant__ const uint8_t kPerm[256] = {
    151, 160, 137, 91, 90, 15, 131, 13, 201, 95, 96, 53, 194, 233, 7, 225, 140, 36, 103, 30, 69, 142, 8, 99,
    37, 240, 21, 10, 23, 190, 6, 148, 247, 120, 234, 75, 0, 26, 197, 62, 94, 252, 219, 203, 117, 35, 11, 32,
    57, 177, 33, 88, 237, 149, 56, 87, 174, 20, 125, 136, 171, 168, 68, 175, 74, 165, 71, 134, 139, 48, 27,
    166, 77, 146, 158, 231, 83, 111, 229, 122, 60, 211, 133, 230, 220, 105, 92, 41, 55, 46, 245, 40, 244, 102,
    143, 54, 65, 25, 63, 161, 1, 216, 80, 73, 209, 76, 132, 187, 208, 89, 18, 169, 200, 196, 135, 130, 116,
    188, 159, 86, 164, 100, 109, 198, 173, 186, 3, 64, 52, 217, 226, 250, 124, 123, 5, 202, 38, 147, 118, 126,
    255, 82, 85, 212, 207, 206, 59, 227, 47, 16, 58, 17, 182, 189, 28, 42, 223, 183, 170, 213, 119, 248, 152,
    2, 44, 154, 163, 70, 221, 153, 101, 155, 167, 43, 172, 9, 129, 22, 39, 253, 19, 98, 108, 110, 79, 113, 224,
    232, 178, 185, 112, 104, 218, 246, 97, 228, 251, 34, 242, 193, 238, 210, 144, 12, 191, 179, 162, 241, 81,
    51, 145, 235, 249, 14, 239, 107, 49, 192, 214, 31, 181, 199, 106, 157, 184, 84, 204, 176, 115, 121, 50, 45,
    127, 4, 150, 254, 138, 236, 205, 93, 222, 114, 67, 29, 24, 72, 243, 141, 128, 195, 78, 66, 215, 61, 156, 180};

__device__ __forceinline__ uint32_t permAt(uint32_t i) { return kPerm[i & 255u]; }

__device__ __forceinline__ float gradDot(uint32_t hash, float x, float y, float z)
{
    const uint32_t h = hash & 15u;
    const float u = h < 8u ? x : y;
    const float v = h < 4u ? y : ((h == 12u || h == 14u) ? x : z);
    return ((h & 1u) == 0u ? u : -u) + ((h & 2u) == 0u ? v : -v);
}

// reference texture/perlin_noise.cpp:103-132
__device__ __forceinline__ float perlinNoise(float px, float py, float pz)
{
    if (!isfinite(px) || !isfinite(py) || !isfinite(pz)) return 0.0f;
    const float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
    float wx = fmodf(fx, 256.0f), wy = fmodf(fy, 256.0f), wz = fmodf(fz, 256.0f);
    if (wx < 0.0f) wx += 256.0f;
    if (wy < 0.0f) wy += 256.0f;
    if (wz < 0.0f) wz += 256.0f;
    const uint32_t xi = uint32_t(wx), yi = uint32_t(wy), zi = uint32_t(wz);
    const float x = px - fx, y = py - fy, z = pz - fz;
    const float u = x * x * x * (x * (x * 6.0f - 15.0f) + 10.0f);
    const float v = y * y * y * (y * (y * 6.0f - 15.0f) + 10.0f);
    const float w = z * z * z * (z * (z * 6.0f - 15.0f) + 10.0f);
    const uint32_t a = permAt(xi) + yi, aa = permAt(a) + zi, ab = permAt(a + 1) + zi;
    const uint32_t b = permAt(xi + 1) + yi, ba = permAt(b) + zi, bb = permAt(b + 1) + zi;
    auto lerp = [](float t, float p, float q) { return p + t * (q - p); };
    return lerp(w,
                lerp(v, lerp(u, gradDot(permAt(aa), x, y, z), gradDot(permAt(ba), x - 1, y, z)),
                     lerp(u, gradDot(permAt(ab), x, y - 1, z), gradDot(permAt(bb), x - 1, y - 1, z))),
                lerp(v, lerp(u, gradDot(permAt(aa + 1), x, y, z - 1), gradDot(permAt(ba + 1), x - 1, y, z - 1)),
                     lerp(u, gradDot(permAt(ab + 1), x, y - 1, z - 1),
                          gradDot(permAt(bb + 1), x - 1, y - 1, z - 1))));
}

// reference perlin_noise.cpp:134-145, perlin_pattern.cpp:11-23
__device__ __forceinline__ float patternWeight(F3 p, float scale, float veins)
{
    const float sx = p.x * scale, sy = p.y * scale, sz = p.z * scale;
    float sum = 0.0f, weight = 1.0f, x = sx, y = sy, z = sz;
#pragma unroll
    for (int o = 0; o < 7; ++o) {
        sum += weight * fabsf(perlinNoise(x, y, z));
        weight *= 0.5f;
        x *= 2.0f;
        y *= 2.0f;
        z *= 2.0f;
    }
    if (veins == 0.0f) return fminf(sum, 1.0f);
    const float wgt = 0.5f * (1.0f + sinf(sx + veins * sum));
    return isnan(wgt) ? 0.5f : wgt;
}

// ------------------------------------------------------------------ textures

// reference texture/checker_texture.cpp:8-17, perlin_texture.cpp:7-11
__device__ __forceinline__ F3 textureAt(const DevMaterial& m, F3 p, float u, float v)
{
    const F3 a = f3(m.a[0], m.a[1], m.a[2]);
    if (m.tex_kind == 1u) {
        const float s = floorf(u * m.frequency) + floorf(v * m.frequency);
        return fabsf(fmodf(s, 2.0f)) == 1.0f ? f3(m.b[0], m.b[1], m.b[2]) : a;
    }
    if (m.tex_kind == 2u) {
        const float w = patternWeight(p, m.frequency, m.veins);
        return a * (1.0f - w) + f3(m.b[0], m.b[1], m.b[2]) * w;
    }
    return a;
}

// Ray-origin offset of Waechter and Binder, "A Fast and Robust Method for
// Avoiding Self-Intersection" (Ray Tracing Gems, ch. 6): the point moves along
// the normal by a fixed number of float ulps (or a fixed distance near the
// origin), which scales with the point's own rounding error. It replaces the
// reference's fixed FP64 offset of 1e-7 (estimator/surface_ray.hpp:18), which
// is below one FP32 ulp at room scale.
__device__ __forceinline__ F3 offsetRay(F3 p, F3 n)
{
    const float kOrigin = 1.0f / 32.0f, kFloatScale = 1.0f / 65536.0f, kIntScale = 256.0f;
    const int ix = int(kIntScale * n.x), iy = int(kIntScale * n.y), iz = int(kIntScale * n.z);
    const float px = __int_as_float(__float_as_int(p.x) + (p.x < 0.0f ? -ix : ix));
    const float py = __int_as_float(__float_as_int(p.y) + (p.y < 0.0f ? -iy : iy));
    const float pz = __int_as_float(__float_as_int(p.z) + (p.z < 0.0f ? -iz : iz));
    return f3(fabsf(p.x) < kOrigin ? p.x + kFloatScale * n.x : px, fabsf(p.y) < kOrigin ? p.y + kFloatScale * n.y : py,
              fabsf(p.z) < kOrigin ? p.z + kFloatScale * n.z : pz);
}

} // namespace ab
