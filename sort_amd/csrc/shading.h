// Device shading: counter-based RNG, sampling routines, the BSDFs, lights and camera of the
// path tracer. Local shading space is y-up (SORT's convention: cosTheta(w) = w.y).
//
// Every function follows a reference routine (paths relative to SORT's src/):
//   u32 -> float mapping                    core/rand.cpp:99-112 (24 random bits / 2^24)
//   UniformSampleDisk / CosSampleHemisphere core/samplemethod.h:37-74
//   UniformSampleSphere                     core/samplemethod.h:117-126
//   Distribution1D::SampleDiscrete          core/samplemethod.h:129-176
//   Lambert                                 scatteringevent/bsdf/lambert.cpp:34-38, bxdf.cpp:54-65
//   GGX D / G1 / sample                     scatteringevent/bsdf/microfacet.cpp:229-284
//   MicroFacetReflection f/sample/pdf       scatteringevent/bsdf/microfacet.cpp:320-356
//   MicroFacetRefraction f/sample/pdf       scatteringevent/bsdf/microfacet.cpp:390-440
//   refract                                 scatteringevent/bsdf/bxdf_utils.h:120-131
//   Dielectric (rough, IOR 1 / 1.5)         scatteringevent/bsdf/dielectric.cpp:32-66
//   Fresnel (dielectric, conductor, Schlick) scatteringevent/bsdf/fresnel.h:30-183
//   DisneyBRDF (subset: no transmission, no thin surface, no SSS, isotropic)
//                                           scatteringevent/bsdf/disney.cpp:77-102,181-434
//   ScatteringEvent frame                   scatteringevent/scatteringevent.cpp:44-61
//   AreaLight / Quad sampling and pdf       light/area.cpp:22-107, shape/quad.cpp:23-44, shape/shape.h:90-100
//   AmbientLight (constant environment)     light/ambientlight.cpp:19-29, light/skylight.cpp:20-50
//   PerspectiveCamera::GenerateRay          camera/perspective.cpp:26-104
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

namespace sortgpu {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kTwoPi = 6.28318530717958647692f;
constexpr float kInvFourPi = 0.07957747154594766788f;

struct V3 {
    float x, y, z;
};
__host__ __device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__host__ __device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ __forceinline__ V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
__host__ __device__ __forceinline__ V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
__host__ __device__ __forceinline__ V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__host__ __device__ __forceinline__ V3 operator*(float s, V3 a) { return v3(a.x * s, a.y * s, a.z * s); }
__host__ __device__ __forceinline__ V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 normalize(V3 a) { return a / length(a); }
__device__ __forceinline__ float luminance(V3 c) {  // RGBSpectrum::GetIntensity (Rec.709 Y)
    return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z;
}
__device__ __forceinline__ float maxc(V3 c) { return fmaxf(c.x, fmaxf(c.y, c.z)); }
__device__ __forceinline__ bool is_black(V3 c) { return c.x == 0.0f && c.y == 0.0f && c.z == 0.0f; }
__device__ __forceinline__ float saturate(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
__device__ __forceinline__ float lerp(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ V3 lerp(V3 a, V3 b, float t) { return a * (1.0f - t) + b * t; }

// ---------------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy as 1, 2, 3",
// SC'11). Key = (seed, pixel); counter = (sample, bounce, call, 0). Each call gives 4 words.
struct U4 {
    uint32_t x, y, z, w;
};
__host__ __device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
    for (int i = 0; i < 10; i++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}
// 24 random bits mapped to [0, 1 - 2^-24], the reference's sort_rand<float> range.
__host__ __device__ __forceinline__ float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

struct Rng4 {
    float a, b, c, d;
};
__host__ __device__ __forceinline__ Rng4 rng4(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce,
                                              uint32_t call) {
    const U4 r = philox4x32_10(U4{sample, bounce, call, 0u}, seed, pixel);
    return Rng4{u01(r.x), u01(r.y), u01(r.z), u01(r.w)};
}
// Draw assignment per vertex (bounce b): call 0 = light pick, light u, light v, NEE bxdf pick;
// call 1 = NEE bsdf (t, u, v), continuation bxdf pick; call 2 = continuation bsdf (t, u, v), RR;
// call 3 / 4 = fresh samples drawn inside the NEE / continuation BSDF sampling (dielectric,
// Disney). Camera: bounce 0xFFFF, call 0 = pixel jitter (img_u, img_v).
constexpr uint32_t kCameraBounce = 0xFFFFu;

// ---------------------------------------------------------------------------------------------
// sampling (core/samplemethod.h)
__device__ __forceinline__ void uniform_sample_disk(float u, float v, float& x, float& y) {
    float r, theta;
    const float su = 2.0f * u - 1.0f;
    const float sv = 2.0f * v - 1.0f;
    const float asu = fabsf(su), asv = fabsf(sv);
    if (asu < asv) {
        r = asv;
        const float factor = (sv > 0.0f) ? -1.0f : 1.0f;
        theta = factor * su / r + 4.0f + 2.0f * factor;
    } else if (asv < asu) {
        r = asu;
        const float factor = (su > 0.0f) ? 1.0f : -1.0f;
        theta = factor * sv / r - 2.0f * factor + 2.0f;
        if (theta < 0.0f) theta += 8.0f;
    } else {
        x = 0.0f;
        y = 0.0f;
        return;
    }
    theta *= kPi / 4.0f;
    x = cosf(theta) * r;
    y = sinf(theta) * r;
}
__device__ __forceinline__ V3 cos_sample_hemisphere(float u, float v) {
    float x, z;
    uniform_sample_disk(u, v, x, z);
    const float y = sqrtf(fmaxf(0.0f, 1.0f - x * x - z * z));
    return v3(x, y, z);
}
__device__ __forceinline__ V3 spherical_vec(float theta, float phi) {
    return v3(sinf(theta) * cosf(phi), cosf(theta), sinf(theta) * sinf(phi));
}
__device__ __forceinline__ V3 uniform_sample_sphere(float u, float v) {
    return spherical_vec(acosf(1.0f - 2.0f * u), kTwoPi * v);
}

// ---------------------------------------------------------------------------------------------
// microfacet distributions (isotropic GGX; clearcoat GGX)
struct Ggx {
    float a2;  // alpha^2, alpha = max(roughness, 1e-3)^2
};
__device__ __forceinline__ Ggx make_ggx(float roughness) {
    const float r = fmaxf(roughness, 1e-3f);
    const float a = r * r;
    return Ggx{a * a};
}
__device__ __forceinline__ float ggx_D(const Ggx& g, V3 h) {
    const float c2 = h.y * h.y;
    if (c2 <= 0.0f) return 0.0f;
    const float beta = c2 + (h.x * h.x / g.a2 + h.z * h.z / g.a2);
    return 1.0f / (kPi * g.a2 * beta * beta);  // alphaU * alphaV = alpha^2
}
__device__ __forceinline__ float ggx_G1(const Ggx& g, V3 v) {
    const float tan2 = 1.0f / (v.y * v.y) - 1.0f;
    if (isinf(tan2)) return 0.0f;
    return 2.0f / (1.0f + sqrtf(1.0f + g.a2 * tan2));
}
__device__ __forceinline__ V3 ggx_sample(const Ggx& g, float u, float v) {
    const float alpha = sqrtf(g.a2);
    const float theta = atanf(alpha * sqrtf(v / (1.0f - v)));
    return spherical_vec(theta, kTwoPi * u);
}
__device__ __forceinline__ float ggx_pdf(const Ggx& g, V3 h) { return ggx_D(g, h) * fabsf(h.y); }

struct Ccggx {  // ClearcoatGGX(roughness): GGX alpha conversion, Burley's D, G1 with alpha 0.25
    float a2;
};
__device__ __forceinline__ Ccggx make_ccggx(float roughness) {
    const Ggx g = make_ggx(roughness);
    return Ccggx{g.a2};
}
__device__ __forceinline__ float cc_D(const Ccggx& g, V3 h) {
    return (g.a2 - 1.0f) / (kPi * logf(g.a2) * (1.0f + (g.a2 - 1.0f) * h.y * h.y));
}
__device__ __forceinline__ float cc_G1(V3 v) {
    if (fabsf(v.y) == 1.0f) return 0.0f;
    const float tan2 = 1.0f / (v.y * v.y) - 1.0f;
    return 1.0f / (1.0f + sqrtf(1.0f + 0.0625f * tan2));
}
__device__ __forceinline__ V3 cc_sample(const Ccggx& g, float u, float v) {
    const float phi = kTwoPi * u;
    const float theta = g.a2 == 1.0f ? acosf(sqrtf(v)) : acosf(sqrtf((expf(logf(g.a2) * v) - 1.0f) / (g.a2 - 1.0f)));
    return spherical_vec(theta, phi);
}

// ---------------------------------------------------------------------------------------------
// Fresnel
__device__ __forceinline__ float schlick_weight(float c) {
    const float m = saturate(1.0f - c);
    const float m2 = m * m;
    return m2 * m2 * m;
}
__device__ __forceinline__ float fresnel_dielectric(float cosI, float eta_i, float eta_t) {
    const bool entering = cosI > 0.0f;
    const float ei = entering ? eta_i : eta_t;
    const float et = entering ? eta_t : eta_i;
    const float sinI = sqrtf(saturate(1.0f - cosI * cosI));
    const float sinT = ei * sinI / et;
    if (sinT >= 1.0f) return 1.0f;
    if (!entering) cosI = -cosI;
    const float cosT = sqrtf(1.0f - sinT * sinT);
    const float t0 = et * cosI, t1 = ei * cosT, t2 = ei * cosI, t3 = et * cosT;
    const float rpl = (t0 - t1) / (t0 + t1);
    const float rpp = (t2 - t3) / (t2 + t3);
    return (rpl * rpl + rpp * rpp) * 0.5f;
}
__device__ __forceinline__ V3 fresnel_conductor(float cosI, V3 eta, V3 k) {
    const float c2 = cosI * cosI;
    const V3 t = 2.0f * eta * cosI;
    const V3 tf = eta * eta + k * k;
    const V3 tmp = tf * c2;
    const V3 one = v3(1, 1, 1);
    const V3 rpl = v3((tmp.x - t.x + 1) / (tmp.x + t.x + 1), (tmp.y - t.y + 1) / (tmp.y + t.y + 1),
                      (tmp.z - t.z + 1) / (tmp.z + t.z + 1));
    const V3 rpp = v3((tf.x - t.x + c2) / (tf.x + t.x + c2), (tf.y - t.y + c2) / (tf.y + t.y + c2),
                      (tf.z - t.z + c2) / (tf.z + t.z + c2));
    (void)one;
    return (rpl + rpp) * 0.5f;
}

// ---------------------------------------------------------------------------------------------
// Materials
enum MatType : uint32_t { MAT_LAMBERT = 0, MAT_GGX = 1, MAT_DIELECTRIC = 2, MAT_DISNEY = 3, MAT_TRANSPARENT = 4 };

struct alignas(16) MatRec {
    uint32_t type;
    float roughness;
    float ior;          // dielectric: interior IOR (exterior 1.0)
    float pad0;
    V3 base; float pad1;      // Lambert R, GGX R, dielectric reflectance, Disney base color
    V3 trans; float pad2;     // dielectric transmittance
    V3 eta; float pad3;       // conductor Fresnel
    V3 k; float pad4;
    float metallic, specular, specular_tint, sheen;
    float sheen_tint, clearcoat, clearcoat_gloss, pad5;
};

struct BsdfSampleOut {
    V3 f;       // f * |cos| (SORT's f includes the cosine)
    V3 wi;
    float pdf;
};

// Lambert, one-sided (doubleSided = false): needs wo and wi above the shading normal.
__device__ __forceinline__ V3 lambert_f(const MatRec& m, V3 wo, V3 wi) {
    if (!((wo.y > 0.0f) == (wi.y > 0.0f))) return v3(0, 0, 0);
    if (!(wo.y > 0.0f)) return v3(0, 0, 0);
    return m.base * (kInvPi * fabsf(wi.y));
}
__device__ __forceinline__ float lambert_pdf(V3 wo, V3 wi) {
    if (!((wo.y > 0.0f) == (wi.y > 0.0f))) return 0.0f;
    if (!(wo.y > 0.0f)) return 0.0f;
    return fabsf(wi.y) * kInvPi;
}

// MicroFacetReflection with GGX; fresnel selected by `fkind` (0 conductor, 1 dielectric 1/1.5,
// 2 Schlick(F0)); doubleSided as given.
struct MfRefl {
    Ggx g;
    V3 R;
    int fkind;
    V3 eta, k;   // conductor
    V3 F0;       // Schlick
    float ei, et;
    bool double_sided;
};
__device__ __forceinline__ V3 mfr_fresnel(const MfRefl& m, float c) {
    if (m.fkind == 0) return fresnel_conductor(c, m.eta, m.k);
    if (m.fkind == 1) {
        const float f = fresnel_dielectric(c, m.ei, m.et);
        return v3(f, f, f);
    }
    const float w = schlick_weight(c);
    return m.F0 + w * (v3(1, 1, 1) - m.F0);
}
__device__ __forceinline__ V3 mfr_f(const MfRefl& m, V3 wo, V3 wi) {
    if (!((wo.y > 0.0f) == (wi.y > 0.0f))) return v3(0, 0, 0);
    if (!m.double_sided && !(wo.y > 0.0f)) return v3(0, 0, 0);
    const float nov = fabsf(wo.y);
    if (nov == 0.0f) return v3(0, 0, 0);
    const V3 wh = normalize(wi + wo);
    const V3 F = mfr_fresnel(m, dot(wo, wh));
    const float G = ggx_G1(m.g, wo) * ggx_G1(m.g, wi);
    return m.R * ggx_D(m.g, wh) * F * G / (4.0f * nov);
}
__device__ __forceinline__ float mfr_pdf(const MfRefl& m, V3 wo, V3 wi) {
    if (!((wo.y > 0.0f) == (wi.y > 0.0f))) return 0.0f;
    if (!m.double_sided && !(wo.y > 0.0f)) return 0.0f;
    const V3 h = normalize(wo + wi);
    const float eoh = fabsf(dot(wo, h));
    return ggx_pdf(m.g, h) / (4.0f * eoh);
}
__device__ __forceinline__ BsdfSampleOut mfr_sample(const MfRefl& m, V3 wo, float u, float v) {
    const V3 wh = ggx_sample(m.g, u, v);
    BsdfSampleOut o;
    o.wi = 2.0f * dot(wo, wh) * wh - wo;
    o.pdf = mfr_pdf(m, wo, o.wi);
    if (!((wo.y > 0.0f) == (o.wi.y > 0.0f)) || (!m.double_sided && !(wo.y > 0.0f))) {
        o.f = v3(0, 0, 0);
        return o;
    }
    o.f = mfr_f(m, wo, o.wi);
    return o;
}

// MicroFacetRefraction (GGX, etaI = 1 outside, etaT inside).
struct MfRefr {
    Ggx g;
    V3 T;
    float etaI, etaT;
};
__device__ __forceinline__ V3 mft_f(const MfRefr& m, V3 wo, V3 wi) {
    if ((wi.y > 0.0f) == (wo.y > 0.0f)) return v3(0, 0, 0);
    const float nov = wo.y;
    if (nov == 0.0f) return v3(0, 0, 0);
    const float eta = wo.y > 0.0f ? (m.etaT / m.etaI) : (m.etaI / m.etaT);
    V3 wh = normalize(wo + wi * eta);
    if (wh.y < 0.0f) wh = -wh;
    const float svoh = dot(wo, wh), sioh = dot(wi, wh);
    const float F = fresnel_dielectric(dot(wh, wo), m.etaI, m.etaT);
    const float sd = svoh + eta * sioh;
    const float t = eta / sd;
    const float G = ggx_G1(m.g, wo) * ggx_G1(m.g, wi);
    return (1.0f - F) * m.T * fabsf(ggx_D(m.g, wh) * G * t * t * sioh * svoh / nov);
}
__device__ __forceinline__ float mft_pdf(const MfRefr& m, V3 wo, V3 wi) {
    if (wo.y * wi.y > 0.0f) return 0.0f;
    const float eta = wo.y > 0.0f ? (m.etaT / m.etaI) : (m.etaI / m.etaT);
    const V3 wh = normalize(wo + wi * eta);
    const float sd = dot(wo, wh) + eta * dot(wi, wh);
    const float dwh = eta * eta * fabsf(dot(wi, wh)) / (sd * sd);
    return ggx_pdf(m.g, wh) * dwh;
}
__device__ __forceinline__ V3 refract(V3 v, V3 n, float in_eta, float ext_eta, bool& tir) {
    const float coso = dot(v, n);
    const float eta = coso > 0.0f ? (ext_eta / in_eta) : (in_eta / ext_eta);
    const float t = 1.0f - eta * eta * fmaxf(0.0f, 1.0f - coso * coso);
    tir = t <= 0.0f;
    if (tir) return v3(0, 0, 0);
    const float scale = coso < 0.0f ? -1.0f : 1.0f;
    return -eta * v + (eta * coso - scale * sqrtf(t)) * n;
}
__device__ __forceinline__ BsdfSampleOut mft_sample(const MfRefr& m, V3 wo, float u, float v) {
    BsdfSampleOut o;
    o.f = v3(0, 0, 0);
    o.pdf = 0.0f;
    o.wi = v3(0, 0, 0);
    if (wo.y == 0.0f) return o;
    const V3 wh = ggx_sample(m.g, u, v);
    bool tir = false;
    o.wi = refract(wo, wh, m.etaT, m.etaI, tir);
    if (tir) return o;
    o.pdf = mft_pdf(m, wo, o.wi);
    o.f = mft_f(m, wo, o.wi);
    return o;
}

// Dielectric: rough reflection + refraction sharing one GGX, interior IOR m.ior.
__device__ __forceinline__ void dielectric_parts(const MatRec& m, MfRefl& r, MfRefr& t) {
    r.g = make_ggx(m.roughness);
    r.R = m.base;
    r.fkind = 1;
    r.ei = 1.0f;
    r.et = m.ior;
    r.double_sided = true;
    t.g = r.g;
    t.T = m.trans;
    t.etaI = 1.0f;
    t.etaT = m.ior;
}
__device__ __forceinline__ float dielectric_ratio(const MatRec& m, V3 wo) {
    const float fr = fresnel_dielectric(wo.y, 1.0f, m.ior);
    const float r = fr * luminance(m.base);
    const float t = (1.0f - fr) * luminance(m.trans);
    return r / (r + t);
}
__device__ __forceinline__ float dielectric_pdf(const MatRec& m, V3 wo, V3 wi) {
    MfRefl r;
    MfRefr t;
    dielectric_parts(m, r, t);
    const float ratio = dielectric_ratio(m, wo);
    return lerp(mft_pdf(t, wo, wi), mfr_pdf(r, wo, wi), ratio);
}

// Disney BRDF subset: diffuse + retro-reflection + sheen, GGX specular with Schlick Fresnel,
// clearcoat (metallic, specular, specularTint, roughness, sheen, sheenTint, clearcoat,
// clearcoatGloss). Interior / exterior IOR fixed at 1.5 / 1 as in the reference.
__device__ __forceinline__ V3 disney_ctint(const MatRec& m) {
    const float lum = luminance(m.base);
    return lum > 0.0f ? m.base * (1.0f / lum) : v3(1, 1, 1);
}
__device__ __forceinline__ V3 disney_cspec0(const MatRec& m) {
    const float r0 = ((1.0f / 1.5f - 1.0f) / (1.0f / 1.5f + 1.0f)) * ((1.0f / 1.5f - 1.0f) / (1.0f / 1.5f + 1.0f));
    return lerp(m.specular * r0 * lerp(v3(1, 1, 1), disney_ctint(m), m.specular_tint), m.base, m.metallic);
}
__device__ __forceinline__ V3 disney_f(const MatRec& m, V3 wo, V3 wi) {
    const float dw = 1.0f - m.metallic;
    const V3 wh = normalize(wo + wi);
    const float hoo = dot(wo, wh);
    const float hoo2r = hoo * hoo * m.roughness;
    const V3 ctint = disney_ctint(m);
    V3 ret = v3(0, 0, 0);
    const bool refl = (wo.y > 0.0f) && (wi.y > 0.0f);
    if (dw > 0.0f && refl) {
        const float cnoi = saturate(wi.y);
        const float fo = schlick_weight(wo.y), fi = schlick_weight(wi.y);
        const V3 diffuse = m.base * (kInvPi * (1.0f - fo * 0.5f) * (1.0f - fi * 0.5f));
        ret = ret + dw * diffuse * cnoi;
        const float rr = 2.0f * hoo2r;
        const V3 frr = m.base * (kInvPi * rr * (fo + fi + fo * fi * (rr - 1.0f)));
        ret = ret + dw * frr * cnoi;
        if (m.sheen > 0.0f) {
            const V3 csheen = lerp(v3(1, 1, 1), ctint, m.sheen_tint);
            const float fh = schlick_weight(hoo);
            ret = ret + dw * (fh * m.sheen * csheen) * cnoi;
        }
    }
    const V3 cspec0 = disney_cspec0(m);
    if (!is_black(cspec0) && refl) {
        MfRefl s;
        s.g = make_ggx(m.roughness);
        s.R = v3(1, 1, 1);
        s.fkind = 2;
        s.F0 = cspec0;
        s.double_sided = false;
        ret = ret + mfr_f(s, wo, wi);
    }
    if (m.clearcoat > 0.0f && refl) {
        const Ccggx cg = make_ccggx(sqrtf(lerp(0.1f, 0.001f, m.clearcoat_gloss)));
        const float nov = fabsf(wo.y);
        if (nov != 0.0f) {
            const float w = schlick_weight(dot(wo, wh));
            const float F = 0.04f + w * (1.0f - 0.04f);
            const float G = cc_G1(wo) * cc_G1(wi);
            ret = ret + m.clearcoat * v3(1, 1, 1) * (cc_D(cg, wh) * F * G / (4.0f * nov));
        }
    }
    return ret;
}
struct DisneyW {
    float cc, sr, dr, total;
};
__device__ __forceinline__ DisneyW disney_weights(const MatRec& m) {
    DisneyW w;
    w.cc = m.clearcoat * 0.04f;
    w.sr = luminance(disney_cspec0(m)) * (8.0f * (1.0f - m.roughness));
    w.dr = luminance(m.base) * (1.0f - m.metallic);
    w.total = w.cc + w.sr + w.dr;
    return w;
}
__device__ __forceinline__ float disney_pdf(const MatRec& m, V3 wo, V3 wi) {
    const DisneyW w = disney_weights(m);
    if (w.total <= 0.0f) return 0.0f;
    float p = 0.0f;
    const V3 wh = normalize(wi + wo);
    if (w.cc > 0.0f) {
        const Ccggx cg = make_ccggx(sqrtf(lerp(0.1f, 0.001f, m.clearcoat_gloss)));
        p += w.cc * (cc_D(cg, wh) * fabsf(wh.y)) / (4.0f * fabsf(dot(wo, wh)));
    }
    if (w.sr > 0.0f) {
        const Ggx g = make_ggx(m.roughness);
        p += w.sr * ggx_pdf(g, wh) / (4.0f * fabsf(dot(wo, wh)));
    }
    if (w.dr > 0.0f) p += w.dr * (fabsf(wi.y) * kInvPi);
    return p / w.total;
}
// r = lobe pick, (u, v) = fresh 2-D sample for the picked lobe.
__device__ __forceinline__ BsdfSampleOut disney_sample(const MatRec& m, V3 wo, float r, float u, float v) {
    BsdfSampleOut o;
    o.f = v3(0, 0, 0);
    o.pdf = 0.0f;
    o.wi = v3(0, 0, 0);
    const DisneyW w = disney_weights(m);
    if (w.total <= 0.0f) return o;
    const float inv = 1.0f / w.total;
    const float cc_w = w.cc * inv;
    const float sr_w = w.sr * inv + cc_w;
    if (r <= cc_w) {
        const Ccggx cg = make_ccggx(sqrtf(lerp(0.1f, 0.001f, m.clearcoat_gloss)));
        const V3 wh = cc_sample(cg, u, v);
        o.wi = 2.0f * dot(wo, wh) * wh - wo;
    } else if (r <= sr_w) {
        const V3 wh = ggx_sample(make_ggx(m.roughness), u, v);
        o.wi = 2.0f * dot(wo, wh) * wh - wo;
    } else {
        o.wi = cos_sample_hemisphere(u, v);
    }
    o.pdf = disney_pdf(m, wo, o.wi);
    o.f = disney_f(m, wo, o.wi);
    return o;
}

// Material dispatch in local space.
__device__ __forceinline__ V3 mat_f(const MatRec& m, V3 wo, V3 wi) {
    switch (m.type) {
        case MAT_LAMBERT: return lambert_f(m, wo, wi);
        case MAT_GGX: {
            MfRefl s;
            s.g = make_ggx(m.roughness);
            s.R = m.base;
            s.fkind = 0;
            s.eta = m.eta;
            s.k = m.k;
            s.double_sided = false;
            return mfr_f(s, wo, wi);
        }
        case MAT_DIELECTRIC: {
            MfRefl r;
            MfRefr t;
            dielectric_parts(m, r, t);
            return mft_f(t, wo, wi) + mfr_f(r, wo, wi);
        }
        case MAT_TRANSPARENT: return v3(0.f, 0.f, 0.f);  // Transparent::f (bsdf/transparent.cpp:24-26)
        default: return disney_f(m, wo, wi);
    }
}
__device__ __forceinline__ float mat_pdf(const MatRec& m, V3 wo, V3 wi) {
    switch (m.type) {
        case MAT_LAMBERT: return lambert_pdf(wo, wi);
        case MAT_GGX: {
            MfRefl s;
            s.g = make_ggx(m.roughness);
            s.double_sided = false;
            return mfr_pdf(s, wo, wi);
        }
        case MAT_DIELECTRIC: return dielectric_pdf(m, wo, wi);
        case MAT_TRANSPARENT: return 0.0f;  // Transparent::pdf (bsdf/transparent.cpp:28-30)
        default: return disney_pdf(m, wo, wi);
    }
}
// bs = (t, u, v) of the BsdfSample; fr = 4 fresh draws (dielectric: t,u,v; Disney: r,u,v).
__device__ __forceinline__ BsdfSampleOut mat_sample(const MatRec& m, V3 wo, float bs_u, float bs_v, Rng4 fr) {
    switch (m.type) {
        case MAT_LAMBERT: {
            BsdfSampleOut o;
            o.wi = cos_sample_hemisphere(bs_u, bs_v);
            o.pdf = lambert_pdf(wo, o.wi);
            o.f = lambert_f(m, wo, o.wi);
            return o;
        }
        case MAT_GGX: {
            MfRefl s;
            s.g = make_ggx(m.roughness);
            s.R = m.base;
            s.fkind = 0;
            s.eta = m.eta;
            s.k = m.k;
            s.double_sided = false;
            return mfr_sample(s, wo, bs_u, bs_v);
        }
        case MAT_DIELECTRIC: {
            MfRefl r;
            MfRefr t;
            dielectric_parts(m, r, t);
            const float ratio = dielectric_ratio(m, wo);
            BsdfSampleOut o;
            if (bs_u < ratio || ratio == 1.0f) o = mfr_sample(r, wo, fr.b, fr.c);
            else o = mft_sample(t, wo, fr.b, fr.c);
            o.pdf = lerp(mft_pdf(t, wo, o.wi), mfr_pdf(r, wo, o.wi), ratio);
            return o;
        }
        case MAT_TRANSPARENT: {  // Transparent::sample_f (bsdf/transparent.cpp:32-37): straight through
            BsdfSampleOut o;
            o.wi = -wo;
            o.pdf = 1.0f;
            o.f = m.base;  // the attenuation A
            return o;
        }
        default: return disney_sample(m, wo, fr.a, fr.b, fr.c);
    }
}

// ---------------------------------------------------------------------------------------------
// Lights: area quads (one-sided, radiance L) and a constant environment.
struct alignas(16) LightRec {
    float mat[12];     // light-to-world rows 0..2 (row 3 is 0 0 0 1)
    float inv[16];     // world-to-light
    V3 L; float area;  // radiance, sizeX * sizeY
    float half_x, half_y;
    uint32_t prim;     // primitive id of the quad
    uint32_t kind;     // 0 quad, 1 environment
};

__device__ __forceinline__ V3 xform_point3(const float* m, V3 p) {
    return v3(p.x * m[0] + p.y * m[1] + p.z * m[2] + m[3], p.x * m[4] + p.y * m[5] + p.z * m[6] + m[7],
              p.x * m[8] + p.y * m[9] + p.z * m[10] + m[11]);
}
__device__ __forceinline__ V3 xform_vec3(const float* m, V3 v) {
    return v3(v.x * m[0] + v.y * m[1] + v.z * m[2], v.x * m[4] + v.y * m[5] + v.z * m[6],
              v.x * m[8] + v.y * m[9] + v.z * m[10]);
}
__device__ __forceinline__ V3 quad_normal(const LightRec& l) {  // TransformNormal(DIR_UP): rows of inv^T
    return v3(l.inv[4], l.inv[5], l.inv[6]);
}
// Quad::GetIntersect in world space for a ray (o, d, tmin 0, tmax FLT_MAX); returns t or -1.
__device__ __forceinline__ float quad_hit(const LightRec& l, V3 o, V3 d) {
    const float* m = l.inv;
    V3 lo = v3(o.x * m[0] + o.y * m[1] + o.z * m[2] + m[3], o.x * m[4] + o.y * m[5] + o.z * m[6] + m[7],
               o.x * m[8] + o.y * m[9] + o.z * m[10] + m[11]);
    const float w = o.x * m[12] + o.y * m[13] + o.z * m[14] + m[15];
    if (!(w == 1.0f)) lo = lo / w;
    const V3 ld = xform_vec3(m, d);
    if (ld.y == 0.0f) return -1.0f;
    const float t = -lo.y / ld.y;
    if (t > FLT_MAX || t <= 0.0f) return -1.0f;
    const float px = lo.x + t * ld.x, pz = lo.z + t * ld.z;
    if (px > l.half_x || px < -l.half_x) return -1.0f;
    if (pz > l.half_y || pz < -l.half_y) return -1.0f;
    return t;
}

}  // namespace sortgpu
