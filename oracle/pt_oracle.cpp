// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of SORT's path tracer for the configurations' scenes, used by the tests as
// the image-parity checker and by bench.py as the path-tracing CPU baseline. Nothing in the
// product links or calls this file.
//
// Structure follows the reference (paths relative to SORT's src/):
//   PathTracing::li bounce loop                   integrator/pathtracing.cpp:44-258
//   EvaluateDirect + MisFactor                     integrator/integratormethod.cpp:28-30,100-192
//   Scene::SampleLight / genLightDistribution      core/scene.cpp:227-257; core/samplemethod.h:129-176
//   AreaLight sample_l / Pdf / Le / Power          light/area.cpp:22-107; shape/quad.cpp:23-94; shape/shape.h:90-100
//   AmbientLight / SkyLight                        light/ambientlight.{h,cpp}; light/skylight.{h,cpp}
//   ScatteringEvent (frame, pick, Sample/Eval/Pdf) scatteringevent/scatteringevent.cpp:24-139
//   Bxdf::sample_f / pdf, Lambert                  scatteringevent/bsdf/bxdf.cpp:54-65; lambert.cpp:34-38
//   GGX, MicroFacetReflection / Refraction         scatteringevent/bsdf/microfacet.cpp:229-440
//   Dielectric                                     scatteringevent/bsdf/dielectric.cpp:32-66
//   Fresnel                                        scatteringevent/bsdf/fresnel.h:30-183
//   DisneyBRDF (subset: no transmission/thin/SSS)  scatteringevent/bsdf/disney.cpp:77-102,181-434
//   sampling                                       core/samplemethod.h:37-126
//   PerspectiveCamera                              camera/perspective.cpp:26-104
//   per-pixel mean of valid samples                work/image_evaluation/image_evaluation.cpp:210-244
// Intersection uses the OBVH restatement in obvh_oracle.cpp (same process, same tree).
//
// Random numbers: the reference seeds MT19937 with time(0) (core/rand.cpp:41-57), so its images
// are only statistically reproducible. The oracle instead draws from Philox4x32-10 keyed by
// (seed, pixel) with counter (sample, bounce, call, 0), 24 bits per float like sort_rand<float>;
// the product uses the same keys, so renders can also be compared path by path.
//
// How this restatement is pinned (tests/test_oracle_pt.py): the BXDF statistical suite of
// unittests/bxdf.cpp (reciprocity, energy <= 1.03, pdf vs sample consistency, E[1/pdf]); a
// furnace scene (Lambert sphere under a constant environment renders rho * L_env); the closed
// form irradiance of a quad light on a diffuse plane.

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <memory>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace pto {

constexpr float PI = 3.14159265358979323846f;
constexpr float INV_PI = 0.31830988618379067154f;
constexpr float TWO_PI = 6.28318530717958647692f;
constexpr float INV_FOUR_PI = 0.07957747154594766788f;

struct Vec {
    float x = 0, y = 0, z = 0;
    Vec() = default;
    Vec(float a, float b, float c) : x(a), y(b), z(c) {}
    explicit Vec(float a) : x(a), y(a), z(a) {}
};
static inline Vec operator+(Vec a, Vec b) { return Vec(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline Vec operator-(Vec a, Vec b) { return Vec(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline Vec operator-(Vec a) { return Vec(-a.x, -a.y, -a.z); }
static inline Vec operator*(Vec a, float s) { return Vec(a.x * s, a.y * s, a.z * s); }
static inline Vec operator*(float s, Vec a) { return Vec(a.x * s, a.y * s, a.z * s); }
static inline Vec operator/(Vec a, float s) { return Vec(a.x / s, a.y / s, a.z / s); }
static inline float dot(Vec a, Vec b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline Vec cross(Vec a, Vec b) { return Vec(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline float len(Vec a) { return std::sqrt(dot(a, a)); }
static inline Vec normalize(Vec a) { return a / len(a); }
static inline float satDot(Vec a, Vec b) { return std::min(std::max(dot(a, b), 0.0f), 1.0f); }
static inline float absDot(Vec a, Vec b) { return std::fabs(dot(a, b)); }
static inline float saturate(float x) { return std::min(std::max(x, 0.0f), 1.0f); }

struct Spectrum {  // spectrum/rgbspectrum.h
    float r = 0, g = 0, b = 0;
    Spectrum() = default;
    Spectrum(float v) : r(v), g(v), b(v) {}
    Spectrum(float a, float c, float d) : r(a), g(c), b(d) {}
    bool IsBlack() const { return r == 0.0f && g == 0.0f && b == 0.0f; }
    float GetIntensity() const { return 0.212671f * r + 0.715160f * g + 0.072169f * b; }
    float GetMaxComponent() const { return std::max(r, std::max(g, b)); }
    bool IsValid() const { return std::isfinite(r) && std::isfinite(g) && std::isfinite(b); }
};
static inline Spectrum operator+(Spectrum a, Spectrum b) { return Spectrum(a.r + b.r, a.g + b.g, a.b + b.b); }
static inline Spectrum operator-(Spectrum a, Spectrum b) { return Spectrum(a.r - b.r, a.g - b.g, a.b - b.b); }
static inline Spectrum operator*(Spectrum a, Spectrum b) { return Spectrum(a.r * b.r, a.g * b.g, a.b * b.b); }
static inline Spectrum operator*(Spectrum a, float s) { return Spectrum(a.r * s, a.g * s, a.b * s); }
static inline Spectrum operator*(float s, Spectrum a) { return Spectrum(a.r * s, a.g * s, a.b * s); }
static inline Spectrum operator/(Spectrum a, float s) { return Spectrum(a.r / s, a.g / s, a.b / s); }
static inline Spectrum slerpS(Spectrum a, Spectrum b, float t) { return a * (1.0f - t) + b * t; }
static inline float slerpF(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// ---- random numbers --------------------------------------------------------------------------
static inline void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int i = 0; i < 10; i++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}
struct Draws {
    float v[4];
};
static inline Draws draws(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t call) {
    uint32_t c[4] = {sample, bounce, call, 0u};
    philox(c, seed, pixel);
    Draws d;
    for (int i = 0; i < 4; i++) d.v[i] = (float)(c[i] >> 8) * (1.0f / 16777216.0f);
    return d;
}
constexpr uint32_t kCameraBounce = 0xFFFFu;

// The per-vertex draws: LightSample (t,u,v), the NEE bxdf pick, BsdfSample for NEE, the
// continuation pick, the continuation BsdfSample, the RR draw, and two blocks of fresh draws
// consumed inside BXDF sampling (dielectric's BsdfSample, Disney's lobe pick + sample).
struct VertexRng {
    Draws a, b, c, nee_fresh, cont_fresh;
};

struct LightSample {  // sampler/sample.h:27-46
    float t, u, v;
};
struct BsdfSample {  // sampler/sample.h:48-66
    float t, u, v;
    const Draws* fresh;  // the BsdfSample(rc) / sort_rand draws a BXDF makes internally
};

// ---- sampling --------------------------------------------------------------------------------
static inline void UniformSampleDisk(float u, float v, float& x, float& y) {
    float r, theta;
    const float su = 2.0f * u - 1.0f, sv = 2.0f * v - 1.0f;
    const float asu = std::fabs(su), asv = std::fabs(sv);
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
    theta *= PI / 4.0f;
    x = std::cos(theta) * r;
    y = std::sin(theta) * r;
}
static inline Vec CosSampleHemisphere(float u, float v) {
    float x, z;
    UniformSampleDisk(u, v, x, z);
    return Vec(x, std::sqrt(std::max(0.0f, 1.0f - x * x - z * z)), z);
}
static inline float CosHemispherePdf(Vec v) { return std::fabs(v.y) * INV_PI; }
static inline Vec sphericalVec(float theta, float phi) {
    return Vec(std::sin(theta) * std::cos(phi), std::cos(theta), std::sin(theta) * std::sin(phi));
}
static inline Vec UniformSampleSphere(float u, float v) { return sphericalVec(std::acos(1 - 2.0f * u), TWO_PI * v); }

// ---- Fresnel ---------------------------------------------------------------------------------
static inline float SchlickWeight(float c) {
    const float m = saturate(1.0f - c);
    return m * m * m * m * m;
}
static inline float DielectricFresnel(float cosI, float eta_i, float eta_t) {
    const bool entering = cosI > 0.0f;
    const float ei = entering ? eta_i : eta_t, et = entering ? eta_t : eta_i;
    const float sinI = std::sqrt(saturate(1.0f - cosI * cosI));
    const float sinT = ei * sinI / et;
    if (sinT >= 1.0f) return 1.0f;
    if (!entering) cosI = -cosI;
    const float cosT = std::sqrt(1.0f - sinT * sinT);
    const float t0 = et * cosI, t1 = ei * cosT, t2 = ei * cosI, t3 = et * cosT;
    const float a = (t0 - t1) / (t0 + t1), b = (t2 - t3) / (t2 + t3);
    return (a * a + b * b) * 0.5f;
}
struct Fresnel {
    virtual ~Fresnel() = default;
    virtual Spectrum Evaluate(float cosI) const = 0;
};
struct FresnelConductor : Fresnel {
    Spectrum eta, k;
    FresnelConductor(Spectrum e, Spectrum kk) : eta(e), k(kk) {}
    Spectrum Evaluate(float cosI) const override {
        const float c2 = cosI * cosI;
        const Spectrum t = 2.0f * eta * cosI;
        const Spectrum tf = eta * eta + k * k;
        const Spectrum tmp = tf * c2;
        const Spectrum one(1.0f);
        const Spectrum a((tmp.r - t.r + 1) / (tmp.r + t.r + 1), (tmp.g - t.g + 1) / (tmp.g + t.g + 1),
                         (tmp.b - t.b + 1) / (tmp.b + t.b + 1));
        const Spectrum b((tf.r - t.r + c2) / (tf.r + t.r + c2), (tf.g - t.g + c2) / (tf.g + t.g + c2),
                         (tf.b - t.b + c2) / (tf.b + t.b + c2));
        (void)one;
        return (a + b) * 0.5f;
    }
};
struct FresnelDielectric : Fresnel {
    float eta_t, eta_i;
    FresnelDielectric(float ei, float et) : eta_t(et), eta_i(ei) {}
    Spectrum Evaluate(float cosI) const override { return DielectricFresnel(cosI, eta_i, eta_t); }
};
struct FresnelSchlick : Fresnel {
    Spectrum F0;
    explicit FresnelSchlick(Spectrum f) : F0(f) {}
    Spectrum Evaluate(float c) const override { return F0 + SchlickWeight(c) * (Spectrum(1.0f) - F0); }
};

// ---- microfacet distributions ----------------------------------------------------------------
static inline float tanTheta2(Vec w) { return 1.0f / (w.y * w.y) - 1.0f; }
struct MicroFacetDistribution {
    virtual ~MicroFacetDistribution() = default;
    virtual float D(Vec h) const = 0;
    virtual float G1(Vec v) const = 0;
    virtual Vec sample_f(const BsdfSample& bs) const = 0;
    float G(Vec wo, Vec wi) const { return G1(wo) * G1(wi); }
    float Pdf(Vec h) const { return D(h) * std::fabs(h.y); }
};
struct GGX : MicroFacetDistribution {  // microfacet.cpp:229-284 (isotropic)
    float alphaU, alphaU2;
    explicit GGX(float roughness) {
        const float r = std::max(roughness, 1e-3f);
        alphaU = r * r;
        alphaU2 = alphaU * alphaU;
    }
    float D(Vec h) const override {
        const float c2 = h.y * h.y;
        if (c2 <= 0.0f) return 0.0f;
        const float beta = c2 + (h.x * h.x / alphaU2 + h.z * h.z / alphaU2);
        return 1.0f / (PI * (alphaU * alphaU) * beta * beta);
    }
    float G1(Vec v) const override {
        const float t2 = tanTheta2(v);
        if (std::isinf(t2)) return 0.0f;
        return 2.0f / (1.0f + std::sqrt(1.0f + alphaU2 * t2));
    }
    Vec sample_f(const BsdfSample& bs) const override {
        const float theta = std::atan(alphaU * std::sqrt(bs.v / (1.0f - bs.v)));
        return sphericalVec(theta, TWO_PI * bs.u);
    }
};
struct ClearcoatGGX : GGX {  // disney.cpp:77-102
    explicit ClearcoatGGX(float r) : GGX(r) {}
    float D(Vec h) const override {
        return (alphaU2 - 1) / (PI * std::log(alphaU2) * (1 + (alphaU2 - 1) * h.y * h.y));
    }
    float G1(Vec v) const override {
        if (std::fabs(v.y) == 1.0f) return 0.0f;
        const float t2 = tanTheta2(v);
        return 1.0f / (1.0f + std::sqrt(1.0f + 0.0625f * t2));
    }
    Vec sample_f(const BsdfSample& bs) const override {
        const float phi = TWO_PI * bs.u;
        const float theta = alphaU2 == 1.0f ? std::acos(std::sqrt(bs.v))
                                            : std::acos(std::sqrt((std::exp(std::log(alphaU2) * bs.v) - 1.0f) / (alphaU2 - 1.0f)));
        return sphericalVec(theta, phi);
    }
};

// ---- BXDFs (local space, y-up; gnormal = DIR_UP without normal maps) --------------------------
struct Bxdf {
    bool doubleSided = false;
    virtual ~Bxdf() = default;
    virtual Spectrum f(Vec wo, Vec wi) const = 0;
    virtual Spectrum sample_f(Vec wo, Vec& wi, const BsdfSample& bs, float* pdf) const {  // bxdf.cpp:54-58
        wi = CosSampleHemisphere(bs.u, bs.v);
        if (pdf) *pdf = this->pdf(wo, wi);
        return f(wo, wi);
    }
    virtual float pdf(Vec wo, Vec wi) const {  // bxdf.cpp:61-65
        if (!SameHemiSphere(wo, wi)) return 0.0f;
        if (!doubleSided && !PointingUp(wo)) return 0.0f;
        return CosHemispherePdf(wi);
    }
    bool PointingUp(Vec v) const { return v.y > 0.0f; }
    bool SameHemiSphere(Vec a, Vec b) const { return !(PointingUp(a) ^ PointingUp(b)); }
};
struct Lambert : Bxdf {
    Spectrum R;
    explicit Lambert(Spectrum r) : R(r) {}
    Spectrum f(Vec wo, Vec wi) const override {
        if (!SameHemiSphere(wo, wi)) return 0.0f;
        if (!doubleSided && !PointingUp(wo)) return 0.0f;
        return R * (INV_PI * std::fabs(wi.y));
    }
};
struct MicroFacetReflection : Bxdf {  // microfacet.cpp:320-356
    Spectrum R;
    const Fresnel* fresnel;
    const MicroFacetDistribution* distribution;
    MicroFacetReflection(Spectrum r, const Fresnel* fr, const MicroFacetDistribution* d, bool ds) : R(r), fresnel(fr), distribution(d) {
        doubleSided = ds;
    }
    Spectrum f(Vec wo, Vec wi) const override {
        if (!SameHemiSphere(wo, wi)) return 0.0f;
        if (!doubleSided && !PointingUp(wo)) return 0.0f;
        const float NoV = std::fabs(wo.y);
        if (NoV == 0.f) return 0.0f;
        const Vec wh = normalize(wi + wo);
        const Spectrum F = fresnel->Evaluate(dot(wo, wh));
        return R * distribution->D(wh) * F * distribution->G(wo, wi) / (4.0f * NoV);
    }
    Spectrum sample_f(Vec wo, Vec& wi, const BsdfSample& bs, float* pPdf) const override {
        const Vec wh = distribution->sample_f(bs);
        wi = 2.0f * dot(wo, wh) * wh - wo;
        if (pPdf) *pPdf = pdf(wo, wi);
        if (!SameHemiSphere(wo, wi)) return 0.0f;
        if (!doubleSided && !PointingUp(wo)) return 0.0f;
        return f(wo, wi);
    }
    float pdf(Vec wo, Vec wi) const override {
        if (!SameHemiSphere(wo, wi)) return 0.0f;
        if (!doubleSided && !PointingUp(wo)) return 0.0f;
        const Vec h = normalize(wo + wi);
        return distribution->Pdf(h) / (4.0f * absDot(wo, h));
    }
};
static inline Vec refract(Vec v, Vec n, float in_eta, float ext_eta, bool& tir) {  // bxdf_utils.h:120-131
    const float coso = dot(v, n);
    const float eta = coso > 0 ? (ext_eta / in_eta) : (in_eta / ext_eta);
    const float t = 1.0f - eta * eta * std::max(0.0f, 1.0f - coso * coso);
    tir = t <= 0.0f;
    if (tir) return Vec(0, 0, 0);
    const float scale = coso < 0.0f ? -1.0f : 1.0f;
    return -eta * v + (eta * coso - scale * std::sqrt(t)) * n;
}
struct MicroFacetRefraction : Bxdf {  // microfacet.cpp:390-440
    Spectrum T;
    float etaI, etaT;
    FresnelDielectric fresnel;
    const MicroFacetDistribution* distribution;
    MicroFacetRefraction(Spectrum t, const MicroFacetDistribution* d, float ei, float et)
        : T(t), etaI(ei), etaT(et), fresnel(ei, et), distribution(d) {
        doubleSided = true;
    }
    Spectrum f(Vec wo, Vec wi) const override {
        if (SameHemiSphere(wi, wo)) return 0.0f;
        const float NoV = wo.y;
        if (NoV == 0.f) return 0.0f;
        const float eta = wo.y > 0 ? (etaT / etaI) : (etaI / etaT);
        Vec wh = normalize(wo + wi * eta);
        if (wh.y < 0.0f) wh = -wh;
        const float sVoH = dot(wo, wh), sIoH = dot(wi, wh);
        const Spectrum F = fresnel.Evaluate(dot(wh, wo));
        const float sd = sVoH + eta * sIoH;
        const float t = eta / sd;
        return (Spectrum(1.f) - F) * T * std::fabs(distribution->D(wh) * distribution->G(wo, wi) * t * t * sIoH * sVoH / NoV);
    }
    Spectrum sample_f(Vec wo, Vec& wi, const BsdfSample& bs, float* pPdf) const override {
        if (wo.y == 0.0f) return 0.0f;
        const Vec wh = distribution->sample_f(bs);
        bool tir = false;
        wi = refract(wo, wh, etaT, etaI, tir);
        if (tir) return 0.0f;
        if (pPdf) *pPdf = pdf(wo, wi);
        return f(wo, wi);
    }
    float pdf(Vec wo, Vec wi) const override {
        if (wo.y * wi.y > 0.f) return 0.0f;
        const float eta = wo.y > 0 ? (etaT / etaI) : (etaI / etaT);
        const Vec wh = normalize(wo + wi * eta);
        const float sd = dot(wo, wh) + eta * dot(wi, wh);
        const float dwh = eta * eta * absDot(wi, wh) / (sd * sd);
        return distribution->Pdf(wh) * dwh;
    }
};
struct Dielectric : Bxdf {  // dielectric.cpp:32-66
    Spectrum R, T;
    GGX ggx;
    FresnelDielectric fresnel;
    MicroFacetReflection mf_reflect;
    MicroFacetRefraction mf_refract;
    Dielectric(Spectrum r, Spectrum t, float roughness, float ior)
        : R(r), T(t), ggx(roughness), fresnel(1.0f, ior), mf_reflect(r, &fresnel, &ggx, true), mf_refract(t, &ggx, 1.0f, ior) {
        doubleSided = true;
    }
    float ratio(Vec wo) const {
        const float fr = fresnel.Evaluate(wo.y).GetIntensity();
        const float r = fr * R.GetIntensity();
        const float t = (1.0f - fr) * T.GetIntensity();
        return r / (r + t);
    }
    Spectrum f(Vec wo, Vec wi) const override { return mf_refract.f(wo, wi) + mf_reflect.f(wo, wi); }
    Spectrum sample_f(Vec wo, Vec& wi, const BsdfSample& bs, float* pPdf) const override {
        const float spec_ratio = ratio(wo);
        const BsdfSample nbs{bs.fresh->v[0], bs.fresh->v[1], bs.fresh->v[2], nullptr};  // BsdfSample new_bs(rc)
        Spectrum ret;
        if (bs.u < spec_ratio || spec_ratio == 1.0f) ret = mf_reflect.sample_f(wo, wi, nbs, pPdf);
        else ret = mf_refract.sample_f(wo, wi, nbs, pPdf);
        if (pPdf) *pPdf = pdf(wo, wi);
        return ret;
    }
    float pdf(Vec wo, Vec wi) const override {
        return slerpF(mf_refract.pdf(wo, wi), mf_reflect.pdf(wo, wi), ratio(wo));
    }
};
struct DisneyBRDF : Bxdf {  // disney.cpp:181-434, subset
    Spectrum basecolor;
    float metallic, specular, specularTint, roughness, sheen, sheenTint, clearcoat, clearcoatGloss;
    DisneyBRDF(Spectrum b, float m, float s, float st, float r, float sh, float sht, float cc, float ccg)
        : basecolor(b), metallic(m), specular(s), specularTint(st), roughness(r), sheen(sh), sheenTint(sht), clearcoat(cc),
          clearcoatGloss(ccg) {
        doubleSided = true;
    }
    Spectrum Ctint() const {
        const float l = basecolor.GetIntensity();
        return l > 0.0f ? basecolor * (1.0f / l) : Spectrum(1.0f);
    }
    Spectrum Cspec0() const {
        const float e = 1.0f / 1.5f;
        const float r0 = ((e - 1.0f) / (e + 1.0f)) * ((e - 1.0f) / (e + 1.0f));  // SchlickR0FromEta(ior_ex / ior_in)
        return slerpS(specular * r0 * slerpS(Spectrum(1.0f), Ctint(), specularTint), basecolor, metallic);
    }
    Spectrum f(Vec wo, Vec wi) const override {
        const float diffuseWeight = 1.0f - metallic;
        const Vec wh = normalize(wo + wi);
        const float HoO = dot(wo, wh);
        const float HoO2ByRoughness = HoO * HoO * roughness;
        Spectrum ret(0.0f);
        const bool evaluate_reflection = PointingUp(wo) && PointingUp(wi);
        if (diffuseWeight > 0.0f && evaluate_reflection) {
            const float NoO = wo.y, NoI = wi.y, cNoI = saturate(NoI);
            const float FO = SchlickWeight(NoO), FI = SchlickWeight(NoI);
            const Spectrum disneyDiffuse = basecolor * (INV_PI * (1.0f - FO * 0.5f) * (1.0f - FI * 0.5f));
            ret = ret + diffuseWeight * disneyDiffuse * cNoI;
            const float Rr = 2.0f * HoO2ByRoughness;
            const Spectrum frr = basecolor * (INV_PI * Rr * (FO + FI + FO * FI * (Rr - 1.0f)));
            ret = ret + diffuseWeight * frr * cNoI;
            if (sheen > 0.0f) {
                const Spectrum Csheen = slerpS(Spectrum(1.0f), Ctint(), sheenTint);
                ret = ret + diffuseWeight * (SchlickWeight(HoO) * sheen * Csheen) * cNoI;
            }
        }
        const GGX ggx(roughness);
        const Spectrum cs = Cspec0();
        if (!cs.IsBlack() && evaluate_reflection) {
            const FresnelSchlick fr(cs);
            const MicroFacetReflection mf(Spectrum(1.0f), &fr, &ggx, false);
            ret = ret + mf.f(wo, wi);
        }
        if (clearcoat > 0.0f && evaluate_reflection) {
            const ClearcoatGGX cggx(std::sqrt(slerpF(0.1f, 0.001f, clearcoatGloss)));
            const FresnelSchlick fr(Spectrum(0.04f));
            const MicroFacetReflection mf(Spectrum(1.0f), &fr, &cggx, false);
            ret = ret + clearcoat * mf.f(wo, wi);
        }
        return ret;
    }
    void weights(float& cc, float& sr, float& dr, float& total) const {
        cc = clearcoat * 0.04f;
        sr = Cspec0().GetIntensity() * (8.0f * (1.0f - roughness));
        dr = basecolor.GetIntensity() * (1.0f - metallic);
        total = cc + sr + dr;
    }
    Spectrum sample_f(Vec wo, Vec& wi, const BsdfSample& bs, float* pPdf) const override {
        float cc, sr, dr, total;
        weights(cc, sr, dr, total);
        if (total <= 0.0f) {
            if (pPdf) *pPdf = 0.0f;
            return 0.0f;
        }
        const float inv = 1.0f / total;
        const float cc_w = cc * inv, sr_w = sr * inv + cc_w;
        const float r = bs.fresh->v[0];                                  // sort_rand<float>(rc)
        const BsdfSample s{0.0f, bs.fresh->v[1], bs.fresh->v[2], nullptr};  // BsdfSample(rc)
        if (r <= cc_w) {
            const ClearcoatGGX cggx(std::sqrt(slerpF(0.1f, 0.001f, clearcoatGloss)));
            const Vec wh = cggx.sample_f(s);
            wi = 2 * dot(wo, wh) * wh - wo;
        } else if (r <= sr_w) {
            const GGX ggx(roughness);
            const Vec wh = ggx.sample_f(s);
            wi = 2 * dot(wo, wh) * wh - wo;
        } else {
            wi = CosSampleHemisphere(s.u, s.v);
        }
        if (pPdf) *pPdf = pdf(wo, wi);
        return f(wo, wi);
    }
    float pdf(Vec wo, Vec wi) const override {
        float cc, sr, dr, total;
        weights(cc, sr, dr, total);
        if (total <= 0.0f) return 0.0f;
        float p = 0.0f;
        const Vec wh = normalize(wi + wo);
        if (cc > 0.0f) {
            const ClearcoatGGX cggx(std::sqrt(slerpF(0.1f, 0.001f, clearcoatGloss)));
            p += cc * cggx.Pdf(wh) / (4.0f * absDot(wo, wh));
        }
        if (sr > 0.0f) {
            const GGX ggx(roughness);
            p += sr * ggx.Pdf(wh) / (4.0f * absDot(wo, wh));
        }
        if (dr > 0.0f) p += dr * CosHemispherePdf(wi);
        return p / total;
    }
};

// Transparent (bsdf/transparent.cpp:24-37): no reflection; sampling passes straight through
// with the attenuation A and pdf 1.
struct Transparent : Bxdf {
    Spectrum A;
    explicit Transparent(Spectrum a) : A(a) {}
    Spectrum f(Vec, Vec) const override { return Spectrum(0.0f); }
    float pdf(Vec, Vec) const override { return 0.0f; }
    Spectrum sample_f(Vec wo, Vec& wi, const BsdfSample&, float* pdf) const override {
        wi = -wo;
        if (pdf) *pdf = 1.0f;
        return A;
    }
};

// Material record (same field order as the product's sortgpu_material).
struct MatIn {
    uint32_t type;
    float base[3], trans[3], roughness, ior, eta[3], k[3];
    float metallic, specular, specular_tint, sheen, sheen_tint, clearcoat, clearcoat_gloss;
};

// One BXDF per material, built per shading point like material->UpdateScatteringEvent.
struct BxdfHolder {
    Lambert lambert{Spectrum(1.0f)};
    FresnelConductor conductor{Spectrum(1.0f), Spectrum(0.0f)};
    GGX ggx{0.5f};
    MicroFacetReflection mfr{Spectrum(1.0f), &conductor, &ggx, false};
    Transparent transparent{Spectrum(1.0f)};
    const Bxdf* bxdf = nullptr;
    std::unique_ptr<Dielectric> diel;
    std::unique_ptr<DisneyBRDF> disney;
};
static void make_bxdf(const MatIn* m, BxdfHolder& h) {
    if (!m) {  // default material: white Lambert
        h.lambert = Lambert(Spectrum(1.0f));
        h.bxdf = &h.lambert;
        return;
    }
    const Spectrum base(m->base[0], m->base[1], m->base[2]);
    switch (m->type) {
        case 0:
            h.lambert = Lambert(base);
            h.bxdf = &h.lambert;
            break;
        case 1:
            h.conductor = FresnelConductor(Spectrum(m->eta[0], m->eta[1], m->eta[2]), Spectrum(m->k[0], m->k[1], m->k[2]));
            h.ggx = GGX(m->roughness);
            h.mfr = MicroFacetReflection(base, &h.conductor, &h.ggx, false);
            h.bxdf = &h.mfr;
            break;
        case 2:
            h.diel.reset(new Dielectric(base, Spectrum(m->trans[0], m->trans[1], m->trans[2]), m->roughness,
                                        m->ior > 0.0f ? m->ior : 1.5f));
            h.bxdf = h.diel.get();
            break;
        case 4:
            h.transparent = Transparent(base);
            h.bxdf = &h.transparent;
            break;
        default:
            h.disney.reset(new DisneyBRDF(base, m->metallic, m->specular, m->specular_tint, m->roughness, m->sheen,
                                          m->sheen_tint, m->clearcoat, m->clearcoat_gloss));
            h.bxdf = h.disney.get();
            break;
    }
}

// ScatteringEvent with one BXDF (scatteringevent.cpp:44-139).
struct ScatteringEvent {
    Vec n, bt, t;
    const Bxdf* bxdf;
    ScatteringEvent(Vec normal, Vec tangent, const Bxdf* b) : bxdf(b) {
        n = normalize(normal);
        bt = normalize(cross(n, tangent));
        t = normalize(cross(bt, n));
    }
    Vec worldToLocal(Vec v) const { return Vec(dot(v, t), dot(v, n), dot(v, bt)); }
    Vec localToWorld(Vec v) const {
        return Vec(v.x * t.x + v.y * n.x + v.z * bt.x, v.x * t.y + v.y * n.y + v.z * bt.y, v.x * t.z + v.y * n.z + v.z * bt.z);
    }
    Spectrum Evaluate_BSDF(Vec wo, Vec wi) const { return bxdf->f(worldToLocal(wo), worldToLocal(wi)); }
    float Pdf_BSDF(Vec wo, Vec wi) const { return bxdf->pdf(worldToLocal(wo), worldToLocal(wi)); }
    Spectrum Sample_BSDF(Vec wo, Vec& wi, const BsdfSample& bs, float& pdf) const {
        pdf = 0.0f;
        const Vec swo = worldToLocal(wo);
        const Spectrum ret = bxdf->sample_f(swo, wi, bs, &pdf);
        if (pdf == 0.0f) return 0.0f;
        wi = localToWorld(wi);
        return ret;
    }
};

// ---- scene, lights -----------------------------------------------------------------------------
struct QuadLight {
    float mat[16], inv[16];
    float sx, sy;
    Spectrum L;
    uint32_t prim;
    Vec transformPoint(Vec p) const {
        const float* m = mat;
        const float x = p.x * m[0] + p.y * m[1] + p.z * m[2] + m[3];
        const float y = p.x * m[4] + p.y * m[5] + p.z * m[6] + m[7];
        const float z = p.x * m[8] + p.y * m[9] + p.z * m[10] + m[11];
        const float w = p.x * m[12] + p.y * m[13] + p.z * m[14] + m[15];
        if (w == 1.0f) return Vec(x, y, z);
        return Vec(x / w, y / w, z / w);
    }
    Vec transformVector(const float* m, Vec v) const {
        return Vec(v.x * m[0] + v.y * m[1] + v.z * m[2], v.x * m[4] + v.y * m[5] + v.z * m[6], v.x * m[8] + v.y * m[9] + v.z * m[10]);
    }
    Vec normal() const { return Vec(inv[4], inv[5], inv[6]); }  // TransformNormal(DIR_UP)
    float area() const { return sx * sy; }
    // Quad::GetIntersect with limit FLT_MAX and ray window (fmin, FLT_MAX]
    bool intersect(Vec o, Vec d, float fmin, float& t_out) const {
        const float* m = inv;
        float x = o.x * m[0] + o.y * m[1] + o.z * m[2] + m[3];
        float y = o.x * m[4] + o.y * m[5] + o.z * m[6] + m[7];
        float z = o.x * m[8] + o.y * m[9] + o.z * m[10] + m[11];
        const float w = o.x * m[12] + o.y * m[13] + o.z * m[14] + m[15];
        if (w != 1.0f) {
            x /= w;
            y /= w;
            z /= w;
        }
        const Vec ld = transformVector(m, d);
        if (ld.y == 0.0f) return false;
        const float t = -y / ld.y;
        if (t > FLT_MAX || t <= fmin || t > FLT_MAX) return false;
        const float px = x + t * ld.x, pz = z + t * ld.z;
        const float hx = sx * 0.5f, hy = sy * 0.5f;
        if (px > hx || px < -hx || pz > hy || pz < -hy) return false;
        t_out = t;
        return true;
    }
};

}  // namespace pto

// Intersection services of obvh_oracle.cpp.
extern "C" {
struct OracleHandle;
int oracle_ray1(OracleHandle* h, const void* ray, void* hit, int mode);
void oracle_atten1(OracleHandle* h, const void* ray, const float* transp, float* out, uint32_t max_steps);
}

namespace pto {

struct PtScene {
    OracleHandle* bvh;
    std::vector<float> pos, nrm, tan;
    std::vector<uint32_t> idx, tri_mat;
    std::vector<MatIn> mats;
    std::vector<QuadLight> quads;
    Spectrum env;
    bool has_env = false;
    // lights: quads then env; Distribution1D over Power().GetIntensity()
    std::vector<float> cdf;
    uint32_t nlights = 0;
    float bbox[6];
    uint32_t ntris = 0;
    float clamp = 0.0f;  // > 0: ImageEvaluation's per-sample clamp (image_evaluation.cpp:222-226)
    // Material::EvaluateTransparency per primitive id (r, g, b, has-transparency), when some
    // material is transparent: shadow rays then take Scene::GetAttenuation (core/scene.cpp:192-208)
    std::vector<float> transp;
};

struct RayRec {
    float ox, oy, oz, tmin, dx, dy, dz, tmax;
};
struct HitRec {
    float t;
    uint32_t prim;
    float u, v;
};

static bool closest(const PtScene& sc, Vec o, Vec d, float tmin, float tmax, HitRec& h) {
    RayRec r{o.x, o.y, o.z, tmin, d.x, d.y, d.z, tmax};
    oracle_ray1(sc.bvh, &r, &h, 0);
    return h.prim != 0xFFFFFFFFu;
}
static bool occluded(const PtScene& sc, Vec o, Vec d, float tmin, float tmax) {
    RayRec r{o.x, o.y, o.z, tmin, d.x, d.y, d.z, tmax};
    return oracle_ray1(sc.bvh, &r, nullptr, 1) != 0;
}
// Visibility::GetAttenuation (light/light.h:56-58): Scene::GetAttenuation along the shadow ray.
static Spectrum attenuation(const PtScene& sc, Vec o, Vec d, float tmin, float tmax) {
    RayRec r{o.x, o.y, o.z, tmin, d.x, d.y, d.z, tmax};
    float out[4];
    oracle_atten1(sc.bvh, &r, sc.transp.data(), out, 4096u);
    return Spectrum(out[0], out[1], out[2]);
}

static int SampleDiscrete(const PtScene& sc, float u, float* pdf) {  // samplemethod.h:158-176
    const float* cdf = sc.cdf.data();
    const uint32_t count = sc.nlights;
    const float* target = std::lower_bound(cdf, cdf + count + 1, u);
    unsigned offset = (u <= 0.0f) ? 0 : (int)(target - cdf - 1);
    if (offset == 0)
        while (offset < count && cdf[offset + 1] == 0.0f) offset++;
    if (offset == count) {
        *pdf = 0.0f;
        return 0;
    }
    *pdf = cdf[offset + 1] - cdf[offset];
    return offset;
}

struct Shading {
    Vec p, n, t;
    const MatIn* mat;  // nullptr = default white Lambert
    int light;         // quad light index or -1
};
static inline Vec load3(const std::vector<float>& a, uint32_t i) { return Vec(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }

static Shading shading(const PtScene& sc, Vec o, Vec d, const HitRec& h) {
    Shading s;
    s.p = o + h.t * d;  // Ray::operator()(t)
    s.light = -1;
    if (h.prim < sc.ntris) {
        const uint32_t i0 = sc.idx[3 * h.prim], i1 = sc.idx[3 * h.prim + 1], i2 = sc.idx[3 * h.prim + 2];
        const float u = h.u, v = h.v, w = 1 - u - v;
        if (!sc.nrm.empty()) s.n = normalize(w * load3(sc.nrm, i0) + u * load3(sc.nrm, i1) + v * load3(sc.nrm, i2));
        else s.n = normalize(cross(load3(sc.pos, i1) - load3(sc.pos, i0), load3(sc.pos, i2) - load3(sc.pos, i0)));
        if (!sc.tan.empty()) s.t = normalize(w * load3(sc.tan, i0) + u * load3(sc.tan, i1) + v * load3(sc.tan, i2));
        else s.t = normalize(load3(sc.pos, i1) - load3(sc.pos, i0));
        s.mat = &sc.mats[sc.tri_mat.empty() ? 0 : sc.tri_mat[h.prim]];
    } else {
        const QuadLight& q = sc.quads[h.prim - sc.ntris];
        s.n = q.normal();
        s.t = Vec(q.mat[2], q.mat[6], q.mat[10]);
        s.mat = nullptr;
        s.light = (int)(h.prim - sc.ntris);
    }
    return s;
}

static inline float MisFactor(float f, float g) { return (f * f) / (f * f + g * g); }

// EvaluateDirect for area / ambient lights (integratormethod.cpp:100-192).
static Spectrum EvaluateDirect(const PtScene& sc, const ScatteringEvent& se, Vec ip, Vec rdir, int light,
                               const LightSample& ls, const BsdfSample& bs, uint64_t& shadow_rays) {
    Spectrum radiance;
    const Vec wo = -rdir;
    const bool is_env = (uint32_t)light == sc.quads.size();
    Vec wi;
    float light_pdf = 0.0f;
    Spectrum li;
    float vis_max = FLT_MAX;
    if (!is_env) {  // AreaLight::sample_l
        const QuadLight& q = sc.quads[light];
        const float u = 2 * ls.u - 1.0f, v = 2 * ls.v - 1.0f;
        const Vec lp = q.transformPoint(Vec(q.sx * 0.5f * u, 0.0f, q.sy * 0.5f * v));
        const Vec n = q.transformVector(q.mat, Vec(0, 1, 0));
        const Vec delta = lp - ip;
        wi = normalize(delta);
        const float dd = dot(-wi, n);
        light_pdf = dd <= 0 ? 0.0f : dot(delta, delta) / (q.area() * dd);
        const float l = len(delta);
        li = light_pdf == 0.0f ? Spectrum(0.0f) : q.L;
        vis_max = l - 0.01f;
    } else {  // SkyLight::sample_l with AmbientLight
        wi = UniformSampleSphere(ls.u, ls.v);
        light_pdf = INV_FOUR_PI;
        li = sc.env;
    }
    if (light_pdf > 0.0f && !li.IsBlack()) {
        const Spectrum f = se.Evaluate_BSDF(wo, wi);
        if (!f.IsBlack()) {
            shadow_rays++;
            if (sc.transp.empty()) {
                if (!occluded(sc, ip, wi, 0.01f, vis_max)) {
                    const float bsdf_pdf = se.Pdf_BSDF(wo, wi);
                    const float weight = MisFactor(light_pdf, bsdf_pdf);
                    radiance = li * f * weight / light_pdf;
                }
            } else {  // the ENABLE_TRANSPARENT_SHADOW branch (integratormethod.cpp:121-145)
                const Spectrum att = attenuation(sc, ip, wi, 0.01f, vis_max);
                if (!att.IsBlack()) {
                    const float bsdf_pdf = se.Pdf_BSDF(wo, wi);
                    const float weight = MisFactor(light_pdf, bsdf_pdf);
                    radiance = att * li * f * weight / light_pdf;
                }
            }
        }
    }
    float bsdf_pdf = 0.0f;
    Vec wi2;
    const Spectrum f = se.Sample_BSDF(wo, wi2, bs, bsdf_pdf);
    if (!f.IsBlack() && bsdf_pdf != 0.0f) {
        float lp2 = 0.0f, t_light = FLT_MAX;
        Spectrum li2;
        bool hit = true;
        if (!is_env) {  // Shape::Pdf, AreaLight::Le(ray)
            const QuadLight& q = sc.quads[light];
            float t;
            hit = q.intersect(ip, wi2, 0.0f, t);
            if (hit) {
                const Vec hp = ip + t * wi2;
                const Vec delta = ip - hp;
                const float dn = satDot(normalize(delta), q.normal());
                lp2 = dn <= 0.0f ? 0.0f : dot(delta, delta) / (q.area() * dn);
                li2 = satDot(-wi2, q.normal()) == 0.0f ? Spectrum(0.0f) : q.L;
                t_light = t;
            }
        } else {
            lp2 = INV_FOUR_PI;
            li2 = sc.env;
        }
        if (lp2 <= 0.0f) return radiance;
        const float weight = MisFactor(bsdf_pdf, lp2);
        if (!hit) return radiance;
        if (!li2.IsBlack()) {
            shadow_rays++;
            if (sc.transp.empty()) {
                if (!occluded(sc, ip, wi2, 0.001f, t_light - 0.001f)) radiance = radiance + li2 * f * weight / bsdf_pdf;
            } else {  // integratormethod.cpp:168-186
                const Spectrum att = attenuation(sc, ip, wi2, 0.001f, t_light - 0.001f);
                if (!att.IsBlack()) radiance = radiance + att * li2 * f * weight / bsdf_pdf;
            }
        }
    }
    return radiance;
}

// Where li records its extension rays and their hits (slot = base + bounce), when asked to.
struct RayDump {
    RayRec* rays;
    HitRec* hits;
    uint8_t* valid;
    size_t base;
};

// PathTracing::li for one camera (or caller) ray.
static Spectrum li(const PtScene& sc, Vec ro, Vec rd, float rtmin, uint32_t seed, uint32_t pixel, uint32_t sample,
                   uint32_t max_depth, uint64_t& ext_rays, uint64_t& shadow_rays, const RayDump* dump = nullptr) {
    Spectrum L(0.0f), throughput(1.0f);
    uint32_t bounces = 0;
    Vec o = ro, d = rd;
    float tmin = rtmin;
    BxdfHolder holder;
    while (true) {
        if (bounces >= max_depth) return L;
        HitRec h;
        ext_rays++;
        const bool hit_any = closest(sc, o, d, tmin, FLT_MAX, h);
        if (dump) {
            const size_t k = dump->base + bounces;
            dump->rays[k] = RayRec{o.x, o.y, o.z, tmin, d.x, d.y, d.z, FLT_MAX};
            dump->hits[k] = h;
            dump->valid[k] = 1;
        }
        if (!hit_any) {
            if (bounces == 0) return sc.has_env ? sc.env : Spectrum(0.0f);
            break;
        }
        const Shading s = shading(sc, o, d, h);
        if (bounces == 0 && s.light >= 0) {  // SurfaceInteraction::Le -> AreaLight::Le
            const float c = satDot(-d, s.n);
            if (c != 0.0f) L = L + sc.quads[s.light].L;
        }
        make_bxdf(s.mat, holder);
        const ScatteringEvent se(s.n, s.t, holder.bxdf);
        VertexRng r;
        r.a = draws(seed, pixel, sample, bounces, 0);
        r.b = draws(seed, pixel, sample, bounces, 1);
        r.c = draws(seed, pixel, sample, bounces, 2);
        r.nee_fresh = draws(seed, pixel, sample, bounces, 3);
        r.cont_fresh = draws(seed, pixel, sample, bounces, 4);
        {
            const LightSample ls{r.a.v[0], r.a.v[1], r.a.v[2]};
            const BsdfSample bs{r.b.v[0], r.b.v[1], r.b.v[2], &r.nee_fresh};
            float light_pdf = 0.0f;
            const int light = sc.nlights ? SampleDiscrete(sc, ls.t, &light_pdf) : 0;
            if (light_pdf > 0.0f)
                L = L + throughput * EvaluateDirect(sc, se, s.p, d, light, ls, bs, shadow_rays) / light_pdf;
        }
        const BsdfSample cbs{r.c.v[0], r.c.v[1], r.c.v[2], &r.cont_fresh};
        float path_pdf;
        Vec wi;
        const Spectrum f = se.Sample_BSDF(-d, wi, cbs, path_pdf);
        if (f.IsBlack() || path_pdf == 0.0f) break;
        throughput = throughput * (f / path_pdf);
        if (0.0f == throughput.GetIntensity()) break;
        o = s.p;
        d = wi;
        tmin = 0.0001f;
        if (bounces > 3 && throughput.GetMaxComponent() < 0.1f) {
            const float q = std::max(0.05f, 1.0f - throughput.GetMaxComponent());
            if (r.c.v[3] < q) break;
            throughput = throughput / (1 - q);
        }
        ++bounces;
    }
    return L;
}

}  // namespace pto

using namespace pto;

extern "C" {

// cam10: eye[3], target[3], up[3], fov_y. mats: nmats MatIn records. quad_rad: nquads x 3.
// quads34: nquads x (mat[16], inv[16], sx, sy) (the same quads the BVH was built with).
void oracle_pt_set_clamp(void* h, float clamp) { static_cast<PtScene*>(h)->clamp = clamp; }

void* oracle_pt_create(OracleHandle* bvh, const float* pos, uint32_t nverts, const uint32_t* idx, uint32_t ntris,
                       const float* nrm, const float* tan, const uint32_t* tri_mat, const void* mats, uint32_t nmats,
                       const float* quads34, uint32_t nquads, const float* quad_rad, const float* env3,
                       const float* bbox6) {
    auto* sc = new PtScene();
    sc->bvh = bvh;
    sc->pos.assign(pos, pos + 3 * (size_t)nverts);
    sc->idx.assign(idx, idx + 3 * (size_t)ntris);
    sc->ntris = ntris;
    if (nrm) sc->nrm.assign(nrm, nrm + 3 * (size_t)nverts);
    if (tan) sc->tan.assign(tan, tan + 3 * (size_t)nverts);
    if (tri_mat) sc->tri_mat.assign(tri_mat, tri_mat + ntris);
    const MatIn* m = static_cast<const MatIn*>(mats);
    sc->mats.assign(m, m + nmats);
    // transparency per primitive id: a Transparent material's attenuation; every other material
    // and the quads' default white Lambert have no transparent node (EvaluateTransparency = 0)
    bool any_transparent = false;
    for (const MatIn& mi : sc->mats) any_transparent |= mi.type == 4;
    if (any_transparent) {
        sc->transp.assign(4 * ((size_t)ntris + nquads), 0.0f);
        for (uint32_t t = 0; t < ntris; t++) {
            const MatIn& mi = sc->mats[tri_mat ? tri_mat[t] : 0];
            if (mi.type != 4) continue;
            for (int c = 0; c < 3; c++) sc->transp[4 * (size_t)t + c] = mi.base[c];
            sc->transp[4 * (size_t)t + 3] = 1.0f;
        }
    }
    std::vector<float> power;
    for (uint32_t q = 0; q < nquads; q++) {
        QuadLight l;
        std::memcpy(l.mat, quads34 + 34 * q, 64);
        std::memcpy(l.inv, quads34 + 34 * q + 16, 64);
        l.sx = quads34[34 * q + 32];
        l.sy = quads34[34 * q + 33];
        l.L = Spectrum(quad_rad[3 * q], quad_rad[3 * q + 1], quad_rad[3 * q + 2]);
        l.prim = ntris + q;
        sc->quads.push_back(l);
        power.push_back(l.area() * l.L.GetIntensity() * TWO_PI);  // AreaLight::Power
    }
    std::memcpy(sc->bbox, bbox6, sizeof(sc->bbox));
    sc->env = Spectrum(env3[0], env3[1], env3[2]);
    sc->has_env = !sc->env.IsBlack();
    if (sc->has_env) {
        const float dx = bbox6[3] - bbox6[0], dy = bbox6[4] - bbox6[1], dz = bbox6[5] - bbox6[2];
        const float r = std::sqrt(dx * dx + dy * dy + dz * dz) * 0.5f;
        power.push_back(r * r * PI * sc->env.GetIntensity());  // SkyLight::Power
    }
    sc->nlights = (uint32_t)power.size();
    sc->cdf.assign(sc->nlights + 1, 0.0f);
    for (uint32_t i = 0; i < sc->nlights; i++) sc->cdf[i + 1] = sc->cdf[i] + power[i];
    const float sum = sc->nlights ? sc->cdf[sc->nlights] : 0.0f;
    for (uint32_t i = 0; i <= sc->nlights && sc->nlights; i++)
        sc->cdf[i] = sum != 0.0f ? sc->cdf[i] / sum : (float)i / (float)sc->nlights;
    return sc;
}

void oracle_pt_destroy(void* p) { delete static_cast<PtScene*>(p); }

static double render_impl(void* p, const float* cam10, uint32_t width, uint32_t height, uint32_t pix0,
                          uint32_t pix1, uint32_t spp, uint32_t max_depth, uint64_t seed64, int nthreads, float* out,
                          uint64_t* stats, RayRec* dump_rays, HitRec* dump_hits, uint8_t* dump_valid);

// Renders with the pixel-major loop of ImageEvaluation (each pixel: spp samples, mean of the
// valid ones). pixel range [pix0, pix1) only (a crop); out: (pix1 - pix0) x 3.
// stats[0..1] += extension rays, shadow rays. Returns seconds.
double oracle_pt_render(void* p, const float* cam10, uint32_t width, uint32_t height, uint32_t pix0, uint32_t pix1,
                        uint32_t spp, uint32_t max_depth, uint64_t seed64, int nthreads, float* out, uint64_t* stats) {
    return render_impl(p, cam10, width, height, pix0, pix1, spp, max_depth, seed64, nthreads, out, stats, nullptr,
                       nullptr, nullptr);
}

// The same render, also recording every extension ray (the camera ray is bounce 0) with its
// closest hit: slot ((pixel - pix0) * spp + sample) * max_depth + bounce of rays8 (8 floats),
// hits4 (t, prim, u, v) and valid (1 where the path traced that bounce; the caller zeroes it).
double oracle_pt_dump(void* p, const float* cam10, uint32_t width, uint32_t height, uint32_t pix0, uint32_t pix1,
                      uint32_t spp, uint32_t max_depth, uint64_t seed64, int nthreads, float* rays8, float* hits4,
                      uint8_t* valid) {
    std::vector<float> out((size_t)(pix1 - pix0) * 3);
    return render_impl(p, cam10, width, height, pix0, pix1, spp, max_depth, seed64, nthreads, out.data(), nullptr,
                       reinterpret_cast<RayRec*>(rays8), reinterpret_cast<HitRec*>(hits4), valid);
}

}  // extern "C"

static double render_impl(void* p, const float* cam10, uint32_t width, uint32_t height, uint32_t pix0,
                          uint32_t pix1, uint32_t spp, uint32_t max_depth, uint64_t seed64, int nthreads, float* out,
                          uint64_t* stats, RayRec* dump_rays, HitRec* dump_hits, uint8_t* dump_valid) {
    const PtScene& sc = *static_cast<PtScene*>(p);
    const uint32_t seed = (uint32_t)(seed64 ^ (seed64 >> 32));
    // camera (perspective.cpp:26-104): ViewLookat basis and Perspective scales
    const Vec eye(cam10[0], cam10[1], cam10[2]);
    const Vec fwd = normalize(Vec(cam10[3], cam10[4], cam10[5]) - eye);
    const Vec up(cam10[6], cam10[7], cam10[8]);
    const Vec right = normalize(cross(up, fwd));
    const Vec up2 = cross(fwd, right);
    const float yScale = 1.0f / std::tan(cam10[9] * 0.5f);
    const float xScale = yScale / ((float)width / (float)height);
    std::atomic<uint32_t> next{pix0};
    std::atomic<uint64_t> ext{0}, shd{0};
    const auto t0 = std::chrono::steady_clock::now();
    auto work = [&]() {
        uint64_t e = 0, s = 0;
        while (true) {
            const uint32_t lo = next.fetch_add(64);
            if (lo >= pix1) break;
            const uint32_t hi = std::min(pix1, lo + 64);
            for (uint32_t pix = lo; pix < hi; pix++) {
                const uint32_t x = pix % width, y = pix / width;
                Spectrum radiance;
                uint32_t valid = spp;
                for (uint32_t k = 0; k < spp; k++) {
                    const Draws j = draws(seed, pix, k, kCameraBounce, 0);
                    const float rx = (float)x + j.v[0], ry = (float)y + j.v[1];
                    const float cx = 2.0f * rx / (float)width - 1.0f;
                    const float cy = 1.0f - 2.0f * ry / (float)height;
                    const Vec dc = normalize(Vec(cx / xScale, cy / yScale, 1.0f));
                    const Vec d = right * dc.x + up2 * dc.y + fwd * dc.z;
                    RayDump dd{dump_rays, dump_hits, dump_valid, ((size_t)(pix - pix0) * spp + k) * max_depth};
                    Spectrum l = li(sc, eye, d, 0.0f, seed, pix, k, max_depth, e, s, dump_rays ? &dd : nullptr);
                    if (sc.clamp > 0.0f) {  // li.Clamp(0, clamping): clamp() of math/utils.h:89-93
                        auto cl = [&](float x) {
                            if (x > sc.clamp) x = sc.clamp;
                            if (x < 0.0f) x = 0.0f;
                            return x;
                        };
                        l = Spectrum(cl(l.r), cl(l.g), cl(l.b));
                    }
                    if (l.IsValid()) radiance = radiance + l;
                    else --valid;
                }
                if (valid > 0) radiance = radiance / (float)valid;
                float* o = out + 3 * (size_t)(pix - pix0);
                o[0] = radiance.r;
                o[1] = radiance.g;
                o[2] = radiance.b;
            }
        }
        ext += e;
        shd += s;
    };
    std::vector<std::thread> th;
    for (int i = 1; i < nthreads; i++) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
    if (stats) {
        stats[0] += ext.load();
        stats[1] += shd.load();
    }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

extern "C" {

// The random-number generator alone: Philox4x32-10 of a counter and key, and the four floats of
// one (seed, pixel, sample, bounce, call) draw (pinned by the generator's published known-answer
// vectors in the tests).
void oracle_philox(const uint32_t* ctr, uint32_t k0, uint32_t k1, uint32_t* out) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    philox(c, k0, k1);
    for (int i = 0; i < 4; i++) out[i] = c[i];
}
void oracle_draws(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t call, float* out) {
    const Draws d = draws(seed, pixel, sample, bounce, call);
    for (int i = 0; i < 4; i++) out[i] = d.v[i];
}

// Batched Li over caller rays: keys = (pixel, sample) per ray; out n x 3.
void oracle_pt_li(void* p, const float* rays8, const uint32_t* keys, uint32_t n, uint32_t max_depth, uint64_t seed64,
                  int nthreads, float* out) {
    const PtScene& sc = *static_cast<PtScene*>(p);
    const uint32_t seed = (uint32_t)(seed64 ^ (seed64 >> 32));
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
        uint64_t e = 0, s = 0;
        while (true) {
            const uint32_t i = next.fetch_add(1);
            if (i >= n) break;
            const float* r = rays8 + 8 * (size_t)i;
            const Spectrum l = li(sc, Vec(r[0], r[1], r[2]), Vec(r[4], r[5], r[6]), r[3], seed, keys[2 * i], keys[2 * i + 1],
                                  max_depth, e, s);
            out[3 * i] = l.r;
            out[3 * i + 1] = l.g;
            out[3 * i + 2] = l.b;
        }
    };
    std::vector<std::thread> th;
    for (int i = 1; i < nthreads; i++) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
}

// BXDF probe for the statistical suite, over n (wo, wi, u) triples in local space (mat == NULL:
// the default white Lambert). mode 0: f(wo, wi) -> out[0..2]; mode 1: pdf(wo, wi) -> out[0];
// mode 2: sample_f(wo, BsdfSample(u[0..2]), fresh draws u[3..6]) -> out = f[3], wi[3], pdf.
// Records: wo / wi 3 floats, u 7 floats, out 7 floats per item.
void oracle_bxdf(const void* mat, int mode, uint32_t n, const float* wo3, const float* wi3, const float* u7, float* out) {
    BxdfHolder h;
    make_bxdf(static_cast<const MatIn*>(mat), h);
    for (uint32_t i = 0; i < n; i++) {
        const Vec wo(wo3[3 * i], wo3[3 * i + 1], wo3[3 * i + 2]);
        float* o = out + 7 * (size_t)i;
        if (mode == 0) {
            const Spectrum f = h.bxdf->f(wo, Vec(wi3[3 * i], wi3[3 * i + 1], wi3[3 * i + 2]));
            o[0] = f.r; o[1] = f.g; o[2] = f.b;
        } else if (mode == 1) {
            o[0] = h.bxdf->pdf(wo, Vec(wi3[3 * i], wi3[3 * i + 1], wi3[3 * i + 2]));
        } else {
            const float* u = u7 + 7 * (size_t)i;
            Draws fresh;
            for (int k = 0; k < 4; k++) fresh.v[k] = u[3 + k];
            const BsdfSample bs{u[0], u[1], u[2], &fresh};
            Vec wi;
            float pdf = 0.0f;
            const Spectrum f = h.bxdf->sample_f(wo, wi, bs, &pdf);
            o[0] = f.r; o[1] = f.g; o[2] = f.b;
            o[3] = wi.x; o[4] = wi.y; o[5] = wi.z;
            o[6] = pdf;
        }
    }
}

}  // extern "C"
