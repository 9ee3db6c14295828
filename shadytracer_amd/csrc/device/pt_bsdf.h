// The layered PBR BSDF: a metal lobe, a dielectric interface (reflection, and refraction for transmissive
// materials) over a diffuse base, with GGX microfacets. Semantics follow the reference's layered model
// (core/materials_pbr.js:74-553 makeLobes / lobesEval / lobesPdf / lobesSample; core/microfacet.js:72-140):
//   f = m cm F_schlick(base) spec + (1-m) F_diel spec + (1-m)(1-T) base (1-Ec(o))(1-Ec(i)) / (pi (1-Ec_avg))
// with Kulla-Conty's metal compensation cm = 1 + F0 (1/E1(o) - 1) and the albedo-coupled diffuse base; the
// tables Ec, Ec_avg and E1 are integrated on the GPU at start-up (pt_mf_lut below). Sampled one lobe at a time
// (metal with p = m, else the diffuse base with p = (1-ps)(1-T), else the interface) with the lobe's own
// f cos / pdf over its pick probability as the weight; lobesPdf is the mixture density, as MIS needs.
// The rough interface's transmission (materials_pbr.js:99-100 lobesThrough, :296-357) is evaluated too where the caller
// asks for it (Lobes.through, false from makeLobes), for a direction below the surface: Walter et al. 2007's BTDF over
// the generalised half vector h = -(eta o + i), in basic radiance (no n^2 -- the engine carries radiance through an
// interface as the sampler's (1 - F) weight does), so that eval, pdf and the sampler's weight agree exactly.
#pragma once

#define PT_ALPHA_MIN 1e-3f

__device__ __forceinline__ float mfLambda(V3 w, float a) {
    float z2 = fmaxf(w.z * w.z, 1e-12f);
    float t2 = a * a * (w.x * w.x + w.y * w.y) / z2;
    return 0.5f * (-1.0f + sqrtf(1.0f + t2));
}
__device__ __forceinline__ float mfG1(V3 w, float a) { return 1.0f / (1.0f + mfLambda(w, a)); }
__device__ __forceinline__ float mfG2(V3 o, V3 i, float a) { return 1.0f / (1.0f + mfLambda(o, a) + mfLambda(i, a)); }
__device__ __forceinline__ float mfD(V3 h, float a) {
    if (h.z <= 0.0f) return 0.0f;
    float ia2 = 1.0f / (a * a);
    float t = (h.x * h.x + h.y * h.y) * ia2 + h.z * h.z;
    return ia2 / (PT_PI * t * t);
}
// Visible-normal sampling by the spherical-cap method (Dupuy & Benyoub 2023).
__device__ __forceinline__ V3 mfSampleVNDF(V3 o, float a, float u1, float u2) {
    V3 os = normalize(v3(o.x * a, o.y * a, o.z));
    float phi = 2.0f * PT_PI * u1;
    float z = (1.0f - u2) * (1.0f + os.z) - os.z;
    float sinT = sqrtf(fminf(fmaxf(1.0f - z * z, 0.0f), 1.0f));
    float sp, cp;
    __sincosf(phi, &sp, &cp);
    V3 hs = v3(sinT * cp, sinT * sp, z) + os;
    return normalize(v3(hs.x * a, hs.y * a, fmaxf(hs.z, 1e-7f)));
}
__device__ __forceinline__ float mfPdfReflect(V3 o, V3 i, float a) {
    if (o.z <= 0.0f || i.z <= 0.0f) return 0.0f;
    V3 h = normalize(o + i);
    return mfG1(o, a) * mfD(h, a) / (4.0f * o.z);
}
// Exact unpolarised dielectric Fresnel reflectance, eta = n_incident / n_transmitted.
__device__ __forceinline__ float fresnelDielectric(float cosI, float eta) {
    cosI = fminf(fmaxf(cosI, 0.0f), 1.0f);
    float sin2T = eta * eta * (1.0f - cosI * cosI);
    if (sin2T >= 1.0f) return 1.0f;
    float cosT = sqrtf(1.0f - sin2T);
    float rs = (eta * cosI - cosT) / (eta * cosI + cosT);
    float rp = (cosI - eta * cosT) / (cosI + eta * cosT);
    return 0.5f * (rs * rs + rp * rp);
}
__device__ __forceinline__ V3 fresnelSchlick3(float cosI, V3 f0) {
    float c = 1.0f - fminf(fmaxf(cosI, 0.0f), 1.0f);
    float c2 = c * c;
    float c5 = c2 * c2 * c;
    return f0 + (v3s(1.0f) - f0) * c5;
}

// ---- the albedo tables (trilinear reads; axes as the reference's: mu*N - 0.5, r*(N-1), (n-1)(N-1)/2)
__device__ __forceinline__ void mfAxis(float x, int n, int& i0, float& f) {
    float c = fminf(fmaxf(x, 0.0f), (float)(n - 1));
    float fl = fminf(floorf(c), (float)(n - 2));
    i0 = (int)fl; f = c - fl;
}
__device__ __forceinline__ float mfEc(const float* __restrict__ T, float mu, float r, float nrel) {
    int im, ir, in_; float fm, fr, fn;
    mfAxis(mu * PT_MF_NMU - 0.5f, PT_MF_NMU, im, fm);
    mfAxis(r * (PT_MF_NR - 1), PT_MF_NR, ir, fr);
    mfAxis((nrel - 1.0f) * (PT_MF_NIOR - 1) * 0.5f, PT_MF_NIOR, in_, fn);
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        int di = c & 1, dk = (c >> 1) & 1, dj = (c >> 2) & 1;
        float w = (di ? fn : 1.0f - fn) * (dk ? fr : 1.0f - fr) * (dj ? fm : 1.0f - fm);
        acc += w * T[PT_MF_EC + ((in_ + di) * PT_MF_NR + ir + dk) * PT_MF_NMU + im + dj];
    }
    return acc;
}
__device__ __forceinline__ float mfEcAvg(const float* __restrict__ T, float r, float nrel) {
    int ir, in_; float fr, fn;
    mfAxis(r * (PT_MF_NR - 1), PT_MF_NR, ir, fr);
    mfAxis((nrel - 1.0f) * (PT_MF_NIOR - 1) * 0.5f, PT_MF_NIOR, in_, fn);
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        int di = c & 1, dk = (c >> 1) & 1;
        float w = (di ? fn : 1.0f - fn) * (dk ? fr : 1.0f - fr);
        acc += w * T[PT_MF_ECAVG + (in_ + di) * PT_MF_NR + ir + dk];
    }
    return acc;
}
__device__ __forceinline__ float mfE1(const float* __restrict__ T, float mu, float r) {
    int im, ir; float fm, fr;
    mfAxis(mu * PT_MF_NMU - 0.5f, PT_MF_NMU, im, fm);
    mfAxis(r * (PT_MF_NR - 1), PT_MF_NR, ir, fr);
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        int dk = c & 1, dj = (c >> 1) & 1;
        float w = (dk ? fr : 1.0f - fr) * (dj ? fm : 1.0f - fm);
        acc += w * T[PT_MF_E1 + (ir + dk) * PT_MF_NMU + im + dj];
    }
    return acc;
}

// ---- the lobes at one vertex
struct Lobes {
    V3 tX, tY, nZ;     // shading frame
    V3 o;              // view direction, local
    float a;           // GGX alpha
    float r;           // roughness
    bool isDelta;
    float metallic, trans;
    V3 base, tint;
    float coatK;       // Fresnel scale of the dielectric interface (2 * specular, 1 for transmissive)
    float eta;         // n_incident / n_beyond at the interface
    float nrel;
    float ps;          // interface pick at the view
    float ecO, ecAvg;
    V3 cm;
    bool through;      // evaluate the transmission lobe below the surface (guided sampling's mixture reads it)
};

__device__ __forceinline__ void frameOf(V3 n, V3& t, V3& b) {
    // Duff et al. 2017 branchless orthonormal basis
    float s = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (s + n.z);
    float bb = n.x * n.y * a;
    t = v3(1.0f + s * n.x * n.x * a, s * bb, -s * n.x);
    b = v3(bb, s + n.y * n.y * a, -n.y);
}
__device__ __forceinline__ V3 toLocal(const Lobes& lb, V3 w) { return v3(dot(w, lb.tX), dot(w, lb.tY), dot(w, lb.nZ)); }
__device__ __forceinline__ V3 toWorld(const Lobes& lb, V3 w) { return normalize(lb.tX * w.x + lb.tY * w.y + lb.nZ * w.z); }

__device__ __forceinline__ float coatF(const Lobes& lb, float cosI) { return fminf(lb.coatK * fresnelDielectric(cosI, lb.eta), 1.0f); }
__device__ __forceinline__ float coatAlb(const Lobes& lb, const float* __restrict__ T, float mu) {
    return fminf(lb.coatK * mfEc(T, mu, lb.r, lb.nrel), 1.0f);
}

// Beer-Lambert through distance d of an absorbing material: base^(d / attenuation) per channel, the colour held to
// [1e-6, 1] (materials.js mediumMu: exp(-mu d), mu = -ln(base) / attenuationDistance)
__device__ __forceinline__ V3 mediumTr(const PtMaterial& m, float d) {
    if (!(m.attenuation > 0.0f)) return v3s(1.0f);
    const float k = d / m.attenuation;
    return v3(exp2f(k * __log2f(fminf(fmaxf(m.base[0], 1e-6f), 1.0f))), exp2f(k * __log2f(fminf(fmaxf(m.base[1], 1e-6f), 1.0f))),
              exp2f(k * __log2f(fminf(fmaxf(m.base[2], 1e-6f), 1.0f))));
}

__device__ __forceinline__ Lobes makeLobes(V3 n, V3 view, const PtMaterial& m, float eta, const float* __restrict__ T) {
    Lobes lb;
    frameOf(n, lb.tX, lb.tY);
    lb.nZ = n;
    lb.o = v3(dot(view, lb.tX), dot(view, lb.tY), fmaxf(dot(view, n), 1e-4f));
    lb.r = saturate(m.roughness);
    float aIso = lb.r * lb.r;
    lb.isDelta = aIso < PT_ALPHA_MIN;
    lb.a = fminf(fmaxf(aIso, PT_ALPHA_MIN), 1.0f);
    lb.metallic = saturate(m.metallic);
    lb.trans = saturate(m.transmission);
    lb.base = ld3(m.base);
    // a material that absorbs inside takes its colour off along the way through it, so its interface passes every
    // channel alike (materials_pbr.js:206-210 transTint)
    lb.tint = m.attenuation > 0.0f ? v3s(1.0f) : lb.base;
    lb.coatK = mixf(2.0f * m.specular, 1.0f, lb.trans);
    lb.eta = eta;
    lb.nrel = fminf(fmaxf(1.0f / eta, 1.0f), 3.0f);
    float fo = coatF(lb, lb.o.z);
    lb.ps = fminf(fo, fresnelDielectric(lb.o.z, lb.eta) >= 1.0f ? 1.0f : 0.999f);
    lb.ecO = coatAlb(lb, T, lb.o.z);
    lb.ecAvg = fminf(lb.coatK * mfEcAvg(T, lb.r, lb.nrel), 1.0f);
    float e1 = lb.metallic > 0.0f ? mfE1(T, lb.o.z, lb.r) : 1.0f;
    lb.cm = v3s(1.0f) + lb.base * (1.0f / fmaxf(e1, 1e-3f) - 1.0f);
    lb.through = false;
    return lb;
}

// Refraction through a rough interface, i below the surface: the facet that takes o to i is h = -(eta o + i)
// normalised (turned to the view's side), a refraction only with o in front of it and i behind; the Jacobian
// |dw_h / dw_i| = |i.h| / (eta o.h + i.h)^2.
__device__ __forceinline__ bool refractFacet(const Lobes& lb, V3 i, V3& h, float& oh, float& ih, float& jac) {
    h = (lb.o * lb.eta + i) * -1.0f;
    const float hh = dot(h, h);
    if (!(hh > 1e-24f)) return false;
    h = h * frsq(hh);
    if (h.z < 0.0f) h = -h;
    oh = dot(lb.o, h);
    ih = dot(i, h);
    if (oh <= 0.0f || ih >= 0.0f) return false;
    const float den = lb.eta * oh + ih;
    jac = -ih / fmaxf(den * den, 1e-20f);
    return true;
}
// f below the surface: (1 - m) T (1 - F(o.h)) tint D(h) G2(o, i) (o.h) |i.h| / (o.z |i.z| (eta o.h + i.h)^2)
__device__ __forceinline__ V3 lobesEvalThrough(const Lobes& lb, V3 i) {
    const float wd = 1.0f - lb.metallic;
    V3 h; float oh, ih, jac;
    if (wd <= 0.0f || lb.trans <= 0.0f || lb.isDelta || !refractFacet(lb, i, h, oh, ih, jac)) return v3s(0.0f);
    const float k = wd * lb.trans * (1.0f - coatF(lb, oh)) * mfD(h, lb.a) * mfG2(lb.o, i, lb.a) * oh * jac / (lb.o.z * -i.z);
    return lb.tint * k;
}

// f(o, i): reflection above the surface; below it, the rough transmission where the caller asked (through).
__device__ __forceinline__ V3 lobesEval(const Lobes& lb, const float* __restrict__ T, V3 l) {
    V3 i = toLocal(lb, l);
    if (i.z < 0.0f && lb.through) return lobesEvalThrough(lb, i);
    if (i.z <= 0.0f) return v3s(0.0f);
    V3 f = v3s(0.0f);
    if (!lb.isDelta) {
        V3 h = normalize(lb.o + i);
        float spec = mfD(h, lb.a) * mfG2(lb.o, i, lb.a) / (4.0f * lb.o.z * i.z);
        float oh = dot(lb.o, h);
        if (lb.metallic > 0.0f) f += lb.cm * fresnelSchlick3(oh, lb.base) * (lb.metallic * spec);
        f += v3s((1.0f - lb.metallic) * coatF(lb, oh) * spec);
    }
    float wDiff = (1.0f - lb.metallic) * (1.0f - lb.trans);
    if (wDiff > 0.0f) {
        float ecI = coatAlb(lb, T, i.z);
        f += lb.base * (wDiff * (1.0f - lb.ecO) * (1.0f - ecI) / (PT_PI * fmaxf(1.0f - lb.ecAvg, 1e-4f)));
    }
    return f;
}
__device__ __forceinline__ float interfaceReflectP(const Lobes& lb, float oh) {
    float fh = coatF(lb, oh);
    float den = fh + (1.0f - fh) * lb.trans;
    return den > 1e-6f ? fh / den : 1.0f;
}
// the refraction's density under the sampler: the dielectric pick, the interface's, the refraction's share of the
// facet, the visible-normal density G1(o) (o.h) D(h) / o.z, and the Jacobian
__device__ __forceinline__ float lobesPdfThrough(const Lobes& lb, V3 i) {
    const float wd = 1.0f - lb.metallic;
    V3 h; float oh, ih, jac;
    if (wd <= 0.0f || lb.trans <= 0.0f || lb.isDelta || !refractFacet(lb, i, h, oh, ih, jac)) return 0.0f;
    const float pd = (1.0f - lb.ps) * (1.0f - lb.trans);
    return wd * (1.0f - pd) * (1.0f - interfaceReflectP(lb, oh)) * mfG1(lb.o, lb.a) * mfD(h, lb.a) * oh / lb.o.z * jac;
}
__device__ __forceinline__ float lobesPdf(const Lobes& lb, V3 l) {
    V3 i = toLocal(lb, l);
    if (i.z < 0.0f && lb.through) return lobesPdfThrough(lb, i);
    if (i.z <= 0.0f) return 0.0f;
    float wd = 1.0f - lb.metallic;
    float pd = (1.0f - lb.ps) * (1.0f - lb.trans);
    float p = wd * pd * i.z * PT_INV_PI;
    if (!lb.isDelta) {
        float pr = mfPdfReflect(lb.o, i, lb.a);
        p += (lb.metallic + wd * (1.0f - pd) * interfaceReflectP(lb, dot(lb.o, normalize(lb.o + i)))) * pr;
    }
    return p;
}
__device__ __forceinline__ bool lobesNone(const PtMaterial& m) {
    float r = saturate(m.roughness);
    return r * r < PT_ALPHA_MIN && (1.0f - saturate(m.metallic)) * (1.0f - saturate(m.transmission)) <= 0.0f;
}

#define PT_EV_NONE 0u
#define PT_EV_DELTA 1u
#define PT_EV_EVAL 2u
#define PT_EV_OTHER 3u
struct LobeSample {
    V3 dir;
    V3 w;
    unsigned int ev;
    bool transmitted;
    bool isDelta;
};
__device__ __forceinline__ V3 reflectLocal(V3 o, V3 h) { return h * (2.0f * dot(o, h)) - o; }
// Refraction of the local view o through the facet h, eta = n_incident / n_beyond; false at total reflection.
__device__ __forceinline__ bool refractLocal(V3 o, V3 h, float eta, V3& t) {
    float c = dot(o, h);
    float k = 1.0f - eta * eta * (1.0f - c * c);
    if (k <= 0.0f) return false;
    t = h * (eta * c - sqrtf(k)) - o * eta;
    return true;
}

__device__ __forceinline__ LobeSample lobesSample(const Lobes& lb, const float* __restrict__ T, V4 u) {
    LobeSample s;
    s.transmitted = false; s.isDelta = false; s.ev = PT_EV_EVAL; s.w = v3s(0.0f); s.dir = lb.nZ;
    V3 o = lb.o;
    if (u.x < lb.metallic) {
        V3 h = lb.isDelta ? v3(0.0f, 0.0f, 1.0f) : mfSampleVNDF(o, lb.a, u.y, u.z);
        V3 i = reflectLocal(o, h);
        if (i.z <= 0.0f) return s;
        float oh = dot(o, h);
        if (lb.isDelta) { s.w = fresnelSchlick3(oh, lb.base); s.ev = PT_EV_DELTA; s.isDelta = true; }
        else { s.w = lb.cm * fresnelSchlick3(oh, lb.base) * (mfG2(o, i, lb.a) / mfG1(o, lb.a)); }
        s.dir = toWorld(lb, i);
        return s;
    }
    float ux = (u.x - lb.metallic) / fmaxf(1.0f - lb.metallic, 1e-7f);
    float pd = (1.0f - lb.ps) * (1.0f - lb.trans);
    if (ux < pd) {
        // the diffuse base: cosine-weighted
        float r = sqrtf(u.y), phi = 2.0f * PT_PI * u.z;
        float sp, cp;
        __sincosf(phi, &sp, &cp);
        V3 i = v3(r * cp, r * sp, sqrtf(fmaxf(0.0f, 1.0f - u.y)));
        float ecI = coatAlb(lb, T, i.z);
        s.w = lb.base * ((1.0f - lb.trans) * (1.0f - lb.ecO) * (1.0f - ecI) / (fmaxf(1.0f - lb.ecAvg, 1e-4f) * pd));
        s.dir = toWorld(lb, i);
        return s;
    }
    // the interface: a facet, then reflection or refraction by that facet's Fresnel
    V3 h = lb.isDelta ? v3(0.0f, 0.0f, 1.0f) : mfSampleVNDF(o, lb.a, u.y, u.z);
    float oh = dot(o, h);
    float pR = interfaceReflectP(lb, oh);
    float fh = coatF(lb, oh);
    float g = lb.isDelta ? 1.0f : 0.0f;
    float uw = (ux - pd) / fmaxf(1.0f - pd, 1e-7f);
    if (uw < pR) {
        V3 i = reflectLocal(o, h);
        if (i.z <= 0.0f) return s;
        float geo = lb.isDelta ? 1.0f : mfG2(o, i, lb.a) / mfG1(o, lb.a);
        s.w = v3s(fh * geo / ((1.0f - pd) * pR));
        s.dir = toWorld(lb, i);
        s.ev = lb.isDelta ? PT_EV_DELTA : PT_EV_EVAL;
        s.isDelta = lb.isDelta;
        (void)g;
        return s;
    }
    V3 t;
    if (!refractLocal(o, h, lb.eta, t) || t.z >= 0.0f) return s;
    float geo = lb.isDelta ? 1.0f : mfG2(o, v3(t.x, t.y, -t.z), lb.a) / mfG1(o, lb.a);
    s.w = lb.tint * ((1.0f - fh) * lb.trans * geo / ((1.0f - pd) * (1.0f - pR)));
    s.dir = toWorld(lb, t);
    s.transmitted = true;
    s.ev = lb.isDelta ? PT_EV_DELTA : PT_EV_OTHER;
    s.isDelta = lb.isDelta;
    return s;
}
