// The scene code's library on the device: what an object's, a group's or a material's WGSL (translated by
// shadytracer_amd/js/wgsl.js over pt_wgsl_rt.h) may call besides the language's builtins.
//   - The Hit and its fold (core/scene_api.js:26-97 add_new, fold, emptyHit; :173 warp): a Hit keeps the nearest
//     surface by |d| and, of the surfaces the point is inside, the least deep one's material.
//   - The shape functions by their names in core/sdf.js and core/shapes.js, over this build's own primitives (the
//     prelude's) where one exists, and written here from the closed forms otherwise.
//   - The solid noise of core/noise.js: value and gradient noise with analytic gradients, fBm, Worley cells.
// Materials are packed with the primitive that made the surface: m = material | primitive << 8.
#pragma once

#include "pt_wgsl_rt.h"

WG_FN V3 toV3(wvec3 a) { return v3(a.x, a.y, a.z); }
WG_FN wvec3 ofV3(V3 a) { return wvec3(a.x, a.y, a.z); }
WG_FN wvec4 ofV4(V4 a) { return wvec4(a.x, a.y, a.z, a.w); }
WG_FN V4 toV4(wvec4 a) { return v4(v3(a.x, a.y, a.z), a.w); }

// ---- the Hit ------------------------------------------------------------------------------------------------------------
struct WHit {
    wvec4 s;               // the nearest surface: outward normal (world frame of the code), signed distance
    unsigned int m;        // its material | primitive << 8
    unsigned int a;        // the material the point is least deep inside of
    WF insideW;            // ...and how deep
};
WG_FN WHit w_emptyHit(int) { WHit h; h.s = wvec4(0.0f, 0.0f, 1.0f, 1.0e7f); h.m = 0u; h.a = 0u; h.insideW = -1.0e7f; return h; }
WG_FN WHit w_fold(WHit C, wvec4 s, unsigned int m) {
    if (s.w < 0.0f && s.w > C.insideW) { C.insideW = s.w; C.a = m; }
    if (fabsf(s.w) < fabsf(C.s.w)) { C.s = s; C.m = m; }
    return C;
}
// The code's frame: M turns its normals to the frame the Hit is in, S scales its distances, K is the product of the
// warps' factors. warp(k) after a change of the point that is not a move or a turn: distances times k from here on,
// normals zero (left to differences).
struct WXf {
    wmat3 M;
    WF S, K;
    unsigned int SUB;
    WG_FN WXf() : S(1.0f), K(1.0f), SUB(0u) { M = mk_wmat3(wvec3(1.0f, 0.0f, 0.0f), wvec3(0.0f, 1.0f, 0.0f), wvec3(0.0f, 0.0f, 1.0f)); }
    WG_FN WHit add_new(WHit C, wvec4 s, unsigned int m) const { return w_fold(C, mk_wvec4(M * wsw3(s, 0, 1, 2), s.w * S), m); }
    WG_FN void warp(WF k) { S *= k; K *= k; M = mk_wmat3(); }
};

// ---- a material as its code sees it (builder/codegen.js emitMaterials: m's fields), and back ---------------------------
struct WMat { wvec3 baseColor; WF metallic, roughness, transmission, ior, specular, specularTint; wvec3 tangent; };
WG_FN WMat wmatOf(const PtMaterial& m) {
    WMat w;
    w.baseColor = wvec3(m.base[0], m.base[1], m.base[2]);
    w.metallic = m.metallic; w.roughness = m.roughness; w.transmission = m.transmission; w.ior = m.ior;
    w.specular = m.specular; w.specularTint = m.specularTint;
    return w;
}
// the fields back, each held to the range the BSDF takes (colours and the weights non-negative, the weights and the
// roughness at most 1, the index at least 1); the tangent has no use without anisotropy, which the engine does not draw
WG_FN void wmatTo(const WMat& w, PtMaterial& m) {
    m.base[0] = fmaxf(w.baseColor.x, 0.0f); m.base[1] = fmaxf(w.baseColor.y, 0.0f); m.base[2] = fmaxf(w.baseColor.z, 0.0f);
    m.metallic = fminf(fmaxf(w.metallic, 0.0f), 1.0f); m.roughness = fminf(fmaxf(w.roughness, 0.0f), 1.0f);
    m.transmission = fminf(fmaxf(w.transmission, 0.0f), 1.0f); m.ior = fmaxf(w.ior, 1.0f);
    m.specular = fmaxf(w.specular, 0.0f); m.specularTint = fminf(fmaxf(w.specularTint, 0.0f), 1.0f);
}

// ---- shapes (core/sdf.js, core/shapes.js names; the reference's parameter meanings) -------------------------------------
WG_FN wvec4 w_sd(WF f) { return wvec4(0.0f, 0.0f, 0.0f, f); }
WG_FN wvec4 w_sdSphere(wvec3 p, WF r) { return w_sd(w_length(p) - r); }
WG_FN wvec4 w_sdBox(wvec3 p, wvec3 b) {
    const wvec3 q = w_abs(p) - b;
    return w_sd(w_length(w_max(q, wvec3(0.0f))) + fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.0f));
}
WG_FN wvec4 w_sdRoundedBox(wvec3 p, wvec3 b, WF r) { wvec4 d = w_sdBox(p, b); d.w -= r; return d; }
WG_FN wvec4 w_sdCylinder(wvec3 p, WF h, WF r) { return w_sd(Dcylinder(toV3(p), h, r)); }
WG_FN wvec4 w_sdCappedCone(wvec3 p, WF h, WF r1, WF r2) { return w_sd(Dcone(toV3(p), h, r1, r2)); }
// The Menger sponge of side 2 about the origin, `iter` levels: the cube, less at each level k the cross of square holes
// of side 2/3^(k+1) through every cell of the 3^k grid. A cell's cross, in the cell's own coordinates a in [-1, 1]^3
// scaled by s = 3^(k+1): the points where two coordinates are within 1/3 of the centre line.
WG_FN wvec4 w_sdMenger(wvec3 p, unsigned int iter) {
    WF d = w_sdBox(p, wvec3(1.0f)).w;
    WF s = 1.0f;
    for (unsigned int i = 0u; i < iter; ++i) {
        const wvec3 a = w_fract(p * (0.5f * s)) * 2.0f - 1.0f;
        s *= 3.0f;
        const wvec3 r = w_abs(a) * 0.5f;
        const WF c = (fminf(fmaxf(r.x, r.y), fminf(fmaxf(r.y, r.z), fmaxf(r.z, r.x))) - 1.0f / 3.0f) / s;
        d = fmaxf(d, c);
    }
    return w_sd(d);
}
WG_FN wvec4 w_Nsphere(wvec3 p, WF r) { return ofV4(Nsphere(toV3(p), r)); }
WG_FN wvec4 w_Nroundedbox(wvec3 p, wvec3 b, WF r) { return ofV4(Nroundedbox(toV3(p), toV3(b), r)); }
WG_FN wvec4 w_Nbox(wvec3 p, wvec3 b) { return ofV4(Nroundedbox(toV3(p), toV3(b), 0.0f)); }
WG_FN wvec4 w_NboxRounded(wvec3 p, wvec3 b, WF r) { return ofV4(NboxRounded(toV3(p), toV3(b), r)); }
WG_FN wvec4 w_Ncylinder(wvec3 p, WF h, WF r) { return ofV4(Ncylinder(toV3(p), h, r)); }
WG_FN wvec4 w_Ncapsule(wvec3 p, WF h, WF r) { return ofV4(Ncapsule(toV3(p), h, r)); }
WG_FN wvec4 w_Ntorus(wvec3 p, WF R, WF r) { return ofV4(Ntorus(toV3(p), R, r)); }
WG_FN wvec4 w_Nplane(wvec3 p) { return ofV4(Nplane(toV3(p))); }
WG_FN wvec4 w_Nrectangle(wvec3 p, wvec2 h) { return ofV4(Nrectangle(toV3(p), h.x, h.y)); }
WG_FN wvec4 w_Ndisc(wvec3 p, WF r) { return ofV4(Ndisc(toV3(p), r)); }
WG_FN wvec4 w_NroundCone(wvec3 p, WF r1, WF r2, WF h) { return ofV4(NroundCone(toV3(p), r1, r2, h)); }
WG_FN wvec4 w_Noctahedron(wvec3 p, WF s) { return ofV4(Noctahedron(toV3(p), s)); }
WG_FN wvec4 w_Nlink(wvec3 p, WF span, WF R, WF r) { return ofV4(Nlink(toV3(p), span, R, r)); }
WG_FN wvec4 w_Ntube(wvec3 p, WF outer, WF inner, WF h) { return ofV4(Ntube(toV3(p), outer, inner, h)); }
WG_FN wvec4 w_Ndome(wvec3 p, WF R, WF h) { return ofV4(Ndome(toV3(p), R, h)); }
WG_FN wvec3 w_safeDir(wvec3 v, wvec3 fb) { const WF l = w_length(v); return l > 1e-12f ? v / l : fb; }
// A 2-D (normal xy, distance) extruded along z, h either side of z = 0: outside both the profile and the slab, the
// corner's distance and direction; else the larger of the two, with its own normal.
WG_FN wvec4 w_g2Extrude(wvec3 p, wvec3 g, WF h) {
    const WF wz = fabsf(p.z) - h, sz = p.z >= 0.0f ? 1.0f : -1.0f;
    if (g.z > 0.0f && wz > 0.0f) {
        const WF l = sqrtf(g.z * g.z + wz * wz);
        return wvec4(g.x * (g.z / l), g.y * (g.z / l), sz * wz / l, l);
    }
    if (g.z > wz) return wvec4(g.x, g.y, 0.0f, g.z);
    return wvec4(0.0f, 0.0f, sz, wz);
}
// A 2-D (normal in (r, z), distance) revolved about z.
WG_FN wvec4 w_g2Revolve(wvec3 p, wvec3 g) {
    const WF r = sqrtf(p.x * p.x + p.y * p.y);
    const WF cx = r > 1e-12f ? p.x / r : 1.0f, cy = r > 1e-12f ? p.y / r : 0.0f;
    return wvec4(g.x * cx, g.x * cy, g.y, g.z);
}

// ---- noise (core/noise.js) -----------------------------------------------------------------------------------------------
// PCG4D of a lattice cell's integer coordinates and a seed; 24-bit uniforms in [0, 1)
WG_FN wuvec4 nzHash(wvec3 i, unsigned int seed) {
    Rng s; s.a = (unsigned int)(int)i.x; s.b = (unsigned int)(int)i.y; s.c = (unsigned int)(int)i.z; s.d = seed;
    pcg4d(s);
    return wuvec4(s.a, s.b, s.c, s.d);
}
WG_FN wvec4 nzUnit(wuvec4 h) { const WF k = 1.0f / 16777216.0f; return wvec4((WF)(h.x >> 8) * k, (WF)(h.y >> 8) * k, (WF)(h.z >> 8) * k, (WF)(h.w >> 8) * k); }
WG_FN wvec3 nzGrad(wvec3 i, unsigned int seed) {
    const wvec4 u = nzUnit(nzHash(i, seed));
    const wvec3 g = wvec3(u.x, u.y, u.z) * 2.0f - 1.0f;
    return g / fmaxf(w_length(g), 1.0f);
}
// The quintic fade 6t^5 - 15t^4 + 10t^3 and its slope
WG_FN WF nzFade(WF t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }
WG_FN WF nzFadeD(WF t) { return 30.0f * t * t * (t - 1.0f) * (t - 1.0f); }
#define NZ_GRAD_SCALE 1.6f
// Value noise (gradient xyz, value w), values in [-1, 1] at the lattice's corners blended by the fade. Gradient noise
// (Perlin's construction): each corner's random gradient g . (f - corner), blended, times 1.6.
template <bool GRAD> WG_FN wvec4 nzLattice(wvec3 p, unsigned int seed) {
    const wvec3 i = w_floor(p), f = p - i;
    const wvec3 u = wvec3(nzFade(f.x), nzFade(f.y), nzFade(f.z)), du = wvec3(nzFadeD(f.x), nzFadeD(f.y), nzFadeD(f.z));
    wvec4 acc;
    for (unsigned int c = 0u; c < 8u; ++c) {
        const wvec3 o = wvec3((WF)(c & 1u), (WF)((c >> 1) & 1u), (WF)((c >> 2) & 1u));
        const wvec3 w = wvec3(o.x > 0.5f ? u.x : 1.0f - u.x, o.y > 0.5f ? u.y : 1.0f - u.y, o.z > 0.5f ? u.z : 1.0f - u.z);
        const wvec3 dw = wvec3(o.x > 0.5f ? du.x : -du.x, o.y > 0.5f ? du.y : -du.y, o.z > 0.5f ? du.z : -du.z);
        const wvec3 gw = wvec3(dw.x * w.y * w.z, w.x * dw.y * w.z, w.x * w.y * dw.z);
        const WF wc = w.x * w.y * w.z;
        if (GRAD) {
            const wvec3 g = nzGrad(i + o, seed);
            const WF v = w_dot(g, f - o);
            acc += mk_wvec4(g * wc + gw * v, wc * v);
        } else {
            const WF v = nzUnit(nzHash(i + o, seed)).x * 2.0f - 1.0f;
            acc += mk_wvec4(gw * v, wc * v);
        }
    }
    return GRAD ? acc * NZ_GRAD_SCALE : acc;
}
WG_FN wvec4 w_vnoise3(wvec3 p, unsigned int seed) { return nzLattice<false>(p, seed); }
WG_FN wvec4 w_gnoise3(wvec3 p, unsigned int seed) { return nzLattice<true>(p, seed); }
// octaves of gradient noise at frequency lac^k, amplitude gain^k, seed + 7919 k
WG_FN wvec4 w_fbm3(wvec3 p, unsigned int seed, unsigned int octaves, WF lac, WF gain) {
    wvec4 acc;
    WF fr = 1.0f, am = 1.0f;
    for (unsigned int k = 0u; k < octaves; ++k) {
        const wvec4 n = w_gnoise3(p * fr, seed + k * 7919u);
        acc += wvec4(n.x * am * fr, n.y * am * fr, n.z * am * fr, n.w * am);
        fr *= lac;
        am *= gain;
    }
    return acc;
}
WG_FN WF w_nzFbmLip(unsigned int octaves, WF lac, WF gain, WF L) {
    WF s = 0.0f, t = 1.0f;
    for (unsigned int k = 0u; k < octaves; ++k) { s += t; t *= gain * lac; }
    return L * s;
}
#define NZ_JITTER 0.46f
// Worley cells: (F1, F2, id1, id2) over the 27 cells about p, each cell's point within the middle 0.46 of it
WG_FN wvec4 w_worley3(wvec3 p, unsigned int seed) {
    const wvec3 i = w_floor(p);
    WF f1 = 1e9f, f2 = 1e9f, id1 = 0.0f, id2 = 0.0f;
    for (int n = 0; n < 27; ++n) {
        const wvec3 c = i + wvec3((WF)(n % 3), (WF)((n / 3) % 3), (WF)(n / 9)) - 1.0f;
        const wvec4 h = nzUnit(nzHash(c, seed));
        const wvec3 q = c + 0.5f + (wvec3(h.x, h.y, h.z) - 0.5f) * NZ_JITTER;
        const WF d = w_length(p - q);
        if (d < f1) { f2 = f1; id2 = id1; f1 = d; id1 = h.w; }
        else if (d < f2) { f2 = d; id2 = h.w; }
    }
    return wvec4(f1, f2, id1, id2);
}
// Two Worley point sets, the second turned and its cells 1.37 times the first's, scaled so the two hold one point per
// unit volume together (1/A^3 + 1/B^3 = 1); the nearest two of their four
WG_FN wvec4 w_worley3x(wvec3 p, unsigned int seed) {
    const WF A = 1.1522346f, B = 1.5785614f;   // cbrt(1 + 1/1.37^3), and times 1.37
    const wmat3 turn = mk_wmat3(0.36f, 0.48f, -0.8f, -0.8f, 0.6f, 0.0f, 0.48f, 0.64f, 0.6f);
    const wvec4 a0 = w_worley3(p / A, seed), b0 = w_worley3((turn * p) / B, seed + 1013u);
    const wvec4 a = wvec4(a0.x * A, a0.y * A, a0.z, a0.w), b = wvec4(b0.x * B, b0.y * B, b0.z, b0.w);
    if (a.x <= b.x) return wvec4(a.x, fminf(a.y, b.x), a.z, a.y <= b.x ? a.w : b.z);
    return wvec4(b.x, fminf(b.y, a.x), b.z, b.y <= a.x ? b.w : a.z);
}
#define NZ_GRAD_LIP_MEASURED 4.3f
#define NZ_GRAD_VALUE_BOUND 1.3856406f
