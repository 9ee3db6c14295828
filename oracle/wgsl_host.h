/* The scene code's library in FP64 for the oracle -- test infrastructure only: tests/ compile an emitted scene's
 * object and group code (shadytracer_amd/js/usercode.js, the `user` list of a scene) against this header into a
 * module the oracle calls; the product never includes it.
 *
 * The language runtime (types, operators, builtins) is the translation's own, pt_wgsl_rt.h, here over double. What
 * the code calls besides is restated here: the Hit and its fold (core/scene_api.js:26-97, :173 warp), the shape
 * functions of core/sdf.js and core/shapes.js (the oracle's own primitives, oracle.c primLocal, reached through
 * or_prim_local, or the closed forms written below), and the noise of core/noise.js (written below from its
 * definitions: PCG4D lattice hashes, quintic fade, gradient and value noise, fBm, Worley F1/F2 over 27 cells). */
#pragma once
#include <cmath>
#include <cstdint>

#define WG_FN inline
typedef double WF;
#define WL(x) x
#include "pt_wgsl_rt.h"

/* the oracle's primitives (oracle.c primLocal), set by the loader: shape index, point, parameters, (normal, d) out */
static void (*g_prim)(int, const double*, const float*, double*) = 0;
extern "C" void wg_set_prim(void* f) { g_prim = (void (*)(int, const double*, const float*, double*))f; }
enum { WS_PLANE = 0, WS_SPHERE = 1, WS_BOX = 2, WS_CYLINDER = 3, WS_CAPSULE = 4, WS_TORUS = 5, WS_RECT = 6, WS_DISC = 7,
       WS_CONE = 15, WS_ROUNDCONE = 16, WS_OCTA = 17, WS_LINK = 19, WS_TUBE = 20, WS_DOME = 21 };
static wvec4 wprim(int shape, wvec3 p, float a = 0, float b = 0, float c = 0, float d = 0) {
    const double q[3] = {p.x, p.y, p.z};
    const float P[4] = {a, b, c, d};
    double o[4];
    g_prim(shape, q, P, o);
    return wvec4(o[0], o[1], o[2], o[3]);
}

struct WHit { wvec4 s; unsigned int m, a; WF insideW; };
WG_FN WHit w_emptyHit(int) { WHit h; h.s = wvec4(0.0, 0.0, 1.0, 1.0e7); h.m = 0u; h.a = 0u; h.insideW = -1.0e7; return h; }
WG_FN WHit w_fold(WHit C, wvec4 s, unsigned int m) {
    if (s.w < 0.0 && s.w > C.insideW) { C.insideW = s.w; C.a = m; }
    if (std::fabs(s.w) < std::fabs(C.s.w)) { C.s = s; C.m = m; }
    return C;
}
struct WXf {
    wmat3 M;
    WF S, K;
    unsigned int SUB;
    WXf() : S(1.0), K(1.0), SUB(0u) { M = mk_wmat3(wvec3(1, 0, 0), wvec3(0, 1, 0), wvec3(0, 0, 1)); }
    WHit add_new(WHit C, wvec4 s, unsigned int m) const { return w_fold(C, wvec4(0, 0, 0, s.w * S) + mk_wvec4(M * wvec3(s.x, s.y, s.z), 0.0), m); }
    void warp(WF k) { S *= k; K *= k; M = mk_wmat3(); }
};

/* a material as its code sees it, and back (each field held to the range the BSDF takes) */
#include "../include/pt_types.h"
struct WMat { wvec3 baseColor; WF metallic, roughness, transmission, ior, specular, specularTint; wvec3 tangent; };
inline WMat wmatOf(const PtMaterial& m) {
    WMat w;
    w.baseColor = wvec3(m.base[0], m.base[1], m.base[2]);
    w.metallic = m.metallic; w.roughness = m.roughness; w.transmission = m.transmission; w.ior = m.ior;
    w.specular = m.specular; w.specularTint = m.specularTint;
    return w;
}
inline float wclampf(WF x, WF lo, WF hi) { return (float)(x < lo ? lo : (x > hi ? hi : x)); }
inline void wmatTo(const WMat& w, PtMaterial& m) {
    m.base[0] = wclampf(w.baseColor.x, 0, 1e30); m.base[1] = wclampf(w.baseColor.y, 0, 1e30); m.base[2] = wclampf(w.baseColor.z, 0, 1e30);
    m.metallic = wclampf(w.metallic, 0, 1); m.roughness = wclampf(w.roughness, 0, 1); m.transmission = wclampf(w.transmission, 0, 1);
    m.ior = wclampf(w.ior, 1, 1e30); m.specular = wclampf(w.specular, 0, 1e30); m.specularTint = wclampf(w.specularTint, 0, 1);
}

WG_FN wvec4 w_sd(WF f) { return wvec4(0, 0, 0, f); }
WG_FN wvec4 w_sdSphere(wvec3 p, WF r) { return w_sd(w_length(p) - r); }
/* a box of half-size b: the distance to its surface, outside along the nearest point, inside the nearest face */
WG_FN wvec4 w_sdBox(wvec3 p, wvec3 b) {
    const WF qx = std::fabs(p.x) - b.x, qy = std::fabs(p.y) - b.y, qz = std::fabs(p.z) - b.z;
    const WF ox = std::fmax(qx, 0.0), oy = std::fmax(qy, 0.0), oz = std::fmax(qz, 0.0);
    return w_sd(std::sqrt(ox * ox + oy * oy + oz * oz) + std::fmin(std::fmax(qx, std::fmax(qy, qz)), 0.0));
}
WG_FN wvec4 w_sdRoundedBox(wvec3 p, wvec3 b, WF r) { wvec4 d = w_sdBox(p, b); d.w -= r; return d; }
WG_FN wvec4 w_sdCylinder(wvec3 p, WF h, WF r) { return w_sd(wprim(WS_CYLINDER, p, (float)h, (float)r, 0.0f).w); }
WG_FN wvec4 w_sdCappedCone(wvec3 p, WF h, WF r1, WF r2) { return w_sd(wprim(WS_CONE, p, (float)h, (float)r1, (float)r2).w); }
/* The Menger sponge of side 2, iter levels: the cube, less at level k (s = 3^k) the holes where two of a point's
 * coordinates lie in the middle third of their cells of the sponge's 3^k grid (cells of side 2/s, centred on the even
 * multiples of 1/s). a = t - 2 floor(t/2) - 1 with t = p s runs -1..1 between two cell centres, so |a| > 2/3 (r = |a|/2
 * > 1/3) for a coordinate within a third of a cell's half side of its centre; two such make a hole, and (the second
 * largest r - 1/3) / 3s bounds the distance to the hole carved at that level. */
WG_FN wvec4 w_sdMenger(wvec3 p, unsigned int iter) {
    WF d = w_sdBox(p, wvec3(1.0)).w;
    WF s = 1.0;
    for (unsigned int k = 0; k < iter; ++k) {
        WF r[3];
        for (int j = 0; j < 3; ++j) {
            const WF t = p[j] * s;
            r[j] = std::fabs(t - 2.0 * std::floor(0.5 * t) - 1.0) * 0.5;
        }
        s *= 3.0;
        const WF m = std::fmin(std::fmax(r[0], r[1]), std::fmin(std::fmax(r[1], r[2]), std::fmax(r[2], r[0])));
        d = std::fmax(d, (m - 1.0 / 3.0) / s);
    }
    return w_sd(d);
}
WG_FN wvec4 w_Nsphere(wvec3 p, WF r) { return wprim(WS_SPHERE, p, (float)r); }
WG_FN wvec4 w_NboxRounded(wvec3 p, wvec3 b, WF r) { return wprim(WS_BOX, p, (float)b.x, (float)b.y, (float)b.z, (float)r); }
WG_FN wvec4 w_Nroundedbox(wvec3 p, wvec3 b, WF r) { return w_NboxRounded(p, b + r, r); }
WG_FN wvec4 w_Nbox(wvec3 p, wvec3 b) { return w_NboxRounded(p, b, 0.0); }
WG_FN wvec4 w_Ncylinder(wvec3 p, WF h, WF r) { return wprim(WS_CYLINDER, p, (float)h, (float)r, 0.0f); }
WG_FN wvec4 w_Ncapsule(wvec3 p, WF h, WF r) { return wprim(WS_CAPSULE, p, (float)h, (float)r); }
WG_FN wvec4 w_Ntorus(wvec3 p, WF R, WF r) { return wprim(WS_TORUS, p, (float)R, (float)r); }
WG_FN wvec4 w_Nplane(wvec3 p) { return wvec4(0, 0, 1, p.z); }
WG_FN wvec4 w_Nrectangle(wvec3 p, wvec2 h) { return wprim(WS_RECT, p, (float)h.x, (float)h.y); }
WG_FN wvec4 w_Ndisc(wvec3 p, WF r) { return wprim(WS_DISC, p, (float)r); }
WG_FN wvec4 w_NroundCone(wvec3 p, WF r1, WF r2, WF h) { return wprim(WS_ROUNDCONE, p, (float)r1, (float)r2, (float)h); }
WG_FN wvec4 w_Noctahedron(wvec3 p, WF s) { return wprim(WS_OCTA, p, (float)s); }
WG_FN wvec4 w_Nlink(wvec3 p, WF span, WF R, WF r) { return wprim(WS_LINK, p, (float)span, (float)R, (float)r); }
WG_FN wvec4 w_Ntube(wvec3 p, WF outer, WF inner, WF h) { return wprim(WS_TUBE, p, (float)outer, (float)inner, (float)h); }
WG_FN wvec4 w_Ndome(wvec3 p, WF R, WF h) { return wprim(WS_DOME, p, (float)R, (float)h); }
WG_FN wvec3 w_safeDir(wvec3 v, wvec3 fb) { const WF l = w_length(v); return l > 1e-12 ? v / l : fb; }
/* a 2-D field (normal xy, distance) extruded h either side of z = 0: the exact extrusion of an exact 2-D distance */
WG_FN wvec4 w_g2Extrude(wvec3 p, wvec3 g, WF h) {
    const WF wz = std::fabs(p.z) - h, sz = p.z >= 0 ? 1.0 : -1.0;
    if (g.z > 0 && wz > 0) { const WF l = std::hypot(g.z, wz); return wvec4(g.x * g.z / l, g.y * g.z / l, sz * wz / l, l); }
    return g.z > wz ? wvec4(g.x, g.y, 0, g.z) : wvec4(0, 0, sz, wz);
}
WG_FN wvec4 w_g2Revolve(wvec3 p, wvec3 g) {
    const WF r = std::hypot(p.x, p.y);
    return r > 1e-12 ? wvec4(g.x * p.x / r, g.x * p.y / r, g.y, g.z) : wvec4(g.x, 0, g.y, g.z);
}
/* ---- noise ---------------------------------------------------------------------------------------------------- */
/* PCG4D (Jarzynski & Olano 2020) of the cell's integer coordinates and the seed */
static void wpcg(uint32_t v[4]) {
    for (int i = 0; i < 4; ++i) v[i] = v[i] * 1664525u + 1013904223u;
    v[0] += v[1] * v[3]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1]; v[3] += v[1] * v[2];
    for (int i = 0; i < 4; ++i) v[i] ^= v[i] >> 16;
    v[0] += v[1] * v[3]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1]; v[3] += v[1] * v[2];
}
static void wcell(const WF c[3], unsigned int seed, WF u[4]) {
    uint32_t v[4] = {(uint32_t)(int32_t)c[0], (uint32_t)(int32_t)c[1], (uint32_t)(int32_t)c[2], seed};
    wpcg(v);
    for (int i = 0; i < 4; ++i) u[i] = (WF)(v[i] >> 8) / 16777216.0;
}
static WF wfade(WF t) { return ((6 * t - 15) * t + 10) * t * t * t; }
static WF wfadeD(WF t) { return 30 * t * t * (1 - t) * (1 - t); }
/* trilinear blend of the eight corners with the fade, and its gradient; grad: Perlin's corner term g . (f - corner),
 * g the corner's random vector shortened to the unit ball, times 1.6; else the corner's value 2u - 1 */
static wvec4 wlattice(wvec3 p, unsigned int seed, bool grad) {
    const WF i[3] = {std::floor(p.x), std::floor(p.y), std::floor(p.z)};
    const WF f[3] = {p.x - i[0], p.y - i[1], p.z - i[2]};
    WF val = 0, g[3] = {0, 0, 0};
    for (int c = 0; c < 8; ++c) {
        const int o[3] = {c & 1, (c >> 1) & 1, (c >> 2) & 1};
        WF w[3], dw[3];
        for (int j = 0; j < 3; ++j) { w[j] = o[j] ? wfade(f[j]) : 1 - wfade(f[j]); dw[j] = o[j] ? wfadeD(f[j]) : -wfadeD(f[j]); }
        const WF cc[3] = {i[0] + o[0], i[1] + o[1], i[2] + o[2]};
        WF u[4];
        wcell(cc, seed, u);
        const WF W = w[0] * w[1] * w[2];
        const WF gw[3] = {dw[0] * w[1] * w[2], w[0] * dw[1] * w[2], w[0] * w[1] * dw[2]};
        if (grad) {
            WF gv[3] = {2 * u[0] - 1, 2 * u[1] - 1, 2 * u[2] - 1};
            const WF l = std::sqrt(gv[0] * gv[0] + gv[1] * gv[1] + gv[2] * gv[2]);
            if (l > 1) for (int j = 0; j < 3; ++j) gv[j] /= l;
            const WF v = gv[0] * (f[0] - o[0]) + gv[1] * (f[1] - o[1]) + gv[2] * (f[2] - o[2]);
            val += W * v;
            for (int j = 0; j < 3; ++j) g[j] += W * gv[j] + v * gw[j];
        } else {
            const WF v = 2 * u[0] - 1;
            val += W * v;
            for (int j = 0; j < 3; ++j) g[j] += v * gw[j];
        }
    }
    const WF s = grad ? 1.6 : 1.0;
    return wvec4(g[0] * s, g[1] * s, g[2] * s, val * s);
}
WG_FN wvec4 w_vnoise3(wvec3 p, unsigned int seed) { return wlattice(p, seed, false); }
WG_FN wvec4 w_gnoise3(wvec3 p, unsigned int seed) { return wlattice(p, seed, true); }
WG_FN wvec4 w_fbm3(wvec3 p, unsigned int seed, unsigned int octaves, WF lac, WF gain) {
    wvec4 acc;
    for (unsigned int k = 0; k < octaves; ++k) {
        const WF fr = std::pow(lac, (WF)k), am = std::pow(gain, (WF)k);
        const wvec4 n = w_gnoise3(p * fr, seed + k * 7919u);
        acc = acc + wvec4(n.x * am * fr, n.y * am * fr, n.z * am * fr, n.w * am);
    }
    return acc;
}
WG_FN WF w_nzFbmLip(unsigned int octaves, WF lac, WF gain, WF L) {
    WF s = 0;
    for (unsigned int k = 0; k < octaves; ++k) s += std::pow(gain * lac, (WF)k);
    return L * s;
}
/* Worley: each cell's feature point at its centre moved by (u - 1/2) 0.46 per axis; F1, F2 over the 27 cells about p
 * (x fastest, then y, then z, so equal distances keep the earlier cell, as the device's loop does); ids u[3] */
WG_FN wvec4 w_worley3(wvec3 p, unsigned int seed) {
    const WF i[3] = {std::floor(p.x), std::floor(p.y), std::floor(p.z)};
    WF f1 = 1e9, f2 = 1e9, id1 = 0, id2 = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const WF c[3] = {i[0] + dx, i[1] + dy, i[2] + dz};
                WF u[4];
                wcell(c, seed, u);
                WF e = 0;
                for (int j = 0; j < 3; ++j) { const WF q = c[j] + 0.5 + (u[j] - 0.5) * 0.46 - p[j]; e += q * q; }
                const WF d = std::sqrt(e);
                if (d < f1) { f2 = f1; id2 = id1; f1 = d; id1 = u[3]; }
                else if (d < f2) { f2 = d; id2 = u[3]; }
            }
    return wvec4(f1, f2, id1, id2);
}
/* two point sets, the second turned (columns (0.36, 0.48, -0.8), (-0.8, 0.6, 0), (0.48, 0.64, 0.6)) and 1.37 times
 * the first's cells, each scaled so together they have one point per unit volume: the nearest two of the four */
WG_FN wvec4 w_worley3x(wvec3 p, unsigned int seed) {
    const WF A = std::cbrt(1 + 1 / (1.37 * 1.37 * 1.37)), B = A * 1.37;
    const wmat3 T = mk_wmat3(0.36, 0.48, -0.8, -0.8, 0.6, 0.0, 0.48, 0.64, 0.6);
    const wvec4 a0 = w_worley3(p / A, seed), b0 = w_worley3((T * p) / B, seed + 1013u);
    const WF a1 = a0.x * A, a2 = a0.y * A, b1 = b0.x * B, b2 = b0.y * B;
    if (a1 <= b1) return wvec4(a1, std::fmin(a2, b1), a0.z, a2 <= b1 ? a0.w : b0.z);
    return wvec4(b1, std::fmin(b2, a1), b0.z, b2 <= a1 ? b0.w : a0.z);
}
