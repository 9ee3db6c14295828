// Device prelude: vector math, the PCG4D generator, the SDF primitives and CSG joins the emitted scene
// code calls, and the closest-surface fold. Compiled by hipRTC together with the emitted scene and the
// kernels (pt_kernels.h). FP32 throughout, as the reference's WGSL.
#pragma once

// The scene's parameter block is read through the constant address space: wave-uniform indices then become scalar
// loads (s_load) into SGPRs instead of per-lane vector loads.
#define PT_CONST __attribute__((address_space(4)))
typedef const PT_CONST float* __restrict__ ParamPtr;
// The scene's functions read their parameters through U, the block pt_set_params updates. A compile with the
// parameters folded in (pt_fold_params) redefines this to point U at a constant copy, so each read is an immediate.
#ifndef PT_SCENE_PARAMS
#define PT_SCENE_PARAMS(U) (void)0
#endif

#define PT_PI 3.14159265358979f
#define PT_INV_PI 0.318309886183791f
#define PT_F32_STEP 1.1920928955078125e-7f

struct V3 { float x, y, z; };
struct V4 { float x, y, z, w; };

__device__ __forceinline__ V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ V3 v3s(float s) { return v3(s, s, s); }
__device__ __forceinline__ V4 v4(V3 a, float w) { V4 r; r.x = a.x; r.y = a.y; r.z = a.z; r.w = w; return r; }
__device__ __forceinline__ V3 xyz(V4 a) { return v3(a.x, a.y, a.z); }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 operator/(V3 a, V3 b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ __forceinline__ V3& operator+=(V3& a, V3 b) { a = a + b; return a; }
__device__ __forceinline__ V3& operator*=(V3& a, V3 b) { a = a * b; return a; }
__device__ __forceinline__ V3& operator*=(V3& a, float s) { a = a * s; return a; }
__device__ __forceinline__ V4 operator-(V4 a) { V4 r; r.x = -a.x; r.y = -a.y; r.z = -a.z; r.w = -a.w; return r; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
// The hardware's square root, reciprocal and reciprocal square root (v_sqrt/v_rcp/v_rsq_f32, 1 ulp): the IEEE-exact
// library forms add a denormal rescale and a Newton step per call, which the distance fields do not need.
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float length(V3 a) { return fsqrt(dot(a, a)); }
__device__ __forceinline__ V3 normalize(V3 a) { return a * frsq(dot(a, a)); }
__device__ __forceinline__ V3 vabs(V3 a) { return v3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
__device__ __forceinline__ V3 vmax(V3 a, V3 b) { return v3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
__device__ __forceinline__ V3 vmin(V3 a, V3 b) { return v3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
__device__ __forceinline__ float max3(V3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
__device__ __forceinline__ float luma(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
__device__ __forceinline__ float saturate(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
__device__ __forceinline__ float signf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
__device__ __forceinline__ float mixf(float a, float b, float t) { return a + (b - a) * t; }
__device__ __forceinline__ V3 vmix(V3 a, V3 b, float t) { return a + (b - a) * t; }
__device__ __forceinline__ V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }

// ---------------------------------------------------------------------------------------------------
// PCG4D (Jarzynski & Olano, "Hash Functions for GPU Rendering", JCGT 9(3) 2020): the reference's state
// update (core/tracebasics.js:369-380), and its output, the top 24 bits of each lane over 2^24.
struct Rng { unsigned int a, b, c, d; };

__device__ __forceinline__ void pcg4d(Rng& s) {
    unsigned int x = s.a * 1664525u + 1013904223u, y = s.b * 1664525u + 1013904223u;
    unsigned int z = s.c * 1664525u + 1013904223u, w = s.d * 1664525u + 1013904223u;
    x += y * w; y += z * x; z += x * y; w += y * z;
    x ^= x >> 16; y ^= y >> 16; z ^= z >> 16; w ^= w >> 16;
    x += y * w; y += z * x; z += x * y; w += y * z;
    s.a = x; s.b = y; s.c = z; s.d = w;
}
__device__ __forceinline__ V4 rngNext(Rng& s) {
    pcg4d(s);
    const float k = 1.0f / 16777216.0f;
    V4 r; r.x = (float)(s.a >> 8) * k; r.y = (float)(s.b >> 8) * k; r.z = (float)(s.c >> 8) * k; r.w = (float)(s.d >> 8) * k;
    return r;
}
// A path's generator at one event: keyed by the sample's seed, the path's slot, its bounce and the kernel,
// never by queue position, so a path draws the same numbers whatever order the queues hold it in.
__device__ __forceinline__ Rng rngAt(const unsigned int* seed, unsigned int slot, unsigned int bounce, unsigned int stream) {
    Rng s; s.a = seed[0] ^ slot; s.b = seed[1] ^ (bounce * 0x9E3779B9u); s.c = seed[2] ^ stream; s.d = seed[3];
    pcg4d(s);
    return s;
}

// ---------------------------------------------------------------------------------------------------
// THE HIT TOLERANCE (core/tolerances.js:10-34): the setting, never under 1e-9 m nor under 128 f32 spacings of
// the largest coordinate at p.
__device__ __forceinline__ float marchTol(float setting, V3 p) {
    return fmaxf(fmaxf(setting, 1e-9f), 128.0f * PT_F32_STEP * fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))));
}

// ---------------------------------------------------------------------------------------------------
// SDF primitives, (outward normal, signed distance) in the primitive's frame; the reference's N* family
// (core/sdf.js:352-628), restated.
__device__ __forceinline__ V4 Nsphere(V3 p, float r) {
    float l = length(p);
    V3 n = l > 1e-30f ? p * frcp(l) : v3(0.0f, 0.0f, 1.0f);
    return v4(n, l - r);
}
__device__ __forceinline__ float Dsphere(V3 p, float r) { return length(p) - r; }

__device__ __forceinline__ V4 Nroundedbox(V3 p, V3 b, float r) {
    V3 q = vabs(p) - b;
    V3 g = vmax(q, v3s(0.0f));
    float d = length(g) + fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.0f) - r;
    V3 n;
    if (g.x > 0.0f || g.y > 0.0f || g.z > 0.0f) {
        n = normalize(g) * v3(signf(p.x), signf(p.y), signf(p.z));
    } else if (q.x > q.y && q.x > q.z) {
        n = v3(signf(p.x), 0.0f, 0.0f);
    } else if (q.y > q.z) {
        n = v3(0.0f, signf(p.y), 0.0f);
    } else {
        n = v3(0.0f, 0.0f, signf(p.z));
    }
    return v4(n, d);
}
__device__ __forceinline__ float Droundedbox(V3 p, V3 b, float r) {
    V3 q = vabs(p) - b;
    return length(vmax(q, v3s(0.0f))) + fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.0f) - r;
}
// A box of half size b whose edges are rounded by r, the rounding inside the box (sdf.js NboxRounded).
__device__ __forceinline__ V4 NboxRounded(V3 p, V3 b, float r) {
    float rr = fminf(fmaxf(r, 0.0f), fminf(b.x, fminf(b.y, b.z)));
    return Nroundedbox(p, b - v3s(rr), rr);
}
__device__ __forceinline__ float DboxRounded(V3 p, V3 b, float r) {
    float rr = fminf(fmaxf(r, 0.0f), fminf(b.x, fminf(b.y, b.z)));
    return Droundedbox(p, b - v3s(rr), rr);
}

// Cylinder along z: half height h, radius r (sdf.js Ncylinder).
__device__ __forceinline__ V4 Ncylinder(V3 p, float h, float r) {
    float qx = fsqrt(p.x * p.x + p.y * p.y), qy = fabsf(p.z);
    float dx = qx - r, dy = qy - h;
    float ox = fmaxf(dx, 0.0f), oy = fmaxf(dy, 0.0f);
    float ol = fsqrt(ox * ox + oy * oy);
    float dist = fminf(fmaxf(dx, dy), 0.0f) + ol;
    float ilxy = frcp(fmaxf(qx, 1e-8f));
    V3 nside = v3(p.x * ilxy, p.y * ilxy, 0.0f);
    V3 ncap = v3(0.0f, 0.0f, signf(p.z));
    V3 n;
    if (dx > 0.0f && dy > 0.0f) {
        float il = frcp(fmaxf(ol, 1e-8f));
        n = normalize(v3(nside.x * ox * il, nside.y * ox * il, 0.0f) + ncap * (oy * il));
    } else if (dx > dy) {
        n = nside;
    } else {
        n = ncap;
    }
    return v4(n, dist);
}
__device__ __forceinline__ float Dcylinder(V3 p, float h, float r) {
    float qx = fsqrt(p.x * p.x + p.y * p.y), qy = fabsf(p.z);
    float dx = qx - r, dy = qy - h;
    float ox = fmaxf(dx, 0.0f), oy = fmaxf(dy, 0.0f);
    return fminf(fmaxf(dx, dy), 0.0f) + fsqrt(ox * ox + oy * oy);
}
// A cylinder whose rim is rounded by rr (the rounded cylinder of the configs: the cylinder shrunk by rr, then
// inflated by it).
__device__ __forceinline__ V4 NroundCylinder(V3 p, float h, float r, float rr) {
    float k = fminf(fmaxf(rr, 0.0f), fminf(h, r));
    V4 c = Ncylinder(p, h - k, r - k);
    c.w -= k;
    return c;
}
__device__ __forceinline__ float DroundCylinder(V3 p, float h, float r, float rr) {
    float k = fminf(fmaxf(rr, 0.0f), fminf(h, r));
    return Dcylinder(p, h - k, r - k) - k;
}

// Torus about z: ring radius R, tube radius r (sdf.js Ntorus).
__device__ __forceinline__ V4 Ntorus(V3 p, float R, float r) {
    float lxy = fsqrt(p.x * p.x + p.y * p.y);
    float il = frcp(lxy);
    float rx = lxy > 1e-12f ? p.x * il : 1.0f, ry = lxy > 1e-12f ? p.y * il : 0.0f;
    float qx = lxy - R, qy = p.z;
    float ql = fsqrt(qx * qx + qy * qy);
    float iq = frcp(ql);
    V3 n = ql > 1e-12f ? v3(rx * qx * iq, ry * qx * iq, qy * iq) : v3(0.0f, 0.0f, 1.0f);
    return v4(n, ql - r);
}
__device__ __forceinline__ float Dtorus(V3 p, float R, float r) {
    float qx = fsqrt(p.x * p.x + p.y * p.y) - R;
    return fsqrt(qx * qx + p.z * p.z) - r;
}

// Capsule along z: half length h of the segment, radius r (sdf.js Ncapsule).
__device__ __forceinline__ V4 Ncapsule(V3 p, float h, float r) {
    V3 v = p - v3(0.0f, 0.0f, fminf(fmaxf(p.z, -h), h));
    float l = length(v);
    V3 n = l > 1e-12f ? v * frcp(l) : v3(1.0f, 0.0f, 0.0f);
    return v4(n, l - r);
}
__device__ __forceinline__ float Dcapsule(V3 p, float h, float r) {
    return length(p - v3(0.0f, 0.0f, fminf(fmaxf(p.z, -h), h))) - r;
}

// The plane z = 0, solid below (sdf.js Nplane).
__device__ __forceinline__ V4 Nplane(V3 p) { return v4(v3(0.0f, 0.0f, 1.0f), p.z); }
__device__ __forceinline__ float Dplane(V3 p) { return p.z; }

// A rectangle sheet in z = 0 of half sizes hx, hy: unsigned distance, normal away from the sheet
// (sdf.js Nrectangle / sheetNormal).
__device__ __forceinline__ V4 Nrectangle(V3 p, float hx, float hy) {
    float ex = fmaxf(fabsf(p.x) - hx, 0.0f), ey = fmaxf(fabsf(p.y) - hy, 0.0f);
    float s = p.z < 0.0f ? -1.0f : 1.0f;
    float d = fsqrt(ex * ex + ey * ey + p.z * p.z);
    if (fmaxf(ex, ey) <= 0.0f || d < 1e-12f) return v4(v3(0.0f, 0.0f, s), fabsf(p.z));
    float id = frcp(d);
    return v4(v3(ex * id * signf(p.x), ey * id * signf(p.y), p.z * id), d);
}
__device__ __forceinline__ float Drectangle(V3 p, float hx, float hy) {
    float ex = fmaxf(fabsf(p.x) - hx, 0.0f), ey = fmaxf(fabsf(p.y) - hy, 0.0f);
    return fsqrt(ex * ex + ey * ey + p.z * p.z);
}
// A disc sheet of radius r in z = 0.
__device__ __forceinline__ V4 Ndisc(V3 p, float r) {
    float l = fsqrt(p.x * p.x + p.y * p.y);
    float e = fmaxf(l - r, 0.0f);
    float s = p.z < 0.0f ? -1.0f : 1.0f;
    float d = fsqrt(e * e + p.z * p.z);
    if (e <= 0.0f || d < 1e-12f) return v4(v3(0.0f, 0.0f, s), fabsf(p.z));
    float il = frcp(fmaxf(l, 1e-12f)), id = frcp(d);
    return v4(v3(e * p.x * il * id, e * p.y * il * id, p.z * id), d);
}
__device__ __forceinline__ float Ddisc(V3 p, float r) {
    float e = fmaxf(fsqrt(p.x * p.x + p.y * p.y) - r, 0.0f);
    return fsqrt(e * e + p.z * p.z);
}

// A spherical lens along z (sdf.js Nlens:463, lensFace:450): rim radius rim, the faces' vertices at z = +-vz, r1 the
// +z face's signed radius and r2 the -z face's (positive bulges out, negative is hollowed in, 0 is flat). The solid is
// the rim's cylinder cut by each face's solid -- a bulging face's ball; a hollow face's ball taken away, closed by the
// plane through the face's rim at z = +-cap -- and its value the largest of the pieces' (exact off the edges, a bound
// near them). s is the face's side (+1, -1); cap matters only for a hollow face.
__device__ __forceinline__ V4 NlensFace(V3 p, float s, float vz, float R, float cap) {
    if (R == 0.0f) return v4(v3(0.0f, 0.0f, s), s * p.z - vz);
    V3 q = v3(p.x, p.y, p.z - s * (vz - R));        // from the face's centre of curvature
    float l = fmaxf(length(q), 1e-12f);
    float sg = R > 0.0f ? 1.0f : -1.0f;
    V4 f = v4(q * (sg * frcp(l)), sg * (l - fabsf(R)));
    float w = s * p.z - cap;
    if (R < 0.0f && w > f.w) f = v4(v3(0.0f, 0.0f, s), w);
    return f;
}
__device__ __forceinline__ V4 Nlens(V3 p, float rim, float vz, float r1, float r2, float cap1, float cap2) {
    float lxy = fmaxf(fsqrt(p.x * p.x + p.y * p.y), 1e-12f);
    float il = frcp(lxy);
    V4 n = v4(v3(p.x * il, p.y * il, 0.0f), lxy - rim);
    V4 a = NlensFace(p, 1.0f, vz, r1, cap1);
    if (a.w > n.w) n = a;
    V4 b = NlensFace(p, -1.0f, vz, r2, cap2);
    if (b.w > n.w) n = b;
    return n;
}
// one face's value, z already times the face's side, l2 = x^2 + y^2
__device__ __forceinline__ float DlensFace(float l2, float z, float vz, float R, float cap) {
    if (R == 0.0f) return z - vz;
    float dz = z - (vz - R);
    float f = fmaxf(fsqrt(l2 + dz * dz), 1e-12f) - fabsf(R);
    return R > 0.0f ? f : fmaxf(-f, z - cap);
}
__device__ __forceinline__ float Dlens(V3 p, float rim, float vz, float r1, float r2, float cap1, float cap2) {
    float l2 = p.x * p.x + p.y * p.y;
    float d = fmaxf(fsqrt(l2), 1e-12f) - rim;
    d = fmaxf(d, DlensFace(l2, p.z, vz, r1, cap1));
    return fmaxf(d, DlensFace(l2, -p.z, vz, r2, cap2));
}

// A prism along z (shapes.js Nprism:285 over g2RegularPolygon and g2Extrude): a regular polygon of n sides, its
// corners at radius R and one flat side facing -y, extruded to half height h. in = 1/n, a = R cos(pi/n) the apothem,
// e = R sin(pi/n) half a side. The nearest side is the one whose outward normal is nearest the point's direction:
// inside, that side has the largest projection; outside, a point in its wedge is nearest that side or one of its ends.
// Side k's normal points at -pi/2 + 2 pi k / n; (u, v) is the point along that normal and along its tangent.
__device__ __forceinline__ void prismSide(float x, float y, float n, float in, float& nx, float& ny, float& u, float& v) {
    float k = rintf((atan2f(y, x) * (0.5f * PT_INV_PI) + 0.25f) * n);
    float s, c;
    sincospif(2.0f * k * in, &s, &c);
    nx = s;
    ny = -c;
    u = x * nx + y * ny;
    v = x * c + y * s;
}
__device__ __forceinline__ V4 Nprism(V3 p, float n, float in, float a, float e, float h) {
    float nx, ny, u, v;
    prismSide(p.x, p.y, n, in, nx, ny, u, v);
    float gx = nx, gy = ny, g = u - a;
    if (g > 0.0f) {
        float ey = v - fminf(fmaxf(v, -e), e);
        float d = fsqrt(g * g + ey * ey);
        if (d > 1e-12f) {
            float id = frcp(d);
            gx = (nx * g - ny * ey) * id;
            gy = (ny * g + nx * ey) * id;
        }
        g = d;
    }
    float wz = fabsf(p.z) - h, sz = p.z >= 0.0f ? 1.0f : -1.0f;
    if (g > 0.0f && wz > 0.0f) {
        float l = fsqrt(g * g + wz * wz), il = frcp(l);
        return v4(v3(gx * g * il, gy * g * il, sz * wz * il), l);
    }
    if (g > wz) return v4(v3(gx, gy, 0.0f), g);
    return v4(v3(0.0f, 0.0f, sz), wz);
}
// A meridian-plane result (gx, gy: the outward normal in (r, z); d the signed distance), as the revolved shapes
// build theirs; ties go to the first (shapes.js g2Max:102).
struct G2 { float x, y, d; };
__device__ __forceinline__ G2 g2(float x, float y, float d) { G2 r; r.x = x; r.y = y; r.d = d; return r; }
__device__ __forceinline__ G2 g2max(G2 a, G2 b) { return a.d >= b.d ? a : b; }
// Lifted about the z axis: the normal's radial part along p.xy's direction (+x on the axis) (shapes.js g2Revolve).
__device__ __forceinline__ V4 g2revolve(V3 p, G2 g) {
    float l = fsqrt(p.x * p.x + p.y * p.y);
    float rx = 1.0f, ry = 0.0f;
    if (l > 1e-12f) { float il = frcp(l); rx = p.x * il; ry = p.y * il; }
    return v4(v3(rx * g.x, ry * g.x, g.y), g.d);
}
// ---- the builder's other primitives (builder/primitives.js:139-236; shapes.js, sdf.js sdCappedCone:633)
__device__ __forceinline__ G2 g2circle(float qx, float qy, float cx, float cy, float r) {
    const float ex = qx - cx, ey = qy - cy, l = fsqrt(ex * ex + ey * ey);
    if (l > 1e-12f) { const float il = frcp(l); return g2(ex * il, ey * il, l - r); }
    return g2(1.0f, 0.0f, l - r);
}
// the 2D box of half sizes (bx, by) about (cx, cy): exact both sides (shapes.js g2Box:46)
__device__ __forceinline__ G2 g2box(float qx, float qy, float cx, float cy, float bx, float by) {
    const float sx = qx >= cx ? 1.0f : -1.0f, sy = qy >= cy ? 1.0f : -1.0f;
    const float ex = fabsf(qx - cx) - bx, ey = fabsf(qy - cy) - by;
    if (ex > 0.0f || ey > 0.0f) {
        const float ox = fmaxf(ex, 0.0f), oy = fmaxf(ey, 0.0f), l = fsqrt(ox * ox + oy * oy), il = frcp(l);
        return g2(ox * il * sx, oy * il * sy, l);
    }
    if (ex > ey) return g2(sx, 0.0f, ex);
    return g2(0.0f, sy, ey);
}
// A capped cone along z, base radius at z = -h, top radius at +h (Quilez's capped cone, sdf.js sdCappedCone:633):
// distance only; the surface kernel takes its normal by differences.
__device__ __forceinline__ float Dcone(V3 p, float h, float r1, float r2) {
    const float qx = fsqrt(p.x * p.x + p.y * p.y), qy = p.z;
    const float k2x = r2 - r1, k2y = 2.0f * h;
    const float cax = qx - fminf(qx, qy < 0.0f ? r1 : r2), cay = fabsf(qy) - h;
    const float t = fminf(fmaxf(((r2 - qx) * k2x + (h - qy) * k2y) / (k2x * k2x + k2y * k2y), 0.0f), 1.0f);
    const float cbx = qx - r2 + k2x * t, cby = qy - h + k2y * t;
    const float s = cbx < 0.0f && cay < 0.0f ? -1.0f : 1.0f;
    return s * fsqrt(fminf(cax * cax + cay * cay, cbx * cbx + cby * cby));
}
__device__ __forceinline__ V4 Ncone(V3 p, float h, float r1, float r2) { return v4(v3s(0.0f), Dcone(p, h, r1, r2)); }
// Two spheres along z, r1 at z = -h/2 and r2 at +h/2, and the cone touching both (shapes.js NroundCone:119); a height
// under the radii's difference leaves the bigger sphere.
__device__ __forceinline__ V4 NroundCone(V3 p, float r1, float r2, float h) {
    const float qx = fsqrt(p.x * p.x + p.y * p.y), qy = p.z + 0.5f * h;
    if (h <= fabsf(r1 - r2)) return g2revolve(p, r1 >= r2 ? g2circle(qx, qy, 0.0f, 0.0f, r1) : g2circle(qx, qy, 0.0f, h, r2));
    const float b = (r1 - r2) / h, a = fsqrt(1.0f - b * b);
    const float k = -b * qx + a * qy;
    if (k < 0.0f) return g2revolve(p, g2circle(qx, qy, 0.0f, 0.0f, r1));
    if (k > a * h) return g2revolve(p, g2circle(qx, qy, 0.0f, h, r2));
    return g2revolve(p, g2(a, b, a * qx + b * qy - r1));
}
// A regular octahedron, size from its centre to each vertex (shapes.js Noctahedron:135): exact both sides -- inside
// the nearest face's plane, outside the nearest point of the face, edge or vertex in the point's octant.
__device__ __forceinline__ V3 safeDir3(V3 v, V3 fb) { const float l = length(v); return l > 1e-12f ? v * frcp(l) : fb; }
__device__ __forceinline__ V4 Noctahedron(V3 p, float s0) {
    const float s = fmaxf(s0, 0.0f);
    const V3 sg = v3(p.x >= 0.0f ? 1.0f : -1.0f, p.y >= 0.0f ? 1.0f : -1.0f, p.z >= 0.0f ? 1.0f : -1.0f);
    const V3 a = vabs(p);
    const float m = a.x + a.y + a.z - s;
    const float c = 0.5773502691896258f;
    const V3 face = sg * c, fb = v3s(c);
    if (3.0f * a.x < m) {
        const float k = fminf(fmaxf(0.5f * (a.z - a.y + s), 0.0f), s);
        const V3 v = v3(a.x, a.y - s + k, a.z - k);
        return v4(sg * safeDir3(v, fb), length(v));
    }
    if (3.0f * a.y < m) {
        const float k = fminf(fmaxf(0.5f * (a.x - a.z + s), 0.0f), s);
        const V3 v = v3(a.y, a.z - s + k, a.x - k);
        return v4(sg * safeDir3(v3(v.z, v.x, v.y), fb), length(v));
    }
    if (3.0f * a.z < m) {
        const float k = fminf(fmaxf(0.5f * (a.y - a.x + s), 0.0f), s);
        const V3 v = v3(a.z, a.x - s + k, a.y - k);
        return v4(sg * safeDir3(v3(v.y, v.z, v.x), fb), length(v));
    }
    return v4(face, m * c);
}
// The builder's pyramid, a frustum along z (shapes.js Nfrustum:204): a base.x by base.y rectangle at z = -h/2 and a
// top.x by top.y one at +h/2 (0 by 0: an apex). Convex and mirror-symmetric in x and y, so after folding to the
// positive quadrant the nearest surface point is on the +x side, the +y side, the top or the base. Inside, the
// nearest of those planes; outside, the nearest of the four faces' nearest points. Exact both sides.
__device__ __forceinline__ V3 segNear(V3 P, V3 a, V3 b) {
    const V3 ab = b - a;
    return a + ab * fminf(fmaxf(dot(P - a, ab) / fmaxf(dot(ab, ab), 1e-30f), 0.0f), 1.0f);
}
// the nearest point of the planar convex quad a b c d (counter-clockwise about its unit normal n; two corners may
// coincide): the point's projection when it lies within all four edges, else the nearest point of the nearest edge
__device__ __forceinline__ V3 quadNear(V3 P, V3 a, V3 b, V3 c, V3 d, V3 n) {
    const V3 q = P - n * dot(P - a, n);
    if (dot(cross(b - a, q - a), n) >= 0.0f && dot(cross(c - b, q - b), n) >= 0.0f &&
        dot(cross(d - c, q - c), n) >= 0.0f && dot(cross(a - d, q - d), n) >= 0.0f) return q;
    V3 m = segNear(P, a, b);
    float dm = dot(P - m, P - m);
    const V3 e[3] = { segNear(P, b, c), segNear(P, c, d), segNear(P, d, a) };
#pragma unroll
    for (int i = 0; i < 3; ++i) { const float de = dot(P - e[i], P - e[i]); if (de < dm) { dm = de; m = e[i]; } }
    return m;
}
__device__ __forceinline__ V4 Nfrustum(V3 p, float bx0, float by0, float tx0, float ty0, float height) {
    const float bx = 0.5f * fabsf(bx0), by = 0.5f * fabsf(by0), tx = 0.5f * fabsf(tx0), ty = 0.5f * fabsf(ty0);
    const float h = fmaxf(height, 1e-6f);
    const float sx = p.x >= 0.0f ? 1.0f : -1.0f, sy = p.y >= 0.0f ? 1.0f : -1.0f;
    const V3 P = v3(fabsf(p.x), fabsf(p.y), p.z + 0.5f * h);
    const V3 nx = safeDir3(v3(h, 0.0f, bx - tx), v3(1.0f, 0.0f, 0.0f));
    const V3 ny = safeDir3(v3(0.0f, h, by - ty), v3(0.0f, 1.0f, 0.0f));
    const float dx = dot(P - v3(bx, 0.0f, 0.0f), nx), dy = dot(P - v3(0.0f, by, 0.0f), ny);
    V3 g;
    float d;
    if (dx <= 0.0f && dy <= 0.0f && P.z >= 0.0f && P.z <= h) {
        d = dx; g = nx;
        if (dy > d) { d = dy; g = ny; }
        if (P.z - h > d) { d = P.z - h; g = v3(0.0f, 0.0f, 1.0f); }
        if (-P.z > d) { d = -P.z; g = v3(0.0f, 0.0f, -1.0f); }
    } else {
        const V3 cx = quadNear(P, v3(bx, -by, 0.0f), v3(bx, by, 0.0f), v3(tx, ty, h), v3(tx, -ty, h), nx);
        const V3 cy = quadNear(P, v3(bx, by, 0.0f), v3(-bx, by, 0.0f), v3(-tx, ty, h), v3(tx, ty, h), ny);
        const V3 ct = v3(fminf(P.x, tx), fminf(P.y, ty), h), cb = v3(fminf(P.x, bx), fminf(P.y, by), 0.0f);
        V3 c = cx, fb = nx;
        float dc = dot(P - cx, P - cx);
        const float d1 = dot(P - cy, P - cy), d2 = dot(P - ct, P - ct), d3 = dot(P - cb, P - cb);
        if (d1 < dc) { dc = d1; c = cy; fb = ny; }
        if (d2 < dc) { dc = d2; c = ct; fb = v3(0.0f, 0.0f, 1.0f); }
        if (d3 < dc) { dc = d3; c = cb; fb = v3(0.0f, 0.0f, -1.0f); }
        g = safeDir3(P - c, fb);
        d = fsqrt(dc);
    }
    return v4(v3(g.x * sx, g.y * sy, g.z), d);
}
// Part of a torus in the xy plane: an arc of radius R, angle degrees of it symmetric about +y, swept by a tube of
// radius r (shapes.js NcappedTorus:335): the nearest point of the arc -- its end past the wedge -- then the tube.
// sc = (sin, cos) of half the angle, from the emitter.
__device__ __forceinline__ V4 NcappedTorus(V3 p, float R, float r, float sn, float cs) {
    const float sx = p.x >= 0.0f ? 1.0f : -1.0f;
    const V3 q = v3(fabsf(p.x), p.y, p.z);
    float cx, cy;
    if (cs * q.x > sn * q.y) { cx = R * sn; cy = R * cs; }
    else {
        const float l = fsqrt(q.x * q.x + q.y * q.y);
        if (l > 1e-12f) { const float il = frcp(l); cx = R * q.x * il; cy = R * q.y * il; } else { cx = 0.0f; cy = R; }
    }
    const V3 v = v3(q.x - cx, q.y - cy, q.z);
    const V3 n = safeDir3(v, v3(0.0f, 0.0f, 1.0f));
    return v4(v3(n.x * sx, n.y, n.z), length(v) - r);
}
// A chain link in the xy plane: a ring of radius R pulled apart along y by span, swept by a tube of radius r (shapes.js
// Nlink:349).
__device__ __forceinline__ V4 Nlink(V3 p, float span, float R, float r) {
    const float le = fmaxf(0.5f * span, 0.0f);
    const float qy = (p.y >= 0.0f ? 1.0f : -1.0f) * fmaxf(fabsf(p.y) - le, 0.0f);
    return Ntorus(v3(p.x, qy, p.z), R, r);
}
// A hollow cylinder along z, centred: the 2D box [inner, outer] x [-h/2, h/2] revolved (shapes.js Ntube:373); an inner
// radius of 0 or less is the solid cylinder.
__device__ __forceinline__ V4 Ntube(V3 p, float outer, float inner, float height) {
    const float qx = fsqrt(p.x * p.x + p.y * p.y), hh = fmaxf(0.5f * height, 0.0f);
    if (inner <= 0.0f) return g2revolve(p, g2box(qx, p.z, 0.0f, 0.0f, fmaxf(outer, 0.0f), hh));
    const float ro = fmaxf(outer, inner + 1e-6f);
    return g2revolve(p, g2box(qx, p.z, 0.5f * (inner + ro), 0.0f, 0.5f * (ro - inner), hh));
}
// The part of a sphere of radius R above z = 0, height tall (shapes.js Ndome:358, Quilez's cut sphere): its centre at
// c = height - R (held below R), exact both sides -- the sphere, the flat base, or the rim's circle.
__device__ __forceinline__ V4 Ndome(V3 p, float R0, float height) {
    const float R = fmaxf(R0, 1e-6f);
    const float c = fminf(fmaxf(height - R, -R), R * (1.0f - 1e-4f));
    const float hc = -c, w = fsqrt(fmaxf(R * R - hc * hc, 0.0f));
    const float qx = fsqrt(p.x * p.x + p.y * p.y), qy = p.z - c;
    const float s = fmaxf((hc - R) * qx * qx + w * w * (hc + R - 2.0f * qy), hc * qx - w * qy);
    if (s < 0.0f) return g2revolve(p, g2circle(qx, qy, 0.0f, 0.0f, R));
    if (qx < w) return g2revolve(p, g2(0.0f, -1.0f, hc - qy));
    return g2revolve(p, g2circle(qx, qy, w, hc, 0.0f));
}

// Extruded along z to half height h, exactly (shapes.js g2Extrude:717): past both the outline and the caps, the
// corner's distance; else the larger of the two.
__device__ __forceinline__ V4 g2extrude(V3 p, G2 g, float h) {
    const float wz = fabsf(p.z) - h, sz = p.z >= 0.0f ? 1.0f : -1.0f;
    if (g.d > 0.0f && wz > 0.0f) {
        const float l = fsqrt(g.d * g.d + wz * wz), il = frcp(l);
        return v4(v3(g.x * g.d * il, g.y * g.d * il, sz * wz * il), l);
    }
    if (g.d > wz) return v4(v3(g.x, g.y, 0.0f), g.d);
    return v4(v3(0.0f, 0.0f, sz), wz);
}
// An outline of segments and fillet arcs (counter-clockwise), walked piece by piece (shapes.js pfStart..pfEnd:684-715):
// the nearest point so far (v = q minus it, d2 its squared distance, n the piece's outward normal for q on it), and
// inside by the even-odd rule over the edges and each arc's chord, flipped within the arc's circular segment. A
// segment that is no surface (a revolved outline's axis edge) still counts for inside.
struct Prof { float vx, vy, d2, nx, ny; bool inside; };
__device__ __forceinline__ Prof pfStart() { Prof A; A.vx = A.vy = 0.0f; A.d2 = 3.0e38f; A.nx = 1.0f; A.ny = 0.0f; A.inside = false; return A; }
__device__ __forceinline__ bool pfCross(float qx, float qy, float ax, float ay, float bx, float by) {
    return (ay > qy) != (by > qy) && qx < ax + (qy - ay) * (bx - ax) / (by - ay);
}
__device__ __forceinline__ void pfDir(float x, float y, float fx, float fy, float& ox, float& oy) {
    const float l = fsqrt(x * x + y * y);
    if (l > 1e-12f) { const float il = frcp(l); ox = x * il; oy = y * il; } else { ox = fx; oy = fy; }
}
__device__ __forceinline__ void pfSegment(Prof& A, float qx, float qy, float ax, float ay, float bx, float by, bool surface) {
    if (pfCross(qx, qy, ax, ay, bx, by)) A.inside = !A.inside;
    if (!surface) return;
    const float ex = bx - ax, ey = by - ay;
    const float t = fminf(fmaxf(((qx - ax) * ex + (qy - ay) * ey) / fmaxf(ex * ex + ey * ey, 1e-30f), 0.0f), 1.0f);
    const float vx = qx - (ax + ex * t), vy = qy - (ay + ey * t), d2 = vx * vx + vy * vy;
    if (d2 < A.d2) { A.vx = vx; A.vy = vy; A.d2 = d2; pfDir(ey, -ex, 1.0f, 0.0f, A.nx, A.ny); }
}
// a fillet: the minor arc from a to b about c, radius r; convex: the shape lies on c's side of it
__device__ __forceinline__ void pfArc(Prof& A, float qx, float qy, float cx, float cy, float r, float ax, float ay, float bx, float by, bool convex) {
    const float wx = qx - cx, wy = qy - cy;
    const float mx = 0.5f * (ax + bx) - cx, my = 0.5f * (ay + by) - cy;
    const bool seg = wx * wx + wy * wy < r * r && wx * mx + wy * my > mx * mx + my * my;
    if (pfCross(qx, qy, ax, ay, bx, by) != seg) A.inside = !A.inside;
    const float eax = ax - cx, eay = ay - cy, ebx = bx - cx, eby = by - cy;
    const float s = eax * eby - eay * ebx >= 0.0f ? 1.0f : -1.0f;
    const bool wedge = s * (eax * wy - eay * wx) >= 0.0f && s * (wx * eby - wy * ebx) >= 0.0f;
    float px, py;
    if (wedge) {
        float mdx, mdy, dx, dy;
        pfDir(mx, my, 1.0f, 0.0f, mdx, mdy);
        pfDir(wx, wy, mdx, mdy, dx, dy);
        px = cx + dx * r; py = cy + dy * r;
    } else {
        const bool toA = (qx - ax) * (qx - ax) + (qy - ay) * (qy - ay) <= (qx - bx) * (qx - bx) + (qy - by) * (qy - by);
        px = toA ? ax : bx; py = toA ? ay : by;
    }
    const float vx = qx - px, vy = qy - py, d2 = vx * vx + vy * vy;
    if (d2 < A.d2) {
        A.vx = vx; A.vy = vy; A.d2 = d2;
        float nx, ny;
        pfDir(px - cx, py - cy, 1.0f, 0.0f, nx, ny);
        A.nx = convex ? nx : -nx; A.ny = convex ? ny : -ny;
    }
}
__device__ __forceinline__ G2 pfEnd(const Prof& A) {
    const float d = fsqrt(A.d2), s = A.inside ? -1.0f : 1.0f;
    if (d > 1e-12f) { const float k = s / fmaxf(d, 1e-30f); return g2(A.vx * k, A.vy * k, s * d); }
    return g2(A.nx, A.ny, s * d);
}
// The signed distance in a meridian plane from (x, y) to the parabola y = k x^2 (k > 0), positive below it, with its
// outward normal (shapes.js g2Parabola:61). The nearest point's abscissa t solves the depressed cubic
// 2k^2 t^3 + (1 - 2ky) t - x = 0: with a = (2ky - 1) / 6k^2 and b = x / 4k^2 its discriminant is b^2 - a^3; above
// zero one real root (Cardano, as 2b / (u^2 - a + w^2) with u = cbrt(b + sqrt(h)), w = a / u, which keeps its
// digits), else the largest of three, 2 sqrt(a) cos(atan2(sqrt(-h), b) / 3).
__device__ __forceinline__ G2 g2parabola(float x, float y, float k) {
    const float sx = x >= 0.0f ? 1.0f : -1.0f, xa = fabsf(x);
    const float ik = 1.0f / k;
    const float a = ik * (y - 0.5f * ik) * (1.0f / 3.0f);
    const float b = 0.25f * ik * ik * xa;
    const float h = b * b - a * a * a;
    float t;
    if (h > 0.0f) {
        const float u = cbrtf(b + fsqrt(h));
        const float w = a / u;
        t = 2.0f * b / (u * u - a + w * w);
    } else {
        t = 2.0f * fsqrt(fmaxf(a, 0.0f)) * cosf(atan2f(fsqrt(-h), b) * (1.0f / 3.0f));
    }
    const float ex = xa - t, ey = y - k * t * t;
    const float side = k * xa * xa - y >= 0.0f ? 1.0f : -1.0f;
    const float in = frsq(1.0f + 4.0f * k * k * t * t);
    return g2(2.0f * k * t * in * sx, -in, side * fsqrt(ex * ex + ey * ey));
}
// ...with its focus moved out to a ring of radius c in the plane of the foci (shapes.js g2ParabolaOffset:93): the two
// half-planes' parabolas shifted by +-c, the larger of the two values for c < 0, the smaller for c > 0.
__device__ __forceinline__ G2 g2parabolaOff(float x, float y, float k, float c) {
    const G2 a = g2parabola(x - c, y, k);
    if (c == 0.0f) return a;
    const G2 b = g2parabola(x + c, y, k);
    if (c < 0.0f) return a.d >= b.d ? a : b;
    return a.d <= b.d ? a : b;
}
// A parabolic reflector (shapes.js NparabolicOffset:418 over NdishOffset:392 and NparaboloidOffset): its focus at the
// origin, opening to +z, the mirror z = (r - c)^2 / 4f - f out to radius A. solid <= 0.5: the mirror with its wall w
// thick behind it (the band 0 <= s <= w of the distance s to the parabola), cut by the cylinder r <= A; else the
// paraboloid's inside, capped at the rim's height. U: f, A, w, solid, c (c held at most A by the emitter).
__device__ __forceinline__ G2 parabolicG2(V3 p, float f, float A, float wall, float solid, float c) {
    const float fo = fmaxf(f, 1e-6f), Aa = fmaxf(A, 0.0f), cc = fminf(c, Aa);
    const float x = fsqrt(p.x * p.x + p.y * p.y), y = p.z + fo;
    const G2 s = g2parabolaOff(x, y, 0.25f / fo, cc);
    if (solid > 0.5f) return g2max(s, g2(0.0f, 1.0f, y - 0.25f * (Aa - cc) * (Aa - cc) / fo));
    const G2 band = g2max(g2(-s.x, -s.y, -s.d), g2(s.x, s.y, s.d - fmaxf(wall, 1e-6f)));
    return g2max(band, g2(1.0f, 0.0f, x - Aa));
}
__device__ __forceinline__ V4 Nparabolic(V3 p, float f, float A, float wall, float solid, float c) {
    return g2revolve(p, parabolicG2(p, f, A, wall, solid, c));
}
__device__ __forceinline__ float Dparabolic(V3 p, float f, float A, float wall, float solid, float c) {
    return parabolicG2(p, f, A, wall, solid, c).d;
}

// An aspheric lens along z (shapes.js Nasphere:658): a round rim of radius rho and two even-asphere faces, their
// vertices at z = +-ht, each face's sag z(r) = c r^2 / (1 + sqrt(1 - (1 + k) c^2 r^2)) + a4 r^4 + a6 r^6 + a8 r^8 +
// a10 r^10 measured into the lens (c = 1 / R; 0 flat). A face's value is its height above the surface,
// s z - ht + sag(r) (r held at rho beyond the rim), over sqrt(1 + m^2), m the steepest the face can be within the
// rim: 1-Lipschitz, so a bound both sides. Its normal is the surface's own. The root is held at 1e-4 where a conic
// too small for the aperture has none. U per face: c, k, a4, a6, a8, a10, 1 / sqrt(1 + m^2).
__device__ __forceinline__ void asphSag(float r, float c, float k, float a4, float a6, float a8, float a10, float& sag, float& slope) {
    const float r2 = r * r;
    const float w = 1.0f - (1.0f + k) * c * c * r2;
    const float u = fsqrt(fmaxf(w, 1e-4f));
    const float poly = r2 * r2 * (a4 + r2 * (a6 + r2 * (a8 + r2 * a10)));
    const float dpoly = r * r2 * (4.0f * a4 + r2 * (6.0f * a6 + r2 * (8.0f * a8 + r2 * 10.0f * a10)));
    sag = c * r2 / (1.0f + u) + poly;
    slope = (w >= 1e-4f ? c * r / u : 2.0f * c * r / (1.0f + u)) + dpoly;
}
__device__ __forceinline__ V4 asphFace(V3 p, float l, float rx, float ry, float rho, float ht, const PT_CONST float* F, float s) {
    float sag, slope;
    asphSag(fminf(l, rho), F[0], F[1], F[2], F[3], F[4], F[5], sag, slope);
    const float g = l < rho ? slope : 0.0f;
    const float in = frsq(g * g + 1.0f);
    return v4(v3(rx * g * in, ry * g * in, s * in), (s * p.z - ht + sag) * F[6]);
}
__device__ __forceinline__ V4 Nasphere(V3 p, const PT_CONST float* P) {
    const float rho = P[0], ht = P[1];
    const float l = fsqrt(p.x * p.x + p.y * p.y);
    float rx = 1.0f, ry = 0.0f;
    if (l > 1e-12f) { const float il = frcp(l); rx = p.x * il; ry = p.y * il; }
    V4 n = v4(v3(rx, ry, 0.0f), l - rho);
    const V4 top = asphFace(p, l, rx, ry, rho, ht, P + 2, 1.0f);
    if (top.w > n.w) n = top;
    const V4 bottom = asphFace(p, l, rx, ry, rho, ht, P + 9, -1.0f);
    if (bottom.w > n.w) n = bottom;
    return n;
}
__device__ __forceinline__ float Dasphere(V3 p, const PT_CONST float* P) {
    const float rho = P[0], ht = P[1];
    const float l = fsqrt(p.x * p.x + p.y * p.y);
    float d = l - rho, sag, slope;
    const float rr = fminf(l, rho);
    asphSag(rr, P[2], P[3], P[4], P[5], P[6], P[7], sag, slope);
    d = fmaxf(d, (p.z - ht + sag) * P[8]);
    asphSag(rr, P[9], P[10], P[11], P[12], P[13], P[14], sag, slope);
    return fmaxf(d, (-p.z - ht + sag) * P[15]);
}

__device__ __forceinline__ float Dprism(V3 p, float n, float in, float a, float e, float h) {
    float nx, ny, u, v;
    prismSide(p.x, p.y, n, in, nx, ny, u, v);
    float g = u - a, ey = v - fminf(fmaxf(v, -e), e);
    g = g > 0.0f ? fsqrt(g * g + ey * ey) : g;
    float wz = fabsf(p.z) - h;
    return g > 0.0f && wz > 0.0f ? fsqrt(g * g + wz * wz) : fmaxf(g, wz);
}

// ---------------------------------------------------------------------------------------------------
// CSG joins (scene_api.js:352-373): a k > 0 blend is the polynomial smooth min/max, the normal the gradient of
// the blend (mix of the two), none where either side has none. ik is 1/k (0 when k is 0), from the parameter block.
__device__ __forceinline__ V3 csgNormal(V3 a, V3 b, float t) {
    if (dot(a, a) < 1e-12f || dot(b, b) < 1e-12f) return v3s(0.0f);
    V3 n = vmix(a, b, t);
    float l = length(n);
    return l > 1e-12f ? n * frcp(l) : v3s(0.0f);
}
__device__ __forceinline__ V4 csgUnion(V4 a, V4 b, float k, float ik) {
    if (k <= 0.0f) return a.w < b.w ? a : b;
    float h = fmaxf(k - fabsf(a.w - b.w), 0.0f);
    float m = a.w < b.w ? 0.0f : 1.0f;
    float t = m + (1.0f - 2.0f * m) * 0.5f * h * ik;
    return v4(csgNormal(xyz(a), xyz(b), t), fminf(a.w, b.w) - h * h * 0.25f * ik);
}
__device__ __forceinline__ V4 csgIntersect(V4 a, V4 b, float k, float ik) {
    if (k <= 0.0f) return a.w > b.w ? a : b;
    float h = fmaxf(k - fabsf(a.w - b.w), 0.0f);
    float m = a.w > b.w ? 0.0f : 1.0f;
    float t = m + (1.0f - 2.0f * m) * 0.5f * h * ik;
    return v4(csgNormal(xyz(a), xyz(b), t), fmaxf(a.w, b.w) + h * h * 0.25f * ik);
}
__device__ __forceinline__ V4 csgSubtract(V4 a, V4 b, float k, float ik) { return csgIntersect(a, -b, k, ik); }
__device__ __forceinline__ float dUnion(float a, float b, float k, float ik) {
    float h = fmaxf(k - fabsf(a - b), 0.0f);
    return fminf(a, b) - h * h * 0.25f * ik;
}
__device__ __forceinline__ float dIntersect(float a, float b, float k, float ik) {
    float h = fmaxf(k - fabsf(a - b), 0.0f);
    return fmaxf(a, b) + h * h * 0.25f * ik;
}
__device__ __forceinline__ float dSubtract(float a, float b, float k, float ik) { return dIntersect(a, -b, k, ik); }

// ---------------------------------------------------------------------------------------------------
// A node's frame: params t[0..2] position, t[3..11] rotation (columns: local axes in world), t[12] scale,
// t[13] its inverse.
// Local point q = R^T (p - pos) / s; a local (n, d) goes back as (R n, d s) (scene_api.js add_new).
__device__ __forceinline__ V3 xfIn(const PT_CONST float* t, V3 p) {
    V3 d = p - v3(t[0], t[1], t[2]);
    float is = t[13];
    return v3(dot(d, v3(t[3], t[4], t[5])) * is, dot(d, v3(t[6], t[7], t[8])) * is, dot(d, v3(t[9], t[10], t[11])) * is);
}
__device__ __forceinline__ V4 xfOut(const PT_CONST float* t, V4 s) {
    V3 n = v3(t[3], t[4], t[5]) * s.x + v3(t[6], t[7], t[8]) * s.y + v3(t[9], t[10], t[11]) * s.z;
    return v4(n, s.w * t[12]);
}

// The closest-surface fold (scene_api.js:39-63): the closest |d| gives the surface and its material, the
// deepest negative d the material the point is inside of.
struct SdfHit {
    V3 n;
    float d;
    unsigned int m;      // material of the closest surface
    unsigned int a;      // material the point is inside of
    float insideW;
    unsigned int prim;   // the closest surface's primitive (its place in the emitter's program)
};
__device__ __forceinline__ SdfHit sdfEmpty() {
    SdfHit h; h.n = v3(0.0f, 0.0f, 1.0f); h.d = 1.0e7f; h.m = 0u; h.a = 0u; h.insideW = -1.0e7f; h.prim = 0xFFFFFFFFu; return h;
}
__device__ __forceinline__ void foldFull(SdfHit& h, V4 s, unsigned int mat, unsigned int prim) {
    if (s.w < 0.0f && s.w > h.insideW) { h.insideW = s.w; h.a = mat; }
    if (fabsf(s.w) < fabsf(h.d)) { h.n = xyz(s); h.d = s.w; h.m = mat; h.prim = prim; }
}
// A node's bounding sphere (centre c, radius r) can beat the closest |d| so far: |p - c| - r < |best|, compared
// squared (|p - c|^2 < (r + |best|)^2; both sides non-negative) so that no square root is taken. The squared test is
// a hair wide (2e-6 relative, above its own rounding), never narrower than the exact inequality: a node it skips is
// one the exact test skips, which can neither be the closest surface nor hold the point.
__device__ __forceinline__ bool boundNear(V3 p, float cx, float cy, float cz, float r, float best) {
    const float dx = p.x - cx, dy = p.y - cy, dz = p.z - cz;
    const float R = r + fabsf(best);
    return dx * dx + dy * dy + dz * dz < R * R * (1.0f + 2e-6f);
}
__device__ __forceinline__ void foldDist(float& best, float d) { if (fabsf(d) < fabsf(best)) best = d; }
