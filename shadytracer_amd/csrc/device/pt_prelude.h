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

// ---- shapes of revolution are drawn in their meridian half-plane: r = |p.xy| >= 0 across, z up. A piece there answers
// a Mer -- the outward unit normal (nr, nz) and the signed distance d -- and merTo3 turns the answer about the z axis.
// An intersection of pieces keeps the piece with the larger d (the first on a tie).
struct Mer { float nr, nz, d; };
__device__ __forceinline__ Mer mer(float nr, float nz, float d) { Mer m; m.nr = nr; m.nz = nz; m.d = d; return m; }
__device__ __forceinline__ Mer merMax(Mer a, Mer b) { return b.d > a.d ? b : a; }
__device__ __forceinline__ Mer merFlip(Mer a) { return mer(-a.nr, -a.nz, -a.d); }
// The normal's r part goes along p.xy's direction; on the axis, where every direction is one, along +x.
__device__ __forceinline__ V4 merTo3(V3 p, Mer m) {
    const float rr = p.x * p.x + p.y * p.y;
    if (!(rr > 1e-24f)) return v4(v3(m.nr, 0.0f, m.nz), m.d);
    const float k = m.nr * frsq(rr);
    return v4(v3(p.x * k, p.y * k, m.nz), m.d);
}
// A disc of radius rad about (cr, cz) (a ball, turned): the distance from its centre less rad, the normal away from
// the centre (+r from the centre itself).
__device__ __forceinline__ Mer merCircle(float r, float z, float cr, float cz, float rad) {
    const float er = r - cr, ez = z - cz, l = fsqrt(er * er + ez * ez);
    if (!(l > 1e-12f)) return mer(1.0f, 0.0f, l - rad);
    const float k = frcp(l);
    return mer(er * k, ez * k, l - rad);
}
// The rectangle of half sizes (hr, hz) about (cr, cz), exact both sides: within it, the nearer pair of sides (the z
// pair on a tie); past it, the distance to its nearest corner or side, measured by how far past each pair it lies.
__device__ __forceinline__ Mer merRect(float r, float z, float cr, float cz, float hr, float hz) {
    const float sr = r < cr ? -1.0f : 1.0f, sz = z < cz ? -1.0f : 1.0f;
    const float pr = fabsf(r - cr) - hr, pz = fabsf(z - cz) - hz;
    if (fmaxf(pr, pz) <= 0.0f) return pr > pz ? mer(sr, 0.0f, pr) : mer(0.0f, sz, pz);
    const float ar = fmaxf(pr, 0.0f), az = fmaxf(pz, 0.0f), l = fsqrt(ar * ar + az * az), k = frcp(l);
    return mer(sr * ar * k, sz * az * k, l);
}

// A spherical lens along z (sdf.js Nlens:463): rim radius rim, the faces' vertices at z = +-vz, r1 the +z face's signed
// radius and r2 the -z face's (positive bulges out, negative is hollowed in, 0 is flat). In the meridian plane the lens
// is the band r <= rim cut by each face's solid. Measured along a face's own outward axis (zs = s z, s = +1 for the +z
// face), the face's sphere has its centre on the axis at vz - R and radius |R|: a bulging face keeps the inside of that
// circle, a hollow one the outside, closed by the line zs = cap through its rim (the rim's height, from the emitter).
// The value is the largest of the pieces' -- exact away from the edges, a bound near them.
__device__ __forceinline__ Mer lensSide(float r, float z, float s, float vz, float R, float cap) {
    const float zs = s * z;
    if (R == 0.0f) return mer(0.0f, s, zs - vz);
    const Mer c = merCircle(r, zs, 0.0f, vz - R, fabsf(R));
    if (R > 0.0f) return mer(c.nr, s * c.nz, c.d);
    const Mer outside = mer(-c.nr, -s * c.nz, -c.d);
    return zs - cap > outside.d ? mer(0.0f, s, zs - cap) : outside;
}
__device__ __forceinline__ V4 Nlens(V3 p, float rim, float vz, float r1, float r2, float cap1, float cap2) {
    const float r = fsqrt(p.x * p.x + p.y * p.y);
    Mer m = mer(1.0f, 0.0f, r - rim);
    m = merMax(m, lensSide(r, p.z, 1.0f, vz, r1, cap1));
    m = merMax(m, lensSide(r, p.z, -1.0f, vz, r2, cap2));
    return merTo3(p, m);
}
// the same value alone, for the march: rr = r^2, zs along the face's axis
__device__ __forceinline__ float lensSideD(float rr, float zs, float vz, float R, float cap) {
    if (R == 0.0f) return zs - vz;
    const float ez = zs - vz + R;
    const float c = fsqrt(rr + ez * ez) - fabsf(R);
    return R > 0.0f ? c : fmaxf(-c, zs - cap);
}
__device__ __forceinline__ float Dlens(V3 p, float rim, float vz, float r1, float r2, float cap1, float cap2) {
    const float rr = p.x * p.x + p.y * p.y;
    const float d = fmaxf(fsqrt(rr) - rim, lensSideD(rr, p.z, vz, r1, cap1));
    return fmaxf(d, lensSideD(rr, -p.z, vz, r2, cap2));
}

// A prism along z (shapes.js Nprism:285): a regular polygon of n sides, its
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
// ---- the builder's other primitives (builder/primitives.js:139-236)
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
// Two balls on the z axis, radius r1 about z = -h/2 and r2 about z = +h/2, and the cone tangent to both: their convex
// hull (builder/primitives.js round cone). In the meridian plane, measured from the lower centre (z' = z + h/2), the
// outer tangent line has unit normal n = (cos a, sin a) with sin a = (r1 - r2) / h, and meets the circles at r1 n and
// (0, h) + r2 n; along the line's direction t = (-sin a, cos a) the first contact is at 0 and the second at h cos a. A
// point whose foot along t comes before the first contact is nearest the lower ball, after the second the upper one,
// else the line, at distance q . n - r1. A height of |r1 - r2| or less leaves the bigger ball holding the other.
__device__ __forceinline__ V4 NroundCone(V3 p, float r1, float r2, float h) {
    const float r = fsqrt(p.x * p.x + p.y * p.y), z = p.z + 0.5f * h;
    Mer m;
    if (h <= fabsf(r1 - r2)) {
        m = r2 > r1 ? merCircle(r, z, 0.0f, h, r2) : merCircle(r, z, 0.0f, 0.0f, r1);
    } else {
        const float sa = (r1 - r2) / h, ca = fsqrt(1.0f - sa * sa);
        const float foot = ca * z - sa * r;
        if (foot < 0.0f) m = merCircle(r, z, 0.0f, 0.0f, r1);
        else if (foot > h * ca) m = merCircle(r, z, 0.0f, h, r2);
        else m = mer(ca, sa, ca * r + sa * z - r1);
    }
    return merTo3(p, m);
}
// A regular octahedron |x| + |y| + |z| <= s (s from the centre to each vertex), exact both sides. Folded into the
// positive octant (a = |p|) its one face there is the triangle (s,0,0), (0,s,0), (0,0,s) on the plane a.1 = s, normal
// (1,1,1)/sqrt 3. Inside the solid, and outside wherever a's foot on that plane (a less e (1,1,1), e its height over
// the plane over sqrt 3) lies in the triangle, the distance is the plane's. Otherwise the nearest point is on one of the
// triangle's edges, each lying in a coordinate plane, and the least of the three edge distances is the distance.
__device__ __forceinline__ V3 safeDir3(V3 v, V3 fb) { const float l = length(v); return l > 1e-12f ? v * frcp(l) : fb; }
__device__ __forceinline__ V4 Noctahedron(V3 p, float s0) {
    const float s = fmaxf(s0, 0.0f), k = 0.5773502691896258f;
    const V3 a = vabs(p), sg = v3(p.x < 0.0f ? -1.0f : 1.0f, p.y < 0.0f ? -1.0f : 1.0f, p.z < 0.0f ? -1.0f : 1.0f);
    const float h = (a.x + a.y + a.z - s) * k;
    if (fminf(a.x, fminf(a.y, a.z)) >= h * k) return v4(sg * k, h);
    // the edges: x = 0 from (0,s,0) to (0,0,s), y = 0 from (0,0,s) to (s,0,0), z = 0 from (s,0,0) to (0,s,0), each
    // point written by its distance u from the edge's first end, clamped to the edge
    const float ux = fminf(fmaxf(0.5f * (a.z - a.y + s), 0.0f), s);
    const float uy = fminf(fmaxf(0.5f * (a.x - a.z + s), 0.0f), s);
    const float uz = fminf(fmaxf(0.5f * (a.y - a.x + s), 0.0f), s);
    const V3 ex = a - v3(0.0f, s - ux, ux), ey = a - v3(uy, 0.0f, s - uy), ez = a - v3(s - uz, uz, 0.0f);
    V3 v = ex;
    float vv = dot(ex, ex);
    const float yy = dot(ey, ey), zz = dot(ez, ez);
    if (yy < vv) { v = ey; vv = yy; }
    if (zz < vv) { v = ez; vv = zz; }
    return v4(sg * safeDir3(v, v3s(k)), fsqrt(vv));
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
// Part of a ring: the arc of radius R in the xy plane spanning `angle` symmetric about +y, swept by a tube of radius r
// (builder/primitives.js capped torus). sn, cs: the sine and cosine of half the angle (from the emitter). Folded to
// x >= 0 the arc runs from +y to its end R (sn, cs). The nearest arc point is the circle's own point in q.xy's direction
// while that direction lies within the arc, else the end; q.xy's component along (cs, -sn), the normal of the end's ray
// pointing past it, tells which. The tube is then the distance from that point less r.
__device__ __forceinline__ V4 NcappedTorus(V3 p, float R, float r, float sn, float cs) {
    const float x = fabsf(p.x), y = p.y;
    float ax = R * sn, ay = R * cs;
    if (x * cs - y * sn <= 0.0f) {
        const float rr = x * x + y * y;
        if (rr > 1e-24f) { const float k = R * frsq(rr); ax = x * k; ay = y * k; } else { ax = 0.0f; ay = R; }
    }
    const V3 v = v3(x - ax, y - ay, p.z);
    const float l = length(v);
    const V3 n = l > 1e-12f ? v * frcp(l) : v3(0.0f, 0.0f, 1.0f);
    return v4(v3(p.x < 0.0f ? -n.x : n.x, n.y, n.z), l - r);
}
// A chain link: the torus of ring R and tube r about z, elongated along y by span -- y first moved to the nearest point
// of [-span/2, span/2], which opens the ring into two half rings joined by straight lengths.
__device__ __forceinline__ V4 Nlink(V3 p, float span, float R, float r) {
    const float e = fmaxf(0.5f * span, 0.0f);
    return Ntorus(v3(p.x, p.y - fminf(fmaxf(p.y, -e), e), p.z), R, r);
}
// A hollow cylinder along z, centred (builder/primitives.js tube): the rectangle [inner, outer] x [-h/2, h/2] of the
// meridian plane, turned; an inner radius of 0 or less is the solid cylinder (the rectangle from the axis out).
__device__ __forceinline__ V4 Ntube(V3 p, float outer, float inner, float height) {
    const float r = fsqrt(p.x * p.x + p.y * p.y), hh = fmaxf(0.5f * height, 0.0f);
    if (inner <= 0.0f) return merTo3(p, merRect(r, p.z, 0.0f, 0.0f, fmaxf(outer, 0.0f), hh));
    const float ro = fmaxf(outer, inner + 1e-6f);
    return merTo3(p, merRect(r, p.z, 0.5f * (inner + ro), 0.0f, 0.5f * (ro - inner), hh));
}
// A dome: the ball of radius R whose centre is at z = c = height - R (held in [-R, R)), cut by its base plane z = 0 and
// kept above it -- the intersection of a ball and a half-space, convex. In the meridian plane: inside, the nearer of
// the sphere (|q - C| - R) and the base (-z). Outside, the nearer of the base disc r <= w (w = sqrt(R^2 - c^2), the
// rim's radius) and the cap: the sphere's own point towards q while that point is above the base, else the rim. Exact
// both sides.
__device__ __forceinline__ V4 Ndome(V3 p, float R0, float height) {
    const float R = fmaxf(R0, 1e-6f);
    const float c = fminf(fmaxf(height - R, -R), R * (1.0f - 1e-4f));
    const float w = fsqrt(fmaxf(R * R - c * c, 0.0f));
    const float r = fsqrt(p.x * p.x + p.y * p.y), z = p.z, ez = z - c;
    const float l = fsqrt(r * r + ez * ez);
    Mer m;
    if (z > 0.0f && l < R) {
        m = -z > l - R ? mer(0.0f, -1.0f, -z) : merCircle(r, z, 0.0f, c, R);
    } else {
        const float b = fminf(r, w);
        const float base2 = (r - b) * (r - b) + z * z;
        // the cap's nearest point: the sphere's towards q when it is not below the base, else the rim (w, 0)
        const bool radial = l > 1e-12f && c * l + R * ez >= 0.0f;
        const float fr = radial ? R * r / l : w, fz = radial ? c + R * ez / l : 0.0f;
        const float cap2 = (r - fr) * (r - fr) + (z - fz) * (z - fz);
        if (base2 < cap2) m = r > w ? merCircle(r, z, w, 0.0f, 0.0f) : mer(0.0f, -1.0f, -z);
        else m = radial ? merCircle(r, z, 0.0f, c, R) : merCircle(r, z, w, 0.0f, 0.0f);
    }
    return merTo3(p, m);
}

// An outline's meridian answer extruded along z to half height h, exactly: past both the outline and the caps, the
// corner's distance; else the larger of the two.
__device__ __forceinline__ V4 merExtrude(V3 p, Mer g, float h) {
    const float wz = fabsf(p.z) - h, sz = p.z >= 0.0f ? 1.0f : -1.0f;
    if (g.d > 0.0f && wz > 0.0f) {
        const float l = fsqrt(g.d * g.d + wz * wz), il = frcp(l);
        return v4(v3(g.nr * g.d * il, g.nz * g.d * il, sz * wz * il), l);
    }
    if (g.d > wz) return v4(v3(g.nr, g.nz, 0.0f), g.d);
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
__device__ __forceinline__ Mer pfEnd(const Prof& A) {
    const float d = fsqrt(A.d2), s = A.inside ? -1.0f : 1.0f;
    if (d > 1e-12f) { const float k = s / fmaxf(d, 1e-30f); return mer(A.vx * k, A.vy * k, s * d); }
    return mer(A.nx, A.ny, s * d);
}
// The parabola z = k r^2 (k > 0) of a meridian plane: the signed distance from (x, y), positive below the curve
// (outside the bowl), and the outward normal at the nearest point. The nearest point (t, k t^2), for |x|, sets the
// derivative of the squared distance to zero: 2 k^2 t^3 + (1 - 2 k y) t - |x| = 0, i.e. t^3 + P t + Q = 0 with
// P = (1 - 2 k y) / 2k^2 and Q = -|x| / 2k^2 <= 0, whose largest real root is the one. Fujiwara's bound puts every root
// within 2 max(sqrt|P|, cbrt(|Q|/2)); Newton's method started there descends on the largest root without overshoot
// (the cubic is convex and rising beyond it), and eight steps reach f32 precision over the reflectors' range.
__device__ __forceinline__ Mer parabolaMer(float x, float y, float k) {
    const float ax = fabsf(x);
    const float ik2 = 1.0f / (2.0f * k * k);
    const float P = (1.0f - 2.0f * k * y) * ik2, Q = -ax * ik2;
    float t = 2.0f * fmaxf(fsqrt(fabsf(P)), cbrtf(-0.5f * Q));
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const float den = 3.0f * t * t + P;
        t = den > 1e-30f ? (2.0f * t * t * t - Q) / den : t;
    }
    const float er = ax - t, ez = y - k * t * t;
    const float slope = 2.0f * k * t, in = frsq(1.0f + slope * slope);
    const float d = fsqrt(er * er + ez * ez);
    return mer((x < 0.0f ? -slope : slope) * in, -in, k * ax * ax >= y ? d : -d);
}
// A parabolic reflector (shapes.js NparabolicOffset:418, builder/primitives.js reflector): focus at the origin, opening
// to +z, the mirror z = (r - c)^2 / 4f - f out to radius A, c the radius of a ring of foci (c 0: a point focus). The
// mirror's profile in the half-plane r >= 0 is the parabola of focal length f centred on r = c; the surface turned
// about the axis also holds its mirror image, centred on r = -c, and of the two the nearer counts where the ring is on
// the point's side (c > 0, their union) and the farther where it is across the axis (c < 0, the two bowls' common
// inside). solid <= 0.5: the mirror with its wall w thick behind it (0 <= s <= w in the distance s to the profile),
// cut by the cylinder r <= A; else the bowl's inside, capped at the rim's height. U: f, A, w, solid, c (c held at most
// A by the emitter).
__device__ __forceinline__ Mer reflectorMer(V3 p, float f, float A, float wall, float solid, float c) {
    const float fo = fmaxf(f, 1e-6f), Aa = fmaxf(A, 0.0f), cc = fminf(c, Aa);
    const float r = fsqrt(p.x * p.x + p.y * p.y), y = p.z + fo, k = 0.25f / fo;
    Mer s = parabolaMer(r - cc, y, k);
    if (cc != 0.0f) {
        const Mer o = parabolaMer(r + cc, y, k);
        if (cc < 0.0f ? o.d > s.d : o.d < s.d) s = o;
    }
    if (solid > 0.5f) return merMax(s, mer(0.0f, 1.0f, y - k * (Aa - cc) * (Aa - cc)));
    const Mer shell = merMax(merFlip(s), mer(s.nr, s.nz, s.d - fmaxf(wall, 1e-6f)));
    return merMax(shell, mer(1.0f, 0.0f, r - Aa));
}
__device__ __forceinline__ V4 Nparabolic(V3 p, float f, float A, float wall, float solid, float c) {
    return merTo3(p, reflectorMer(p, f, A, wall, solid, c));
}
__device__ __forceinline__ float Dparabolic(V3 p, float f, float A, float wall, float solid, float c) {
    return reflectorMer(p, f, A, wall, solid, c).d;
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

// ---- shaped joins (the builder's joinShape, model version 28; scene_api.js:393-406). For a join of radius r a round
// is a circular fillet where the two surfaces meet square, a chamfer the flat cut between the same two tangent lines,
// written in Mercury's hg_sdf forms (fOpIntersectionRound, fOpIntersectionChamfer, 2016) on the two signed distances:
//     round:   with A = a + r, B = b + r: |(A, B)| - r where both are positive, else max(a, b)
//     chamfer: max(max(a, b), (a + b + r) / sqrt 2)
// A union is the intersection of the two complements, a cut the intersection with the second's complement. The forms
// are exact for two square planes; for curved surfaces, or ones meeting at other angles, they can read more than the
// true distance away from the edge, which the march must never be told. So each is held under the hard join's value h
// plus an allowance: the most the form adds anywhere with h <= 0 -- at the sharp corner, (sqrt 2 - 1) r for the
// round, r / sqrt 2 for the chamfer -- which leaves the surface and everything inside the form's exactly, falling
// quadratically to nothing at h = 2r, past which the hard join's value (never above the true distance) is all there
// is. The allowance falls with slope at most its size over r, under 1, so the value stays 1-Lipschitz.
#define PT_ROUND_MOST 0.41421356f
#define PT_CHAMFER_MOST 0.70710678f
__device__ __forceinline__ float shapedAllowance(float h, float r, float most) {
    const float f = saturate(1.0f - 0.5f * h / fmaxf(r, 1e-30f));
    return most * r * f * f;
}
__device__ __forceinline__ float dRoundI(float a, float b, float r) {
    const float h = fmaxf(a, b), A = a + r, B = b + r;
    if (!(A > 0.0f && B > 0.0f)) return h;
    return fminf(fsqrt(A * A + B * B) - r, h + shapedAllowance(h, r, PT_ROUND_MOST));
}
__device__ __forceinline__ float dChamferI(float a, float b, float r) {
    const float h = fmaxf(a, b);
    return fmaxf(h, fminf((a + b + r) * PT_CHAMFER_MOST, h + shapedAllowance(h, r, PT_CHAMFER_MOST)));
}
__device__ __forceinline__ float dRoundU(float a, float b, float r) { return -dRoundI(-a, -b, r); }
__device__ __forceinline__ float dRoundS(float a, float b, float r) { return dRoundI(a, -b, r); }
__device__ __forceinline__ float dChamferU(float a, float b, float r) { return -dChamferI(-a, -b, r); }
__device__ __forceinline__ float dChamferS(float a, float b, float r) { return dChamferI(a, -b, r); }
// With normals: where the form's value is used, its gradient (A na + B nb for the round, na + nb for the chamfer,
// normalised); where the hard join's or its allowance, the harder side's normal. Either side without a normal leaves
// the result without one (the surface kernel then takes it by differences), as for the smooth blends.
__device__ __forceinline__ V4 shapedResult(V4 a, V4 b, V4 hard, float d, float capped, V3 grad) {
    const bool both = dot(xyz(a), xyz(a)) >= 1e-12f && dot(xyz(b), xyz(b)) >= 1e-12f;
    if (d >= capped) return v4(both ? xyz(hard) : v3s(0.0f), capped);
    const float l = length(grad);
    return v4(both && l > 1e-12f ? grad * frcp(l) : v3s(0.0f), d);
}
__device__ __forceinline__ V4 csgRoundI(V4 a, V4 b, float r) {
    const V4 hard = a.w > b.w ? a : b;
    const float A = a.w + r, B = b.w + r;
    if (!(A > 0.0f && B > 0.0f)) return hard;
    return shapedResult(a, b, hard, fsqrt(A * A + B * B) - r, hard.w + shapedAllowance(hard.w, r, PT_ROUND_MOST),
                        xyz(a) * A + xyz(b) * B);
}
__device__ __forceinline__ V4 csgChamferI(V4 a, V4 b, float r) {
    const V4 hard = a.w > b.w ? a : b;
    const float d = (a.w + b.w + r) * PT_CHAMFER_MOST;
    if (d <= hard.w) return hard;
    return shapedResult(a, b, hard, d, hard.w + shapedAllowance(hard.w, r, PT_CHAMFER_MOST), xyz(a) + xyz(b));
}
__device__ __forceinline__ V4 csgRoundU(V4 a, V4 b, float r) { return -csgRoundI(-a, -b, r); }
__device__ __forceinline__ V4 csgRoundS(V4 a, V4 b, float r) { return csgRoundI(a, -b, r); }
__device__ __forceinline__ V4 csgChamferU(V4 a, V4 b, float r) { return -csgChamferI(-a, -b, r); }
__device__ __forceinline__ V4 csgChamferS(V4 a, V4 b, float r) { return csgChamferI(a, -b, r); }

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

// THE BUILTIN MODIFIER GROUPS (builder/modifiers.js): the point a group's children are evaluated at, in the group's
// frame, and what happens to their value. A twist turns the point about z by rate z; a bend turns it in the xz plane by
// amount x (both scale the distance by their Lipschitz bound, applied by the emitter); a repeat folds it into the cell
// of copies `spacing` apart, at most `count` each way along each axis (0 spacing: none along it); a mirror reflects it
// across x = 0. A round grows the children's surface by r; a shell makes it a skin of thickness t (the normal turned
// outward on both sides).
__device__ __forceinline__ V3 mTwist(V3 p, float rate) {
    float s, c;
    __sincosf(rate * p.z, &s, &c);
    return v3(c * p.x - s * p.y, s * p.x + c * p.y, p.z);
}
__device__ __forceinline__ V3 mBend(V3 p, float amount) {
    float s, c;
    __sincosf(amount * p.x, &s, &c);
    return v3(c * p.x - s * p.z, p.y, s * p.x + c * p.z);
}
__device__ __forceinline__ float mRepeat1(float x, float s, float n) {
    return s != 0.0f ? x - s * fminf(fmaxf(rintf(x / s), -n), n) : x;
}
__device__ __forceinline__ V3 mRepeat(V3 p, V3 s, V3 n) { return v3(mRepeat1(p.x, s.x, n.x), mRepeat1(p.y, s.y, n.y), mRepeat1(p.z, s.z, n.z)); }
__device__ __forceinline__ V3 mMirror(V3 p) { return v3(fabsf(p.x), p.y, p.z); }
__device__ __forceinline__ V4 mGrow(V4 f, float r) { return v4(v3(f.x, f.y, f.z), f.w - r); }
__device__ __forceinline__ V4 mShell(V4 f, float t) {
    const float s = f.w < 0.0f ? -1.0f : 1.0f;
    return v4(v3(f.x, f.y, f.z) * s, fabsf(f.w) - t);
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

// The scene code's runtime (pt_wgsl_rt.h, pt_wgsl_lib.h) on the device: FP32, inlined device functions.
#define WG_FN __device__ __forceinline__
typedef float WF;
#define WL(x) x##f
