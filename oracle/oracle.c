/*
 * THE ORACLE -- test infrastructure only: a CPU restatement, in FP64, of the estimators of ShadyTracer's
 * wavefront path tracer (lib/pathtracer), against which the HIP engine is checked. Nothing in the product
 * links or calls it; tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do, through ctypes.
 *
 * It reads a scene as the emitter's postfix program (shadytracer_amd/js/hipgen.js `program`), the same
 * parameter block, materials, light faces and uniforms (include/pt_types.h), and renders it path by path
 * (depth first, a shadow ray traced where it is made) rather than in wavefronts. Each function cites the
 * reference file:line it restates. Random numbers: PCG4D keyed like the engine's (sample seed, ray, bounce,
 * stream), so a camera ray's jitter -- and with it its first march -- is the engine's own.
 *
 * Pinned by (tests/test_oracle.py): closed forms (Lambertian floor = rho L; white furnace = 1; haze furnace =
 * 1; Fresnel at normal incidence and total internal reflection; Cauchy through n_d and V_d), brute-force
 * ray-primitive intersection on small inputs, quadrature invariants of the BSDF (pdf integrates to its
 * non-delta share; sampling weights agree with eval/pdf), the reference's own measured vectors
 * (lib/pathtracer/test/fixtures/run-0.77.0-pascal.json via tests/golden/reference_rows.json: the furnace rows, the haze
 * rows, the slab's transmissions), known answers of PCG4D from the JCGT 2020 definition, and the error-bar identity of
 * core/measure.js (two renders from disjoint samples differ per pixel by their combined error bars).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/pt_types.h"

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double x, y, z; } d3;
typedef struct { double x, y, z, w; } d4;

static inline d3 v(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static inline d3 add(d3 a, d3 b) { return v(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline d3 sub(d3 a, d3 b) { return v(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline d3 mul(d3 a, d3 b) { return v(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline d3 scl(d3 a, double s) { return v(a.x * s, a.y * s, a.z * s); }
static inline double dotp(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double len(d3 a) { return sqrt(dotp(a, a)); }
static inline d3 unit(d3 a) { double l = len(a); return l > 0 ? scl(a, 1.0 / l) : a; }
static inline d3 neg(d3 a) { return v(-a.x, -a.y, -a.z); }
static inline d3 crossp(d3 a, d3 b) { return v(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline double maxd(double a, double b) { return a > b ? a : b; }
static inline double mind(double a, double b) { return a < b ? a : b; }
static inline double clampd(double x, double a, double b) { return x < a ? a : (x > b ? b : x); }
static inline double sgn(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); }
static inline double max3d(d3 a) { return maxd(a.x, maxd(a.y, a.z)); }
static inline double lumad(d3 c) { return 0.2126 * c.x + 0.7152 * c.y + 0.0722 * c.z; }
static inline d3 f3(const float* p) { return v(p[0], p[1], p[2]); }
static inline d4 d4of(d3 n, double w) { d4 r = {n.x, n.y, n.z, w}; return r; }

#define PI_D 3.14159265358979323846

/* ---- PCG4D (Jarzynski & Olano, JCGT 2020), as core/tracebasics.js:369-380 uses it ---- */
typedef struct { uint32_t a, b, c, d; } OrRng;
static void pcg4d(OrRng* s) {
    uint32_t x = s->a * 1664525u + 1013904223u, y = s->b * 1664525u + 1013904223u;
    uint32_t z = s->c * 1664525u + 1013904223u, w = s->d * 1664525u + 1013904223u;
    x += y * w; y += z * x; z += x * y; w += y * z;
    x ^= x >> 16; y ^= y >> 16; z ^= z >> 16; w ^= w >> 16;
    x += y * w; y += z * x; z += x * y; w += y * z;
    s->a = x; s->b = y; s->c = z; s->d = w;
}
static void rng4(OrRng* s, double* u) {
    pcg4d(s);
    u[0] = (double)(s->a >> 8) / 16777216.0; u[1] = (double)(s->b >> 8) / 16777216.0;
    u[2] = (double)(s->c >> 8) / 16777216.0; u[3] = (double)(s->d >> 8) / 16777216.0;
}
static OrRng rngAt(const uint32_t* seed, uint32_t ray, uint32_t bounce, uint32_t stream) {
    OrRng s = {seed[0] ^ ray, seed[1] ^ (bounce * 0x9E3779B9u), seed[2] ^ stream, seed[3]};
    pcg4d(&s);
    return s;
}
static void sampleSeed(uint32_t base, uint32_t sample, uint32_t* out) {
    OrRng r = {base, sample, 0x5EED5EEDu, 0x2026F00Du};
    pcg4d(&r); pcg4d(&r);
    out[0] = r.a; out[1] = r.b; out[2] = r.c; out[3] = r.d;
}

/* ---- the scene program (hipgen.js `program`) ---- */
enum { OP_PRIM = 0, OP_JOIN = 1, OP_FOLD = 2, OP_NODE = 3, OP_MOD = 4, OP_MODEND = 5, OP_USER = 6, OP_UGROUP = 7, OP_UFOLD = 8,
       OP_UGROUPEND = 9, OP_RELIEF = 10 };
/* the builtin modifier groups (builder/modifiers.js; hipgen.js MODIFIERS), in oracle.py's MODS order */
enum { MOD_TWIST = 0, MOD_BEND, MOD_REPEAT, MOD_MIRROR, MOD_CIRCULAR, MOD_ROUND, MOD_SHELL };
enum { SH_PLANE = 0, SH_SPHERE, SH_BOX, SH_CYLINDER, SH_CAPSULE, SH_TORUS, SH_RECT, SH_DISC, SH_LENS, SH_PRISM, SH_REFLECTOR, SH_ASPHERE,
       SH_REVOLUTION, SH_FRESNEL, SH_EXTRUSION, SH_CONE, SH_ROUNDCONE, SH_OCTA, SH_CTORUS, SH_LINK, SH_TUBE, SH_DOME, SH_FRUSTUM };
enum { J_UNION = 0, J_SUBTRACT, J_INTERSECT };
/* prim: shape, xf, p, mat, prim; join: join, k, and its shape (0 smooth, 1 round, 2 chamfer) in shape; node (a top-level node's start): xf = its bounding sphere's offset in
 * U (-1: unbounded), join = its mask bit (-1: none), k = the index of its fold; mod (a modifier group's start): shape =
 * its kind, xf = the group's frame, p = its parameters, k = the index of its modend; user (an object's own code):
 * shape = the index of its compiled function, xf its frame, mat, prim; ugroup (a group's own code): shape = its
 * function, xf its frame, k = the index of its ugroupend, each child's value ops followed by a ufold */
typedef struct { int op, shape, xf, p, mat, prim, join, k; } OrOp;

typedef struct {
    const OrOp* ops;
    int nops;
    const float* U;
    const PtMaterial* mats;
    int nmats;
    const PtLightFace* faces;
    int nfaces;
    double faceTotal;
    const PtUniforms* u;
} OrScene;

/* One face of a spherical lens (core/sdf.js lensFace:450): s the side (+1 the +z face), its vertex at s*vz, R its
 * signed radius (0 flat, positive bulging out, negative hollowed in, closed by the plane through its rim). */
static d4 lensFace(d3 p, double s, double vz, double R, double rim) {
    if (R == 0) return d4of(v(0, 0, s), s * p.z - vz);
    d3 q = sub(p, v(0, 0, s * (vz - R)));
    double l = maxd(len(q), 1e-12), sg = R > 0 ? 1 : -1;
    d4 f = d4of(scl(q, sg / l), sg * (l - fabs(R)));
    if (R < 0) {
        double sag = fabs(R) - sqrt(maxd(R * R - rim * rim, 0)), w = s * p.z - (vz + sag);
        if (w > f.w) f = d4of(v(0, 0, s), w);
    }
    return f;
}

/* The nearest point of the parabola y = k x^2 (k > 0) to (x, y) in a meridian plane, by Newton's method on
 * g(t) = 2k^2 t^3 + (1 - 2ky) t - x from t = x (g' > 0 wherever the minimum is unique, g is increasing past the
 * largest root), polished in FP64 -- a different route to the root the GPU takes by Cardano (shapes.js g2Parabola:61).
 * Returns the signed distance (positive below the curve) and the outward normal (gx, gy). */
static double parabolaD(double x, double y, double k, double* gx, double* gy) {
    double sx = x >= 0 ? 1 : -1, xa = fabs(x);
    double t = maxd(xa, sqrt(maxd(y / k, 0)));   /* at or past the largest root: g(t) >= 0 there */
    for (int i = 0; i < 200; ++i) {
        double g = 2 * k * k * t * t * t + (1 - 2 * k * y) * t - xa, dg = 6 * k * k * t * t + (1 - 2 * k * y);
        if (dg <= 0) { t *= 1.5; continue; }
        double nt = t - g / dg;
        if (fabs(nt - t) <= 1e-15 * maxd(1, fabs(t))) { t = nt; break; }
        t = nt;
    }
    double ex = xa - t, ey = y - k * t * t, side = k * xa * xa - y >= 0 ? 1 : -1;
    double in = 1 / sqrt(1 + 4 * k * k * t * t);
    *gx = 2 * k * t * in * sx; *gy = -in;
    return side * sqrt(ex * ex + ey * ey);
}
typedef struct { double x, y, d; } g2v;
static g2v g2max(g2v a, g2v b) { return a.d >= b.d ? a : b; }
static g2v parabolaOff(double x, double y, double k, double c) {
    g2v a, b;
    a.d = parabolaD(x - c, y, k, &a.x, &a.y);
    if (c == 0) return a;
    b.d = parabolaD(x + c, y, k, &b.x, &b.y);
    if (c < 0) return a.d >= b.d ? a : b;
    return a.d <= b.d ? a : b;
}
/* One face of an aspheric lens (core/shapes.js asphSag, asphSlopeMax, asphFace: 658-700): c the curvature (0 flat), k
 * the conic constant, a[4] the even terms; the height above the surface over the steepest slope within the rim. */
static double asphSagD(double r, double c, double k, const float* a, double* slope) {
    double r2 = r * r, w = 1 - (1 + k) * c * c * r2, u = sqrt(maxd(w, 1e-4));
    double poly = r2 * r2 * (a[0] + r2 * (a[1] + r2 * (a[2] + r2 * a[3])));
    double dpoly = r * r2 * (4 * a[0] + r2 * (6 * a[1] + r2 * (8 * a[2] + r2 * 10 * a[3])));
    *slope = (w >= 1e-4 ? c * r / u : 2 * c * r / (1 + u)) + dpoly;
    return c * r2 / (1 + u) + poly;
}
static d4 asphFaceD(d3 p, double l, double rx, double ry, double rho, double ht, double R, double k, const float* a, double s) {
    double c = R == 0 ? 0 : 1 / R, slope;
    double sag = asphSagD(mind(l, rho), c, k, a, &slope);
    double r2 = rho * rho, u = sqrt(maxd(1 - (1 + k) * c * c * r2, 1e-4));
    double m = fabs(c) * rho / u + rho * r2 * (4 * fabs(a[0]) + r2 * (6 * fabs(a[1]) + r2 * (8 * fabs(a[2]) + r2 * 10 * fabs(a[3]))));
    double g = l < rho ? slope : 0, in = 1 / sqrt(g * g + 1);
    return d4of(v(rx * g * in, ry * g * in, s * in), (s * p.z - ht + sag) / sqrt(1 + m * m));
}

/* An outline walked piece by piece (core/shapes.js pfStart..pfEnd:684-715): the nearest point of its surface pieces
 * and inside by the even-odd rule over the edges and arcs' chords. U: count, then 9 numbers a piece -- kind 0 a
 * segment (a, b, surface) or 1 an arc (centre, radius, a, b, convex). Returns (normal, signed distance) in the plane. */
static g2v dir2(double x, double y, double fx, double fy) { g2v r; double l = sqrt(x * x + y * y); if (l > 1e-12) { r.x = x / l; r.y = y / l; } else { r.x = fx; r.y = fy; } r.d = 0; return r; }
static int cross2q(double qx, double qy, double ax, double ay, double bx, double by) {
    return (ay > qy) != (by > qy) && qx < ax + (qy - ay) * (bx - ax) / (by - ay);
}
static g2v outlineAt(const float* P, double qx, double qy) {
    int n = (int)P[0];
    double bvx = 0, bvy = 0, bd2 = 3e38, nx = 1, ny = 0;
    int inside = 0;
    for (int i = 0; i < n; ++i) {
        const float* k = P + 1 + 9 * i;
        double px, py, nnx, nny;
        if (k[0] == 0) {
            double ax = k[1], ay = k[2], bx = k[3], by = k[4];
            if (cross2q(qx, qy, ax, ay, bx, by)) inside = !inside;
            if (k[5] == 0) continue;
            double ex = bx - ax, ey = by - ay, t = clampd(((qx - ax) * ex + (qy - ay) * ey) / maxd(ex * ex + ey * ey, 1e-30), 0, 1);
            px = ax + ex * t; py = ay + ey * t;
            g2v nn = dir2(ey, -ex, 1, 0); nnx = nn.x; nny = nn.y;
        } else {
            double cx = k[1], cy = k[2], r = k[3], ax = k[4], ay = k[5], bx = k[6], by = k[7];
            double wx = qx - cx, wy = qy - cy, mx = 0.5 * (ax + bx) - cx, my = 0.5 * (ay + by) - cy;
            int segIn = wx * wx + wy * wy < r * r && wx * mx + wy * my > mx * mx + my * my;
            if (cross2q(qx, qy, ax, ay, bx, by) != segIn) inside = !inside;
            double eax = ax - cx, eay = ay - cy, ebx = bx - cx, eby = by - cy, s = eax * eby - eay * ebx >= 0 ? 1 : -1;
            if (s * (eax * wy - eay * wx) >= 0 && s * (wx * eby - wy * ebx) >= 0) {
                g2v md = dir2(mx, my, 1, 0), d = dir2(wx, wy, md.x, md.y);
                px = cx + d.x * r; py = cy + d.y * r;
            } else if ((qx - ax) * (qx - ax) + (qy - ay) * (qy - ay) <= (qx - bx) * (qx - bx) + (qy - by) * (qy - by)) { px = ax; py = ay; }
            else { px = bx; py = by; }
            g2v nn = dir2(px - cx, py - cy, 1, 0);
            nnx = k[8] != 0 ? nn.x : -nn.x; nny = k[8] != 0 ? nn.y : -nn.y;
        }
        double vx = qx - px, vy = qy - py, d2 = vx * vx + vy * vy;
        if (d2 < bd2) { bvx = vx; bvy = vy; bd2 = d2; nx = nnx; ny = nny; }
    }
    double d = sqrt(bd2), s = inside ? -1 : 1;
    g2v out;
    if (d > 1e-12) { out.x = bvx * s / d; out.y = bvy * s / d; } else { out.x = nx; out.y = ny; }
    out.d = s * d;
    return out;
}

/* The builder's other primitives (core/shapes.js:119-373, core/sdf.js sdCappedCone:633), in a meridian plane where they
 * are revolved: r = |p.xy|, z; the result lifted about the axis by revolveD. */
static d4 revolveD(d3 p, g2v g) {
    double l = sqrt(p.x * p.x + p.y * p.y), rx = l > 1e-12 ? p.x / l : 1, ry = l > 1e-12 ? p.y / l : 0;
    return d4of(v(rx * g.x, ry * g.x, g.y), g.d);
}
static g2v circle2(double qx, double qy, double cx, double cy, double r) {
    g2v g = dir2(qx - cx, qy - cy, 1, 0);
    g.d = sqrt((qx - cx) * (qx - cx) + (qy - cy) * (qy - cy)) - r;
    return g;
}
static g2v box2(double qx, double qy, double cx, double cy, double bx, double by) {
    double sx = qx >= cx ? 1 : -1, sy = qy >= cy ? 1 : -1, ex = fabs(qx - cx) - bx, ey = fabs(qy - cy) - by;
    g2v g;
    if (ex > 0 || ey > 0) { double ox = maxd(ex, 0), oy = maxd(ey, 0), l = sqrt(ox * ox + oy * oy); g.x = ox / l * sx; g.y = oy / l * sy; g.d = l; return g; }
    if (ex > ey) { g.x = sx; g.y = 0; g.d = ex; return g; }
    g.x = 0; g.y = sy; g.d = ey; return g;
}
static d3 safeDirD(d3 a, d3 fb) { double l = len(a); return l > 1e-12 ? scl(a, 1 / l) : fb; }
/* The nearest point of the planar convex polygon c[0..3] (either winding; corners may coincide) with unit normal n:
   the projection when the turn about n has one sign along every edge, else the nearest point of the nearest edge. */
static d3 polyNearD(d3 p, const d3* c, d3 n) {
    d3 q = sub(p, scl(n, dotp(sub(p, c[0]), n)));
    int pos = 0, neg = 0;
    for (int i = 0; i < 4; ++i) { double t = dotp(crossp(sub(c[(i + 1) & 3], c[i]), sub(q, c[i])), n); pos += t > 0; neg += t < 0; }
    if ((pos == 0) != (neg == 0)) return q;      /* one sign only: inside (none at all: no area) */
    d3 best = c[0]; double bd = 1e300;
    for (int i = 0; i < 4; ++i) {
        d3 a = c[i], ab = sub(c[(i + 1) & 3], a);
        double t = clampd(dotp(sub(p, a), ab) / maxd(dotp(ab, ab), 1e-300), 0, 1);
        d3 e = add(a, scl(ab, t)); double de = dotp(sub(p, e), sub(p, e));
        if (de < bd) { bd = de; best = e; }
    }
    return best;
}

/* Primitive (outward normal, distance) in its own frame: core/sdf.js:352-628 restated; the lens core/sdf.js:463, the
 * prism core/shapes.js:263-287 (a regular polygon extruded), the reflector core/shapes.js:392-421, the aspheric lens
 * core/shapes.js:658. */
static d4 primLocal(int shape, d3 p, const float* P) {
    switch (shape) {
    case SH_PLANE: return d4of(v(0, 0, 1), p.z);
    case SH_SPHERE: { double l = len(p); return d4of(l > 1e-300 ? scl(p, 1 / l) : v(0, 0, 1), l - P[0]); }
    case SH_BOX: {
        double rr = clampd(P[3], 0, mind(P[0], mind(P[1], P[2])));
        d3 b = v(P[0] - rr, P[1] - rr, P[2] - rr);
        d3 q = v(fabs(p.x) - b.x, fabs(p.y) - b.y, fabs(p.z) - b.z);
        d3 g = v(maxd(q.x, 0), maxd(q.y, 0), maxd(q.z, 0));
        double d = len(g) + mind(maxd(q.x, maxd(q.y, q.z)), 0) - rr;
        d3 n;
        if (g.x > 0 || g.y > 0 || g.z > 0) n = mul(unit(g), v(sgn(p.x), sgn(p.y), sgn(p.z)));
        else if (q.x > q.y && q.x > q.z) n = v(sgn(p.x), 0, 0);
        else if (q.y > q.z) n = v(0, sgn(p.y), 0);
        else n = v(0, 0, sgn(p.z));
        return d4of(n, d);
    }
    case SH_CYLINDER: {
        double k = clampd(P[2], 0, mind(P[0], P[1]));
        double h = P[0] - k, r = P[1] - k;
        double qx = sqrt(p.x * p.x + p.y * p.y), qy = fabs(p.z);
        double dx = qx - r, dy = qy - h, ox = maxd(dx, 0), oy = maxd(dy, 0), ol = sqrt(ox * ox + oy * oy);
        double dist = mind(maxd(dx, dy), 0) + ol - k;
        double l = maxd(qx, 1e-8);
        d3 side = v(p.x / l, p.y / l, 0), cap = v(0, 0, sgn(p.z)), n;
        if (dx > 0 && dy > 0) { double il = 1 / maxd(ol, 1e-8); n = unit(add(scl(side, ox * il), scl(cap, oy * il))); }
        else if (dx > dy) n = side;
        else n = cap;
        return d4of(n, dist);
    }
    case SH_CAPSULE: {
        d3 w = sub(p, v(0, 0, clampd(p.z, -P[0], P[0])));
        double l = len(w);
        return d4of(l > 1e-12 ? scl(w, 1 / l) : v(1, 0, 0), l - P[1]);
    }
    case SH_TORUS: {
        double lxy = sqrt(p.x * p.x + p.y * p.y);
        double rx = lxy > 1e-12 ? p.x / lxy : 1, ry = lxy > 1e-12 ? p.y / lxy : 0;
        double qx = lxy - P[0], qy = p.z, ql = sqrt(qx * qx + qy * qy);
        d3 n = ql > 1e-12 ? v(rx * qx / ql, ry * qx / ql, qy / ql) : v(0, 0, 1);
        return d4of(n, ql - P[1]);
    }
    case SH_RECT: {
        double ex = maxd(fabs(p.x) - P[0], 0), ey = maxd(fabs(p.y) - P[1], 0);
        double s = p.z < 0 ? -1 : 1, d = sqrt(ex * ex + ey * ey + p.z * p.z);
        if (maxd(ex, ey) <= 0 || d < 1e-12) return d4of(v(0, 0, s), fabs(p.z));
        return d4of(v(ex / d * sgn(p.x), ey / d * sgn(p.y), p.z / d), d);
    }
    case SH_DISC: {
        double l = sqrt(p.x * p.x + p.y * p.y), e = maxd(l - P[0], 0);
        double s = p.z < 0 ? -1 : 1, d = sqrt(e * e + p.z * p.z);
        if (e <= 0 || d < 1e-12) return d4of(v(0, 0, s), fabs(p.z));
        double il = 1 / maxd(l, 1e-12);
        return d4of(v(e * p.x * il / d, e * p.y * il / d, p.z / d), d);
    }
    case SH_LENS: {     /* P: rim, vz, r1, r2 (the planes after them are the emitter's, recomputed here) */
        double rim = P[0], vz = P[1], lxy = maxd(sqrt(p.x * p.x + p.y * p.y), 1e-12);
        d4 n = d4of(v(p.x / lxy, p.y / lxy, 0), lxy - rim);
        d4 a = lensFace(p, 1, vz, P[2], rim);
        if (a.w > n.w) n = a;
        d4 b = lensFace(p, -1, vz, P[3], rim);
        if (b.w > n.w) n = b;
        return n;
    }
    case SH_PRISM: {    /* P[4] the half height, P[5] the document's sides, P[6] its corner radius */
        double sides = P[5], nn = clampd(nearbyint(sides), 3, 64), R = maxd(P[6], 0), h = P[4];
        double an = PI_D / nn, a = R * cos(an), he = R * sin(an);
        double k = nearbyint((atan2(p.y, p.x) + 0.5 * PI_D) / (2 * an)), phi = -0.5 * PI_D + 2 * an * k;
        double nx = cos(phi), ny = sin(phi), u = p.x * nx + p.y * ny, w = -p.x * ny + p.y * nx;
        double gx = nx, gy = ny, g = u - a;
        if (g > 0) {
            double ey = w - clampd(w, -he, he), d = sqrt(g * g + ey * ey);
            if (d > 1e-12) { gx = (nx * g - ny * ey) / d; gy = (ny * g + nx * ey) / d; }
            g = d;
        }
        double wz = fabs(p.z) - h, sz = p.z >= 0 ? 1 : -1;
        if (g > 0 && wz > 0) { double l = sqrt(g * g + wz * wz); return d4of(v(gx * g / l, gy * g / l, sz * wz / l), l); }
        if (g > wz) return d4of(v(gx, gy, 0), g);
        return d4of(v(0, 0, sz), wz);
    }
    case SH_REFLECTOR: {    /* P: focal, aperture, wall, solid, focus ring radius */
        double fo = maxd(P[0], 1e-6), A = maxd(P[1], 0), c = mind(P[4], A);
        double x = sqrt(p.x * p.x + p.y * p.y), y = p.z + fo, l = x > 1e-12 ? x : 1;
        g2v s = parabolaOff(x, y, 0.25 / fo, c), r;
        if (P[3] > 0.5) {
            g2v cap = {0, 1, y - 0.25 * (A - c) * (A - c) / fo};
            r = g2max(s, cap);
        } else {
            g2v ns = {-s.x, -s.y, -s.d}, off = {s.x, s.y, s.d - maxd(P[2], 1e-6)}, cyl = {1, 0, x - A};
            r = g2max(g2max(ns, off), cyl);
        }
        double rx = x > 1e-12 ? p.x / l : 1, ry = x > 1e-12 ? p.y / l : 0;
        return d4of(v(rx * r.x, ry * r.x, r.y), r.d);
    }
    case SH_CONE: {         /* P: half height, base radius, top radius; distance only */
        double h = P[0], r1 = P[1], r2 = P[2], qx = sqrt(p.x * p.x + p.y * p.y), qy = p.z;
        double k2x = r2 - r1, k2y = 2 * h, cax = qx - mind(qx, qy < 0 ? r1 : r2), cay = fabs(qy) - h;
        double t = clampd(((r2 - qx) * k2x + (h - qy) * k2y) / (k2x * k2x + k2y * k2y), 0, 1);
        double cbx = qx - r2 + k2x * t, cby = qy - h + k2y * t, s = cbx < 0 && cay < 0 ? -1 : 1;
        return d4of(v(0, 0, 0), s * sqrt(mind(cax * cax + cay * cay, cbx * cbx + cby * cby)));
    }
    case SH_ROUNDCONE: {    /* P: bottom radius, top radius, height */
        double r1 = P[0], r2 = P[1], h = P[2], qx = sqrt(p.x * p.x + p.y * p.y), qy = p.z + 0.5 * h;
        if (h <= fabs(r1 - r2)) return revolveD(p, r1 >= r2 ? circle2(qx, qy, 0, 0, r1) : circle2(qx, qy, 0, h, r2));
        double b = (r1 - r2) / h, a = sqrt(1 - b * b), k = -b * qx + a * qy;
        if (k < 0) return revolveD(p, circle2(qx, qy, 0, 0, r1));
        if (k > a * h) return revolveD(p, circle2(qx, qy, 0, h, r2));
        g2v g = {a, b, a * qx + b * qy - r1};
        return revolveD(p, g);
    }
    case SH_OCTA: {         /* P: size */
        double s = maxd(P[0], 0), c = 0.5773502691896258;
        d3 sg = v(p.x >= 0 ? 1 : -1, p.y >= 0 ? 1 : -1, p.z >= 0 ? 1 : -1), a = v(fabs(p.x), fabs(p.y), fabs(p.z)), fb = v(c, c, c);
        double m = a.x + a.y + a.z - s;
        if (3 * a.x < m) { double k = clampd(0.5 * (a.z - a.y + s), 0, s); d3 w = v(a.x, a.y - s + k, a.z - k); return d4of(mul(sg, safeDirD(w, fb)), len(w)); }
        if (3 * a.y < m) { double k = clampd(0.5 * (a.x - a.z + s), 0, s); d3 w = v(a.y, a.z - s + k, a.x - k); return d4of(mul(sg, safeDirD(v(w.z, w.x, w.y), fb)), len(w)); }
        if (3 * a.z < m) { double k = clampd(0.5 * (a.y - a.x + s), 0, s); d3 w = v(a.z, a.x - s + k, a.y - k); return d4of(mul(sg, safeDirD(v(w.y, w.z, w.x), fb)), len(w)); }
        return d4of(scl(sg, c), m * c);
    }
    case SH_CTORUS: {       /* P: ring, tube, sin and cos of half the angle */
        double R = P[0], sn = P[2], cs = P[3], sx = p.x >= 0 ? 1 : -1, qx = fabs(p.x), cx, cy;
        if (cs * qx > sn * p.y) { cx = R * sn; cy = R * cs; }
        else { g2v d = dir2(qx, p.y, 0, 1); cx = R * d.x; cy = R * d.y; }
        d3 w = v(qx - cx, p.y - cy, p.z), n = safeDirD(w, v(0, 0, 1));
        return d4of(v(n.x * sx, n.y, n.z), len(w) - P[1]);
    }
    case SH_LINK: {         /* P: span, ring, tube: the torus with its y pulled apart */
        double le = maxd(0.5 * P[0], 0), qy = (p.y >= 0 ? 1 : -1) * maxd(fabs(p.y) - le, 0);
        double lxy = sqrt(p.x * p.x + qy * qy), rx = lxy > 1e-12 ? p.x / lxy : 1, ry = lxy > 1e-12 ? qy / lxy : 0;
        double ax = lxy - P[1], ay = p.z, ql = sqrt(ax * ax + ay * ay);
        d3 n = ql > 1e-12 ? v(rx * ax / ql, ry * ax / ql, ay / ql) : v(0, 0, 1);
        return d4of(n, ql - P[2]);
    }
    case SH_TUBE: {         /* P: outer, inner, height */
        double qx = sqrt(p.x * p.x + p.y * p.y), hh = maxd(0.5 * P[2], 0);
        if (P[1] <= 0) return revolveD(p, box2(qx, p.z, 0, 0, maxd(P[0], 0), hh));
        double ro = maxd(P[0], P[1] + 1e-6);
        return revolveD(p, box2(qx, p.z, 0.5 * (P[1] + ro), 0, 0.5 * (ro - P[1]), hh));
    }
    case SH_FRUSTUM: {      /* P: base x, y, top x, y, height -- all six faces, no folding */
        double bx = 0.5 * fabs(P[0]), by = 0.5 * fabs(P[1]), tx = 0.5 * fabs(P[2]), ty = 0.5 * fabs(P[3]), hh = 0.5 * maxd(P[4], 1e-6);
        d3 Bc[4] = {v(bx, -by, -hh), v(bx, by, -hh), v(-bx, by, -hh), v(-bx, -by, -hh)};
        d3 Tc[4] = {v(tx, -ty, hh), v(tx, ty, hh), v(-tx, ty, hh), v(-tx, -ty, hh)};
        d3 nrm[6] = {unit(v(2 * hh, 0, bx - tx)), unit(v(0, 2 * hh, by - ty)), unit(v(-2 * hh, 0, bx - tx)), unit(v(0, -2 * hh, by - ty)), v(0, 0, 1), v(0, 0, -1)};
        d3 F[6][4];
        for (int i = 0; i < 4; ++i) { F[i][0] = Bc[i]; F[i][1] = Bc[(i + 1) & 3]; F[i][2] = Tc[(i + 1) & 3]; F[i][3] = Tc[i]; F[4][i] = Tc[i]; F[5][i] = Bc[i]; }
        double best = -1e300; int bi = 0, inside = 1;
        for (int f = 0; f < 6; ++f) { double d = dotp(sub(p, F[f][0]), nrm[f]); if (d > 0) inside = 0; if (d > best) { best = d; bi = f; } }
        if (inside) return d4of(nrm[bi], best);
        double d2 = 1e300; d3 c = p;
        for (int f = 0; f < 6; ++f) { d3 q = polyNearD(p, F[f], nrm[f]); double e = dotp(sub(p, q), sub(p, q)); if (e < d2) { d2 = e; c = q; } }
        return d4of(safeDirD(sub(p, c), nrm[bi]), sqrt(d2));
    }
    case SH_DOME: {         /* P: radius, height */
        double R = maxd(P[0], 1e-6), c = clampd(P[1] - R, -R, R * (1 - 1e-4)), hc = -c, w = sqrt(maxd(R * R - hc * hc, 0));
        double qx = sqrt(p.x * p.x + p.y * p.y), qy = p.z - c;
        double s = maxd((hc - R) * qx * qx + w * w * (hc + R - 2 * qy), hc * qx - w * qy);
        if (s < 0) return revolveD(p, circle2(qx, qy, 0, 0, R));
        if (qx < w) { g2v g = {0, -1, hc - qy}; return revolveD(p, g); }
        return revolveD(p, circle2(qx, qy, w, hc, 0));
    }
    case SH_REVOLUTION: case SH_FRESNEL: case SH_EXTRUSION: {   /* P: mode (0 revolve, 1 extrude), half height, the outline */
        if (P[0] == 0) {
            double l = sqrt(p.x * p.x + p.y * p.y), rx = l > 1e-12 ? p.x / l : 1, ry = l > 1e-12 ? p.y / l : 0;
            g2v g = outlineAt(P + 2, l, p.z);
            return d4of(v(rx * g.x, ry * g.x, g.y), g.d);
        }
        g2v g = outlineAt(P + 2, p.x, p.y);
        double h = P[1], wz = fabs(p.z) - h, sz = p.z >= 0 ? 1 : -1;
        if (g.d > 0 && wz > 0) { double l = sqrt(g.d * g.d + wz * wz); return d4of(v(g.x * g.d / l, g.y * g.d / l, sz * wz / l), l); }
        if (g.d > wz) return d4of(v(g.x, g.y, 0), g.d);
        return d4of(v(0, 0, sz), wz);
    }
    case SH_ASPHERE: {      /* P: rho, ht, 7 per face (the emitter's), then the faces' radii R1, R2 */
        double rho = P[0], ht = P[1], l = sqrt(p.x * p.x + p.y * p.y);
        double rx = l > 1e-12 ? p.x / l : 1, ry = l > 1e-12 ? p.y / l : 0;
        d4 n = d4of(v(rx, ry, 0), l - rho);
        d4 a = asphFaceD(p, l, rx, ry, rho, ht, P[16], P[3], P + 4, 1);
        if (a.w > n.w) n = a;
        d4 b = asphFaceD(p, l, rx, ry, rho, ht, P[17], P[10], P + 11, -1);
        if (b.w > n.w) n = b;
        return n;
    }
    }
    return d4of(v(0, 0, 1), 1e30);
}

/* a node's frame: U[xf..xf+12] = position, rotation columns, scale (hipgen.js xfParams) */
static d4 primWorld(const OrScene* S, const OrOp* o, d3 p) {
    const float* t = S->U + o->xf;
    d3 c0 = v(t[3], t[4], t[5]), c1 = v(t[6], t[7], t[8]), c2 = v(t[9], t[10], t[11]);
    d3 d = sub(p, v(t[0], t[1], t[2]));
    double s = t[12];
    d3 q = v(dotp(d, c0) / s, dotp(d, c1) / s, dotp(d, c2) / s);
    d4 l = primLocal(o->shape, q, S->U + o->p);
    d3 n = add(add(scl(c0, l.x), scl(c1, l.y)), scl(c2, l.z));
    return d4of(n, l.w * s);
}

/* CSG joins (core/scene_api.js:352-373): polynomial smooth min/max for k > 0, the normal mixed */
static d3 csgN(d3 a, d3 b, double t) {
    if (dotp(a, a) < 1e-12 || dotp(b, b) < 1e-12) return v(0, 0, 0);
    d3 n = add(a, scl(sub(b, a), t));
    double l = len(n);
    return l > 1e-12 ? scl(n, 1 / l) : v(0, 0, 0);
}
static d4 joinU(d4 a, d4 b, double k) {
    if (k <= 0) return a.w < b.w ? a : b;
    double h = maxd(k - fabs(a.w - b.w), 0), m = a.w < b.w ? 0 : 1, t = m + (1 - 2 * m) * 0.5 * h / k;
    return d4of(csgN(v(a.x, a.y, a.z), v(b.x, b.y, b.z), t), mind(a.w, b.w) - h * h * 0.25 / k);
}
static d4 joinI(d4 a, d4 b, double k) {
    if (k <= 0) return a.w > b.w ? a : b;
    double h = maxd(k - fabs(a.w - b.w), 0), m = a.w > b.w ? 0 : 1, t = m + (1 - 2 * m) * 0.5 * h / k;
    return d4of(csgN(v(a.x, a.y, a.z), v(b.x, b.y, b.z), t), maxd(a.w, b.w) + h * h * 0.25 / k);
}

/* Shaped joins (core/scene_api.js:393-406, model version 28 joinShape): hg_sdf's round and chamfer intersections of
 * radius r, their value held at most the hard join's h plus csgShapedCap -- the most the form adds where h <= 0
 * ((sqrt 2 - 1) r round, r / sqrt 2 chamfer) faded linearly to 0 at h = 2r. Unions and cuts through complements. */
static double shapedCap(double h, double r, double most) {
    double f = 1 - 0.5 * h / maxd(r, 1e-30);
    return most * r * (f < 0 ? 0 : f > 1 ? 1 : f);
}
static d4 shapedOut(d4 a, d4 b, d4 hard, double d, double cap, d3 g) {
    int both = dotp(v(a.x, a.y, a.z), v(a.x, a.y, a.z)) >= 1e-12 && dotp(v(b.x, b.y, b.z), v(b.x, b.y, b.z)) >= 1e-12;
    if (d - hard.w >= cap) return both ? d4of(v(hard.x, hard.y, hard.z), hard.w + cap) : d4of(v(0, 0, 0), hard.w + cap);
    double l = len(g);
    return d4of(both && l > 1e-12 ? scl(g, 1 / l) : v(0, 0, 0), d);
}
static d4 roundI(d4 a, d4 b, double r) {
    d4 hard = a.w > b.w ? a : b;
    double A = a.w + r, B = b.w + r;
    if (!(A > 0 && B > 0)) return hard;
    return shapedOut(a, b, hard, sqrt(A * A + B * B) - r, shapedCap(hard.w, r, 0.41421356237309515),
                     add(scl(v(a.x, a.y, a.z), A), scl(v(b.x, b.y, b.z), B)));
}
static d4 chamferI(d4 a, d4 b, double r) {
    d4 hard = a.w > b.w ? a : b;
    double d = (a.w + b.w + r) * 0.7071067811865476;
    if (d <= hard.w) return hard;
    return shapedOut(a, b, hard, d, shapedCap(hard.w, r, 0.7071067811865476), add(v(a.x, a.y, a.z), v(b.x, b.y, b.z)));
}
static d4 neg4(d4 a) { d4 r = {-a.x, -a.y, -a.z, -a.w}; return r; }
/* one join: op J_*, shape 0 smooth / 1 round / 2 chamfer, k its blend or radius */
static d4 joinAny(d4 a, d4 b, int op, int shape, double k) {
    if (shape == 0) {
        if (op == J_UNION) return joinU(a, b, k);
        return joinI(a, op == J_SUBTRACT ? neg4(b) : b, k);
    }
    d4 (*I)(d4, d4, double) = shape == 1 ? roundI : chamferI;
    if (op == J_UNION) return neg4(I(neg4(a), neg4(b), k));
    return I(a, op == J_SUBTRACT ? neg4(b) : b, k);
}

/* THE MODIFIER GROUPS (builder/modifiers.js twist, bend, repeat, mirror, circular, round, shell): the group's
 * children, ops [i + 1, end), evaluated at the group-frame point the modifier makes, joined as they say, then the
 * modifier's change to the value and the group's frame back to its parent's. A twist's or bend's value is the children's
 * times its Lipschitz bound 1 / sqrt(1 + (rate extent)^2) (U[p + 1]), with no normal (differences); a mirror's likewise
 * without a factor; the circular array is the least of its copies, each's normal turned with it. */
static d4 evalRange(const OrScene* S, int lo, int hi, d3 p, int* mat, int* prim);

/* THE SCENE'S OWN CODE (an object's or a group's WGSL), compiled in FP64 by the tests from the emitter's translation
 * against oracle/wgsl_host.h: by index, an object's function (local point, metres -> normal, distance, material |
 * primitive << 8) or a group's (which asks the oracle for its children's values at a point through the callback). */
typedef void (*OrUserObj)(const float* U, const double* q, double* out, unsigned int* m);
typedef void (*OrUserGroup)(const float* U, const double* q, void* self, int (*kids)(void*, const double*, double*, unsigned int*),
                            double* out, unsigned int* m);
static void* g_user[64];
static int g_nuser = 0;
void or_set_user(void* const* fns, int n) {
    g_nuser = n < 64 ? n : 64;
    for (int i = 0; i < g_nuser; ++i) g_user[i] = fns[i];
}
/* Beer-Lambert through distance d inside an absorbing material (materials.js mediumMu): exp(-mu d) per channel,
 * mu = -ln(base held to [1e-6, 1]) / attenuation */
static d3 mediumTrD(const PtMaterial* m, double d) {
    if (!(m->attenuation > 0)) return v(1, 1, 1);
    double t[3];
    for (int k = 0; k < 3; ++k) t[k] = exp(log(clampd(m->base[k], 1e-6, 1)) * d / m->attenuation);
    return v(t[0], t[1], t[2]);
}
/* the medium after a crossing: into the surface's material, or out to the one beyond (absorbing ones only), + 1 */
static int mediumAfter(const OrScene* S, int inside, int hitMat, int beyond) {
    const int into = inside ? beyond : hitMat;
    return (into >= 0 && into < S->nmats && S->mats[into].attenuation > 0) ? into + 1 : 0;
}
/* the materials' own code: (U, material, hit point, outward normal, the hit's primitive, the material to change) */
typedef void (*OrUserMat)(const float* U, unsigned int mat, const double* p, const double* n, int prim, PtMaterial* m);
/* a material's relief: (U, the object's point in real millimetres, the scene's point in millimetres) -> height, slope */
typedef void (*OrRelief)(const float* U, const double* q, const double* world, double* out);
static void* g_relief[64];
static int g_nrelief = 0;
void or_set_reliefs(void* const* fns, int n) {
    g_nrelief = n < 64 ? n : 64;
    for (int i = 0; i < g_nrelief; ++i) g_relief[i] = fns[i];
}
/* the scene's point the evaluation started from (a relief reads it), per thread */
static _Thread_local d3 g_world;
static OrUserMat g_umat = 0;
void or_set_material(void* fn) { g_umat = (OrUserMat)fn; }
/* the oracle's primitives for the scene code's shape functions: shape index, local point, parameters, (normal, d) */
void or_prim_local(int shape, const double* p, const float* P, double* out) {
    d4 r = primLocal(shape, v(p[0], p[1], p[2]), P);
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
static d3 frameIn(const float* t, d3 pw) {
    d3 c0 = v(t[3], t[4], t[5]), c1 = v(t[6], t[7], t[8]), c2 = v(t[9], t[10], t[11]);
    d3 d = sub(pw, v(t[0], t[1], t[2]));
    return v(dotp(d, c0) / t[12], dotp(d, c1) / t[12], dotp(d, c2) / t[12]);
}
static d4 frameOut(const float* t, const double* l) {
    d3 c0 = v(t[3], t[4], t[5]), c1 = v(t[6], t[7], t[8]), c2 = v(t[9], t[10], t[11]);
    return d4of(add(add(scl(c0, l[0]), scl(c1, l[1])), scl(c2, l[2])), l[3] * t[12]);
}
static int evalOp(const OrScene* S, int i, d3 p, d4* st, int* sm, int* sk, int* top);
typedef struct { const OrScene* S; int lo, hi; } OrKids;
/* a coded group's children at q (the group's frame, metres): each child's value, in order */
static int orKids(void* self, const double* q, double* vals, unsigned int* ms) {
    const OrKids* K = (const OrKids*)self;
    d3 p = v(q[0], q[1], q[2]);
    d4 st[64]; int sm[64], sk[64]; int top = 0, n = 0;
    for (int i = K->lo; i < K->hi; ++i) {
        if (K->S->ops[i].op == OP_UFOLD) {
            d4 r = st[--top];
            vals[4 * n] = r.x; vals[4 * n + 1] = r.y; vals[4 * n + 2] = r.z; vals[4 * n + 3] = r.w;
            ms[n] = (unsigned int)sm[top] | ((unsigned int)sk[top] << 8);
            if (++n == 64) break;
        } else {
            i = evalOp(K->S, i, p, st, sm, sk, &top);
        }
    }
    return n;
}
static d4 userEval(const OrScene* S, int i, d3 pw, int* mat, int* prim) {
    const OrOp* o = &S->ops[i];
    const float* t = S->U + o->xf;
    d3 p = frameIn(t, pw);
    double q[3] = {p.x, p.y, p.z}, out[4];
    unsigned int m = 0;
    if (o->shape >= g_nuser || !g_user[o->shape]) { out[0] = 0; out[1] = 0; out[2] = 1; out[3] = NAN; }
    else if (o->op == OP_USER) ((OrUserObj)g_user[o->shape])(S->U, q, out, &m);
    else {
        OrKids K = {S, i + 1, o->k};
        ((OrUserGroup)g_user[o->shape])(S->U, q, &K, orKids, out, &m);
    }
    *mat = (int)(m & 0xFFu); *prim = (int)(m >> 8);
    return frameOut(t, out);
}
static double repeat1(double x, double s, double n) {
    if (s == 0) return x;
    double k = nearbyint(x / s);      /* round half to even, as the device's rintf */
    return x - s * (k < -n ? -n : k > n ? n : k);
}
static d4 modEval(const OrScene* S, int i, d3 pw, int* mat, int* prim) {
    const OrOp* o = &S->ops[i];
    const float* t = S->U + o->xf;
    const float* P = S->U + o->p;
    d3 c0 = v(t[3], t[4], t[5]), c1 = v(t[6], t[7], t[8]), c2 = v(t[9], t[10], t[11]);
    d3 d = sub(pw, v(t[0], t[1], t[2]));
    double sc = t[12];
    d3 p = v(dotp(d, c0) / sc, dotp(d, c1) / sc, dotp(d, c2) / sc);
    d4 r;
    switch (o->shape) {
    case MOD_TWIST: case MOD_BEND: {
        double a = P[0] * (o->shape == MOD_TWIST ? p.z : p.x), ca = cos(a), sa = sin(a);
        d3 q = o->shape == MOD_TWIST ? v(ca * p.x - sa * p.y, sa * p.x + ca * p.y, p.z) : v(ca * p.x - sa * p.z, p.y, sa * p.x + ca * p.z);
        d4 u = evalRange(S, i + 1, o->k, q, mat, prim);
        r = d4of(v(0, 0, 0), u.w * P[1]);
        break;
    }
    case MOD_REPEAT: {
        d3 q = v(repeat1(p.x, P[0], P[3]), repeat1(p.y, P[1], P[4]), repeat1(p.z, P[2], P[5]));
        r = evalRange(S, i + 1, o->k, q, mat, prim);
        break;
    }
    case MOD_MIRROR: {
        d4 u = evalRange(S, i + 1, o->k, v(fabs(p.x), p.y, p.z), mat, prim);
        r = d4of(v(0, 0, 0), u.w);
        break;
    }
    case MOD_CIRCULAR: {
        int n = (int)P[0];
        r = d4of(v(0, 0, 0), 1e300);
        for (int c = 0; c < n; ++c) {
            double a = c * (double)P[1], ca = cos(a), sa = sin(a);
            int m, k;
            d4 u = evalRange(S, i + 1, o->k, v(ca * p.x + sa * p.y, ca * p.y - sa * p.x, p.z), &m, &k);
            if (u.w < r.w) { r = d4of(v(ca * u.x - sa * u.y, sa * u.x + ca * u.y, u.z), u.w); *mat = m; *prim = k; }
        }
        break;
    }
    case MOD_ROUND: {
        d4 u = evalRange(S, i + 1, o->k, p, mat, prim);
        r = d4of(v(u.x, u.y, u.z), u.w - P[0]);
        break;
    }
    default: {
        d4 u = evalRange(S, i + 1, o->k, p, mat, prim);
        double sg = u.w < 0 ? -1 : 1;
        r = d4of(v(u.x * sg, u.y * sg, u.z * sg), fabs(u.w) - P[0]);
        break;
    }
    }
    d3 n = add(add(scl(c0, r.x), scl(c1, r.y)), scl(c2, r.z));
    return d4of(n, r.w * sc);
}
/* one value op at i (a prim, a modifier or coded group, a join) on the stack; returns the index it ended at */
static int evalOp(const OrScene* S, int i, d3 p, d4* st, int* sm, int* sk, int* top) {
    const OrOp* o = &S->ops[i];
    if (o->op == OP_PRIM) { st[*top] = primWorld(S, o, p); sm[*top] = o->mat; sk[*top] = o->prim; ++*top; }
    else if (o->op == OP_MOD) { st[*top] = modEval(S, i, p, &sm[*top], &sk[*top]); ++*top; i = o->k; }
    else if (o->op == OP_USER) { st[*top] = userEval(S, i, p, &sm[*top], &sk[*top]); ++*top; }
    else if (o->op == OP_UGROUP) { st[*top] = userEval(S, i, p, &sm[*top], &sk[*top]); ++*top; i = o->k; }
    else if (o->op == OP_RELIEF) {
        /* the relief of the object just evaluated (builder/codegen.js:1735-1760): beyond two amplitudes its distance moved
         * towards zero by one, within them less the height at its point and divided by one plus the slope bound; no normal */
        d4* v = &st[*top - 1];
        const double a = S->U[o->p];
        if (fabs(v->w) > 2 * a) v->w -= v->w < 0 ? -a : a;
        else {
            const float* t = S->U + o->xf;
            d3 q = scl(frameIn(t, p), t[12] * 1000.0);
            double qa[3] = {q.x, q.y, q.z}, wa[3] = {g_world.x * 1000.0, g_world.y * 1000.0, g_world.z * 1000.0}, hs[2] = {0, 0};
            if (o->shape < g_nrelief && g_relief[o->shape]) ((OrRelief)g_relief[o->shape])(S->U, qa, wa, hs);
            else hs[0] = NAN;
            v->w = (v->w - 0.001 * hs[0]) / (1 + hs[1]);
        }
        v->x = v->y = v->z = 0;
    }
    else if (o->op == OP_JOIN) {
        int t = --*top;
        d4 b = st[t], a = st[t - 1];
        double k = S->U[o->k];
        int wins = o->join == J_UNION ? b.w < a.w : o->join == J_SUBTRACT ? -b.w > a.w : b.w > a.w;
        if (wins) { sm[t - 1] = sm[t]; sk[t - 1] = sk[t]; }
        st[t - 1] = joinAny(a, b, o->join, o->shape, k);
    }
    return i;
}
/* a node's value ops in [lo, hi) (prims, joins, modifier and coded groups) on a stack; the value, its material and
 * primitive */
static d4 evalRange(const OrScene* S, int lo, int hi, d3 p, int* mat, int* prim) {
    d4 st[64]; int sm[64], sk[64]; int top = 0;
    for (int i = lo; i < hi; ++i) i = evalOp(S, i, p, st, sm, sk, &top);
    *mat = top ? sm[0] : 0; *prim = top ? sk[0] : -1;
    return top ? st[0] : d4of(v(0, 0, 1), 1e30);
}

typedef struct { d3 n; double d; int m, a, prim; double insideW; } OrHit;
/* a hit's material, its code run on it when the scene has some */
static const PtMaterial* hitMaterial(const OrScene* S, const OrHit* h, d3 p, d3 ns, PtMaterial* own) {
    const int k = h->m < S->nmats ? h->m : 0;
    if (!g_umat) return &S->mats[k];
    *own = S->mats[k];
    const double pw[3] = {p.x, p.y, p.z}, nw[3] = {ns.x, ns.y, ns.z};
    g_umat(S->U, (unsigned int)k, pw, nw, h->prim, own);
    return own;
}

/* The closest-surface fold over the program (scene_api.js:39-63 fold; codegen.js:1821-1828 joins). A top-level node
 * is skipped when the ray's mask leaves it out, or when its bounding sphere cannot beat the closest |d| so far (which
 * changes nothing: such a node is neither the closest surface nor one p is inside). */
static OrHit sceneAtM(const OrScene* S, d3 p, uint64_t mask) {
    d4 st[64]; int sm[64], sk[64]; int top = 0;
    g_world = p;
    OrHit h; h.n = v(0, 0, 1); h.d = 1e7; h.m = 0; h.a = 0; h.prim = -1; h.insideW = -1e7;
    for (int i = 0; i < S->nops; ++i) {
        const OrOp* o = &S->ops[i];
        if (o->op == OP_NODE) {
            int skip = o->join >= 0 && !((mask >> o->join) & 1ull);
            if (!skip && o->xf >= 0) {
                const float* b = S->U + o->xf;
                skip = len(sub(p, v(b[0], b[1], b[2]))) - b[3] >= fabs(h.d);
            }
            if (skip) i = o->k;
        } else if (o->op == OP_PRIM || o->op == OP_MOD || o->op == OP_USER || o->op == OP_UGROUP || o->op == OP_JOIN || o->op == OP_RELIEF) {
            i = evalOp(S, i, p, st, sm, sk, &top);
        } else if (o->op == OP_MODEND) {
        } else {
            d4 s = st[--top];
            if (s.w < 0 && s.w > h.insideW) { h.insideW = s.w; h.a = sm[top]; }
            if (fabs(s.w) < fabs(h.d)) { h.n = v(s.x, s.y, s.z); h.d = s.w; h.m = sm[top]; h.prim = sk[top]; }
        }
    }
    return h;
}
static OrHit sceneAt(const OrScene* S, d3 p) { return sceneAtM(S, p, ~0ull); }
/* The nodes a ray from o along d can reach within tmax: those whose bounding sphere, widened by margin, its segment
 * meets. A node left out can neither be hit nor stop the ray (it never comes within a tolerance of it); leaving it
 * out only lengthens the march's steps. */
static uint64_t sceneMask(const OrScene* S, d3 o, d3 d, double tmax, double margin) {
    uint64_t m = ~0ull;
    for (int i = 0; i < S->nops; ++i) {
        const OrOp* op = &S->ops[i];
        if (op->op != OP_NODE || op->join < 0 || op->xf < 0) continue;
        const float* b = S->U + op->xf;
        d3 oc = sub(v(b[0], b[1], b[2]), o);
        double t = clampd(dotp(oc, d), 0, tmax);
        d3 q = sub(oc, scl(d, t));
        double rr = b[3] + margin;
        if (dotp(q, q) > rr * rr) m &= ~(1ull << op->join);
    }
    return m;
}
static double maskMargin(const OrScene* S) {
    return maxd(S->u->eps > 0 ? S->u->eps : S->u->epsAuto, 1e-9) + 128.0 * pow(2.0, -23) * (S->u->envRadius + 1.0);
}
static double sceneDist(const OrScene* S, d3 p) { return sceneAt(S, p).d; }
static double sceneDistM(const OrScene* S, d3 p, uint64_t mask) { return sceneAtM(S, p, mask).d; }

/* THE HIT TOLERANCE (core/tolerances.js:10-34) */
static double tolSetting(const OrScene* S) { return S->u->eps > 0 ? S->u->eps : S->u->epsAuto; }
static double marchTol(const OrScene* S, d3 p) {
    return maxd(maxd(tolSetting(S), 1e-9), 128.0 * pow(2.0, -23) * maxd(fabs(p.x), maxd(fabs(p.y), fabs(p.z))));
}
static double offsetTol(const OrScene* S, d3 p) { return 10.0 * marchTol(S, p); }

/* ---- GGX, Fresnel, albedo tables (core/microfacet.js:72-140, :245-330) ---- */
#define ALPHA_MIN 1e-3
static double mfLambda(d3 w, double a) { double z2 = maxd(w.z * w.z, 1e-12); return 0.5 * (-1 + sqrt(1 + a * a * (w.x * w.x + w.y * w.y) / z2)); }
static double mfG1(d3 w, double a) { return 1 / (1 + mfLambda(w, a)); }
static double mfG2(d3 o, d3 i, double a) { return 1 / (1 + mfLambda(o, a) + mfLambda(i, a)); }
static double mfD(d3 h, double a) {
    if (h.z <= 0) return 0;
    double t = (h.x * h.x + h.y * h.y) / (a * a) + h.z * h.z;
    return 1 / (PI_D * a * a * t * t);
}
static d3 mfVNDF(d3 o, double a, double u1, double u2) {
    d3 os = unit(v(o.x * a, o.y * a, o.z));
    double phi = 2 * PI_D * u1, z = (1 - u2) * (1 + os.z) - os.z, s = sqrt(clampd(1 - z * z, 0, 1));
    d3 hs = add(v(s * cos(phi), s * sin(phi), z), os);
    return unit(v(hs.x * a, hs.y * a, maxd(hs.z, 1e-7)));
}
static double mfPdfR(d3 o, d3 i, double a) {
    if (o.z <= 0 || i.z <= 0) return 0;
    return mfG1(o, a) * mfD(unit(add(o, i)), a) / (4 * o.z);
}
double or_fresnel(double cosI, double eta) {
    cosI = clampd(cosI, 0, 1);
    double s2 = eta * eta * (1 - cosI * cosI);
    if (s2 >= 1) return 1;
    double ct = sqrt(1 - s2), rs = (eta * cosI - ct) / (eta * cosI + ct), rp = (cosI - eta * ct) / (cosI + eta * ct);
    return 0.5 * (rs * rs + rp * rp);
}
static d3 schlick3(double c, d3 f0) { double m = pow(1 - clampd(c, 0, 1), 5); return add(f0, scl(sub(v(1, 1, 1), f0), m)); }
static d3 reflL(d3 o, d3 h) { return sub(scl(h, 2 * dotp(o, h)), o); }

static double g_mf[PT_MF_WORDS];
static int g_mfReady = 0;
/* The tables by stratified visible-normal quadrature (64 x 64 per entry), FP64. */
static double lutAt(d3 o, double a, int e1, double nrel, double u1, double u2, int* above) {
    d3 h = mfVNDF(o, a, u1, u2), i = reflL(o, h);
    *above = i.z > 0;
    if (!*above) return 0;
    return (e1 ? 1 : or_fresnel(dotp(o, h), 1 / nrel)) * mfG2(o, i, a) / mfG1(o, a);
}
/* the tables as computed before (a cache the Python side keeps on disk: they take seconds) */
void or_mf_load(const double* in) { memcpy(g_mf, in, sizeof g_mf); g_mfReady = 1; }
int or_mf_words(void) { return (int)(sizeof g_mf / sizeof g_mf[0]); }
void or_mf_tables(double* out) {
    if (!g_mfReady) {
        #pragma omp parallel for schedule(dynamic)
        for (int t = 0; t < PT_MF_NIOR * PT_MF_NR * PT_MF_NMU + PT_MF_NR * PT_MF_NMU; ++t) {
            int e1 = t >= PT_MF_NIOR * PT_MF_NR * PT_MF_NMU;
            int k = e1 ? t - PT_MF_NIOR * PT_MF_NR * PT_MF_NMU : t;
            int im = k % PT_MF_NMU, ir = (k / PT_MF_NMU) % PT_MF_NR, ii = e1 ? 0 : k / (PT_MF_NMU * PT_MF_NR);
            double mu = (im + 0.5) / PT_MF_NMU, r = (double)ir / (PT_MF_NR - 1), nrel = 1 + 2.0 * ii / (PT_MF_NIOR - 1);
            double a = clampd(r * r, ALPHA_MIN, 1);
            d3 o = v(sqrt(maxd(0, 1 - mu * mu)), 0, mu);
            /* the mean over the visible-normal square of F G2/G1 above the horizon: midpoints on 128 x 128, a cell
             * the horizon's cut crosses (a corner disagreeing with the centre) taken again as 8 x 8 midpoints */
            const int N = 128, S = 8;
            double acc = 0;
            for (int ca = 0; ca < N; ++ca) for (int cb = 0; cb < N; ++cb) {
                int above, c, cut = 0;
                double mid = lutAt(o, a, e1, nrel, (ca + 0.5) / N, (cb + 0.5) / N, &above);
                for (int k = 0; k < 4; ++k) { lutAt(o, a, e1, nrel, (double)(ca + (k & 1)) / N, (double)(cb + (k >> 1)) / N, &c); cut |= c != above; }
                if (!cut) { acc += mid; continue; }
                double sub = 0;
                for (int sa = 0; sa < S; ++sa) for (int sb = 0; sb < S; ++sb)
                    sub += lutAt(o, a, e1, nrel, (ca + (sa + 0.5) / S) / N, (cb + (sb + 0.5) / S) / N, &c);
                acc += sub / (S * S);
            }
            acc /= (double)N * N;
            if (e1) g_mf[PT_MF_E1 + ir * PT_MF_NMU + im] = acc;
            else g_mf[PT_MF_EC + (ii * PT_MF_NR + ir) * PT_MF_NMU + im] = acc;
        }
        for (int t = 0; t < PT_MF_NIOR * PT_MF_NR; ++t) {
            double acc = 0;
            for (int im = 0; im < PT_MF_NMU; ++im) acc += g_mf[PT_MF_EC + t * PT_MF_NMU + im] * (im + 0.5) / PT_MF_NMU;
            g_mf[PT_MF_ECAVG + t] = 2 * acc / PT_MF_NMU;
        }
        g_mfReady = 1;
    }
    if (out) memcpy(out, g_mf, sizeof g_mf);
}
static void axis(double x, int n, int* i0, double* f) {
    double c = clampd(x, 0, n - 1), fl = mind(floor(c), n - 2);
    *i0 = (int)fl; *f = c - fl;
}
static double lutEc(double mu, double r, double nrel) {
    int im, ir, in; double fm, fr, fn;
    axis(mu * PT_MF_NMU - 0.5, PT_MF_NMU, &im, &fm); axis(r * (PT_MF_NR - 1), PT_MF_NR, &ir, &fr);
    axis((nrel - 1) * (PT_MF_NIOR - 1) * 0.5, PT_MF_NIOR, &in, &fn);
    double acc = 0;
    for (int c = 0; c < 8; ++c) {
        int di = c & 1, dk = (c >> 1) & 1, dj = (c >> 2) & 1;
        acc += (di ? fn : 1 - fn) * (dk ? fr : 1 - fr) * (dj ? fm : 1 - fm) * g_mf[PT_MF_EC + ((in + di) * PT_MF_NR + ir + dk) * PT_MF_NMU + im + dj];
    }
    return acc;
}
static double lutEcAvg(double r, double nrel) {
    int ir, in; double fr, fn;
    axis(r * (PT_MF_NR - 1), PT_MF_NR, &ir, &fr); axis((nrel - 1) * (PT_MF_NIOR - 1) * 0.5, PT_MF_NIOR, &in, &fn);
    double acc = 0;
    for (int c = 0; c < 4; ++c) { int di = c & 1, dk = (c >> 1) & 1; acc += (di ? fn : 1 - fn) * (dk ? fr : 1 - fr) * g_mf[PT_MF_ECAVG + (in + di) * PT_MF_NR + ir + dk]; }
    return acc;
}
static double lutE1(double mu, double r) {
    int im, ir; double fm, fr;
    axis(mu * PT_MF_NMU - 0.5, PT_MF_NMU, &im, &fm); axis(r * (PT_MF_NR - 1), PT_MF_NR, &ir, &fr);
    double acc = 0;
    for (int c = 0; c < 4; ++c) { int dk = c & 1, dj = (c >> 1) & 1; acc += (dk ? fr : 1 - fr) * (dj ? fm : 1 - fm) * g_mf[PT_MF_E1 + (ir + dk) * PT_MF_NMU + im + dj]; }
    return acc;
}

/* ---- the layered BSDF (core/materials_pbr.js:74-553) ---- */
typedef struct {
    d3 tX, tY, nZ, o, base, tint, cm;
    double a, r, metallic, trans, coatK, eta, nrel, ps, ecO, ecAvg;
    int isDelta, through;
} OrLobes;
static void basis(d3 n, d3* t, d3* b) {
    /* any orthonormal frame will do for isotropic lobes; Frisvad's, restated */
    if (n.z < -0.9999999) { *t = v(0, -1, 0); *b = v(-1, 0, 0); return; }
    double a = 1 / (1 + n.z), c = -n.x * n.y * a;
    *t = v(1 - n.x * n.x * a, c, -n.x);
    *b = v(c, 1 - n.y * n.y * a, -n.y);
}
static double coatF(const OrLobes* L, double c) { return mind(L->coatK * or_fresnel(c, L->eta), 1); }
static double coatAlb(const OrLobes* L, double mu) { return mind(L->coatK * lutEc(mu, L->r, L->nrel), 1); }
static OrLobes makeLobes(d3 n, d3 view, const PtMaterial* m, double eta) {
    OrLobes L;
    basis(n, &L.tX, &L.tY); L.nZ = n;
    L.o = v(dotp(view, L.tX), dotp(view, L.tY), maxd(dotp(view, n), 1e-4));
    L.r = clampd(m->roughness, 0, 1);
    L.isDelta = L.r * L.r < ALPHA_MIN;
    L.a = clampd(L.r * L.r, ALPHA_MIN, 1);
    L.metallic = clampd(m->metallic, 0, 1);
    L.trans = clampd(m->transmission, 0, 1);
    L.base = f3(m->base); L.tint = m->attenuation > 0 ? v(1, 1, 1) : L.base;   /* an absorbing body tints along the way, not at its faces */
    L.coatK = 2 * m->specular + (1 - 2 * m->specular) * L.trans;
    L.eta = eta; L.nrel = clampd(1 / eta, 1, 3);
    double fo = coatF(&L, L.o.z);
    L.ps = mind(fo, or_fresnel(L.o.z, L.eta) >= 1 ? 1 : 0.999);
    L.ecO = coatAlb(&L, L.o.z);
    L.ecAvg = mind(L.coatK * lutEcAvg(L.r, L.nrel), 1);
    double e1 = L.metallic > 0 ? lutE1(L.o.z, L.r) : 1;
    L.cm = add(v(1, 1, 1), scl(L.base, 1 / maxd(e1, 1e-3) - 1));
    L.through = 0;
    return L;
}
static d3 toLoc(const OrLobes* L, d3 w) { return v(dotp(w, L->tX), dotp(w, L->tY), dotp(w, L->nZ)); }
static d3 toWld(const OrLobes* L, d3 w) { return unit(add(add(scl(L->tX, w.x), scl(L->tY, w.y)), scl(L->nZ, w.z))); }
static double reflP(const OrLobes* L, double oh);
/* The rough interface's transmission below the surface (materials_pbr.js:296-357 lobesThroughEval/Pdf; Walter et al.
 * 2007), read only where the caller sets `through`: the generalised half vector h = -(eta o + i), turned to the view's
 * side, refracting only with o.h > 0 > i.h; f = (1-m) T (1-F(o.h)) tint D G2 (o.h)|i.h| / (o.z |i.z| (eta o.h + i.h)^2),
 * in basic radiance; its density the sampler's picks times the visible-normal density times |i.h| / (eta o.h + i.h)^2. */
static int throughHalf(const OrLobes* L, d3 i, d3* h, double* oh, double* ih, double* jac) {
    d3 g = neg(add(scl(L->o, L->eta), i));
    double gl = len(g);
    if (!(gl > 1e-12)) return 0;
    g = scl(g, 1 / gl);
    if (g.z < 0) g = neg(g);
    *h = g; *oh = dotp(L->o, g); *ih = dotp(i, g);
    if (*oh <= 0 || *ih >= 0) return 0;
    double den = L->eta * *oh + *ih;
    *jac = -*ih / maxd(den * den, 1e-300);
    return 1;
}
static d3 lobesEvalThrough(const OrLobes* L, d3 i) {
    double wd = 1 - L->metallic, oh, ih, jac;
    d3 h;
    if (wd <= 0 || L->trans <= 0 || L->isDelta || !throughHalf(L, i, &h, &oh, &ih, &jac)) return v(0, 0, 0);
    return scl(L->tint, wd * L->trans * (1 - coatF(L, oh)) * mfD(h, L->a) * mfG2(L->o, i, L->a) * oh * jac / (L->o.z * -i.z));
}
static double lobesPdfThrough(const OrLobes* L, d3 i) {
    double wd = 1 - L->metallic, oh, ih, jac, pd = (1 - L->ps) * (1 - L->trans);
    d3 h;
    if (wd <= 0 || L->trans <= 0 || L->isDelta || !throughHalf(L, i, &h, &oh, &ih, &jac)) return 0;
    return wd * (1 - pd) * (1 - reflP(L, oh)) * mfG1(L->o, L->a) * mfD(h, L->a) * oh / L->o.z * jac;
}
static d3 lobesEval(const OrLobes* L, d3 l) {
    d3 i = toLoc(L, l);
    if (i.z < 0 && L->through) return lobesEvalThrough(L, i);
    if (i.z <= 0) return v(0, 0, 0);
    d3 f = v(0, 0, 0);
    if (!L->isDelta) {
        d3 h = unit(add(L->o, i));
        double spec = mfD(h, L->a) * mfG2(L->o, i, L->a) / (4 * L->o.z * i.z), oh = dotp(L->o, h);
        if (L->metallic > 0) f = add(f, scl(mul(L->cm, schlick3(oh, L->base)), L->metallic * spec));
        double c = (1 - L->metallic) * coatF(L, oh) * spec;
        f = add(f, v(c, c, c));
    }
    double wd = (1 - L->metallic) * (1 - L->trans);
    if (wd > 0) f = add(f, scl(L->base, wd * (1 - L->ecO) * (1 - coatAlb(L, i.z)) / (PI_D * maxd(1 - L->ecAvg, 1e-4))));
    return f;
}
static double reflP(const OrLobes* L, double oh) { double fh = coatF(L, oh), den = fh + (1 - fh) * L->trans; return den > 1e-6 ? fh / den : 1; }
static double lobesPdf(const OrLobes* L, d3 l) {
    d3 i = toLoc(L, l);
    if (i.z < 0 && L->through) return lobesPdfThrough(L, i);
    if (i.z <= 0) return 0;
    double wd = 1 - L->metallic, pd = (1 - L->ps) * (1 - L->trans), p = wd * pd * i.z / PI_D;
    if (!L->isDelta) p += (L->metallic + wd * (1 - pd) * reflP(L, dotp(L->o, unit(add(L->o, i))))) * mfPdfR(L->o, i, L->a);
    return p;
}
static int lobesNone(const PtMaterial* m) {
    double r = clampd(m->roughness, 0, 1);
    return r * r < ALPHA_MIN && (1 - clampd(m->metallic, 0, 1)) * (1 - clampd(m->transmission, 0, 1)) <= 0;
}
typedef struct { d3 dir, w; int delta, transmitted, eval; } OrSample;
static OrSample lobesSample(const OrLobes* L, const double* u) {
    OrSample s; s.w = v(0, 0, 0); s.dir = L->nZ; s.delta = 0; s.transmitted = 0; s.eval = 1;
    d3 o = L->o;
    if (u[0] < L->metallic) {
        d3 h = L->isDelta ? v(0, 0, 1) : mfVNDF(o, L->a, u[1], u[2]);
        d3 i = reflL(o, h);
        if (i.z <= 0) return s;
        double oh = dotp(o, h);
        if (L->isDelta) { s.w = schlick3(oh, L->base); s.delta = 1; s.eval = 0; }
        else s.w = scl(mul(L->cm, schlick3(oh, L->base)), mfG2(o, i, L->a) / mfG1(o, L->a));
        s.dir = toWld(L, i);
        return s;
    }
    double ux = (u[0] - L->metallic) / maxd(1 - L->metallic, 1e-7), pd = (1 - L->ps) * (1 - L->trans);
    if (ux < pd) {
        double r = sqrt(u[1]), ph = 2 * PI_D * u[2];
        d3 i = v(r * cos(ph), r * sin(ph), sqrt(maxd(0, 1 - u[1])));
        s.w = scl(L->base, (1 - L->trans) * (1 - L->ecO) * (1 - coatAlb(L, i.z)) / (maxd(1 - L->ecAvg, 1e-4) * pd));
        s.dir = toWld(L, i);
        return s;
    }
    d3 h = L->isDelta ? v(0, 0, 1) : mfVNDF(o, L->a, u[1], u[2]);
    double oh = dotp(o, h), pR = reflP(L, oh), fh = coatF(L, oh), uw = (ux - pd) / maxd(1 - pd, 1e-7);
    if (uw < pR) {
        d3 i = reflL(o, h);
        if (i.z <= 0) return s;
        double geo = L->isDelta ? 1 : mfG2(o, i, L->a) / mfG1(o, L->a);
        double w = fh * geo / ((1 - pd) * pR);
        s.w = v(w, w, w); s.dir = toWld(L, i); s.delta = L->isDelta; s.eval = !L->isDelta;
        return s;
    }
    double c = dotp(o, h), k = 1 - L->eta * L->eta * (1 - c * c);
    if (k <= 0) return s;
    d3 t = sub(scl(h, L->eta * c - sqrt(k)), scl(o, L->eta));
    if (t.z >= 0) return s;
    double geo = L->isDelta ? 1 : mfG2(o, v(t.x, t.y, -t.z), L->a) / mfG1(o, L->a);
    s.w = scl(L->tint, (1 - fh) * L->trans * geo / ((1 - pd) * (1 - pR)));
    s.dir = toWld(L, t); s.transmitted = 1; s.delta = L->isDelta; s.eval = 0;
    return s;
}

/* ---- spectral (core/spectral.js:121-135) ---- */
double or_cauchy(double nd, double abbe, double lam) {
    double iF = 1 / (0.4861 * 0.4861), iC = 1 / (0.6563 * 0.6563), id = 1 / (0.5876 * 0.5876);
    double B = (nd - 1) / (abbe * (iF - iC)), A = nd - B * id;
    return A + B / (lam * lam);
}
static double wss(double l, double mu, double s1, double s2) { double t = (l - mu) / (l < mu ? s1 : s2); return exp(-0.5 * t * t); }
void or_spectral_weight(double lamUm, double* rgb) {
    double l = lamUm * 1000;
    double x = 1.056 * wss(l, 599.8, 37.9, 31.0) + 0.362 * wss(l, 442.0, 16.0, 26.7) - 0.065 * wss(l, 501.1, 20.4, 26.2);
    double y = 0.821 * wss(l, 568.8, 46.9, 40.5) + 0.286 * wss(l, 530.9, 16.3, 31.1);
    double z = 1.217 * wss(l, 437.0, 11.8, 36.0) + 0.681 * wss(l, 459.0, 26.0, 13.8);
    rgb[0] = maxd(3.2406 * x - 1.5372 * y - 0.4986 * z, 0) / 0.44045557;
    rgb[1] = maxd(-0.9689 * x + 1.8758 * y + 0.0415 * z, 0) / 0.28846547;
    rgb[2] = maxd(0.0557 * x - 0.2040 * y + 1.0570 * z, 0) / 0.27323972;
}

/* ---- environment, sun, light faces ---- */
static int envLit(const PtEnv* e) { return (e->type != 0 && e->L > 0) || e->sunE > 0; }
static double sunOmega(const PtEnv* e) { return e->sunOmega; }
/* in the disc: |d x s|^2 <= sin^2 on the sun's side (sdf.js inSunDisc:88) */
static int inSun(const PtEnv* e, d3 d) { d3 s = f3(e->sunDir), c = crossp(d, s); return e->sunE > 0 && dotp(d, s) > 0 && dotp(c, c) <= e->sunSin2; }

/* ---- the sun & sky (environment.js SKY_BAND_TABLE:66, SUNSKY_FN:110-170), in double: sixteen bands from 380 to 730
 * nm, each with its Rayleigh depth (Hansen & Travis 1974 eq. 2.29), haze factor (Angstrom, 1.3), ozone depth (300 DU of
 * the band's Chappuis cross-section) and its share of 5778 K sunlight in linear Rec.709 (Wyman-Sloan-Shirley's 2013
 * fit to the CIE 1931 functions, summed at 1 nm, normalised to total luminance 1). */
static double skyTauR[16], skyHaze[16], skyOzone[16], skyRgb[16][3];
static int skyReady = 0;
static double wssLobe(double x, double mu, double s1, double s2) { double t = (x - mu) / (x < mu ? s1 : s2); return exp(-0.5 * t * t); }
static void skyTable(void) {
    static const double sigma[16] = {6.2648e-24, 3.1169e-23, 1.0195e-22, 2.4886e-22, 5.4883e-22, 1.1517e-21, 2.0211e-21, 3.0497e-21,
                                     4.2480e-21, 4.5252e-21, 4.7153e-21, 3.4726e-21, 2.3885e-21, 1.5936e-21, 9.8370e-22, 6.3061e-22};
    /* linear Rec.709 from XYZ: the inverse of IEC 61966-2-1's four-decimal matrix */
    double M[3][3] = {{0.4124, 0.3576, 0.1805}, {0.2126, 0.7152, 0.0722}, {0.0193, 0.1192, 0.9505}}, I[3][3];
    double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) + M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) {
        int a = (j + 1) % 3, b = (j + 2) % 3, c = (i + 1) % 3, d = (i + 2) % 3;
        I[i][j] = (M[a][c] * M[b][d] - M[a][d] * M[b][c]) / det;
    }
    double step = (730.0 - 380.0) / 16, tot[3] = {0, 0, 0};
    for (int k = 0; k < 16; ++k) {
        double X = 0, Y = 0, Z = 0, wl = 0, ws = 0;
        for (double l = 380 + k * step; l < 380 + (k + 1) * step; l += 1) {
            double n = floor(l + 0.5), m = l * 1e-9;
            double x = 1.056 * wssLobe(n, 599.8, 37.9, 31.0) + 0.362 * wssLobe(n, 442.0, 16.0, 26.7) - 0.065 * wssLobe(n, 501.1, 20.4, 26.2);
            double y = 0.821 * wssLobe(n, 568.8, 46.9, 40.5) + 0.286 * wssLobe(n, 530.9, 16.3, 31.1);
            double z = 1.217 * wssLobe(n, 437.0, 11.8, 36.0) + 0.681 * wssLobe(n, 459.0, 26.0, 13.8);
            double w = 1 / (pow(m, 5) * (exp(1.438776877e-2 / (m * 5778)) - 1));
            X += w * x; Y += w * y; Z += w * z; wl += w * (x + y + z) * l; ws += w * (x + y + z);
        }
        double nm = wl / ws, um = nm / 1000;
        skyTauR[k] = 0.008569 * pow(um, -4) * (1 + 0.0113 * pow(um, -2) + 0.00013 * pow(um, -4));
        skyHaze[k] = pow(nm / 550, -1.3);
        skyOzone[k] = sigma[k] * 300 * 2.6867e16;
        for (int i = 0; i < 3; ++i) { skyRgb[k][i] = I[i][0] * X + I[i][1] * Y + I[i][2] * Z; tot[i] += skyRgb[k][i]; }
    }
    double Yt = 0.2126 * tot[0] + 0.7152 * tot[1] + 0.0722 * tot[2];
    for (int k = 0; k < 16; ++k) for (int i = 0; i < 3; ++i) skyRgb[k][i] /= Yt;
#pragma omp flush
    skyReady = 1;
}
/* Kasten-Young relative air mass (sdf.js airMass:96) */
static double airMassD(double cz) {
    double c = clampd(cz, 0, 1), z = acos(c) * 180 / PI_D;
    return 1 / (c + 0.15 * pow(maxd(93.885 - z, 0.1), -1.253));
}
static void skyInit(void) {
    if (skyReady) return;
#pragma omp critical(skytable)
    { if (!skyReady) skyTable(); }
}
static d3 sunSkyD(const PtEnv* e, d3 d) {
    skyInit();
    d3 s = f3(e->sunDir);
    double mv = airMassD(d.z), ms = airMassD(s.z), dm = mv - ms, c = dotp(d, s), g = 0.7;
    double pR = 3 / (16 * PI_D) * (1 + c * c), pM = (1 - g * g) / (4 * PI_D * pow(1 + g * g - 2 * g * c, 1.5));
    double up = clampd((s.z + 0.05) / 0.05, 0, 1), aer = e->skyAerosol, L[3] = {0, 0, 0};
    for (int k = 0; k < 16; ++k) {
        double ta = aer * skyHaze[k], tau = skyTauR[k] + ta + skyOzone[k];
        double k1 = fabs(dm) < 1e-3 ? tau * mv * exp(-tau * mv) : mv * (exp(-tau * ms) - exp(-tau * mv)) / dm;
        double w = (skyTauR[k] * pR + 0.9 * ta * pM) / tau * k1 * up;
        for (int i = 0; i < 3; ++i) L[i] += skyRgb[k][i] * w;
    }
    /* col: the sky's gain per channel (the reference's envSky) */
    return v(maxd(e->L * L[0] * e->col[0], 0), maxd(e->L * L[1] * e->col[1], 0), maxd(e->L * L[2] * e->col[2], 0));
}
/* the scene's own sky code (its background(dir)), from the tests' compile of it */
typedef void (*OrSky)(const double* d, double* out);
static OrSky g_sky = 0;
void or_set_sky(void* fn) { g_sky = (OrSky)fn; }
static d3 envBg(const PtEnv* e, d3 d) {
    if (e->type == 5) {
        double in[3] = {d.x, d.y, d.z}, out[3] = {0, 0, 0};
        if (g_sky) g_sky(in, out);
        return v(out[0], out[1], out[2]);
    }
    if (e->type == 1 || (e->type == 2 && d.z > 0)) return scl(f3(e->col), e->L);
    if (e->type == 3) return scl(f3(e->col), e->L * maxd(1 + e->top * d.z, 0));   /* environment.js studioLight:168 */
    if (e->type == 4) return inSun(e, d) ? v(0, 0, 0) : sunSkyD(e, d);
    return v(0, 0, 0);
}
/* the sun through its aperture: the lines crossing the disc of radius sunRadius about sunPoint, perpendicular to the
 * sun (the disc the photon pass emits the beam from) */
static d3 envSun(const PtEnv* e, d3 p, d3 d) {
    if (!inSun(e, d)) return v(0, 0, 0);
    if (e->sunRadius > 0) {
        d3 a = f3(e->sunDir), c = f3(e->sunPoint);
        double t = dotp(sub(c, p), a) / dotp(d, a);
        d3 q = sub(add(p, scl(d, t)), c);
        if (!(dotp(q, q) <= (double)e->sunRadius * e->sunRadius)) return v(0, 0, 0);
    }
    if (e->type == 4) return f3(e->sunL);
    double L = e->sunE / sunOmega(e);
    return v(L, L, L);
}
static double sunPick(const PtEnv* e) { return e->sunE > 0 ? 0.5 : 0; }
static double envPdf(const PtEnv* e, d3 L, d3 n) {
    double ps = sunPick(e);
    return ps * (inSun(e, L) ? 1 / sunOmega(e) : 0) + (1 - ps) * maxd(dotp(L, n), 0) / PI_D;
}
static d3 cosDir(d3 n, double u1, double u2) {
    d3 t, b; basis(n, &t, &b);
    double r = sqrt(u1), ph = 2 * PI_D * u2;
    return unit(add(add(scl(t, r * cos(ph)), scl(b, r * sin(ph))), scl(n, sqrt(maxd(0, 1 - u1)))));
}
static d3 sunDir(const PtEnv* e, double u1, double u2) {
    d3 a = f3(e->sunDir), t, b; basis(a, &t, &b);
    double omc = u1 * e->sunOmega / (2 * PI_D), c = 1 - omc, s = sqrt(maxd(0, omc * (2 - omc))), ph = 2 * PI_D * u2;
    return unit(add(add(scl(t, s * cos(ph)), scl(b, s * sin(ph))), scl(a, c)));
}
static double misW(double a, double b) { a *= a; b *= b; return a / maxd(a + b, 1e-20); }
static double faceArea(const PtLightFace* f) { return f->shape == 1 ? PI_D * f->halfx * f->halfx : 4.0 * f->halfx * f->halfy; }
static double faceShare(const OrScene* S, int bounces) {
    if (bounces == 0 || S->nfaces == 0 || S->faceTotal <= 0) return 0;
    return envLit(&S->u->env) ? S->u->faceShare : 1;
}
static int facePick(const OrScene* S, double u) {
    double acc = 0; int last = 0;
    for (int k = 0; k < S->nfaces; ++k) if (S->faces[k].weight > 0) { last = k; acc += S->faces[k].weight; if (u < acc) return k; }
    return last;
}
static double faceDensity(const OrScene* S, double share, const PtLightFace* f, double dist, double cosY) {
    if (share <= 0 || f->weight <= 0) return 0;
    return share * (f->weight / S->faceTotal) * dist * dist / maxd(faceArea(f) * fabs(cosY), 1e-20);
}

/* ---- haze (core/haze.js:191-290) ---- */
static int hazeOn(const OrScene* S) { return S->u->haze.sigmaS > 0; }
static int hazeSpan(const PtHaze* h, d3 o, d3 d, double* t0, double* t1) {
    double lo = 0, hi = 1e300, oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    for (int k = 0; k < 3; ++k) {
        double c = h->centre[k], e = h->half[k];
        if (dd[k] == 0) { if (oo[k] < c - e || oo[k] > c + e) return 0; continue; }
        double a = (c - e - oo[k]) / dd[k], b = (c + e - oo[k]) / dd[k];
        lo = maxd(lo, mind(a, b)); hi = mind(hi, maxd(a, b));
    }
    *t0 = lo; *t1 = hi;
    return hi > lo;
}
static double hazeTr(const PtHaze* h, d3 o, d3 d, double lenMax) {
    double t0, t1;
    if (!hazeSpan(h, o, d, &t0, &t1)) return 1;
    return exp(-h->sigmaS * maxd(mind(t1, lenMax) - t0, 0));
}
static double hazeDraw(const PtHaze* h, d3 o, d3 d, double u) {
    double t0, t1;
    if (!hazeSpan(h, o, d, &t0, &t1)) return 1e300;
    double t = t0 - log(1 - u) / h->sigmaS;
    return t < t1 ? t : 1e300;
}
static double hg(double g, double c) { double d = 1 + g * g - 2 * g * c; return (1 - g * g) / (4 * PI_D * d * sqrt(d)); }
static d3 hgDir(double g, d3 w, double u1, double u2) {
    double c;
    if (fabs(g) < 1e-3) c = 1 - 2 * u1;
    else { double s = (1 - g * g) / (1 - g + 2 * g * u1); c = (1 + g * g - s * s) / (2 * g); }
    c = clampd(c, -1, 1);
    double sn = sqrt(maxd(0, 1 - c * c)), ph = 2 * PI_D * u2;
    d3 t, b; basis(w, &t, &b);
    return unit(add(add(scl(t, sn * cos(ph)), scl(b, sn * sin(ph))), scl(w, c)));
}

/* ---- march (core/wavefront.js:834-1005): returns the state; moves p, ltot, steps ---- */
enum { ST_ACTIVE = 0, ST_ESCAPED = 1, ST_MAXITER = 2, ST_SCATTER = 4 };
static int march(const OrScene* S, d3* p, d3 dir, double n, double* ltot, uint32_t* steps, double collide) {
    const PtUniforms* u = S->u;
    double R2 = (double)u->envRadius * u->envRadius;
    uint64_t mask = sceneMask(S, *p, dir, 1e300, maskMargin(S));
    for (;;) {
        if (*ltot >= collide) { *p = sub(*p, scl(dir, (*ltot - collide) / n)); *ltot = collide; return ST_SCATTER; }
        double d = sceneDistM(S, *p, mask), ea = fabs(d), eps = marchTol(S, *p);
        int st = ST_ACTIVE;
        if (*ltot > 1000 || (*steps & 0xFFFFu) > u->maxiter) st = ST_MAXITER;
        if (dotp(*p, dir) > 0 && dotp(*p, *p) > R2) st = ST_ESCAPED;
        if (ea < eps || st != ST_ACTIVE) return st;
        *steps += 1;
        double t = ea * u->stepscale;
        *p = add(*p, scl(dir, t));
        *ltot += t * n;
    }
}
/* a shadow ray (core/wavefront.js:1481-1692): 1 reached, 0 blocked */
static int occluded(const OrScene* S, d3 o, d3 dir, double tmax, long long* steps) {
    const PtUniforms* u = S->u;
    double R2 = (double)u->envRadius * u->envRadius, trav = 0;
    d3 p = o;
    uint32_t k = 0;
    uint64_t mask = sceneMask(S, o, dir, tmax, maskMargin(S));
    for (;;) {
        double d = sceneDistM(S, p, mask), ea = fabs(d);
        if (ea < marchTol(S, p)) return 0;
        if (ea >= tmax - trav || (dotp(p, dir) > 0 && dotp(p, p) > R2)) return 1;
        /* cut short by maxiter: unoccluded, as wavefront.js:1640-1645 counts it (occluded would bias dark) */
        if (k > u->maxiter) return 1;
        double t = ea * u->stepscale;
        p = add(p, scl(dir, t)); trav += t; ++k; ++*steps;
    }
}

static d3 normalAt(const OrScene* S, d3 p, double h) {
    double a = sceneDist(S, add(p, v(h, -h, -h))), b = sceneDist(S, add(p, v(-h, -h, h)));
    double c = sceneDist(S, add(p, v(-h, h, -h))), d = sceneDist(S, add(p, v(h, h, h)));
    return unit(v(a - b - c + d, -a - b + c + d, -a + b - c + d));
}

/* ---- camera (core/genrays.js:36-107) ---- */
typedef struct { uint32_t tilesX, tilesY; double invTX, invTY; } OrTiles;
static OrTiles tilesOf(uint32_t N, uint32_t w, uint32_t h) {
    OrTiles t;
    double r = floor(sqrt((double)N * ((double)w / (double)(h ? h : 1))) + 0.5);
    t.tilesX = r < 1 ? 1 : (uint32_t)r;
    t.tilesY = (N + t.tilesX - 1) / t.tilesX;
    t.invTX = 1.0 / t.tilesX; t.invTY = 1.0 / t.tilesY;
    return t;
}
static void cameraRay(const OrScene* S, const OrTiles* T, const uint32_t* seed, uint32_t i, d3* o, d3* d, uint32_t* pix, double* uhaze) {
    const PtCamera* c = &S->u->cam;
    OrRng r = rngAt(seed, i, 0, 0xCA3E8Au);
    double u4[4]; rng4(&r, u4);
    uint32_t tx = i % T->tilesX, ty = i / T->tilesX;
    double uu = clampd((tx + 0.5) * T->invTX + (u4[0] - 0.5) * T->invTX, 0, 1);
    double vv = clampd((ty + 0.5) * T->invTY + (u4[1] - 0.5) * T->invTY, 0, 1);
    double x = 2 * uu - 1, y = 2 * vv - 1;
    d3 right = f3(c->right), up = f3(c->up), fwd = f3(c->fwd);
    if (c->type == 0) {
        d3 dc = sub(add(scl(right, x * c->fx), scl(up, y * c->fy)), fwd);
        *o = f3(c->pos); *d = unit(dc);
        if (c->lensR > 0) {          /* the thin lens: a point of its disc (lens.js lensDisc:111), aimed at the focus plane */
            double l2[4]; rng4(&r, l2);
            double rr = c->lensR * sqrt(u4[3]), ph = 2 * PI_D * l2[0];
            d3 off = add(scl(right, rr * cos(ph)), scl(up, rr * sin(ph)));
            *o = add(*o, off); *d = unit(sub(scl(dc, c->focus), off));
        }
    }
    else { *o = add(f3(c->pos), add(scl(right, x * c->fx), scl(up, y * c->fy))); *d = neg(fwd); }
    double fw = S->u->width, fh = S->u->height;
    uint32_t px = (uint32_t)clampd(uu * fw, 0, fw - 1), py = (uint32_t)clampd(vv * fh, 0, fh - 1);
    *pix = px + py * S->u->width;
    *uhaze = u4[2];
}

typedef struct { long long segments, shadow, steps, shadowSteps, maxiter, nonfinite, photonSegs, stored, gathers; } OrStats;

/* ---- SPPM (core/sppm.js:79-192 map and schedule, :487-520 gather; core/wavefront.js:542-626 photon emission) ----
 * The oracle's map is exact: records sorted by cell (side 2r), every record of the 8 cells about a gather point
 * walked (the engine subsamples crowded cells by sub-lists, exact in expectation). */
typedef struct { double p[3], pw[3], l[3], n[3]; long long key; } OrPhoton;
typedef struct { OrPhoton* rec; size_t n, cap; double r; } OrMap;

static long long cellKey(long long i, long long j, long long k) {
    return ((i + (1LL << 20)) << 42) | ((j + (1LL << 20)) << 21) | (k + (1LL << 20));
}
static int photonCmp(const void* a, const void* b) {
    long long x = ((const OrPhoton*)a)->key, y = ((const OrPhoton*)b)->key;
    return x < y ? -1 : (x > y ? 1 : 0);
}
static void mapPush(OrMap* m, d3 p, d3 pw, d3 l, d3 n) {
    if (m->n == m->cap) { m->cap = m->cap ? 2 * m->cap : 1024; m->rec = (OrPhoton*)realloc(m->rec, m->cap * sizeof(OrPhoton)); }
    OrPhoton* r = &m->rec[m->n++];
    r->p[0] = p.x; r->p[1] = p.y; r->p[2] = p.z; r->pw[0] = pw.x; r->pw[1] = pw.y; r->pw[2] = pw.z;
    r->l[0] = l.x; r->l[1] = l.y; r->l[2] = l.z; r->n[0] = n.x; r->n[1] = n.y; r->n[2] = n.z;
    double inv = 1 / (2 * m->r);
    r->key = cellKey((long long)floor(p.x * inv), (long long)floor(p.y * inv), (long long)floor(p.z * inv));
}
/* r_k = r_1 sqrt(prod_{i=1}^{k-1} (i + alpha) / (i + 1)) (sppm.js:177-192), by the product itself */
double or_sppm_radius(double r1, double alpha, uint32_t k) {
    double s = 1;
    for (uint32_t i = 1; i < (k > 1 ? k : 1); ++i) s *= (i + alpha) / (i + 1.0);
    return r1 * sqrt(s);
}
static double faceFlux(const PtLightFace* f) {
    return f->weight > 0 ? lumad(f3(f->le)) * PI_D * faceArea(f) * (f->twoSided ? 2 : 1) : 0;
}
static double sunFlux(const PtEnv* e) { return (e->sunE > 0 && e->sunRadius > 0) ? e->sunE * PI_D * (double)e->sunRadius * e->sunRadius : 0; }
static int evaluable(const PtMaterial* m) { return !lobesNone(m); }

/* ---- LIGHT TRACING (core/materials_pbr.js:554-700 ltConnect, camFootprint; core/haze.js:404-466 ltConnectMedium) ----
 * A sun photon that came through a delta event connects to the camera where it first meets a surface (by the non-delta
 * lobes, seen from the camera) or the haze (the phase function, the albedo); a camera path whose first vertex left it
 * by a non-delta lobe leaves the sun it meets through delta events only to these. Pinhole and orthographic cameras,
 * one ray a pixel, not the illuminance view. */
static int ltOn(const OrScene* S) {
    const PtUniforms* u = S->u;
    OrTiles T = tilesOf(u->width * u->height, u->width, u->height);
    return u->lt && u->measureMode != 1 && !(u->cam.type == 0 && u->cam.lensR > 0) && T.tilesX == u->width;
}
/* The camera's view of P: its pixel, the unit way to the camera and its length, and G over a pixel's area on the image
 * plane at distance 1 (4 fx fy / (w h)): G = |n.v| dist / z^3 for a pinhole, |n.v| orthographic; in a volume the same
 * without |n.v|. 0 where the camera does not see P (behind it, off the picture). */
static int camSees(const OrScene* S, d3 P, d3 nrm, int surface, uint32_t* pix, d3* vv, double* dist, double* g) {
    const PtUniforms* u = S->u;
    const PtCamera* c = &u->cam;
    d3 right = f3(c->right), up = f3(c->up), fwd = f3(c->fwd), rel = sub(P, f3(c->pos));
    double z = -dotp(rel, fwd), x, y, G;
    if (!(z > 1e-4)) return 0;
    if (c->type == 0) {
        *dist = len(rel);
        *vv = scl(rel, -1 / *dist);
        x = dotp(rel, right) / (z * c->fx);
        y = dotp(rel, up) / (z * c->fy);
        G = (surface ? fabs(dotp(nrm, *vv)) : 1) * *dist / (z * z * z);
    } else {
        *dist = z;
        *vv = fwd;
        x = dotp(rel, right) / c->fx;
        y = dotp(rel, up) / c->fy;
        G = surface ? fabs(dotp(nrm, fwd)) : 1;
    }
    if (!(fabs(x) < 1 && fabs(y) < 1)) return 0;
    double fw = u->width, fh = u->height;
    uint32_t px = (uint32_t)clampd((x + 1) * 0.5 * fw, 0, fw - 1), py = (uint32_t)clampd((y + 1) * 0.5 * fh, 0, fh - 1);
    *pix = px + py * u->width;
    *g = G * fw * fh / (4.0 * c->fx * c->fy);
    return 1;
}
/* A connection from o towards the camera, stopping two offsets short: its light into its pixel if nothing is in the way,
 * dimmed by the haze it crosses (the engine's occlusion pass). */
static void ltDeposit(const OrScene* S, d3 o, d3 w, double l, d3 c, uint32_t pix, double* frame, OrStats* st) {
    double tmax = l - 2 * offsetTol(S, o);
    st->shadow++;
    if (occluded(S, o, w, tmax, &st->shadowSteps)) {
        if (hazeOn(S)) c = scl(c, hazeTr(&S->u->haze, o, w, tmax));
        #pragma omp atomic
        frame[4 * pix + 0] += c.x;
        #pragma omp atomic
        frame[4 * pix + 1] += c.y;
        #pragma omp atomic
        frame[4 * pix + 2] += c.z;
    }
}

/* One photon: drawn from the lights by flux (the sun's beam through its aperture, the light faces), marched and
 * scattered through delta events only, stored where it lands on an evaluable surface after one at least (out: the
 * map, NULL without one); with light tracing a sun photon's light there, or where it scatters in the haze after a
 * delta event, goes to the camera (frame). */
static void tracePhoton(const OrScene* S, const uint32_t* seed, uint32_t i, uint32_t N, OrMap* out, OrStats* st, double* frame) {
    const PtUniforms* u = S->u;
    const PtEnv* env = &u->env;
    double total = sunFlux(env), fsun = total;
    for (int k = 0; k < S->nfaces; ++k) total += faceFlux(&S->faces[k]);
    if (!(total > 0)) return;
    OrRng r = rngAt(seed, i, 0, 0x94070000u);
    double u0[4], u1[4]; rng4(&r, u0); rng4(&r, u1);
    double pick = u0[0] * total, each = total / N;
    d3 o, d, beta;
    const int sunPh = pick < fsun, lt = ltOn(S);
    if (sunPh) {
        d3 a = f3(env->sunDir), t, b; basis(a, &t, &b);
        double rr = env->sunRadius * sqrt(u0[1]), ph = 2 * PI_D * u0[2];
        o = add(f3(env->sunPoint), add(scl(t, rr * cos(ph)), scl(b, rr * sin(ph))));
        d = neg(sunDir(env, u1[0], u1[1]));
        double w = each * maxd(-dotp(d, a), 0);   /* L cos(theta) Omega A with L = E / Omega */
        beta = v(w, w, w);
    } else {
        pick -= fsun;
        int k = 0;
        for (; k + 1 < S->nfaces; ++k) { double ff = faceFlux(&S->faces[k]); if (pick < ff) break; pick -= ff; }
        const PtLightFace* f = &S->faces[k];
        d3 y = f3(f->centre);
        if (f->shape == 1) { double rr = f->halfx * sqrt(u0[1]), ph = 2 * PI_D * u0[2]; y = add(y, scl(add(scl(f3(f->u), cos(ph)), scl(f3(f->v), sin(ph))), rr)); }
        else y = add(y, add(scl(f3(f->u), (2 * u0[1] - 1) * f->halfx), scl(f3(f->v), (2 * u0[2] - 1) * f->halfy)));
        d3 side = (f->twoSided && u1[2] < 0.5) ? neg(f3(f->n)) : f3(f->n);
        d = cosDir(side, u1[0], u1[1]);
        o = add(y, scl(side, 2 * offsetTol(S, y)));
        d3 le = f3(f->le);
        beta = scl(le, each / maxd(lumad(le), 1e-30));
    }
    d3 p = o, dir = d;
    double n = 1, ltot = 0, lambda = 0, vtxL = 0;
    int med = 0;   /* the absorbing medium the photon travels in: material + 1, 0 none */
    uint32_t steps = 0, bounces = 0;
    int hadDelta = 0;
    double collide = hazeOn(S) ? hazeDraw(&u->haze, p, dir, u1[3]) : 1e300;
    for (;;) {
        int state = march(S, &p, dir, n, &ltot, &steps, collide);
        st->photonSegs++;
        if (state != ST_ACTIVE) {             /* escaped, out of steps, or collided in the haze */
            if (state == ST_SCATTER && lt && hadDelta) {
                uint32_t cp; d3 vv; double dist, g;
                if (camSees(S, p, v(0, 0, 0), 0, &cp, &vv, &dist, &g)) {
                    d3 c = scl(beta, u->haze.albedo * hg(u->haze.g, dotp(dir, vv)) * g);
                    if (max3d(c) > 0) ltDeposit(S, p, vv, dist, c, cp, frame, st);
                }
            }
            return;
        }
        OrRng rng = rngAt(seed, i, bounces, (0x5AADEu ^ ((steps & 0xFFFFu) << 12)) ^ 0x80000000u);
        OrHit h = sceneAt(S, p);
        d3 ns = h.n;
        if (dotp(ns, ns) < 0.25) ns = normalAt(S, p, marchTol(S, p));
        int inside = h.d < 0;
        d3 nrm = inside ? neg(ns) : ns;
        if (dotp(dir, nrm) > 0) { double off = offsetTol(S, p); p = add(p, scl(dir, off)); ltot += off * n; continue; }
        bounces++;
        if (med) beta = mul(beta, mediumTrD(&S->mats[med - 1], maxd(ltot - vtxL, 0) / maxd(n, 1e-4)));
        PtMaterial mown;
        const PtMaterial* m = hitMaterial(S, &h, p, ns, &mown);
        if (max3d(f3(m->emissive)) > 0) return;
        double n0 = n, n1 = maxd(m->ior, 1);
        int beyond = -1;   /* the material beyond the surface when leaving one, if any */
        if (inside) {
            n1 = 1;
            if (m->transmission > 0) {
                OrHit hb = sceneAt(S, add(p, scl(dir, 2 * offsetTol(S, p))));
                if (hb.insideW > -1e6 && hb.a < S->nmats) { n1 = maxd(S->mats[hb.a].ior, 1); beyond = hb.a; }   /* inside something (-1e7: nothing, air) */
            }
        }
        double ud[4]; rng4(&rng, ud);
        int draw = 0; double lambda0 = lambda; double sw[3] = {1, 1, 1};
        if (m->abbe > 0 && m->transmission > 0 && !inside) {
            if (lambda <= 0) { draw = 1; lambda = 0.38 + 0.40 * ud[3]; or_spectral_weight(lambda, sw); }
            n1 = or_cauchy(n1, m->abbe, lambda);
        }
        OrLobes L = makeLobes(nrm, neg(dir), m, n0 / n1);
        if (evaluable(m) && hadDelta) {
            if (out) { mapPush(out, p, beta, neg(dir), nrm); st->stored++; }
            if (lt && fabs(n0 - 1) < 1e-3) {
                uint32_t cp; d3 vv; double dist, g;
                if (camSees(S, p, nrm, 1, &cp, &vv, &dist, &g) && dotp(nrm, vv) > 0) {
                    OrLobes Lc = makeLobes(nrm, vv, m, n0 / n1);
                    d3 c = scl(mul(beta, lobesEval(&Lc, neg(dir))), g);
                    if (max3d(c) > 0) {
                        d3 o2 = add(p, scl(nrm, 2 * offsetTol(S, p)));
                        d3 toCam = u->cam.type == 0 ? sub(f3(u->cam.pos), o2) : scl(vv, dist);
                        double l = len(toCam);
                        ltDeposit(S, o2, scl(toCam, 1 / l), l, c, cp, frame, st);
                    }
                }
            }
        }
        double us[4]; rng4(&rng, us);
        OrSample s = lobesSample(&L, us);
        if (!s.delta) return;                 /* the map holds L S+ D paths: a photon goes on through delta lobes only */
        dir = s.dir;
        beta = mul(beta, s.w);
        if (s.transmitted) { n = n1; if (draw) beta = mul(beta, v(sw[0], sw[1], sw[2])); med = mediumAfter(S, inside, h.m < S->nmats ? h.m : 0, beyond); }
        else if (draw) lambda = lambda0;
        vtxL = ltot;
        if (!(max3d(s.w) > 0)) return;
        hadDelta = 1;
        if (bounces > u->maxbounce) return;
        if (u->rr && bounces >= u->rrStart) {
            double pr = clampd(max3d(beta), 0.05, 0.95);
            if (ud[0] > pr) return; else beta = scl(beta, 1 / pr);
        }
        double off = offsetTol(S, p);
        p = add(p, scl(dir, off)); ltot += off * n;
        collide = (hazeOn(S) && n <= 1) ? ltot + hazeDraw(&u->haze, p, dir, ud[1]) * n : 1e300;
    }
}

/* sum f(l_p) P_p / (pi r^2) over the map's photons within r of p whose normal agrees (dot >= 0.9) */
static d3 mapGather(const OrMap* m, const OrLobes* L, d3 p, d3 nrm) {
    double r = m->r, inv = 1 / (2 * r);
    long long i0 = (long long)floor(p.x * inv - 0.5), j0 = (long long)floor(p.y * inv - 0.5), k0 = (long long)floor(p.z * inv - 0.5);
    d3 sum = v(0, 0, 0);
    for (int c = 0; c < 8; ++c) {
        long long key = cellKey(i0 + (c & 1), j0 + ((c >> 1) & 1), k0 + ((c >> 2) & 1));
        size_t lo = 0, hi = m->n;
        while (lo < hi) { size_t mid = (lo + hi) / 2; if (m->rec[mid].key < key) lo = mid + 1; else hi = mid; }
        for (size_t q = lo; q < m->n && m->rec[q].key == key; ++q) {
            const OrPhoton* ph = &m->rec[q];
            d3 dd = sub(v(ph->p[0], ph->p[1], ph->p[2]), p);
            if (dotp(dd, dd) >= r * r || dotp(v(ph->n[0], ph->n[1], ph->n[2]), nrm) < 0.9) continue;
            sum = add(sum, mul(lobesEval(L, v(ph->l[0], ph->l[1], ph->l[2])), v(ph->pw[0], ph->pw[1], ph->pw[2])));
        }
    }
    return scl(sum, 1 / (PI_D * r * r));
}

/* One camera path to its end (core/wavefront.js:1244-1442 shade; core/tracebasics.js:834-1108 path phases;
 * core/materials_pbr.js:1014-1070 emitterHit, :1110-1454 processMaterial_advanced). Adds to rgb. */
static void tracePath(const OrScene* S, const OrTiles* T, const uint32_t* seed, uint32_t ray, const OrMap* map,
                      double* frame, OrStats* st) {
    const PtUniforms* u = S->u;
    /* SPPM bookkeeping: deltaOnly (every event so far delta), caustic (a gather was left by a non-delta lobe),
     * sdelta (every event since it delta): light met with caustic && sdelta is the map's. With light tracing, ltcls
     * (the first vertex was left by a non-delta lobe): the sun met with ltcls && sdelta is the photons' connections' */
    int deltaOnly = 1, caustic = 0, sdelta = 0, ltcls = 0;
    const int lt = ltOn(S);
    const PtEnv* env = &u->env;
    d3 p, dir; uint32_t pix; double uh;
    cameraRay(S, T, seed, ray, &p, &dir, &pix, &uh);
    d3 beta = v(1, 1, 1), rad = v(0, 0, 0);
    double n = 1, ltot = 0, prevPdf = -1, vtxLtot = 0, sunmis = 1, lambda = 0;
    int med = 0;   /* the absorbing medium the path travels in: material + 1, 0 none */
    uint32_t steps = 0, bounces = 0;
    double collide = hazeOn(S) ? hazeDraw(&u->haze, p, dir, uh) : 1e300;
    int direct = 0;
    for (;;) {
        uint32_t s0 = steps;
        int state = march(S, &p, dir, n, &ltot, &steps, collide);
        st->steps += steps - s0;
        st->segments++;
        if (state == ST_ESCAPED) {
            int mapped = (caustic || ltcls) && sdelta && sunFlux(env) > 0;
            /* the illuminance view (uniforms.schema.js measureMode 1): a camera ray that meets nothing shows 0 */
            if (!(u->measureMode == 1 && bounces == 0))
                rad = add(rad, scl(mul(beta, add(envBg(env, dir), mapped ? v(0, 0, 0) : envSun(env, p, dir))), sunmis));
            direct = bounces == 0;
            break;
        }
        if (state == ST_MAXITER) { st->maxiter++; break; }
        OrRng rng = rngAt(seed, ray, bounces, 0x5AADEu ^ ((steps & 0xFFFFu) << 12));
        if (state == ST_SCATTER) {
            const PtHaze* hz = &u->haze;
            bounces++;
            beta = scl(beta, hz->albedo);
            double fs = faceShare(S, bounces), u0[4], u1[4];
            rng4(&rng, u0); rng4(&rng, u1);
            if (u->nee) {
                d3 L = v(0, 0, 0), le = v(0, 0, 0); double pS = 0, tmax = 1e300;
                if (u0[0] < fs) {
                    const PtLightFace* f = &S->faces[facePick(S, u1[3] * S->faceTotal)];
                    d3 y = f3(f->centre);
                    if (f->shape == 1) { double rr = f->halfx * sqrt(u0[1]), ph = 2 * PI_D * u0[2]; y = add(y, scl(add(scl(f3(f->u), cos(ph)), scl(f3(f->v), sin(ph))), rr)); }
                    else y = add(y, add(scl(f3(f->u), (2 * u0[1] - 1) * f->halfx), scl(f3(f->v), (2 * u0[2] - 1) * f->halfy)));
                    d3 dy = sub(y, p); double dist = len(dy); L = scl(dy, 1 / maxd(dist, 1e-20));
                    double cosY = -dotp(f3(f->n), L);
                    if (dist > 0 && fabs(cosY) > 1e-6 && (cosY > 0 || f->twoSided)) { pS = faceDensity(S, fs, f, dist, cosY); le = f3(f->le); tmax = dist - 2 * offsetTol(S, y); }
                } else if (envLit(env)) {
                    double pick = (u0[0] - fs) / maxd(1 - fs, 1e-6);
                    L = pick < sunPick(env) ? sunDir(env, u0[1], u0[2]) : cosDir(v(0, 0, 1), u0[1], u0[2]);
                    pS = (1 - fs) * envPdf(env, L, v(0, 0, 1));
                    le = add(envBg(env, L), envSun(env, p, L));
                }
                if (pS > 0) {
                    double ph = hg(hz->g, dotp(dir, L));
                    d3 c = scl(mul(beta, le), ph * misW(pS, ph) / pS);
                    if (max3d(c) > 1e-7) {
                        st->shadow++;
                        if (occluded(S, p, L, tmax, &st->shadowSteps)) {
                            c = scl(c, hazeTr(hz, p, L, tmax < 1e299 ? tmax : 1e30));
                            #pragma omp atomic
                            frame[4 * pix + 0] += c.x;
                            #pragma omp atomic
                            frame[4 * pix + 1] += c.y;
                            #pragma omp atomic
                            frame[4 * pix + 2] += c.z;
                        }
                    }
                }
            }
            d3 nd = hgDir(hz->g, dir, u1[0], u1[1]);
            double ph = hg(hz->g, dotp(dir, nd));
            double pLout = (1 - fs) * (envLit(env) ? envPdf(env, nd, v(0, 0, 1)) : 0);
            sunmis = (u->nee && pLout > 0) ? misW(ph, pLout) : 1;
            prevPdf = u->nee ? ph : -1;
            vtxLtot = ltot;
            dir = nd;
            deltaOnly = caustic = sdelta = ltcls = 0;
            if (lt && bounces == 1) ltcls = 1;   /* the first vertex in the haze */
            if (bounces > u->maxbounce) break;
            collide = ltot + hazeDraw(hz, p, dir, u1[2]) * n;
            continue;
        }
        OrHit h = sceneAt(S, p);
        d3 ns = h.n;
        if (dotp(ns, ns) < 0.25) ns = normalAt(S, p, marchTol(S, p));
        int inside = h.d < 0;
        d3 nrm = inside ? neg(ns) : ns;
        if (dotp(dir, nrm) > 0) { double off = offsetTol(S, p); p = add(p, scl(dir, off)); ltot += off * n; continue; }
        bounces++;
        if (med) beta = mul(beta, mediumTrD(&S->mats[med - 1], maxd(ltot - vtxLtot, 0) / maxd(n, 1e-4)));
        PtMaterial mown;
        const PtMaterial* m = hitMaterial(S, &h, p, ns, &mown);
        /* ...and the camera's first hit is a white Lambertian for that bounce: the pixel reads E / pi there */
        PtMaterial white;
        if (u->measureMode == 1 && bounces == 1) {
            memset(&white, 0, sizeof white);
            white.base[0] = white.base[1] = white.base[2] = 1; white.roughness = 1; white.ior = 1.5; white.face = -1;
            m = &white;
        }
        d3 Le = f3(m->emissive);
        if (max3d(Le) > 0 && (caustic || ltcls) && sdelta && m->face >= 0 && faceFlux(&S->faces[m->face]) > 0) break;
        if (max3d(Le) > 0) {
            double w = 1;
            if (m->face >= 0) {
                const PtLightFace* f = &S->faces[m->face];
                double cosY = -dotp(f3(f->n), dir);
                if (!(cosY > 0 || f->twoSided)) Le = v(0, 0, 0);
                else if (prevPdf >= 0) {
                    double dist = maxd(ltot - vtxLtot, 0) / maxd(n, 1e-4);
                    w = misW(prevPdf, faceDensity(S, faceShare(S, bounces - 1), f, dist, cosY));
                }
            }
            rad = add(rad, scl(mul(beta, Le), w));
            direct = bounces == 1;
            break;
        }
        double n0 = n, n1 = maxd(m->ior, 1);
        int beyond = -1;   /* the material beyond the surface when leaving one, if any */
        if (inside) {
            n1 = 1;
            if (m->transmission > 0) {
                OrHit hb = sceneAt(S, add(p, scl(dir, 2 * offsetTol(S, p))));
                if (hb.insideW > -1e6 && hb.a < S->nmats) { n1 = maxd(S->mats[hb.a].ior, 1); beyond = hb.a; }   /* inside something (-1e7: nothing, air) */
            }
        }
        double ud[4]; rng4(&rng, ud);
        int draw = 0; double lambda0 = lambda; double sw[3] = {1, 1, 1};
        if (m->abbe > 0 && m->transmission > 0 && !inside) {
            if (lambda <= 0) { draw = 1; lambda = 0.38 + 0.40 * ud[3]; or_spectral_weight(lambda, sw); }
            n1 = or_cauchy(n1, m->abbe, lambda);
        }
        OrLobes L = makeLobes(nrm, neg(dir), m, n0 / n1);
        double fs = faceShare(S, bounces);
        int gathered = 0;
        /* sppm 2: at every non-delta camera vertex; with light tracing not at the first (the connections') */
        if (map && (u->sppm >= 2 || deltaOnly) && evaluable(m) && !(lt && bounces == 1)) {
            rad = add(rad, mul(beta, mapGather(map, &L, p, nrm)));
            gathered = 1;
            st->gathers++;
        }
        if (u->nee && !lobesNone(m)) {
            double u0[4], u1[4]; rng4(&rng, u0); rng4(&rng, u1);
            d3 Ld = v(0, 0, 0), le = v(0, 0, 0); double pS = 0, tmax = 1e300;
            if (u0[0] < fs) {
                const PtLightFace* f = &S->faces[facePick(S, u1[0] * S->faceTotal)];
                d3 y = f3(f->centre);
                if (f->shape == 1) { double rr = f->halfx * sqrt(u0[1]), ph = 2 * PI_D * u0[2]; y = add(y, scl(add(scl(f3(f->u), cos(ph)), scl(f3(f->v), sin(ph))), rr)); }
                else y = add(y, add(scl(f3(f->u), (2 * u0[1] - 1) * f->halfx), scl(f3(f->v), (2 * u0[2] - 1) * f->halfy)));
                d3 dy = sub(y, p); double dist = len(dy); Ld = scl(dy, 1 / maxd(dist, 1e-20));
                double cosY = -dotp(f3(f->n), Ld);
                if (dist > 0 && dotp(nrm, Ld) > 0 && fabs(cosY) > 1e-6 && (cosY > 0 || f->twoSided)) {
                    pS = faceDensity(S, fs, f, dist, cosY);
                    le = scl(f3(f->le), 1 / maxd(n0 * n0, 1));
                    d3 o = add(p, scl(nrm, 2 * offsetTol(S, p)));
                    tmax = len(sub(y, o)) - 2 * offsetTol(S, p);
                }
            } else if (envLit(env)) {
                double pick = (u0[0] - fs) / maxd(1 - fs, 1e-6);
                Ld = pick < sunPick(env) ? sunDir(env, u0[1], u0[2]) : cosDir(nrm, u0[1], u0[2]);
                if (dotp(nrm, Ld) > 0) { pS = (1 - fs) * envPdf(env, Ld, nrm); le = add(envBg(env, Ld), envSun(env, p, Ld)); }
            }
            if (pS > 0 && max3d(le) > 0) {
                d3 f = lobesEval(&L, Ld);
                double pB = lobesPdf(&L, Ld), NoL = dotp(nrm, Ld);
                d3 c = scl(mul(mul(beta, f), le), NoL * misW(pS, pB) / pS);
                if (max3d(c) > 1e-7) {
                    d3 o = add(p, scl(nrm, 2 * offsetTol(S, p)));
                    if (tmax < 1e299) Ld = unit(sub(add(o, scl(Ld, tmax + 2 * offsetTol(S, p))), o));
                    st->shadow++;
                    if (occluded(S, o, Ld, tmax, &st->shadowSteps)) {
                        if (hazeOn(S)) c = scl(c, hazeTr(&u->haze, o, Ld, tmax < 1e299 ? tmax : 1e30));
                        #pragma omp atomic
                        frame[4 * pix + 0] += c.x;
                        #pragma omp atomic
                        frame[4 * pix + 1] += c.y;
                        #pragma omp atomic
                        frame[4 * pix + 2] += c.z;
                    }
                }
            }
        }
        double us[4]; rng4(&rng, us);
        OrSample s = lobesSample(&L, us);
        dir = s.dir;
        beta = mul(beta, s.w);
        if (s.transmitted) { n = n1; if (draw) beta = mul(beta, v(sw[0], sw[1], sw[2])); med = mediumAfter(S, inside, h.m < S->nmats ? h.m : 0, beyond); }
        else if (draw) lambda = lambda0;
        int terminate = !(max3d(s.w) > 0);
        int misOut = u->nee && !s.delta;
        double pOut = misOut ? lobesPdf(&L, dir) : 0;
        int above = dotp(nrm, dir) > 0 && !s.delta;
        if (misOut) {
            double pLout = above && envLit(env) ? (1 - fs) * envPdf(env, dir, nrm) : 0;
            sunmis = pLout > 0 ? misW(pOut, pLout) : 1;
            prevPdf = above ? pOut : -1;
        } else { sunmis = 1; prevPdf = -1; }
        vtxLtot = ltot;
        if (s.eval) steps += 0x10000u;
        if (!s.delta) deltaOnly = 0;
        /* the classes kept through delta events, closed at any other; opened where a non-delta lobe leaves the vertex on
         * its own side: after a gather (the map's), at the first vertex with light tracing (the connections') */
        if (s.delta) { if (caustic || ltcls) sdelta = 1; }
        else caustic = ltcls = sdelta = 0;
        if (!s.delta && !s.transmitted) {
            if (gathered) caustic = 1;
            else if (lt && bounces == 1 && evaluable(m)) ltcls = 1;
        }
        if (bounces > u->maxbounce) terminate = 1;
        if (!terminate && u->diffuseCap && (steps >> 16) > u->diffuseCap) terminate = 1;
        if (!terminate && u->rr && bounces >= u->rrStart) {
            double pr = clampd(max3d(beta), 0.05, 0.95);
            if (ud[0] > pr) terminate = 1; else beta = scl(beta, 1 / pr);
        }
        if (terminate) break;
        double off = offsetTol(S, p);
        p = add(p, scl(dir, off)); ltot += off * n;
        collide = (hazeOn(S) && n <= 1) ? ltot + hazeDraw(&u->haze, p, dir, ud[1]) * n : 1e300;
    }
    (void)direct;
    if (!(isfinite(rad.x) && isfinite(rad.y) && isfinite(rad.z)) || lumad(rad) < 0) { st->nonfinite++; rad = v(0, 0, 0); }
    #pragma omp atomic
    frame[4 * pix + 0] += rad.x;
    #pragma omp atomic
    frame[4 * pix + 1] += rad.y;
    #pragma omp atomic
    frame[4 * pix + 2] += rad.z;
    #pragma omp atomic
    frame[4 * pix + 3] += 1;
}

static void sceneInit(OrScene* S, const int* ops, int nops, const float* U, const PtMaterial* mats, int nmats,
                      const PtLightFace* faces, int nfaces, const PtUniforms* u, OrOp* buf) {
    for (int i = 0; i < nops; ++i) {
        const int* o = ops + 8 * i;
        buf[i].op = o[0]; buf[i].shape = o[1]; buf[i].xf = o[2]; buf[i].p = o[3]; buf[i].mat = o[4]; buf[i].prim = o[5];
        buf[i].join = o[6]; buf[i].k = o[7];
    }
    S->ops = buf; S->nops = nops; S->U = U; S->mats = mats; S->nmats = nmats; S->faces = faces; S->nfaces = nfaces; S->u = u;
    S->faceTotal = 0;
    for (int k = 0; k < nfaces; ++k) if (faces[k].weight > 0) S->faceTotal += faces[k].weight;
}

/* Renders `samples` samples from sample index `first`, FP64 accumulation as the engine's resolve
 * (core/wavefront.js:414-494, core/measure.js:36-57); accum [w*h*4] mean rgb + count, moments [w*h*2].
 * With u->sppm each sample first runs its photon pass (u->sppmPhotons photons, radius r_k for sample k = first + s + 1).
 * ops: 8 ints per program entry (op, shape, xf, p, mat, prim, join, k). stats: 9 counters. */
int or_render(const int* ops, int nops, const float* U, const PtMaterial* mats, int nmats, const PtLightFace* faces,
              int nfaces, const PtUniforms* u, uint32_t first, uint32_t samples, double* accum, double* moments,
              long long* stats, int threads) {
    skyInit();
    OrOp* buf = (OrOp*)malloc(sizeof(OrOp) * (nops > 0 ? nops : 1));
    OrScene S; sceneInit(&S, ops, nops, U, mats, nmats, faces, nfaces, u, buf);
    or_mf_tables(NULL);
    uint32_t npx = u->width * u->height, N = npx;
    OrTiles T = tilesOf(N, u->width, u->height);
    double* frame = (double*)calloc((size_t)npx * 4, sizeof(double));
    OrStats tot; memset(&tot, 0, sizeof tot);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    OrMap map; memset(&map, 0, sizeof map);
    for (uint32_t s = 0; s < samples; ++s) {
        uint32_t seed[4]; sampleSeed(u->seed, first + s, seed);
        memset(frame, 0, sizeof(double) * npx * 4);
        if (u->sppm || u->lt) {
            /* the sample's photon pass: per-thread maps merged, then sorted by cell; light tracing's connections into
             * the frame */
            map.n = 0;
            map.r = or_sppm_radius(u->sppmRadius, u->sppmAlpha, first + s + 1);
            uint32_t NP = u->sppmPhotons;
            #pragma omp parallel
            {
                OrStats st; memset(&st, 0, sizeof st);
                OrMap mine; memset(&mine, 0, sizeof mine); mine.r = map.r;
                #pragma omp for schedule(dynamic, 256)
                for (uint32_t i = 0; i < NP; ++i) tracePhoton(&S, seed, i, NP, u->sppm ? &mine : NULL, &st, frame);
                #pragma omp critical
                {
                    for (size_t q = 0; q < mine.n; ++q) {
                        if (map.n == map.cap) { map.cap = map.cap ? 2 * map.cap : 4096; map.rec = (OrPhoton*)realloc(map.rec, map.cap * sizeof(OrPhoton)); }
                        map.rec[map.n++] = mine.rec[q];
                    }
                    tot.photonSegs += st.photonSegs; tot.stored += st.stored; tot.shadow += st.shadow; tot.shadowSteps += st.shadowSteps;
                }
                free(mine.rec);
            }
            qsort(map.rec, map.n, sizeof(OrPhoton), photonCmp);
        }
        #pragma omp parallel
        {
            OrStats st; memset(&st, 0, sizeof st);
            #pragma omp for schedule(dynamic, 64)
            for (uint32_t i = 0; i < N; ++i) tracePath(&S, &T, seed, i, u->sppm ? &map : NULL, frame, &st);
            #pragma omp critical
            { tot.segments += st.segments; tot.shadow += st.shadow; tot.steps += st.steps; tot.shadowSteps += st.shadowSteps;
              tot.maxiter += st.maxiter; tot.nonfinite += st.nonfinite; tot.gathers += st.gathers; }
        }
        for (uint32_t i = 0; i < npx; ++i) {
            double cnt = frame[4 * i + 3];
            if (cnt == 0) continue;
            double s0 = frame[4 * i] / cnt, s1 = frame[4 * i + 1] / cnt, s2 = frame[4 * i + 2] / cnt;
            double* a = accum + 4 * (size_t)i;
            double prevn = a[3];
            if (u->maxaverages > 0) prevn = mind(prevn, u->maxaverages);
            double c = prevn + 1, y = 0.2126 * s0 + 0.7152 * s1 + 0.0722 * s2, yPrev = 0.2126 * a[0] + 0.7152 * a[1] + 0.0722 * a[2];
            a[0] = (a[0] * prevn + s0) / c; a[1] = (a[1] * prevn + s1) / c; a[2] = (a[2] * prevn + s2) / c; a[3] = c;
            double w = 1 / c, dl = prevn > 0 ? y - yPrev : 0;
            double* mo = moments + 2 * (size_t)i;
            mo[0] = (1 - w) * (mo[0] + w * dl * dl);
            mo[1] = (1 - w) * (1 - w) * mo[1] + w * w;
        }
    }
    if (stats) { stats[0] = tot.segments; stats[1] = tot.shadow; stats[2] = tot.steps; stats[3] = tot.shadowSteps; stats[4] = tot.maxiter; stats[5] = tot.nonfinite;
                 stats[6] = tot.photonSegs; stats[7] = tot.stored; stats[8] = tot.gathers; }
    free(frame); free(buf); free(map.rec);
    return 0;
}

/* Each camera ray of sample `sample` to its first event (the engine's pt_pick_rays): pos4 (x, y, z, ltot),
 * info (state | material << 8, primitive or 0xFFFFFFFF). Rays [first, first + count). */
int or_pick(const int* ops, int nops, const float* U, const PtUniforms* u, uint32_t sample, uint32_t first, uint32_t count,
            double* pos4, uint32_t* info) {
    OrOp* buf = (OrOp*)malloc(sizeof(OrOp) * (nops > 0 ? nops : 1));
    OrScene S; sceneInit(&S, ops, nops, U, NULL, 0, NULL, 0, u, buf);
    uint32_t N = u->width * u->height;
    OrTiles T = tilesOf(N, u->width, u->height);
    uint32_t seed[4]; sampleSeed(u->seed, sample, seed);
    #pragma omp parallel for schedule(dynamic, 64)
    for (uint32_t k = 0; k < count; ++k) {
        uint32_t i = first + k;
        d3 p, d; uint32_t pix; double uh;
        cameraRay(&S, &T, seed, i, &p, &d, &pix, &uh);
        double ltot = 0; uint32_t steps = 0;
        double collide = hazeOn(&S) ? hazeDraw(&u->haze, p, d, uh) : 1e300;
        int st = march(&S, &p, d, 1, &ltot, &steps, collide);
        uint32_t prim = 0xFFFFFFFFu, mat = 0xFFFFFFu;
        if (st == ST_ACTIVE) { OrHit h = sceneAt(&S, p); prim = (uint32_t)h.prim; mat = (uint32_t)h.m; }
        pos4[4 * k] = p.x; pos4[4 * k + 1] = p.y; pos4[4 * k + 2] = p.z; pos4[4 * k + 3] = ltot;
        info[2 * k] = (uint32_t)st | (mat << 8); info[2 * k + 1] = prim;
    }
    free(buf);
    return 0;
}

/* The scene's signed distance and closest primitive at points (for brute-force checks). */
int or_scene_at(const int* ops, int nops, const float* U, const double* pts, int n, double* d, double* nrm, int* prim) {
    OrOp* buf = (OrOp*)malloc(sizeof(OrOp) * (nops > 0 ? nops : 1));
    PtUniforms u; memset(&u, 0, sizeof u);
    OrScene S; sceneInit(&S, ops, nops, U, NULL, 0, NULL, 0, &u, buf);
    for (int i = 0; i < n; ++i) {
        OrHit h = sceneAt(&S, v(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
        d[i] = h.d; if (nrm) { nrm[3 * i] = h.n.x; nrm[3 * i + 1] = h.n.y; nrm[3 * i + 2] = h.n.z; }
        if (prim) prim[i] = h.prim;
    }
    free(buf);
    return 0;
}

/* The BSDF for quadrature tests: eval (rgb), pdf, and a sample from u (dir, weight, flags). view/l/n world. through:
 * evaluate the transmission below the surface too. */
int or_bsdf(const PtMaterial* m, const double* n3, const double* view3, double eta, const double* l3, double* f, double* pdf,
            const double* u4, double* sdir, double* sw, int* flags, int through) {
    or_mf_tables(NULL);
    OrLobes L = makeLobes(v(n3[0], n3[1], n3[2]), v(view3[0], view3[1], view3[2]), m, eta);
    L.through = through;
    if (l3 && f) { d3 e = lobesEval(&L, v(l3[0], l3[1], l3[2])); f[0] = e.x; f[1] = e.y; f[2] = e.z; }
    if (l3 && pdf) *pdf = lobesPdf(&L, v(l3[0], l3[1], l3[2]));
    if (u4 && sdir) {
        OrSample s = lobesSample(&L, u4);
        sdir[0] = s.dir.x; sdir[1] = s.dir.y; sdir[2] = s.dir.z; sw[0] = s.w.x; sw[1] = s.w.y; sw[2] = s.w.z;
        *flags = s.delta | (s.transmitted << 1) | (s.eval << 2);
    }
    return 0;
}

/* Haze helpers for closed-form checks: transmittance along a ray, and the free-flight draw. */
/* The environment's radiance in each unit direction: the sky (envBg) and the sun's disc (envSun, from the origin). */
int or_env(const PtUniforms* u, const double* dirs, int n, double* out) {
    skyInit();
    for (int i = 0; i < n; ++i) {
        d3 d = unit(v(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
        d3 L = add(envBg(&u->env, d), envSun(&u->env, v(0, 0, 0), d));
        out[3 * i] = L.x; out[3 * i + 1] = L.y; out[3 * i + 2] = L.z;
    }
    return 0;
}
double or_haze_tr(const PtHaze* h, const double* o, const double* d, double lenMax) { return hazeTr(h, v(o[0], o[1], o[2]), v(d[0], d[1], d[2]), lenMax); }
double or_haze_draw(const PtHaze* h, const double* o, const double* d, double u) { return hazeDraw(h, v(o[0], o[1], o[2]), v(d[0], d[1], d[2]), u); }
double or_hg(double g, double c) { return hg(g, c); }
void or_hg_sample(double g, const double* w, double u1, double u2, double* out) { d3 r = hgDir(g, v(w[0], w[1], w[2]), u1, u2); out[0] = r.x; out[1] = r.y; out[2] = r.z; }
void or_pcg4d(const uint32_t* in, uint32_t* out) { OrRng s = {in[0], in[1], in[2], in[3]}; pcg4d(&s); out[0] = s.a; out[1] = s.b; out[2] = s.c; out[3] = s.d; }
void or_tiles(uint32_t N, uint32_t w, uint32_t h, uint32_t* tx, uint32_t* ty) { OrTiles t = tilesOf(N, w, h); *tx = t.tilesX; *ty = t.tilesY; }
