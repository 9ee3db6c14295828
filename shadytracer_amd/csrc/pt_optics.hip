// THE OPTICS BENCH: many luminaire scenes traced in one persistent launch, their far fields (candela) and lux maps
// tallied in FP64, and their numbers -- beam and field angles, CCT and Duv -- worked out on the GPU in FP64.
//
// The reference traces its optics through the renderer's SDF march and material code (lib/pathtracer/optics/kernels.js:
// prog_opt_gen, prog_opt_shade), tallying flux in fixed point, and reads the numbers on the host (optics/farfield.js
// intensity, farStats; optics/detmap.js detectorMap; core/colorimetry.js xyToCctDuv). Here every surface of a bench
// luminaire is a quadric -- the lens's conic faces, its rim, the paraboloid reflector, the board -- so a ray meets each
// in closed form, exactly, where a march would stop within a tolerance after tens of steps.
//
// WORK: the scenes' photon budgets are uneven (2^18..2^24), so the launch is persistent: the photons of all scenes are
// cut into chunks of PT_OPT_CHUNK, numbered scene after scene, and each wave takes the next chunk from one counter
// until none is left. A lane walks its photons one at a time (a walk is a handful of surfaces).
// TALLIES: escaped flux by direction into the type-C far field (181 x 360, FP64 atomics), by where it crosses z = 1 m
// into the lux map (256 x 256); the totals (X, Y, Z escaped, Y absorbed) summed per chunk by its wave and stored.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pt.h"

#define PT_OPT_CHUNK 4096u
#define PT_OPT_BOUNCES 32
#define PT_OPT_TOT 8          // per chunk: X, Y, Z escaped, Y absorbed, photons traced, surface tests, FP32 ops, spare

namespace {

struct OptDev {
    const PtOpticsScene* sc;
    unsigned int n;
    const unsigned long long* chunkStart;   // n + 1: the first chunk of each scene, then the total
    unsigned int* next;                     // the chunk counter
    double* far;                            // n x NG x NC, lm
    double* lux;                            // n x LUX x LUX, lm (lux after the stats)
    double* part;                           // chunks x PT_OPT_TOT: each chunk's sums
    PtOpticsResult* res;                    // n
};

// PCG4D (Jarzynski and Olano 2020): 24-bit uniforms, keyed by (scene seed, photon, bounce)
struct ORng { unsigned int a, b, c, d; };
__device__ __forceinline__ void opcg(ORng& s) {
    s.a = s.a * 1664525u + 1013904223u; s.b = s.b * 1664525u + 1013904223u;
    s.c = s.c * 1664525u + 1013904223u; s.d = s.d * 1664525u + 1013904223u;
    s.a += s.b * s.d; s.b += s.c * s.a; s.c += s.a * s.b; s.d += s.b * s.c;
    s.a ^= s.a >> 16; s.b ^= s.b >> 16; s.c ^= s.c >> 16; s.d ^= s.d >> 16;
    s.a += s.b * s.d; s.b += s.c * s.a; s.c += s.a * s.b; s.d += s.b * s.c;
}
__device__ __forceinline__ float4 onext(ORng& s) {
    opcg(s);
    const float k = 1.0f / 16777216.0f;
    return make_float4((float)(s.a >> 8) * k, (float)(s.b >> 8) * k, (float)(s.c >> 8) * k, (float)(s.d >> 8) * k);
}
__device__ __forceinline__ ORng orngAt(unsigned int seed, unsigned int photon, unsigned int bounce) {
    ORng s = {seed ^ 0x0F71C5u, photon, bounce * 0x9E3779B9u, 0x5EED0F71u};
    opcg(s);
    return s;
}

struct F3 { float x, y, z; };
__device__ __forceinline__ F3 f3(float x, float y, float z) { F3 r = {x, y, z}; return r; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 norm3(F3 a) { return a * (1.0f / sqrtf(dot3(a, a))); }

// sag of a conic face and its radial slope
__device__ __forceinline__ float sag(float c, float k, float r2) {
    float q = 1.0f - (1.0f + k) * c * c * r2;
    return c * r2 / (1.0f + sqrtf(fmaxf(q, 0.0f)));
}
// WORK COUNT: each step below adds the FP32 operations its source performs to the lane's tally `fl` (+, -, *, /, sqrt
// and each transcendental one apiece; comparisons, min/max, abs, negation and selects none). The bench's roofline
// divides their sum by the walk's time.
#define OPF_PLANE 12.0f       // hitConic, a flat face
#define OPF_CONIC 38.0f       // hitConic, a curved face: the quadratic's set-up and roots
#define OPF_ROOT 9.0f         //   each root tried: the point and its radius
#define OPF_SAG 14.0f         //   each root within the rim: the sag test
#define OPF_RIM 16.0f         // the rim cylinder's quadratic
#define OPF_RIM_ROOT 5.0f     //   each of its roots tried (one sqrt before the first)
#define OPF_REFL 24.0f        // the paraboloid's quadratic and roots
#define OPF_REFL_ROOT 2.0f
#define OPF_BOARD 11.0f
#define OPF_MOVE 6.0f
#define OPF_MIRROR 40.0f      // the paraboloid's normal, the reflection, the weights
#define OPF_NORMAL 19.0f      // a curved face's normal
#define OPF_FRESNEL 30.0f     // cos, eta, Fresnel, the uniform
#define OPF_REFLECT 17.0f
#define OPF_REFRACT 21.0f
#define OPF_ESCAPE 8.0f       // gamma, C and their bins
#define OPF_LUX 10.0f         // the crossing of z = 1 m and its cell
#define OPF_EMIT 45.0f        // the LED's draw, the weights, the faces' rim heights

// The nearest t > tmin where the ray meets the conic face (vertex zv, within radius rmax), on the branch through the
// vertex; 1e30 for none. The face: c (x^2 + y^2) - 2 (z - zv) + c (1 + k) (z - zv)^2 = 0.
__device__ __forceinline__ float hitConic(F3 o, F3 d, float zv, float c, float k, float rmax, float tmin, float& fl) {
    const float R2 = rmax * rmax;
    if (c == 0.0f) {
        fl += OPF_PLANE;
        if (fabsf(d.z) < 1e-12f) return 1e30f;
        float t = (zv - o.z) / d.z;
        F3 p = o + d * t;
        return (t > tmin && p.x * p.x + p.y * p.y <= R2) ? t : 1e30f;
    }
    fl += OPF_CONIC;
    const float Z = o.z - zv, ck = c * (1.0f + k);
    const float A = c * (d.x * d.x + d.y * d.y) + ck * d.z * d.z;
    const float B = 2.0f * (c * (o.x * d.x + o.y * d.y) - d.z + ck * Z * d.z);
    const float C = c * (o.x * o.x + o.y * o.y) - 2.0f * Z + ck * Z * Z;
    float t0, t1;
    if (fabsf(A) < 1e-12f) {
        if (fabsf(B) < 1e-20f) return 1e30f;
        t0 = t1 = -C / B;
    } else {
        const float disc = B * B - 4.0f * A * C;
        if (disc < 0.0f) return 1e30f;
        const float sq = sqrtf(disc);
        const float q = -0.5f * (B + (B >= 0.0f ? sq : -sq));   // the stable pair
        t0 = q / A; t1 = fabsf(q) > 0.0f ? C / q : t0;
        if (t0 > t1) { float s = t0; t0 = t1; t1 = s; }
    }
    for (int i = 0; i < 2; ++i) {
        const float t = i == 0 ? t0 : t1;
        if (!(t > tmin)) continue;
        fl += OPF_ROOT;
        F3 p = o + d * t;
        const float r2 = p.x * p.x + p.y * p.y;
        if (r2 > R2) continue;
        fl += OPF_SAG;
        // on the branch through the vertex
        if (fabsf((p.z - zv) - sag(c, k, r2)) > 1e-4f * (1.0f + fabsf(1.0f / c))) continue;
        return t;
    }
    return 1e30f;
}
// the face's normal at p, pointing to +z
__device__ __forceinline__ F3 conicNormal(F3 p, float zv, float c, float k) {
    if (c == 0.0f) return f3(0.0f, 0.0f, 1.0f);
    const float Z = p.z - zv;
    // gradient of c r^2 - 2 Z + c (1 + k) Z^2, negated so that it points to +z near the vertex
    return norm3(f3(-2.0f * c * p.x, -2.0f * c * p.y, 2.0f - 2.0f * c * (1.0f + k) * Z));
}
// Fresnel reflectance of unpolarised light, cosI > 0, eta = n_incident / n_transmitted; 1 at total reflection
__device__ __forceinline__ float fresnelU(float cosI, float eta, float& cosT) {
    const float s2 = eta * eta * (1.0f - cosI * cosI);
    if (s2 >= 1.0f) { cosT = 0.0f; return 1.0f; }
    cosT = sqrtf(1.0f - s2);
    const float rs = (eta * cosI - cosT) / (eta * cosI + cosT);
    const float rp = (cosI - eta * cosT) / (cosI + eta * cosT);
    return 0.5f * (rs * rs + rp * rp);
}

__device__ __forceinline__ unsigned int laneId() { return threadIdx.x & 63u; }
__device__ __forceinline__ double waveSumD(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// One photon of scene `s`: emitted from the LED, walked through lens, rim, reflector and board, tallied where it
// leaves. acc: X, Y, Z escaped, Y absorbed (this lane's running sums); ev, fl: its surface tests and FP32 operations.
__device__ void tracePhoton(const OptDev& D, unsigned int s, const PtOpticsScene& S, unsigned int i, double* acc,
                            float& ev, float& fl) {
    ORng r = orngAt(S.seed, i, 0u);
    const float4 u = onext(r);
    F3 o = f3(S.ledHalf * (2.0f * u.x - 1.0f), S.ledHalf * (2.0f * u.y - 1.0f), 0.0f);
    const float sr = sqrtf(u.z), phi = 2.0f * 3.14159265f * u.w;
    float sp, cp;
    sincosf(phi, &sp, &cp);
    F3 d = f3(sr * cp, sr * sp, sqrtf(fmaxf(0.0f, 1.0f - u.z)));
    const float each = S.flux / (float)S.photons;
    float wX = S.xyz[0] * each, wY = S.xyz[1] * each, wZ = S.xyz[2] * each;
    const float zv1 = S.lensZ, zv2 = S.lensZ + S.lensT, rim = 0.5f * S.lensD;
    const float zr1 = zv1 + sag(S.c1, S.k1, rim * rim), zr2 = zv2 + sag(S.c2, S.k2, rim * rim);
    const bool refl = S.reflF > 0.0f;
    const float f4 = 4.0f * S.reflF, zb = refl ? S.reflR0 * S.reflR0 / f4 - S.reflF : 0.0f;
    bool inLens = false;
    const float eps = 1e-6f;
    fl += OPF_EMIT;
    for (int b = 0; b < PT_OPT_BOUNCES; ++b) {
        // the nearest surface
        float tBest = 1e30f;
        int what = -1;   // 0 front face, 1 back face, 2 rim, 3 reflector, 4 board
        ev += 1.0f;
        float t = hitConic(o, d, zv1, S.c1, S.k1, rim, eps, fl);
        if (t < tBest) { tBest = t; what = 0; }
        t = hitConic(o, d, zv2, S.c2, S.k2, rim, eps, fl);
        if (t < tBest) { tBest = t; what = 1; }
        {   // the rim: x^2 + y^2 = rim^2 between the faces' edges
            const float A = d.x * d.x + d.y * d.y, B = 2.0f * (o.x * d.x + o.y * d.y), C = o.x * o.x + o.y * o.y - rim * rim;
            const float disc = B * B - 4.0f * A * C;
            fl += OPF_RIM;
            if (A > 1e-12f && disc >= 0.0f) {
                const float sq = sqrtf(disc);
                fl += 1.0f;
                for (int k = 0; k < 2; ++k) {
                    const float tt = (-B + (k == 0 ? -sq : sq)) / (2.0f * A);
                    const float z = o.z + d.z * tt;
                    fl += OPF_RIM_ROOT;
                    if (tt > eps && tt < tBest && z >= fminf(zr1, zr2) && z <= fmaxf(zr1, zr2)) { tBest = tt; what = 2; break; }
                }
            }
        }
        if (refl && !inLens) {   // the paraboloid x^2 + y^2 = 4 f (z + f), between its base and its mouth
            const float A = d.x * d.x + d.y * d.y;
            const float B = 2.0f * (o.x * d.x + o.y * d.y) - f4 * d.z;
            const float C = o.x * o.x + o.y * o.y - f4 * (o.z + S.reflF);
            float t0 = 1e30f, t1 = 1e30f;
            fl += OPF_REFL;
            if (fabsf(A) < 1e-12f) { if (fabsf(B) > 1e-20f) t0 = t1 = -C / B; }
            else {
                const float disc = B * B - 4.0f * A * C;
                if (disc >= 0.0f) {
                    const float sq = sqrtf(disc), q = -0.5f * (B + (B >= 0.0f ? sq : -sq));
                    t0 = q / A; t1 = fabsf(q) > 0.0f ? C / q : t0;
                    if (t0 > t1) { float x = t0; t0 = t1; t1 = x; }
                }
            }
            for (int k = 0; k < 2; ++k) {
                const float tt = k == 0 ? t0 : t1;
                const float z = o.z + d.z * tt;
                fl += OPF_REFL_ROOT;
                if (tt > eps && tt < tBest && z >= zb && z <= S.reflZ1) { tBest = tt; what = 3; break; }
            }
        }
        if (d.z < 0.0f && !inLens) {   // the board at z = 0
            const float tt = -o.z / d.z;
            const F3 p = o + d * tt;
            fl += OPF_BOARD;
            if (tt > eps && tt < tBest && p.x * p.x + p.y * p.y <= S.boardR * S.boardR) { tBest = tt; what = 4; }
        }
        if (what < 0) {
            // ESCAPED: the far field by direction (farfield.js farBin: gamma from +z, C about it from +x towards +y)
            const float g = acosf(fminf(fmaxf(d.z, -1.0f), 1.0f));
            float c = atan2f(d.y, d.x);
            if (c < 0.0f) c += 2.0f * 3.14159265f;
            const int ig = min((int)floorf(g * (1.0f / 3.14159265f) * (PT_OPT_NG - 1) + 0.5f), PT_OPT_NG - 1);
            const int ic = (int)floorf(c * (1.0f / (2.0f * 3.14159265f)) * PT_OPT_NC + 0.5f) % PT_OPT_NC;
            atomicAdd(&D.far[(size_t)s * PT_OPT_NG * PT_OPT_NC + ig * PT_OPT_NC + ic], (double)wY);
            fl += OPF_ESCAPE;
            // the lux map: where the line crosses z = 1 m
            if (d.z > 0.0f && o.z < 1.0f) {
                const float tt = (1.0f - o.z) / d.z;
                const float x = o.x + d.x * tt, y = o.y + d.y * tt;
                fl += OPF_LUX;
                if (fabsf(x) < 1.0f && fabsf(y) < 1.0f) {
                    const int cx = min((int)((x + 1.0f) * (0.5f * PT_OPT_LUX)), PT_OPT_LUX - 1);
                    const int cy = min((int)((y + 1.0f) * (0.5f * PT_OPT_LUX)), PT_OPT_LUX - 1);
                    atomicAdd(&D.lux[(size_t)s * PT_OPT_LUX * PT_OPT_LUX + cy * PT_OPT_LUX + cx], (double)wY);
                }
            }
            acc[0] += wX; acc[1] += wY; acc[2] += wZ;
            return;
        }
        o = o + d * tBest;
        fl += OPF_MOVE;
        if (what == 2 || what == 4) { acc[3] += wY; return; }   // rim and board absorb
        if (what == 3) {
            // specular reflection off the paraboloid, tinted
            const F3 n = norm3(f3(2.0f * o.x, 2.0f * o.y, -f4));
            d = norm3(d - n * (2.0f * dot3(d, n)));
            acc[3] += wY * (1.0f - S.rho[1]);   // what the mirror keeps
            wX *= S.rho[0]; wY *= S.rho[1]; wZ *= S.rho[2];
            fl += OPF_MIRROR;
            continue;
        }
        // a lens face: Fresnel's choice between reflection and refraction
        const float cf = what == 0 ? S.c1 : S.c2;
        F3 n = what == 0 ? conicNormal(o, zv1, S.c1, S.k1) : conicNormal(o, zv2, S.c2, S.k2);
        // the outward normal: the front face's points away from the lens (-z), the back face's to +z
        if (what == 0) n = n * -1.0f;
        float cosI = -dot3(d, n);
        const bool entering = cosI > 0.0f;       // arriving from outside
        if (!entering) { n = n * -1.0f; cosI = -cosI; }
        const float eta = entering ? 1.0f / S.ior : S.ior;
        float cosT;
        const float F = fresnelU(cosI, eta, cosT);
        ORng rr = orngAt(S.seed, i, (unsigned int)b + 1u);
        const float4 uf = onext(rr);
        fl += OPF_FRESNEL + (cf != 0.0f ? OPF_NORMAL : 0.0f);
        if (uf.x < F) {
            d = norm3(d + n * (2.0f * cosI));
            fl += OPF_REFLECT;
        } else {
            d = norm3(d * eta + n * (eta * cosI - cosT));
            inLens = entering;
            fl += OPF_REFRACT;
        }
    }
    acc[3] += wY;   // out of bounces: counted as absorbed
}

// Each wave claims chunks until none is left. The claim has no branch on the lane: every lane executes the atomic
// (lane 0 adds 1, the others 0) and the chunk is lane 0's return read into a scalar register, so the loop's exit is a
// scalar branch that the whole wave takes together. (The form `if (lane == 0) c = atomicAdd(..); c = __shfl(c, 0)`
// lets the compiler thread lane 0's branch round the loop apart from the other lanes', whose broadcast then reads
// lane 0 while it is masked off: a chunk index below the total, forever.) A chunk's sums go to its own row of
// `part` with plain stores; the numbers kernel adds a scene's rows in a fixed order.
extern "C" __global__ __launch_bounds__(256) void pt_optics_walk(OptDev D) {
    const unsigned long long total = D.chunkStart[D.n];
    double acc[6];
    for (;;) {
        const unsigned int mine = atomicAdd(D.next, laneId() == 0u ? 1u : 0u);
        const unsigned int chunk = __builtin_amdgcn_readfirstlane(mine);
        if ((unsigned long long)chunk >= total) break;
        // the chunk's scene
        unsigned int lo = 0u, hi = D.n;
        while (hi - lo > 1u) { unsigned int mid = (lo + hi) / 2u; if (D.chunkStart[mid] <= chunk) lo = mid; else hi = mid; }
        const unsigned int s = lo;
        const PtOpticsScene S = D.sc[s];
        const unsigned int first = (unsigned int)(chunk - D.chunkStart[s]) * PT_OPT_CHUNK;
        const unsigned int last = min(first + PT_OPT_CHUNK, S.photons);
        acc[0] = acc[1] = acc[2] = acc[3] = acc[4] = acc[5] = 0.0;
        for (unsigned int i = first + laneId(); i < last; i += 64u) {
            float ev = 0.0f, fl = 0.0f;
            tracePhoton(D, s, S, i, acc, ev, fl);
            acc[4] += ev; acc[5] += fl;
        }
        double v[6];
        for (int k = 0; k < 6; ++k) v[k] = waveSumD(acc[k]);
        // lanes 0..6 store the chunk's row: X, Y, Z escaped, Y absorbed, photons, surface tests, FP32 operations
        const unsigned int l = laneId();
        const double mineV = l == 0u ? v[0] : l == 1u ? v[1] : l == 2u ? v[2] : l == 3u ? v[3]
                           : l == 4u ? (double)(last - first) : l == 5u ? v[4] : v[5];
        if (l < 7u) D.part[(size_t)chunk * PT_OPT_TOT + l] = mineV;
    }
}

// ---------------------------------------------------------------------------------------------------
// THE NUMBERS, FP64, one block a scene (farfield.js intensity and farStats; detmap.js; colorimetry.js xyToCctDuv).
// The bench's luminaires are round about the axis, so the intensity is read with the reference's rotational symmetry
// (farfield.js intensity, symmetry 'rotational'): every gamma row's C bins pooled, its flux over its solid angle --
// a 1 x 1 degree bin near the axis holds a few photons of a 2^20 budget, and its maximum would be the noise's.
__device__ double rowOmega(int j) {
    const double dG = 180.0 / (PT_OPT_NG - 1), rad = 3.14159265358979323846 / 180.0;
    const double lo = fmax(0.0, (j - 0.5) * dG) * rad, hi = fmin(180.0, (j + 0.5) * dG) * rad;
    return 2.0 * 3.14159265358979323846 * (cos(lo) - cos(hi));
}
// The plane curve (gamma -180..180; the same in every C plane once pooled) falls to `level` walking out from its peak:
// the full width, or -1 when it never does
__device__ double planeWidth(const double* I, double level) {
    const int np = 2 * PT_OPT_NG - 1;
    auto pt = [&](int k, double& a, double& v) {
        const int j = k < PT_OPT_NG - 1 ? PT_OPT_NG - 1 - k : k - (PT_OPT_NG - 1);
        a = k < PT_OPT_NG - 1 ? -(double)j : (double)j;
        v = I[j];
    };
    int p = 0;
    double ap, Ip;
    pt(0, ap, Ip);
    for (int k = 1; k < np; ++k) { double a, v; pt(k, a, v); if (v > Ip) { Ip = v; p = k; } }
    if (!(Ip >= level) || !(level > 0.0)) return -1.0;
    double ends[2];
    for (int side = 0; side < 2; ++side) {
        const int dir = side == 0 ? -1 : 1;
        ends[side] = 1e30;
        double aA, IA;
        pt(p, aA, IA);
        for (int k = p; k + dir >= 0 && k + dir < np; k += dir) {
            double aB, IB;
            pt(k + dir, aB, IB);
            if (IB < level) {
                const double t = IA == IB ? 0.0 : (IA - level) / (IA - IB);
                ends[side] = aA + (aB - aA) * t;
                break;
            }
            aA = aB; IA = IB;
        }
        if (ends[side] > 1e29) return -1.0;
    }
    return ends[1] - ends[0];
}
// The Planckian locus in CIE 1960 uv by Krystek's 1985 fit, each coordinate a ratio of quadratics in T over
// 1000..15000 K, with its direction dP/dT from the ratios' derivatives (N'D - N D') / D^2.
__device__ void krystek(const double* c, double t, double& val, double& slope) {
    const double n = c[0] + t * (c[1] + t * c[2]), d = c[3] + t * (c[4] + t * c[5]);
    val = n / d;
    slope = ((c[1] + 2.0 * t * c[2]) * d - n * (c[4] + 2.0 * t * c[5])) / (d * d);
}
__device__ void locusAt(double T, double& u, double& v, double& su, double& sv) {
    const double KU[6] = {0.860117757, 1.54118254e-4, 1.28641212e-7, 1.0, 8.42420234e-4, 7.08145163e-7};
    const double KV[6] = {0.317398726, 4.22806245e-5, 4.20481691e-8, 1.0, -2.89741816e-5, 1.61456053e-7};
    const double t = fmin(15000.0, fmax(1000.0, T));
    krystek(KU, t, u, su);
    krystek(KV, t, v, sv);
}
// CCT and Duv of a chromaticity. The locus point nearest (u, v) is where the offset from it is square to the locus:
// g(T) = (P(T) - uv) . P'(T) = 0, g rising through zero. g is sampled at 48 temperatures even in mired, each bracket
// where it turns from negative to positive is closed by bisection to double precision, and the nearest of those feet
// and the two ends of the range is the CCT. Duv is the offset along the locus's unit normal, the tangent turned a
// quarter clockwise: positive above the locus (towards green).
__device__ void cctDuv(double x, double y, double& cct, double& duv) {
    const double den = -2.0 * x + 12.0 * y + 3.0;
    const double u = 4.0 * x / den, v = 6.0 * y / den;
    auto g = [&](double T) { double a, b, sa, sb; locusAt(T, a, b, sa, sb); return (a - u) * sa + (b - v) * sb; };
    auto dist2 = [&](double T) { double a, b, sa, sb; locusAt(T, a, b, sa, sb); return (a - u) * (a - u) + (b - v) * (b - v); };
    const double Mhot = 1e6 / 15000.0, Mcold = 1e6 / 1000.0;
    const int K = 48;
    double bestT = 15000.0, bestD = dist2(15000.0);
    if (dist2(1000.0) < bestD) { bestT = 1000.0; bestD = dist2(1000.0); }
    // walking from hot to cold, T falls: a foot is where g turns from positive to negative along the walk
    double Tprev = 15000.0, gprev = g(Tprev);
    for (int k = 1; k <= K; ++k) {
        const double Tk = 1e6 / (Mhot + (Mcold - Mhot) * k / K), gk = g(Tk);
        if (gprev > 0.0 && gk <= 0.0) {
            double lo = Tk, hi = Tprev;            // g(lo) <= 0 < g(hi)
            for (int it = 0; it < 64; ++it) {
                const double mid = 0.5 * (lo + hi);
                if (g(mid) > 0.0) hi = mid; else lo = mid;
            }
            const double T = 0.5 * (lo + hi), dd = dist2(T);
            if (dd < bestD) { bestD = dd; bestT = T; }
        }
        Tprev = Tk; gprev = gk;
    }
    double pu, pv, su, sv;
    locusAt(bestT, pu, pv, su, sv);
    const double l = sqrt(su * su + sv * sv);
    cct = bestT;
    duv = ((u - pu) * sv - (v - pv) * su) / l;
}

extern "C" __global__ __launch_bounds__(256) void pt_optics_numbers(OptDev D) {
    const unsigned int s = blockIdx.x;
    if (s >= D.n) return;
    const double* far = D.far + (size_t)s * PT_OPT_NG * PT_OPT_NC;
    double* lux = D.lux + (size_t)s * PT_OPT_LUX * PT_OPT_LUX;
    __shared__ double rowI[PT_OPT_NG];
    __shared__ double bestL[256];
    __shared__ double sums[7][256];
    // the lux map: lm per cell over its area, in m^2 (cells of 2 / LUX metres)
    const double area = (2.0 / PT_OPT_LUX) * (2.0 / PT_OPT_LUX);
    double lm = 0.0;
    for (int k = threadIdx.x; k < PT_OPT_LUX * PT_OPT_LUX; k += blockDim.x) {
        lux[k] /= area;
        lm = fmax(lm, lux[k]);
    }
    bestL[threadIdx.x] = lm;
    // the scene's chunk sums, each thread a strided share of its rows
    {
        double t[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (unsigned long long c = D.chunkStart[s] + threadIdx.x; c < D.chunkStart[s + 1]; c += blockDim.x)
            for (int k = 0; k < 7; ++k) t[k] += D.part[c * PT_OPT_TOT + k];
        for (int k = 0; k < 7; ++k) sums[k][threadIdx.x] = t[k];
    }
    // each gamma row's intensity, its C bins pooled
    for (int j = threadIdx.x; j < PT_OPT_NG; j += blockDim.x) {
        double f = 0.0;
        for (int q = 0; q < PT_OPT_NC; ++q) f += far[j * PT_OPT_NC + q];
        rowI[j] = f / rowOmega(j);
    }
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            bestL[threadIdx.x] = fmax(bestL[threadIdx.x], bestL[threadIdx.x + w]);
            for (int k = 0; k < 7; ++k) sums[k][threadIdx.x] += sums[k][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    PtOpticsResult R;
    memset(&R, 0, sizeof R);
    const PtOpticsScene S = D.sc[s];
    const double tot[7] = {sums[0][0], sums[1][0], sums[2][0], sums[3][0], sums[4][0], sums[5][0], sums[6][0]};
    R.emitted = S.flux;
    R.escaped = tot[1];
    R.absorbed = tot[3];
    R.efficiency = tot[1] / S.flux;
    int bj = 0;
    for (int j = 1; j < PT_OPT_NG; ++j) if (rowI[j] > rowI[bj]) bj = j;
    R.Imax = rowI[bj];
    R.gammaMax = (double)bj * (180.0 / (PT_OPT_NG - 1));
    R.cMax = 0.0;
    R.beam[0] = R.beam[1] = planeWidth(rowI, 0.5 * R.Imax);
    R.field[0] = R.field[1] = planeWidth(rowI, 0.1 * R.Imax);
    R.X = tot[0]; R.Y = tot[1]; R.Z = tot[2];
    const double sum = R.X + R.Y + R.Z;
    R.x = sum > 0.0 ? R.X / sum : 0.0;
    R.y = sum > 0.0 ? R.Y / sum : 0.0;
    if (sum > 0.0) cctDuv(R.x, R.y, R.cct, R.duv);
    R.luxMax = bestL[0];
    const int h = PT_OPT_LUX / 2;
    R.luxCentre = 0.25 * (lux[(h - 1) * PT_OPT_LUX + h - 1] + lux[(h - 1) * PT_OPT_LUX + h] + lux[h * PT_OPT_LUX + h - 1] + lux[h * PT_OPT_LUX + h]);
    R.photons = tot[4];
    R.events = tot[5];
    R.flops = tot[6];
    D.res[s] = R;
}

thread_local std::string g_optErr;

int ofail(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_optErr = buf;
    return PT_ERR_ARG;
}

// ---- the optics bench without allocation (pt_optics_trace_static): every device buffer pt_optics_trace allocates,
// laid out in one array of device memory the library holds for the process (each GPU its own copy), for callers that
// cannot allocate. Offsets 256-byte aligned; sizes for PT_OPT_STATIC_SCENES scenes and PT_OPT_STATIC_CHUNKS chunks.
constexpr size_t alignUp(size_t x) { return (x + 255) & ~size_t(255); }
constexpr size_t kStSc = 0;
constexpr size_t kStStart = alignUp(kStSc + PT_OPT_STATIC_SCENES * sizeof(PtOpticsScene));
constexpr size_t kStNext = alignUp(kStStart + (PT_OPT_STATIC_SCENES + 1) * 8);
constexpr size_t kStFar = alignUp(kStNext + 4);
constexpr size_t kStLux = alignUp(kStFar + size_t(PT_OPT_STATIC_SCENES) * PT_OPT_NG * PT_OPT_NC * 8);
constexpr size_t kStPart = alignUp(kStLux + size_t(PT_OPT_STATIC_SCENES) * PT_OPT_LUX * PT_OPT_LUX * 8);
constexpr size_t kStRes = alignUp(kStPart + size_t(PT_OPT_STATIC_CHUNKS) * PT_OPT_TOT * 8);
constexpr size_t kStBytes = alignUp(kStRes + PT_OPT_STATIC_SCENES * sizeof(PtOpticsResult));
static_assert(kStBytes <= PT_OPT_STATIC_BYTES, "the static optics buffer outgrew its stated size");

}  // namespace

// the array itself (host code sees only a stand-in of its name; its device address comes from hipGetSymbolAddress)
static __device__ __align__(256) unsigned char scratch[PT_OPT_STATIC_BYTES];

namespace {
// its address on each device, taken once with that device current; calls on one device take the lock, so that two
// callers in this process never share it at once
constexpr int kMaxDevices = 64;
std::mutex g_scratchLock[kMaxDevices];
unsigned char* g_scratchAt[kMaxDevices] = {};
}  // namespace

extern "C" {

const char* pt_optics_error(void) { return g_optErr.c_str(); }

int pt_optics_trace_static(int device, const PtOpticsScene* scenes, uint32_t n, PtOpticsResult* out, double* far, double* lux) {
    if (device < 0 || device >= kMaxDevices) return ofail("device %d not in 0..%d", device, kMaxDevices - 1);
    if (!scenes || !out || n == 0 || n > PT_OPT_STATIC_SCENES) return ofail("scenes/out NULL or n %u not in 1..%u", n, PT_OPT_STATIC_SCENES);
    unsigned long long start[PT_OPT_STATIC_SCENES + 1];
    start[0] = 0;
    for (uint32_t s = 0; s < n; ++s) {
        const PtOpticsScene& S = scenes[s];
        if (S.photons == 0 || S.photons > (1u << 30)) return ofail("scene %u: photons %u not in 1..2^30", s, S.photons);
        if (!(S.flux > 0.0f) || !(S.ledHalf > 0.0f) || !(S.ior >= 1.0f) || !(S.lensD > 0.0f) || !(S.lensT > 0.0f))
            return ofail("scene %u: flux, LED size, lens diameter and thickness must be positive, ior >= 1", s);
        start[s + 1] = start[s] + (S.photons + PT_OPT_CHUNK - 1) / PT_OPT_CHUNK;
    }
    if (start[n] > PT_OPT_STATIC_CHUNKS)
        return ofail("%llu chunks of %u photons over the scenes, more than the %u the static buffer holds", start[n], PT_OPT_CHUNK,
                     PT_OPT_STATIC_CHUNKS);
    std::lock_guard<std::mutex> hold(g_scratchLock[device]);
    hipError_t e;
#define OCHK(x) do { if ((e = (x)) != hipSuccess) { g_optErr = std::string(#x) + ": " + hipGetErrorString(e); return PT_ERR_HIP; } } while (0)
    OCHK(hipSetDevice(device));
    if (!g_scratchAt[device]) {
        void* p = nullptr;
        OCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratchAt[device] = static_cast<unsigned char*>(p);
    }
    unsigned char* base = g_scratchAt[device];
    int cus = 256;
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    const size_t farW = (size_t)n * PT_OPT_NG * PT_OPT_NC, luxW = (size_t)n * PT_OPT_LUX * PT_OPT_LUX;
    const size_t partW = (size_t)start[n] * PT_OPT_TOT;
    OptDev D = {};
    D.sc = reinterpret_cast<const PtOpticsScene*>(base + kStSc); D.n = n;
    D.chunkStart = reinterpret_cast<const unsigned long long*>(base + kStStart);
    D.next = reinterpret_cast<unsigned int*>(base + kStNext);
    D.far = reinterpret_cast<double*>(base + kStFar); D.lux = reinterpret_cast<double*>(base + kStLux);
    D.part = reinterpret_cast<double*>(base + kStPart); D.res = reinterpret_cast<PtOpticsResult*>(base + kStRes);
    // the last call's tallies are still there: what is summed into starts at zero
    OCHK(hipMemcpy(base + kStSc, scenes, n * sizeof(PtOpticsScene), hipMemcpyHostToDevice));
    OCHK(hipMemcpy(base + kStStart, start, (n + 1) * 8, hipMemcpyHostToDevice));
    OCHK(hipMemsetAsync(base + kStNext, 0, 4, 0));
    OCHK(hipMemsetAsync(base + kStFar, 0, farW * 8, 0));
    OCHK(hipMemsetAsync(base + kStLux, 0, luxW * 8, 0));
    OCHK(hipMemsetAsync(base + kStPart, 0, partW * 8, 0));
    hipLaunchKernelGGL(pt_optics_walk, dim3(cus * 8), dim3(256), 0, 0, D);
    OCHK(hipGetLastError());
    hipLaunchKernelGGL(pt_optics_numbers, dim3(n), dim3(256), 0, 0, D);
    OCHK(hipGetLastError());
    OCHK(hipMemcpy(out, base + kStRes, n * sizeof(PtOpticsResult), hipMemcpyDeviceToHost));
    if (far) OCHK(hipMemcpy(far, base + kStFar, farW * 8, hipMemcpyDeviceToHost));
    if (lux) OCHK(hipMemcpy(lux, base + kStLux, luxW * 8, hipMemcpyDeviceToHost));
    return PT_OK;
#undef OCHK
}

int pt_optics_trace(int device, const PtOpticsScene* scenes, uint32_t n, PtOpticsResult* out, double* far, double* lux,
                    double* ms) {
    if (!scenes || !out || n == 0 || n > 4096) return ofail("scenes/out NULL or n %u not in 1..4096", n);
    std::vector<unsigned long long> start(n + 1, 0);
    for (uint32_t s = 0; s < n; ++s) {
        const PtOpticsScene& S = scenes[s];
        if (S.photons == 0 || S.photons > (1u << 30)) return ofail("scene %u: photons %u not in 1..2^30", s, S.photons);
        if (!(S.flux > 0.0f) || !(S.ledHalf > 0.0f) || !(S.ior >= 1.0f) || !(S.lensD > 0.0f) || !(S.lensT > 0.0f))
            return ofail("scene %u: flux, LED size, lens diameter and thickness must be positive, ior >= 1", s);
        start[s + 1] = start[s] + (S.photons + PT_OPT_CHUNK - 1) / PT_OPT_CHUNK;
    }
    hipError_t e;
#define OCHK(x) do { if ((e = (x)) != hipSuccess) { g_optErr = std::string(#x) + ": " + hipGetErrorString(e); goto done; } } while (0)
    int rc = PT_ERR_HIP;
    OptDev D = {};
    void *dsc = nullptr, *dstart = nullptr, *dnext = nullptr, *dfar = nullptr, *dlux = nullptr, *dpart = nullptr, *dres = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    const size_t farW = (size_t)n * PT_OPT_NG * PT_OPT_NC, luxW = (size_t)n * PT_OPT_LUX * PT_OPT_LUX;
    const size_t partW = (size_t)start[n] * PT_OPT_TOT;
    int cus = 256;
    OCHK(hipSetDevice(device));
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    OCHK(hipMalloc(&dsc, n * sizeof(PtOpticsScene)));
    OCHK(hipMalloc(&dstart, (n + 1) * 8));
    OCHK(hipMalloc(&dnext, 4));
    OCHK(hipMalloc(&dfar, farW * 8));
    OCHK(hipMalloc(&dlux, luxW * 8));
    OCHK(hipMalloc(&dpart, partW * 8));
    OCHK(hipMalloc(&dres, n * sizeof(PtOpticsResult)));
    OCHK(hipMemcpy(dsc, scenes, n * sizeof(PtOpticsScene), hipMemcpyHostToDevice));
    OCHK(hipMemcpy(dstart, start.data(), (n + 1) * 8, hipMemcpyHostToDevice));
    OCHK(hipMemset(dnext, 0, 4));
    OCHK(hipMemset(dfar, 0, farW * 8));
    OCHK(hipMemset(dlux, 0, luxW * 8));
    OCHK(hipMemset(dpart, 0, partW * 8));
    D.sc = (const PtOpticsScene*)dsc; D.n = n; D.chunkStart = (const unsigned long long*)dstart; D.next = (unsigned int*)dnext;
    D.far = (double*)dfar; D.lux = (double*)dlux; D.part = (double*)dpart; D.res = (PtOpticsResult*)dres;
    OCHK(hipEventCreate(&e0));
    OCHK(hipEventCreate(&e1));
    OCHK(hipEventCreate(&e2));
    OCHK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(pt_optics_walk, dim3(cus * 8), dim3(256), 0, 0, D);
    OCHK(hipGetLastError());
    OCHK(hipEventRecord(e1, 0));
    hipLaunchKernelGGL(pt_optics_numbers, dim3(n), dim3(256), 0, 0, D);
    OCHK(hipGetLastError());
    OCHK(hipEventRecord(e2, 0));
    OCHK(hipEventSynchronize(e2));
    if (ms) {
        float t = 0.0f;
        hipEventElapsedTime(&t, e0, e1); ms[0] = t;
        hipEventElapsedTime(&t, e1, e2); ms[1] = t;
    }
    OCHK(hipMemcpy(out, dres, n * sizeof(PtOpticsResult), hipMemcpyDeviceToHost));
    if (far) OCHK(hipMemcpy(far, dfar, farW * 8, hipMemcpyDeviceToHost));
    if (lux) OCHK(hipMemcpy(lux, dlux, luxW * 8, hipMemcpyDeviceToHost));
    rc = PT_OK;
done:
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (e2) hipEventDestroy(e2);
    for (void* p : {dsc, dstart, dnext, dfar, dlux, dpart, dres}) if (p) hipFree(p);
    return rc;
#undef OCHK
}

}  // extern "C"
