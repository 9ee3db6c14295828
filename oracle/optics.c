/*
 * THE OPTICS ORACLE -- test infrastructure only: an FP64 CPU restatement of the optics bench (the engine's
 * pt_optics_trace), used by tests/ and bench.py's cpu_baseline leg as the checker; nothing in the product calls it.
 *
 * It walks each photon as the engine does -- the same PCG4D keying (scene seed, photon, bounce), so the same photons
 * -- in double precision: a Lambertian square LED, a lens of two conic faces with an absorbing rim, an optional
 * paraboloid reflector, an absorbing board (the reference's optics/kernels.js:43-87 prog_opt_gen / prog_opt_shade
 * over the renderer's surfaces). It then reads the numbers off the tallies (any tallies: its own, or the engine's):
 * intensity per gamma row with the C bins pooled (the luminaires are round: optics/farfield.js:45-96 farGeometry,
 * intensity with symmetry 'rotational'), the brightest row, beam and field angles as full widths at 50 % and 10 % of
 * I max on the plane curve walked out from its peak (farfield.js:98-168 planeCurve, crossing, farStats), lux as flux over cell area (optics/detmap.js), CCT by the
 * nearest point of the Planckian locus in CIE 1960 uv and Duv as the signed distance (core/colorimetry.js:58-104
 * xyToCctDuv, with Krystek 1985's rational locus).
 *
 * Pinned by (tests/test_optics_oracle.py): a bare Lambertian LED's far field, I(gamma) = (flux / pi) cos(gamma), its
 * beam angle 120 deg and field angle 2 acos(0.1); the lux on axis at 1 m, I0 / 1 m^2; flux balance (escaped +
 * absorbed = emitted); a thin plano-convex lens's focus (f = R / (n - 1)) collimating a point source; published CCTs
 * of CIE illuminants A (2856 K) and D65 (6504 K) from their chromaticities; Fresnel at normal incidence.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/pt_types.h"

#ifdef _OPENMP
#include <omp.h>
#endif

#define OPT_BOUNCES 32
#define OPT_PI 3.14159265358979323846

typedef struct { uint32_t a, b, c, d; } ORng;
static void opcg(ORng* s) {
    s->a = s->a * 1664525u + 1013904223u; s->b = s->b * 1664525u + 1013904223u;
    s->c = s->c * 1664525u + 1013904223u; s->d = s->d * 1664525u + 1013904223u;
    s->a += s->b * s->d; s->b += s->c * s->a; s->c += s->a * s->b; s->d += s->b * s->c;
    s->a ^= s->a >> 16; s->b ^= s->b >> 16; s->c ^= s->c >> 16; s->d ^= s->d >> 16;
    s->a += s->b * s->d; s->b += s->c * s->a; s->c += s->a * s->b; s->d += s->b * s->c;
}
static void onext(ORng* s, double* u) {
    opcg(s);
    u[0] = (double)(s->a >> 8) / 16777216.0; u[1] = (double)(s->b >> 8) / 16777216.0;
    u[2] = (double)(s->c >> 8) / 16777216.0; u[3] = (double)(s->d >> 8) / 16777216.0;
}
static ORng orngAt(uint32_t seed, uint32_t photon, uint32_t bounce) {
    ORng s = {seed ^ 0x0F71C5u, photon, bounce * 0x9E3779B9u, 0x5EED0F71u};
    opcg(&s);
    return s;
}

typedef struct { double x, y, z; } D3;
static D3 d3(double x, double y, double z) { D3 r = {x, y, z}; return r; }
static D3 add3(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
static D3 scl3(D3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
static double dt3(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static D3 nrm3(D3 a) { return scl3(a, 1.0 / sqrt(dt3(a, a))); }

static double sag(double c, double k, double r2) {
    double q = 1.0 - (1.0 + k) * c * c * r2;
    return c * r2 / (1.0 + sqrt(q > 0 ? q : 0));
}
static double hitConic(D3 o, D3 d, double zv, double c, double k, double rmax, double tmin) {
    const double R2 = rmax * rmax;
    if (c == 0.0) {
        if (fabs(d.z) < 1e-15) return 1e300;
        double t = (zv - o.z) / d.z;
        D3 p = add3(o, scl3(d, t));
        return (t > tmin && p.x * p.x + p.y * p.y <= R2) ? t : 1e300;
    }
    const double Z = o.z - zv, ck = c * (1.0 + k);
    const double A = c * (d.x * d.x + d.y * d.y) + ck * d.z * d.z;
    const double B = 2.0 * (c * (o.x * d.x + o.y * d.y) - d.z + ck * Z * d.z);
    const double C = c * (o.x * o.x + o.y * o.y) - 2.0 * Z + ck * Z * Z;
    double t0, t1;
    if (fabs(A) < 1e-15) {
        if (fabs(B) < 1e-30) return 1e300;
        t0 = t1 = -C / B;
    } else {
        double disc = B * B - 4.0 * A * C;
        if (disc < 0) return 1e300;
        double sq = sqrt(disc), q = -0.5 * (B + (B >= 0 ? sq : -sq));
        t0 = q / A; t1 = fabs(q) > 0 ? C / q : t0;
        if (t0 > t1) { double s = t0; t0 = t1; t1 = s; }
    }
    for (int i = 0; i < 2; ++i) {
        double t = i == 0 ? t0 : t1;
        if (!(t > tmin)) continue;
        D3 p = add3(o, scl3(d, t));
        double r2 = p.x * p.x + p.y * p.y;
        if (r2 > R2) continue;
        if (fabs((p.z - zv) - sag(c, k, r2)) > 1e-4 * (1.0 + fabs(1.0 / c))) continue;
        return t;
    }
    return 1e300;
}
static D3 conicNormal(D3 p, double zv, double c, double k) {
    if (c == 0.0) return d3(0, 0, 1);
    double Z = p.z - zv;
    return nrm3(d3(-2 * c * p.x, -2 * c * p.y, 2 - 2 * c * (1 + k) * Z));
}
double or_fresnel_unpolarised(double cosI, double eta, double* cosT) {
    double s2 = eta * eta * (1 - cosI * cosI);
    if (s2 >= 1) { *cosT = 0; return 1; }
    *cosT = sqrt(1 - s2);
    double rs = (eta * cosI - *cosT) / (eta * cosI + *cosT), rp = (cosI - eta * *cosT) / (cosI + eta * *cosT);
    return 0.5 * (rs * rs + rp * rp);
}

/* one photon; far (NG x NC) and lux (LUX x LUX) in lm, acc: X, Y, Z escaped, Y absorbed, surface tests */
static void tracePhoton(const PtOpticsScene* S, uint32_t i, double* far, double* lux, double* acc) {
    double u[4];
    ORng r = orngAt(S->seed, i, 0);
    onext(&r, u);
    /* the engine draws in float: the same 24-bit uniforms */
    D3 o = d3(S->ledHalf * (2 * u[0] - 1), S->ledHalf * (2 * u[1] - 1), 0);
    double sr = sqrt(u[2]), phi = 2 * OPT_PI * u[3];
    D3 d = d3(sr * cos(phi), sr * sin(phi), sqrt(1 - u[2] > 0 ? 1 - u[2] : 0));
    double each = (double)S->flux / S->photons;
    double wX = S->xyz[0] * each, wY = S->xyz[1] * each, wZ = S->xyz[2] * each;
    double zv1 = S->lensZ, zv2 = (double)S->lensZ + S->lensT, rim = 0.5 * S->lensD;
    double zr1 = zv1 + sag(S->c1, S->k1, rim * rim), zr2 = zv2 + sag(S->c2, S->k2, rim * rim);
    int refl = S->reflF > 0;
    double f4 = 4.0 * S->reflF, zb = refl ? (double)S->reflR0 * S->reflR0 / f4 - S->reflF : 0;
    int inLens = 0;
    const double eps = 1e-6;
    for (int b = 0; b < OPT_BOUNCES; ++b) {
        double tBest = 1e300, t;
        int what = -1;
        acc[4] += 1;
        t = hitConic(o, d, zv1, S->c1, S->k1, rim, eps); if (t < tBest) { tBest = t; what = 0; }
        t = hitConic(o, d, zv2, S->c2, S->k2, rim, eps); if (t < tBest) { tBest = t; what = 1; }
        {
            double A = d.x * d.x + d.y * d.y, B = 2 * (o.x * d.x + o.y * d.y), C = o.x * o.x + o.y * o.y - rim * rim;
            double disc = B * B - 4 * A * C;
            if (A > 1e-15 && disc >= 0) {
                double sq = sqrt(disc);
                for (int k = 0; k < 2; ++k) {
                    double tt = (-B + (k == 0 ? -sq : sq)) / (2 * A), z = o.z + d.z * tt;
                    double zlo = zr1 < zr2 ? zr1 : zr2, zhi = zr1 < zr2 ? zr2 : zr1;
                    if (tt > eps && tt < tBest && z >= zlo && z <= zhi) { tBest = tt; what = 2; break; }
                }
            }
        }
        if (refl && !inLens) {
            double A = d.x * d.x + d.y * d.y, B = 2 * (o.x * d.x + o.y * d.y) - f4 * d.z;
            double C = o.x * o.x + o.y * o.y - f4 * (o.z + S->reflF);
            double t0 = 1e300, t1 = 1e300;
            if (fabs(A) < 1e-15) { if (fabs(B) > 1e-30) t0 = t1 = -C / B; }
            else {
                double disc = B * B - 4 * A * C;
                if (disc >= 0) {
                    double sq = sqrt(disc), q = -0.5 * (B + (B >= 0 ? sq : -sq));
                    t0 = q / A; t1 = fabs(q) > 0 ? C / q : t0;
                    if (t0 > t1) { double x = t0; t0 = t1; t1 = x; }
                }
            }
            for (int k = 0; k < 2; ++k) {
                double tt = k == 0 ? t0 : t1, z = o.z + d.z * tt;
                if (tt > eps && tt < tBest && z >= zb && z <= S->reflZ1) { tBest = tt; what = 3; break; }
            }
        }
        if (d.z < 0 && !inLens) {
            double tt = -o.z / d.z;
            D3 p = add3(o, scl3(d, tt));
            if (tt > eps && tt < tBest && p.x * p.x + p.y * p.y <= (double)S->boardR * S->boardR) { tBest = tt; what = 4; }
        }
        if (what < 0) {
            double g = acos(d.z < -1 ? -1 : (d.z > 1 ? 1 : d.z));
            double c = atan2(d.y, d.x);
            if (c < 0) c += 2 * OPT_PI;
            int ig = (int)floor(g / OPT_PI * (PT_OPT_NG - 1) + 0.5);
            if (ig > PT_OPT_NG - 1) ig = PT_OPT_NG - 1;
            int ic = (int)floor(c / (2 * OPT_PI) * PT_OPT_NC + 0.5) % PT_OPT_NC;
            #pragma omp atomic
            far[ig * PT_OPT_NC + ic] += wY;
            if (d.z > 0 && o.z < 1) {
                double tt = (1 - o.z) / d.z, x = o.x + d.x * tt, y = o.y + d.y * tt;
                if (fabs(x) < 1 && fabs(y) < 1) {
                    int cx = (int)((x + 1) * (0.5 * PT_OPT_LUX)), cy = (int)((y + 1) * (0.5 * PT_OPT_LUX));
                    if (cx > PT_OPT_LUX - 1) cx = PT_OPT_LUX - 1;
                    if (cy > PT_OPT_LUX - 1) cy = PT_OPT_LUX - 1;
                    #pragma omp atomic
                    lux[cy * PT_OPT_LUX + cx] += wY;
                }
            }
            acc[0] += wX; acc[1] += wY; acc[2] += wZ;
            return;
        }
        o = add3(o, scl3(d, tBest));
        if (what == 2 || what == 4) { acc[3] += wY; return; }
        if (what == 3) {
            D3 n = nrm3(d3(2 * o.x, 2 * o.y, -f4));
            d = nrm3(add3(d, scl3(n, -2 * dt3(d, n))));
            acc[3] += wY * (1 - S->rho[1]);
            wX *= S->rho[0]; wY *= S->rho[1]; wZ *= S->rho[2];
            continue;
        }
        D3 n = what == 0 ? conicNormal(o, zv1, S->c1, S->k1) : conicNormal(o, zv2, S->c2, S->k2);
        if (what == 0) n = scl3(n, -1);
        double cosI = -dt3(d, n);
        int entering = cosI > 0;
        if (!entering) { n = scl3(n, -1); cosI = -cosI; }
        double eta = entering ? 1.0 / S->ior : S->ior, cosT;
        double F = or_fresnel_unpolarised(cosI, eta, &cosT);
        double uf[4];
        ORng rr = orngAt(S->seed, i, (uint32_t)b + 1u);
        onext(&rr, uf);
        if (uf[0] < F) d = nrm3(add3(d, scl3(n, 2 * cosI)));
        else { d = nrm3(add3(scl3(d, eta), scl3(n, eta * cosI - cosT))); inLens = entering; }
    }
    acc[3] += wY;
}

/* ---- the numbers (farfield.js, detmap.js, colorimetry.js), the intensity read with rotational symmetry (every
 * gamma row's C bins pooled: farfield.js intensity, symmetry 'rotational') ---- */
static double rowOmega(int j) {
    double dG = 180.0 / (PT_OPT_NG - 1), rad = OPT_PI / 180.0;
    double lo = ((j - 0.5) * dG > 0 ? (j - 0.5) * dG : 0) * rad, hi = ((j + 0.5) * dG < 180 ? (j + 0.5) * dG : 180) * rad;
    return 2 * OPT_PI * (cos(lo) - cos(hi));
}
/* I of each gamma row (NG doubles) from a far field of lm per bin */
void or_optics_rows(const double* far, double* I) {
    for (int j = 0; j < PT_OPT_NG; ++j) {
        double f = 0;
        for (int q = 0; q < PT_OPT_NC; ++q) f += far[j * PT_OPT_NC + q];
        I[j] = f / rowOmega(j);
    }
}
static void planePt(const double* I, int k, double* a, double* v) {
    int j = k < PT_OPT_NG - 1 ? PT_OPT_NG - 1 - k : k - (PT_OPT_NG - 1);
    *a = k < PT_OPT_NG - 1 ? -(double)j : (double)j;
    *v = I[j];
}
static double planeWidth(const double* I, double level) {
    int np = 2 * PT_OPT_NG - 1, p = 0;
    double ap, Ip;
    planePt(I, 0, &ap, &Ip);
    for (int k = 1; k < np; ++k) { double a, v; planePt(I, k, &a, &v); if (v > Ip) { Ip = v; p = k; } }
    if (!(Ip >= level) || !(level > 0)) return -1;
    double ends[2];
    for (int side = 0; side < 2; ++side) {
        int dir = side == 0 ? -1 : 1;
        ends[side] = 1e300;
        double aA, IA;
        planePt(I, p, &aA, &IA);
        for (int k = p; k + dir >= 0 && k + dir < np; k += dir) {
            double aB, IB;
            planePt(I, k + dir, &aB, &IB);
            if (IB < level) {
                double t = IA == IB ? 0 : (IA - level) / (IA - IB);
                ends[side] = aA + (aB - aA) * t;
                break;
            }
            aA = aB; IA = IB;
        }
        if (ends[side] > 1e299) return -1;
    }
    return ends[1] - ends[0];
}
static void planckUv(double T, double* u, double* v) {
    double t = T < 1000 ? 1000 : (T > 15000 ? 15000 : T);
    *u = (0.860117757 + 1.54118254e-4 * t + 1.28641212e-7 * t * t) / (1 + 8.42420234e-4 * t + 7.08145163e-7 * t * t);
    *v = (0.317398726 + 4.22806245e-5 * t + 4.20481691e-8 * t * t) / (1 - 2.89741816e-5 * t + 1.61456053e-7 * t * t);
}
static double planckD2(double T, double u, double v) { double a, b; planckUv(T, &a, &b); return (a - u) * (a - u) + (b - v) * (b - v); }
void or_cct_duv(double x, double y, double* cct, double* duv) {
    double den = -2 * x + 12 * y + 3, u = 4 * x / den, v = 6 * y / den;
    double M0 = 1e6 / 15000.0, M1 = 1e6 / 1000.0;
    int N = 200, best = 0;
    double bd = 1e300;
    for (int k = 0; k <= N; ++k) { double dd = planckD2(1e6 / (M0 + (M1 - M0) * k / N), u, v); if (dd < bd) { bd = dd; best = k; } }
    double lo = M0 + (M1 - M0) * (best - 1 > 0 ? best - 1 : 0) / N, hi = M0 + (M1 - M0) * (best + 1 < N ? best + 1 : N) / N;
    double g = (sqrt(5.0) - 1) / 2;
    for (int k = 0; k < 80; ++k) {
        double a = hi - g * (hi - lo), b = lo + g * (hi - lo);
        if (planckD2(1e6 / a, u, v) < planckD2(1e6 / b, u, v)) hi = b; else lo = a;
    }
    double T = 1e6 / ((lo + hi) / 2), pu, pv, u0, v0, u1, v1;
    planckUv(T, &pu, &pv);
    double h = T * 1e-5 > 1e-3 ? T * 1e-5 : 1e-3;
    planckUv(T - h, &u0, &v0);
    planckUv(T + h, &u1, &v1);
    double du = u1 - u0, dv = v1 - v0, l = sqrt(du * du + dv * dv);
    *cct = T;
    *duv = (u - pu) * (dv / l) + (v - pv) * (-du / l);
}
/* The numbers of one scene from its tallies: far (lm a bin), lux (lx a cell), tot (X, Y, Z escaped, Y absorbed,
 * photons). */
void or_optics_numbers(const PtOpticsScene* S, const double* far, const double* lux, const double* tot, PtOpticsResult* R) {
    memset(R, 0, sizeof *R);
    R->emitted = S->flux; R->escaped = tot[1]; R->absorbed = tot[3]; R->efficiency = tot[1] / S->flux;
    double I[PT_OPT_NG];
    or_optics_rows(far, I);
    int bj = 0;
    for (int j = 1; j < PT_OPT_NG; ++j) if (I[j] > I[bj]) bj = j;
    R->Imax = I[bj]; R->gammaMax = (double)bj * (180.0 / (PT_OPT_NG - 1)); R->cMax = 0;
    R->beam[0] = R->beam[1] = planeWidth(I, 0.5 * R->Imax);
    R->field[0] = R->field[1] = planeWidth(I, 0.1 * R->Imax);
    R->X = tot[0]; R->Y = tot[1]; R->Z = tot[2];
    double s = R->X + R->Y + R->Z;
    R->x = s > 0 ? R->X / s : 0; R->y = s > 0 ? R->Y / s : 0;
    if (s > 0) or_cct_duv(R->x, R->y, &R->cct, &R->duv);
    double m = 0;
    for (int k = 0; k < PT_OPT_LUX * PT_OPT_LUX; ++k) if (lux[k] > m) m = lux[k];
    R->luxMax = m;
    int h = PT_OPT_LUX / 2;
    R->luxCentre = 0.25 * (lux[(h - 1) * PT_OPT_LUX + h - 1] + lux[(h - 1) * PT_OPT_LUX + h] + lux[h * PT_OPT_LUX + h - 1] + lux[h * PT_OPT_LUX + h]);
    R->photons = tot[4];
}

/* The bench in FP64 on the CPU: results, and optionally the tallies (far lm, lux lx) of every scene. */
int or_optics_trace(const PtOpticsScene* scenes, int n, PtOpticsResult* out, double* farOut, double* luxOut, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    const size_t FW = (size_t)PT_OPT_NG * PT_OPT_NC, LW = (size_t)PT_OPT_LUX * PT_OPT_LUX;
    double* far = (double*)calloc(FW, sizeof(double));
    double* lux = (double*)calloc(LW, sizeof(double));
    for (int s = 0; s < n; ++s) {
        const PtOpticsScene* S = &scenes[s];
        memset(far, 0, FW * sizeof(double));
        memset(lux, 0, LW * sizeof(double));
        double tot[5] = {0, 0, 0, 0, (double)S->photons}, events = 0;
        #pragma omp parallel
        {
            double acc[5] = {0, 0, 0, 0, 0};
            #pragma omp for schedule(dynamic, 1024)
            for (int64_t i = 0; i < (int64_t)S->photons; ++i) tracePhoton(S, (uint32_t)i, far, lux, acc);
            #pragma omp critical
            { for (int k = 0; k < 4; ++k) tot[k] += acc[k]; events += acc[4]; }
        }
        double area = (2.0 / PT_OPT_LUX) * (2.0 / PT_OPT_LUX);
        for (size_t k = 0; k < LW; ++k) lux[k] /= area;
        or_optics_numbers(S, far, lux, tot, &out[s]);
        out[s].events = events;
        if (farOut) memcpy(farOut + FW * s, far, FW * sizeof(double));
        if (luxOut) memcpy(luxOut + LW * s, lux, LW * sizeof(double));
    }
    free(far); free(lux);
    return 0;
}
