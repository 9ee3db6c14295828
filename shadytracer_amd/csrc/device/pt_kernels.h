// The wavefront kernels. The emitted scene (namespace-free functions sceneDist / sceneFull and the constant
// PT_SCENE_MATERIALS) is compiled in front of this file by hipRTC.
//
// Per sample: pt_genrays, then maxbounce + 1 rounds of pt_march -> pt_shade -> pt_occlude, then pt_resolve.
// The reference's dispatch chain (core/raytrace_main.js:2043-2179) has more passes per round because its march
// stops after 32 steps and re-queues; here march and occlude are persistent kernels whose waves refill idle lanes
// from the queue, so a ray marches to its event in one launch and a round is one segment of every live path.
//
// Path state lives in structure-of-arrays by slot (16-byte lanes, coalesced when the queue holds ascending
// slots); queues hold slots; every append is wave-aggregated (one atomic per wave, ballot + prefix).
#pragma once

#define PT_BLOCK 256

// ---------------------------------------------------------------------------------------------------
// wave helpers
__device__ __forceinline__ unsigned int laneId() { return __lane_id(); }
__device__ __forceinline__ unsigned int rankIn(unsigned long long m) {
    unsigned int l = laneId();
    return (unsigned int)__popcll(m & ((1ull << l) - 1ull));
}
// An atomic add on device memory through a global (address space 1) pointer: the compiler's atomic optimizer rewrites
// these, not flat ones, into one atomic per wave (and broadcasts the result itself).
typedef __attribute__((address_space(1))) unsigned int PtGlobalU32;
__device__ __forceinline__ unsigned int gAtomicAdd(unsigned int* p, unsigned int v) {
    return __hip_atomic_fetch_add((PtGlobalU32*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Reserve one slot per predicated lane at *counter; returns this lane's index (meaningless where !pred).
// The atomic optimizer makes the predicated lanes' increments one atomic per wave, the slots in lane order; without it
// each lane's own atomic still gives it a slot of its own. No branch on a lane id precedes a cross-lane read here:
// with `if (lane == leader) base = atomicAdd(..); base = __shfl(base, leader)`, the compiler may thread the leader's
// branch round an enclosing loop apart from the other lanes, whose broadcast then reads the leader while it is
// masked off.
__device__ __forceinline__ unsigned int waveAppend(unsigned int* counter, bool pred) {
    unsigned int slot = 0u;
    if (pred) slot = gAtomicAdd(counter, 1u);
    return slot;
}
__device__ __forceinline__ unsigned long long waveSum64(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
__device__ __forceinline__ float waveSum64f(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
__device__ __forceinline__ void statAdd(const PtDev* __restrict__ D, int which, unsigned long long v) {
    v = waveSum64(v);
    if (laneId() == 0 && v) atomicAdd(&D->stats[which], v);
}

// The engine's buffers are device memory, read and written with global instructions, not flat ones: a flat access
// counts against the same counter as the scalar loads of the scene's parameters, so one in flight would hold up
// every wait of the distance function on those (a ray read ahead would stall the march it was read ahead for).
#define PT_GLOBAL __attribute__((address_space(1)))
__device__ __forceinline__ float4 ld4(const float* p, unsigned int i) { return ((const PT_GLOBAL float4*)p)[i]; }
__device__ __forceinline__ uint4 ld4u(const unsigned int* p, unsigned int i) { return ((const PT_GLOBAL uint4*)p)[i]; }
__device__ __forceinline__ float ldf(const float* p, unsigned int i) { return ((const PT_GLOBAL float*)p)[i]; }
__device__ __forceinline__ unsigned int ldu(const unsigned int* p, unsigned int i) { return ((const PT_GLOBAL unsigned int*)p)[i]; }
__device__ __forceinline__ unsigned long long ldu64(const unsigned long long* p, unsigned int i) {
    return ((const PT_GLOBAL unsigned long long*)p)[i];
}
__device__ __forceinline__ void st4(float* p, unsigned int i, float a, float b, float c, float d) {
    ((PT_GLOBAL float4*)p)[i] = make_float4(a, b, c, d);
}
__device__ __forceinline__ void st4u(unsigned int* p, unsigned int i, uint4 v) { ((PT_GLOBAL uint4*)p)[i] = v; }
__device__ __forceinline__ void stu(unsigned int* p, unsigned int i, unsigned int v) { ((PT_GLOBAL unsigned int*)p)[i] = v; }

// A wave's share of a queue, for the persistent kernels: batches of PT_GRAB consecutive entries taken by one atomic
// from one of PT_PARTS partitions of the queue; a wave starts on its own partition and moves to the next when that
// is empty, reading the head first so that an empty partition costs no atomic. Wave-uniform state.
#define PT_GRAB 64u
struct WaveFetch { unsigned int cur, end, part, left; bool done; };
__device__ __forceinline__ WaveFetch fetchInit() {
    WaveFetch f;
    unsigned int wave = (blockIdx.x * PT_BLOCK + threadIdx.x) >> 6;
    f.cur = 0u; f.end = 0u; f.part = wave % PT_PARTS; f.left = PT_PARTS; f.done = false;
    return f;
}
// Every lane calls it; lanes with `want` may receive a queue index (else 0xFFFFFFFF).
__device__ __forceinline__ unsigned int fetchNext(WaveFetch& f, unsigned int* heads, unsigned int count, bool want) {
    unsigned int got = 0xFFFFFFFFu;
    unsigned long long need = __ballot(want);
    const unsigned int chunk = (count + PT_PARTS - 1u) / PT_PARTS;
    while (need != 0ull && !f.done) {
        if (f.cur >= f.end) {
            const unsigned int part = __builtin_amdgcn_readfirstlane(f.part);   // uniform: a scalar address
            unsigned int lo = min(part * chunk, count), hi = min(lo + chunk, count);
            unsigned int base = __builtin_amdgcn_readfirstlane(__atomic_load_n(&heads[part], __ATOMIC_RELAXED));
            if (lo + base < hi) {
                // every active lane executes the grab, the first adding PT_GRAB and the rest 0, and the wave takes
                // the first's return: no branch on the lane before the broadcast (see waveAppend)
                const unsigned int first = (unsigned int)__ffsll((long long)__ballot(1)) - 1u;
                base = __builtin_amdgcn_readfirstlane(gAtomicAdd(&heads[part], laneId() == first ? PT_GRAB : 0u));
            }
            if (lo + base < hi) { f.cur = lo + base; f.end = min(lo + base + PT_GRAB, hi); }
            else { f.part = (f.part + 1u) % PT_PARTS; if (--f.left == 0u) f.done = true; }
            continue;
        }
        unsigned int take = min((unsigned int)__popcll(need), f.end - f.cur);
        bool mine = (need >> laneId()) & 1ull;
        unsigned int r = rankIn(need);
        bool served = mine && r < take;
        if (served) got = f.cur + r;
        need &= ~__ballot(served);
        f.cur += take;
    }
    return got;
}

// This sample's seed: the base seed and the number of samples resolved so far, through PCG4D.
__device__ __forceinline__ void sampleSeed(const PtDev* __restrict__ D, unsigned int* s) {
    Rng r; r.a = D->u.seed; r.b = D->counters[PT_C_SAMPLE]; r.c = 0x5EED5EEDu; r.d = 0x2026F00Du;
    pcg4d(r); pcg4d(r);
    s[0] = r.a; s[1] = r.b; s[2] = r.c; s[3] = r.d;
}

__device__ __forceinline__ float tolSetting(const PtDev* __restrict__ D) { return D->u.eps > 0.0f ? D->u.eps : D->u.epsAuto; }
__device__ __forceinline__ float offsetTol(const PtDev* __restrict__ D, V3 p) { return 10.0f * marchTol(tolSetting(D), p); }

// How far a node may be from a ray and still matter to its march: the largest hit tolerance anywhere the ray can be
// (inside the sky sphere), so a node left out of the ray's mask can never come within a tolerance of it.
__device__ __forceinline__ float maskMargin(const PtDev* __restrict__ D) {
    return fmaxf(tolSetting(D), 1e-9f) + 128.0f * PT_F32_STEP * (D->u.envRadius + 1.0f);
}

// meta.z: bounces (8 bits) | state << 8 | flags << 16
#define PT_F_DELTAONLY 0x1u     // every scatter so far was a delta event (SPPM gathers at the next diffuse hit)
#define PT_F_NEEPREV 0x2u       // the previous vertex ran light sampling (prevPdf meaningful)
#define PT_F_PHOTON 0x4u        // the slot carries a photon of an SPPM pass, not a camera path
#define PT_F_HADDELTA 0x8u      // a photon that has passed at least one delta event (a caustic photon)
#define PT_F_CAUSTIC 0x10u      // a camera path that gathered: light reaching its gather vertex through delta
                                // events only is the photon map's
#define PT_F_SDELTA 0x20u       // ... and it has scattered by a delta event since that vertex
#define PT_F_SUNPH 0x80u        // a photon from the sun (light tracing connects these)
#define PT_F_LTCLS 0x100u       // a camera path whose first vertex left it by a non-delta lobe: with SDELTA, the light it
                                // meets through delta events only is the light tracer's
#define PT_F_GPEND 0x40u        // a recorded guide vertex waits for its S (the radiance so far, at the next event)
#define PT_HR_INSIDE 0x100u     // hit record: the ray is inside the surface's object
#define PT_HR_BEYOND 0x200u     // hit record: the medium beyond is the material in bits 16..
__device__ __forceinline__ unsigned int mzBounces(unsigned int z) { return z & 0xFFu; }
__device__ __forceinline__ unsigned int mzState(unsigned int z) { return (z >> 8) & 0xFFu; }
__device__ __forceinline__ unsigned int mzFlags(unsigned int z) { return z >> 16; }
__device__ __forceinline__ unsigned int mzPack(unsigned int b, unsigned int s, unsigned int f) { return (b & 0xFFu) | ((s & 0xFFu) << 8) | (f << 16); }

// ---------------------------------------------------------------------------------------------------
// environment (uniform sphere or hemisphere, and a sun disc), its light-sampling draw and density
// The disc's solid angle comes from the host in double: 1 - cos of the real sun's radius is 1.07e-5, which f32 holds
// only to 0.3 %. Membership is |d x s|^2 <= sin^2 on the sun's side (sdf.js inSunDisc:88), which keeps its relative
// precision near the axis where a cosine test does not.
__device__ __forceinline__ float sunSolidAngle(const PtEnv& e) { return e.sunOmega; }
__device__ __forceinline__ bool inSun(const PtEnv& e, V3 d) {
    const V3 s = ld3(e.sunDir), c = cross(d, s);
    return e.sunE > 0.0f && dot(d, s) > 0.0f && dot(c, c) <= e.sunSin2;
}
// Kasten-Young relative air mass (sdf.js airMass:96): 1 straight up, about 38 at the horizon
__device__ __forceinline__ float airMass(float cz) {
    const float c = fminf(fmaxf(cz, 0.0f), 1.0f);
    const float zdeg = acosf(c) * (180.0f / PT_PI);
    return 1.0f / (c + 0.15f * __powf(fmaxf(93.885f - zdeg, 0.1f), -1.253f));
}
// The sun & sky's sky (environment.js SUNSKY_FN:130-168): the sunlight scattered once towards the eye by the air
// (Rayleigh phase) and the haze (Henyey-Greenstein g 0.7, albedo 0.9) over the whole column, each band attenuated by
// the extinction along the sun's path and the view's, the ozone absorbing only.
__device__ __forceinline__ V3 sunSky(const PtEnv& e, V3 d) {
    const V3 s = ld3(e.sunDir);
    const float mv = airMass(d.z), ms = e.skyMs, dm = mv - ms;
    const float c = dot(d, s), g = 0.7f;
    const float pR = (3.0f / (16.0f * PT_PI)) * (1.0f + c * c);
    const float x = 1.0f + g * g - 2.0f * g * c;
    const float pM = (1.0f - g * g) / (4.0f * PT_PI * x * sqrtf(x));
    const bool same = fabsf(dm) < 1e-3f;
    const float idm = same ? 0.0f : 1.0f / dm;
    float r = 0.0f, gr = 0.0f, b = 0.0f;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const float tr = PT_SKY_TAUR[k], ta = e.skyAerosol * PT_SKY_HAZE[k];
        const float tau = tr + ta + PT_SKY_OZONE[k];
        const float ev = __expf(-tau * mv);
        const float k1 = same ? tau * mv * ev : mv * (__expf(-tau * ms) - ev) * idm;
        const float v = (tr * pR + 0.9f * ta * pM) / tau * k1;
        r += PT_SKY_RGB[3 * k] * v; gr += PT_SKY_RGB[3 * k + 1] * v; b += PT_SKY_RGB[3 * k + 2] * v;
    }
    const float f = e.L * e.skyUp;     // col: the sky's gain per channel (envSky)
    return v3(fmaxf(f * r * e.col[0], 0.0f), fmaxf(f * gr * e.col[1], 0.0f), fmaxf(f * b * e.col[2], 0.0f));
}
__device__ __forceinline__ V3 envBackground(const PtEnv& e, V3 d) {
    if (e.type == 1u || (e.type == 2u && d.z > 0.0f)) return ld3(e.col) * e.L;
    if (e.type == 3u) return ld3(e.col) * (e.L * fmaxf(1.0f + e.top * d.z, 0.0f));
    if (e.type == 4u) return inSun(e, d) ? v3s(0.0f) : sunSky(e, d);      // the disc is envSun's
#if PT_SCENE_SKY
    if (e.type == 5u) return sceneSky(d);                                   // the scene's own sky code
#endif
    return v3s(0.0f);
}
// The sun's radiance along the ray from p towards d. With an aperture the sun is a beam: the lines that cross the disc
// of radius sunRadius about sunPoint, perpendicular to the sun's direction -- the disc the photon pass emits the beam
// from, so both see one light field, and any point of a ray (its vertex, or where it left the scene) gives the same
// answer. The disc is placed above whatever the beam lights: nothing above it is lit by photons.
__device__ __forceinline__ V3 envSun(const PtEnv& e, V3 p, V3 d) {
    if (!inSun(e, d)) return v3s(0.0f);
    if (e.sunRadius > 0.0f) {
        V3 a = ld3(e.sunDir), c = ld3(e.sunPoint);
        float t = dot(c - p, a) / dot(d, a);
        V3 q = p + d * t - c;
        if (!(dot(q, q) <= e.sunRadius * e.sunRadius)) return v3s(0.0f);
    }
    return e.type == 4u ? ld3(e.sunL) : v3s(e.sunE / sunSolidAngle(e));
}
__device__ __forceinline__ float sunPick(const PtEnv& e) { return e.sunE > 0.0f ? 0.5f : 0.0f; }
__device__ __forceinline__ V3 sampleCosine(V3 n, float u1, float u2) {
    V3 t, b; frameOf(n, t, b);
    float r = sqrtf(u1), phi = 2.0f * PT_PI * u2;
    float sp, cp; __sincosf(phi, &sp, &cp);
    return normalize(t * (r * cp) + b * (r * sp) + n * sqrtf(fmaxf(0.0f, 1.0f - u1)));
}
__device__ __forceinline__ V3 sampleSunCone(const PtEnv& e, float u1, float u2) {
    V3 a = ld3(e.sunDir);
    V3 t, b; frameOf(a, t, b);
    // in 1 - cos, whose relative precision f32 keeps: sin = sqrt(omc (2 - omc)) (sdf.js sampleSunCone:132)
    float omc = u1 * e.sunOmega * (0.5f * PT_INV_PI);
    float c = 1.0f - omc;
    float s = sqrtf(fmaxf(0.0f, omc * (2.0f - omc)));
    float phi = 2.0f * PT_PI * u2;
    float sp, cp; __sincosf(phi, &sp, &cp);
    return normalize(t * (s * cp) + b * (s * sp) + a * c);
}
__device__ __forceinline__ float envPdf(const PtEnv& e, V3 L, V3 n) {
    float ps = sunPick(e);
    float pSun = inSun(e, L) ? 1.0f / sunSolidAngle(e) : 0.0f;
    return ps * pSun + (1.0f - ps) * fmaxf(dot(L, n), 0.0f) * PT_INV_PI;
}
__device__ __forceinline__ bool envLit(const PtEnv& e) { return (e.type != 0u && e.L > 0.0f) || e.sunE > 0.0f; }
__device__ __forceinline__ float misWeight(float a, float b) { a *= a; b *= b; return a / fmaxf(a + b, 1e-20f); }

// ---------------------------------------------------------------------------------------------------
// light faces (NEE towards emissive rectangles and discs): area sampling, density in solid angle
__device__ __forceinline__ float faceArea(const PtLightFace& f) {
    return f.shape == 1u ? PT_PI * f.halfx * f.halfx : 4.0f * f.halfx * f.halfy;
}
// Emitted flux (luminance-weighted, lm) of a light face (Lambertian: pi L A a side) and of the sun's beam through
// its aperture (E pi R^2; none without an aperture): the photon pass draws its lights in proportion to these.
__device__ __forceinline__ float faceFlux(const PtLightFace& f) {
    return f.weight > 0.0f ? luma(ld3(f.le)) * PT_PI * faceArea(f) * (f.twoSided ? 2.0f : 1.0f) : 0.0f;
}
__device__ __forceinline__ float sunFlux(const PtEnv& e) {
    return (e.sunE > 0.0f && e.sunRadius > 0.0f) ? e.sunE * PT_PI * e.sunRadius * e.sunRadius : 0.0f;
}
__device__ __forceinline__ float faceShareAt(const PtDev* __restrict__ D, unsigned int bounces) {
    if (bounces == 0u || D->nfaces == 0u || D->faceTotal <= 0.0f) return 0.0f;
    return envLit(D->u.env) ? D->u.faceShare : 1.0f;
}
__device__ __forceinline__ unsigned int facePick(const PtDev* __restrict__ D, float u) {
    float acc = 0.0f;
    unsigned int last = 0u;
    for (unsigned int k = 0; k < D->nfaces; ++k) {
        float w = D->faces[k].weight;
        if (w > 0.0f) { last = k; acc += w; if (u < acc) return k; }
    }
    return last;
}
__device__ __forceinline__ float faceDensity(const PtDev* __restrict__ D, float share, const PtLightFace& f, float dist, float cosY) {
    if (share <= 0.0f || f.weight <= 0.0f) return 0.0f;
    return share * (f.weight / D->faceTotal) * dist * dist / fmaxf(faceArea(f) * fabsf(cosY), 1e-20f);
}
// Where a ray from x along d meets face f's plane inside the face, the distance; < 0 if it does not.
__device__ __forceinline__ float faceHitDist(const PtLightFace& f, V3 x, V3 d, float& cosY) {
    V3 n = ld3(f.n);
    cosY = -dot(n, d);
    if (fabsf(cosY) < 1e-8f) return -1.0f;
    float t = dot(ld3(f.centre) - x, n) / (-cosY);
    if (t <= 0.0f) return -1.0f;
    V3 q = x + d * t - ld3(f.centre);
    float a = dot(q, ld3(f.u)), b = dot(q, ld3(f.v));
    bool in = f.shape == 1u ? (a * a + b * b <= f.halfx * f.halfx) : (fabsf(a) <= f.halfx && fabsf(b) <= f.halfy);
    return in ? t : -1.0f;
}

// ---------------------------------------------------------------------------------------------------
// deposits into the per-sample frame: refused (and counted) when not finite or of negative luminance
// (core/wavefront.js:202-245 depositWGSL)
__device__ __forceinline__ bool badDeposit(V3 c) {
    return !(isfinite(c.x) && isfinite(c.y) && isfinite(c.z)) || luma(c) < 0.0f;
}
__device__ __forceinline__ void depositRGB(const PtDev* __restrict__ D, unsigned int pix, V3 c, bool direct) {
    if (D->u.fireflyClamp > 0.0f && !direct) {
        float cap = fmaxf(D->u.fireflyClamp, 1.0f) * D->u.fireflyWhite;
        float y = luma(c);
        if (y > cap) c = c * (cap / y);
    }
    float* f = D->frame + 4ull * pix;
    if (c.x != 0.0f) unsafeAtomicAdd(f + 0, c.x);
    if (c.y != 0.0f) unsafeAtomicAdd(f + 1, c.y);
    if (c.z != 0.0f) unsafeAtomicAdd(f + 2, c.z);
}

// ---------------------------------------------------------------------------------------------------
// SPECTRAL: Cauchy's n(lambda) = A + B / lambda^2 through n_d (587.6 nm) and the Abbe number
// V_d = (n_d - 1) / (n_F - n_C) (core/spectral.js:121-128), lambda in micrometres.
__device__ __forceinline__ float cauchyN(float nd, float abbe, float lam) {
    const float iF = 1.0f / (0.4861f * 0.4861f), iC = 1.0f / (0.6563f * 0.6563f), id = 1.0f / (0.5876f * 0.5876f);
    float B = (nd - 1.0f) / (abbe * (iF - iC));
    float A = nd - B * id;
    return A + B / (lam * lam);
}
// The weight of one wavelength drawn uniformly over 380-780 nm, per linear sRGB channel: the CIE 1931
// observer (Wyman, Sloan & Shirley 2013 multi-lobe fit) to XYZ to sRGB, negative lobes clipped, normalised so
// that its mean over the draw is (1, 1, 1): white light stays white on average.
__device__ __forceinline__ float wssLobe(float l, float mu, float s1, float s2) {
    float t = (l - mu) / (l < mu ? s1 : s2);
    return __expf(-0.5f * t * t);
}
__device__ __forceinline__ V3 spectralWeight(float lamUm) {
    float l = lamUm * 1000.0f;
    float x = 1.056f * wssLobe(l, 599.8f, 37.9f, 31.0f) + 0.362f * wssLobe(l, 442.0f, 16.0f, 26.7f) - 0.065f * wssLobe(l, 501.1f, 20.4f, 26.2f);
    float y = 0.821f * wssLobe(l, 568.8f, 46.9f, 40.5f) + 0.286f * wssLobe(l, 530.9f, 16.3f, 31.1f);
    float z = 1.217f * wssLobe(l, 437.0f, 11.8f, 36.0f) + 0.681f * wssLobe(l, 459.0f, 26.0f, 13.8f);
    float r = fmaxf(3.2406f * x - 1.5372f * y - 0.4986f * z, 0.0f);
    float g = fmaxf(-0.9689f * x + 1.8758f * y + 0.0415f * z, 0.0f);
    float b = fmaxf(0.0557f * x - 0.2040f * y + 1.0570f * z, 0.0f);
    return v3(r * (1.0f / 0.44045557f), g * (1.0f / 0.28846547f), b * (1.0f / 0.27323972f));
}

// ---------------------------------------------------------------------------------------------------
// HAZE: one homogeneous axis-aligned box (core/haze.js:191-290): free flight drawn where a leg starts, over
// the leg's overlap with the box; Henyey-Greenstein phase function.
__device__ __forceinline__ bool hazeOn(const PtDev* __restrict__ D) { return D->u.haze.sigmaS > 0.0f; }
// The ray's span [t0, t1] inside the box; false if none ahead.
__device__ __forceinline__ bool hazeSpan(const PtHaze& h, V3 o, V3 d, float& t0, float& t1) {
    float lo = 0.0f, hi = 3.0e38f;
    const float* c = h.centre;
    const float* e = h.half;
    float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float inv = 1.0f / dd[k];
        float a = (c[k] - e[k] - oo[k]) * inv, b = (c[k] + e[k] - oo[k]) * inv;
        float mn = fminf(a, b), mx = fmaxf(a, b);
        if (dd[k] == 0.0f) { if (oo[k] < c[k] - e[k] || oo[k] > c[k] + e[k]) return false; continue; }
        lo = fmaxf(lo, mn); hi = fminf(hi, mx);
    }
    t0 = lo; t1 = hi;
    return hi > lo;
}
// Transmittance of the box along [0, len] of a ray.
__device__ __forceinline__ float hazeTr(const PtHaze& h, V3 o, V3 d, float len) {
    float t0, t1;
    if (!hazeSpan(h, o, d, t0, t1)) return 1.0f;
    float l = fmaxf(fminf(t1, len) - t0, 0.0f);
    return __expf(-h.sigmaS * l);
}
// Draw where the leg from o along d collides: the distance (3e38: no collision before the box is left).
__device__ __forceinline__ float hazeDraw(const PtHaze& h, V3 o, V3 d, float u) {
    float t0, t1;
    if (!hazeSpan(h, o, d, t0, t1)) return 3.0e38f;
    float t = t0 - __logf(1.0f - u) / h.sigmaS;
    return t < t1 ? t : 3.0e38f;
}
__device__ __forceinline__ float hgEval(float g, float cosT) {
    float d = 1.0f + g * g - 2.0f * g * cosT;
    return (1.0f - g * g) / (4.0f * PT_PI * d * sqrtf(d));
}
// Sample HG about the travel direction w (cosT between w and the new direction).
__device__ __forceinline__ V3 hgSample(float g, V3 w, float u1, float u2) {
    float c;
    if (fabsf(g) < 1e-3f) c = 1.0f - 2.0f * u1;
    else { float s = (1.0f - g * g) / (1.0f - g + 2.0f * g * u1); c = (1.0f + g * g - s * s) / (2.0f * g); }
    c = fminf(fmaxf(c, -1.0f), 1.0f);
    float sn = sqrtf(fmaxf(0.0f, 1.0f - c * c));
    V3 t, b; frameOf(w, t, b);
    float sp, cp; __sincosf(2.0f * PT_PI * u2, &sp, &cp);
    return normalize(t * (sn * cp) + b * (sn * sp) + w * c);
}

// ---------------------------------------------------------------------------------------------------
// SPPM (core/sppm.js:79-192, :350-520): photons stored at non-delta landings in a hash grid of PT_SPPM_HEADS
// linked lists (cell 2r, Teschner et al. 2003's spatial hash), each record also on one of PT_SPPM_SUBS sub-lists of
// its cell; camera paths after delta-only chains gather sum f(l_p) P_p / (pi r^2) over photons within r whose
// surface normal agrees (dot >= 0.9).
// Octahedral unit vectors (Meyer et al. 2010) in two signed fields of `bits` bits each.
__device__ __forceinline__ unsigned int octEncode(V3 v, unsigned int bits = 16u) {
    float s = fabsf(v.x) + fabsf(v.y) + fabsf(v.z);
    float ex = v.x / s, ey = v.y / s;
    if (v.z < 0.0f) {
        float ox = (1.0f - fabsf(ey)) * (ex >= 0.0f ? 1.0f : -1.0f);
        float oy = (1.0f - fabsf(ex)) * (ey >= 0.0f ? 1.0f : -1.0f);
        ex = ox; ey = oy;
    }
    const float q = (float)((1u << (bits - 1u)) - 1u);
    const unsigned int mask = (1u << bits) - 1u;
    int ix = (int)rintf(fminf(fmaxf(ex, -1.0f), 1.0f) * q), iy = (int)rintf(fminf(fmaxf(ey, -1.0f), 1.0f) * q);
    return ((unsigned int)ix & mask) | (((unsigned int)iy & mask) << bits);
}
__device__ __forceinline__ V3 octDecode(unsigned int w, unsigned int bits = 16u) {
    const float q = (float)((1u << (bits - 1u)) - 1u);
    const unsigned int mask = (1u << bits) - 1u, sh = 32u - bits;
    int ix = (int)((w & mask) << sh) >> sh, iy = (int)(((w >> bits) & mask) << sh) >> sh;
    float ex = (float)ix / q, ey = (float)iy / q;
    V3 v = v3(ex, ey, 1.0f - fabsf(ex) - fabsf(ey));
    if (v.z < 0.0f) {
        v.x = (1.0f - fabsf(ey)) * (ex >= 0.0f ? 1.0f : -1.0f);
        v.y = (1.0f - fabsf(ex)) * (ey >= 0.0f ? 1.0f : -1.0f);
    }
    return normalize(v);
}
__device__ __forceinline__ unsigned int sppmHashRaw(int i, int j, int k) {
    return ((unsigned int)i * 73856093u) ^ ((unsigned int)j * 19349663u) ^ ((unsigned int)k * 83492791u);
}
__device__ __forceinline__ unsigned int sppmHash(int i, int j, int k) { return sppmHashRaw(i, j, k) & (PT_SPPM_HEADS - 1u); }
// a cell's sub-list s: the cell's hash mixed with s (a golden-ratio multiple), in a table of its own
__device__ __forceinline__ unsigned int sppmHashSub(int i, int j, int k, unsigned int s) {
    return (sppmHashRaw(i, j, k) ^ ((s + 1u) * 0x9E3779B1u)) & (PT_SPPM_HEADS - 1u);
}
#define PT_SPPM_SUBS 32u       // sub-lists a cell's records are spread over
#define PT_SPPM_BUDGET 512u    // records a gather walks in the ordinary case, shared by its eight cells
#define PT_SPPM_HARD 8u        // ...and the bound on its records visited, in budgets
#define PT_SPPM_NORMAL_DOT 0.9f
