// The wavefront kernels proper (genrays, march, shade, occlude, resolve) and the start-up table integration.
#pragma once

// ---------------------------------------------------------------------------------------------------
// GENRAYS (core/genrays.js:36-107): one jittered pinhole ray per tile of the raygen grid (raygenTiles, rounded
// not floored: genrays.js:22-34), pix from the jittered uv. Also resets the queues for the sample.
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_genrays(const PtDev* __restrict__ D) {
    unsigned int i = blockIdx.x * PT_BLOCK + threadIdx.x;
    unsigned int* C = D->counters;
    if (i == 0) { C[PT_C_MARCH0] = D->nrays; C[PT_C_MARCH1] = 0u; C[PT_C_SHADOW] = 0u; }
    if (i < 2u * PT_PARTS) C[PT_C_HEADS_MARCH + i] = 0u;
    if (D->u.guide != 0u) {
        // the guide's training data starts empty each sample
        if (i == 0) C[PT_C_GUIDE_N] = 0u;
        if (i < D->g.steps) { D->gSumF[i] = 0.0; D->gLoss[i] = 0.0; }
    }
    if (i >= D->nrays) return;
    unsigned int gr;
    if (guideRecorded(D, i, gr)) {
        D->gVerts[gr] = 0u;
        reinterpret_cast<float4*>(D->gPath)[2 * gr] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        reinterpret_cast<float4*>(D->gPath)[2 * gr + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    unsigned int seed[4];
    sampleSeed(D, seed);
    Rng r = rngAt(seed, i, 0u, 0xCA3E8Au);
    const PtCamera& cam = D->u.cam;
    unsigned int tx = i % D->tilesX, ty = i / D->tilesX;
    V4 rnd = rngNext(r);
    float u = ((float)tx + 0.5f) * D->invTX + (rnd.x - 0.5f) * D->invTX;
    float v = ((float)ty + 0.5f) * D->invTY + (rnd.y - 0.5f) * D->invTY;
    u = fminf(fmaxf(u, 0.0f), 1.0f);
    v = fminf(fmaxf(v, 0.0f), 1.0f);
    float x = u * 2.0f - 1.0f, y = v * 2.0f - 1.0f;
    V3 right = ld3(cam.right), up = ld3(cam.up), fwd = ld3(cam.fwd);
    V3 o, d;
    if (cam.type == 0u) {
        o = ld3(cam.pos);
        const V3 dc = right * (x * cam.fx) + up * (y * cam.fy) - fwd;
        d = normalize(dc);
        if (cam.lensR > 0.0f) {
            // through the thin lens (genrays.js:78-86, lens.js lensDisc:111): from a point of its disc, uniform in
            // area, towards the point where the pinhole's ray meets the plane in focus
            const V4 l2 = rngNext(r);
            const float rr = cam.lensR * sqrtf(rnd.w), phi = 2.0f * PT_PI * l2.x;
            float sp, cp; __sincosf(phi, &sp, &cp);
            const V3 off = right * (rr * cp) + up * (rr * sp);
            o = o + off;
            d = normalize(dc * cam.focus - off);
        }
    } else {
        o = ld3(cam.pos) + right * (x * cam.fx) + up * (y * cam.fy);
        d = -fwd;
    }
    // the pixel: the tile's own when the tile grid is the pixel grid (the jittered u * width can round up onto the
    // next pixel's edge in FP32, which would give that pixel two rays of the sample and this one none); else where
    // the jittered point falls
    float fw = (float)D->u.width, fh = (float)D->u.height;
    unsigned int px, py;
    if (D->tilesX == D->u.width && D->nrays == D->u.width * D->u.height) {
        px = tx; py = ty;
    } else {
        px = (unsigned int)fminf(fmaxf(u * fw, 0.0f), fw - 1.0f);
        py = (unsigned int)fminf(fmaxf(v * fh, 0.0f), fh - 1.0f);
    }
    float collide = 3.0e38f;
    if (hazeOn(D)) collide = hazeDraw(D->u.haze, o, d, rnd.z);
    st4(D->pos, i, o.x, o.y, o.z, 0.0f);
    st4(D->dir, i, d.x, d.y, d.z, 1.0f);
    st4(D->beta, i, 1.0f, 1.0f, 1.0f, -1.0f);
    st4(D->rad, i, 0.0f, 0.0f, 0.0f, 0.0f);
    st4u(D->meta, i, make_uint4(px + py * D->u.width, 0u, mzPack(0u, PT_ACTIVE, PT_F_DELTAONLY), 0u));
    st4(D->aux, i, 1.0f, collide, 0.0f, 0.0f);
    D->qMarch0[i] = i;
}

// ---------------------------------------------------------------------------------------------------
// BINNING: the round's march queue reordered so that rays needing the same scene nodes share waves. Each ray's node
// mask (sceneMask: the nodes its segment can reach) is hashed to one of PT_BINS keys and the queue counting-sorted by
// key: per-block histograms (LDS), one scan over (bin, block), a stable-per-block scatter. A wave of the march then
// evaluates the union of few masks instead of sixty-four different ones. Random numbers are keyed by slot, so the
// order changes no path's draws.
__device__ __forceinline__ unsigned int maskKey(unsigned long long m) {
    m ^= m >> 33; m *= 0xff51afd7ed558ccdull; m ^= m >> 33; m *= 0xc4ceb9fe1a85ec53ull; m ^= m >> 33;
    return (unsigned int)(m >> 54);
}
__device__ __forceinline__ void binChunk(unsigned int count, unsigned int& lo, unsigned int& hi) {
    unsigned int chunk = (count + PT_BIN_BLOCKS - 1) / PT_BIN_BLOCKS;
    lo = min(blockIdx.x * chunk, count);
    hi = min(lo + chunk, count);
}
// qsel 0/1: the march queue of that parity; 2: the shadow queue (its segments end at tmax)
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_bin_count(const PtDev* __restrict__ D, unsigned int qsel) {
    __shared__ unsigned int hist[PT_BINS];
    for (unsigned int k = threadIdx.x; k < PT_BINS; k += PT_BLOCK) hist[k] = 0u;
    __syncthreads();
    const bool shadow = qsel == 2u;
    const unsigned int count = D->counters[shadow ? PT_C_SHADOW : (qsel ? PT_C_MARCH1 : PT_C_MARCH0)];
    const unsigned int* __restrict__ q = shadow ? D->qShadow : (qsel ? D->qMarch1 : D->qMarch0);
    const float* __restrict__ org = shadow ? D->shOrig : D->pos;
    const float* __restrict__ dir = shadow ? D->shDir : D->dir;
    ParamPtr U = (const PT_CONST float*)D->params;
    const float margin = maskMargin(D);
    unsigned int lo, hi;
    binChunk(count, lo, hi);
    for (unsigned int i = lo + threadIdx.x; i < hi; i += PT_BLOCK) {
        unsigned int slot = q[i];
        float4 P = ld4(org, slot), Dd = ld4(dir, slot);
        float tmax = shadow ? P.w : 3.0e38f;
        unsigned long long m = sceneMask(U, v3(P.x, P.y, P.z), v3(Dd.x, Dd.y, Dd.z), tmax, margin);
        D->maskBuf[slot] = m;
        atomicAdd(&hist[maskKey(m)], 1u);
    }
    __syncthreads();
    for (unsigned int k = threadIdx.x; k < PT_BINS; k += PT_BLOCK) D->binHist[blockIdx.x * PT_BINS + k] = hist[k];
}
// one block of PT_BINS threads: bin k's start is the sum of all bins before it; block b's offset within bin k the sum
// of the blocks before b
extern "C" __global__ __launch_bounds__(PT_BINS) void pt_bin_scan(const PtDev* __restrict__ D) {
    __shared__ unsigned int s[PT_BINS];
    const unsigned int t = threadIdx.x;
    unsigned int tot = 0u;
    for (unsigned int b = 0; b < PT_BIN_BLOCKS; ++b) tot += D->binHist[b * PT_BINS + t];
    s[t] = tot;
    __syncthreads();
    for (unsigned int off = 1; off < PT_BINS; off <<= 1) {
        unsigned int x = t >= off ? s[t - off] : 0u;
        __syncthreads();
        s[t] += x;
        __syncthreads();
    }
    unsigned int run = s[t] - tot;
    for (unsigned int b = 0; b < PT_BIN_BLOCKS; ++b) {
        unsigned int c = D->binHist[b * PT_BINS + t];
        D->binOff[b * PT_BINS + t] = run;
        run += c;
    }
}
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_bin_scatter(const PtDev* __restrict__ D, unsigned int qsel) {
    __shared__ unsigned int rank[PT_BINS];
    for (unsigned int k = threadIdx.x; k < PT_BINS; k += PT_BLOCK) rank[k] = 0u;
    __syncthreads();
    const bool shadow = qsel == 2u;
    const unsigned int count = D->counters[shadow ? PT_C_SHADOW : (qsel ? PT_C_MARCH1 : PT_C_MARCH0)];
    const unsigned int* __restrict__ q = shadow ? D->qShadow : (qsel ? D->qMarch1 : D->qMarch0);
    unsigned int lo, hi;
    binChunk(count, lo, hi);
    for (unsigned int i = lo + threadIdx.x; i < hi; i += PT_BLOCK) {
        unsigned int slot = q[i];
        unsigned int key = maskKey(D->maskBuf[slot]);
        unsigned int r = atomicAdd(&rank[key], 1u);
        D->qSorted[D->binOff[blockIdx.x * PT_BINS + key] + r] = slot;
    }
}

// ---------------------------------------------------------------------------------------------------
// MARCH (core/wavefront.js:834-1005): sphere tracing by |d| * stepscale; a hit where |d| < marchTol(p); escaped
// when outside envRadius and heading out; out of steps past maxiter or an optical length of 1000; a haze
// collision when the leg reaches its drawn free flight. Persistent: each wave refills its idle lanes from the
// (binned) queue in batches (fetchNext). Every ray ends its march in exactly one event, so shade reads the same
// queue: no event queue, no append.
//
// OVER-RELAXATION (Keinert, Schaefer, Korndoerfer, Ganse, Stamminger, "Enhanced Sphere Tracing", 2014), when the
// uniforms' relax w > 1: a step of w r (r = stepscale |d|, the plain step) goes past the free ball it stands on. It
// is kept if the free ball at the new point overlaps the old one (r + r_prev > the step): the two cover the segment,
// so no surface lies on it. Else the march goes back to where the plain step would have put it (inside the old
// ball) and marches plainly for the rest of the ray. A point reached by a gap is never tested for an event, so every
// event is one the plain march could meet; a ray with a haze collision pending marches plainly.
#define PT_MARCH_STEPS 4
#define PT_STEP_FLOPS 30u   // a step's own FP32 work beside the SDF: tolerance, escape test, the move
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_march(const PtDev* __restrict__ D, unsigned int qsel) {
    unsigned int* C = D->counters;
    // the shadow queue this round's shade fills, and its fetch heads, start empty
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) C[PT_C_SHADOW] = 0u;
        if (threadIdx.x < PT_PARTS) C[PT_C_HEADS_SHADOW + threadIdx.x] = 0u;
    }
    const unsigned int count = C[qsel ? PT_C_MARCH1 : PT_C_MARCH0];
    const unsigned int* __restrict__ q = D->qSorted;
    const float setting = tolSetting(D);
    const float R2 = D->u.envRadius * D->u.envRadius;
    const float stepscale = D->u.stepscale;
    const unsigned int maxiter = D->u.maxiter;
    const unsigned int refill = D->refillMarch;
    const float relax = D->u.relax > 1.0f ? D->u.relax : 1.0f;
    ParamPtr U = (const PT_CONST float*)D->params;

    int slot = -1;
    V3 p = v3s(0.0f), dir = v3s(0.0f);
    float ltot = 0.0f, n = 1.0f, collide = 3.0e38f;
    float om = 1.0f, rPrev = 0.0f, sPrev = 0.0f;    // the relaxation, the last plain step, the last step taken
    unsigned int steps = 0u, z = 0u;
    unsigned long long mask = ~0ull;
    WaveFetch f = fetchInit();
    unsigned long long nsteps = 0ull, nmax = 0ull, nflops = 0ull;
    for (;;) {
        unsigned long long idle = __ballot(slot < 0);
        if (!f.done && (unsigned int)__popcll(idle) >= refill) {
            unsigned int k = fetchNext(f, C + PT_C_HEADS_MARCH, count, slot < 0);
            if (slot < 0 && k != 0xFFFFFFFFu) {
                slot = (int)ldu(q, k);
                float4 P = ld4(D->pos, slot), Dd = ld4(D->dir, slot);
                uint4 M = ld4u(D->meta, slot);
                p = v3(P.x, P.y, P.z); ltot = P.w;
                dir = v3(Dd.x, Dd.y, Dd.z); n = Dd.w;
                steps = M.y; z = M.z;
                collide = ldf(D->aux, 4u * slot + 1u);
                mask = ldu64(D->maskBuf, slot);
                om = collide >= 3.0e38f ? relax : 1.0f;
                rPrev = sPrev = 0.0f;
            }
        }
        if (__ballot(slot >= 0) == 0ull) {
            if (f.done) break;
            continue;
        }
        unsigned int state = PT_ACTIVE;
        bool event = false;
        if (slot >= 0) {
            #pragma unroll 1
            for (int k = 0; k < PT_MARCH_STEPS; ++k) {
                if (ltot >= collide) {
                    // back to the collision point, to be shaded as a point in the haze
                    float back = (ltot - collide) / n;
                    p = p - dir * back;
                    ltot = collide;
                    state = PT_SCATTER;
                    event = true;
                    break;
                }
                unsigned int w = PT_STEP_FLOPS;
                float d = sceneDistM(U, p, mask, w);
                nflops += w;
                float ea = fabsf(d);
                float r = ea * stepscale;
                if (om > 1.0f && r + rPrev <= sPrev) {
                    // a gap between the free balls: back to the plain step's point, plainly from here
                    float back = sPrev - rPrev;
                    p = p - dir * back;
                    ltot -= back * n;
                    om = 1.0f;
                    steps += 1u;
                    nsteps += 1ull;
                    continue;
                }
                float eps = marchTol(setting, p);
                if (ltot > 1000.0f || (steps & 0xFFFFu) > maxiter) state = PT_MAXITER;
                if (dot(p, dir) > 0.0f && dot(p, p) > R2) state = PT_ESCAPED;
                if (ea < eps || state != PT_ACTIVE) { event = true; break; }
                steps += 1u;
                nsteps += 1ull;
                float t = r * om;
                p = p + dir * t;
                ltot += t * n;
                rPrev = r;
                sPrev = t;
            }
        }
        if (event) {
            if (state == PT_MAXITER) nmax += 1ull;
            st4(D->pos, slot, p.x, p.y, p.z, ltot);
            stu(D->meta, 4u * slot + 1u, steps);
            stu(D->meta, 4u * slot + 2u, mzPack(mzBounces(z), state, mzFlags(z)));
            slot = -1;
        }
    }
    statAdd(D, PT_S_STEPS, nsteps);
    statAdd(D, PT_S_MAXITER, nmax);
    statAdd(D, PT_S_FLOPS, nflops);
}

__device__ __forceinline__ V3 normalAt(ParamPtr U, V3 p, float h) {
    const float a = sceneDist(U, p + v3(h, -h, -h)), b = sceneDist(U, p + v3(-h, -h, h));
    const float c = sceneDist(U, p + v3(-h, h, -h)), d = sceneDist(U, p + v3(h, h, h));
    V3 g = v3(a - b - c + d, -a - b + c + d, -a + b - c + d);
    float l = length(g);
    return l > 0.0f ? g * (1.0f / l) : v3(0.0f, 0.0f, 1.0f);
}

// ---------------------------------------------------------------------------------------------------
// SURFACE (core/wavefront.js:1037-1147 prog_surface): for each ray of the round that stopped at a surface, the one
// full-SDF look at the hit -- its normal (the SDF's own, else a tetrahedral difference), whether the ray is inside,
// its material -- and, for a transmissive material left from inside, the medium a hair beyond (the innermost
// object there, or air). Written as a 16-byte hit record: the normal, then material | inside << 8 | beyond-known
// << 9 | beyond material << 16 (PT_HR_*). Shade reads it; the guide network reads it too.
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_surface(const PtDev* __restrict__ D, unsigned int qsel) {
    const unsigned int count = D->counters[qsel ? PT_C_MARCH0 : PT_C_MARCH1];
    ParamPtr U = (const PT_CONST float*)D->params;
    for (unsigned int i = blockIdx.x * PT_BLOCK + threadIdx.x; i < count; i += gridDim.x * PT_BLOCK) {
        const unsigned int slot = D->qSorted[i];
        const uint4 M = ld4u(D->meta, slot);
        if (mzState(M.z) != PT_ACTIVE) continue;
        const float4 P = ld4(D->pos, slot), Dd = ld4(D->dir, slot);
        const V3 p = v3(P.x, P.y, P.z), dir = v3(Dd.x, Dd.y, Dd.z);
        SdfHit h = sceneFull(U, p);
        V3 ns = h.n;
        if (dot(ns, ns) < 0.25f) ns = normalAt(U, p, marchTol(tolSetting(D), p));
        const bool inside = h.d < 0.0f;
        const unsigned int mi = h.m < D->nmats ? h.m : 0u;
        unsigned int w = mi | (inside ? PT_HR_INSIDE : 0u);
        if (inside && D->mats[mi].transmission > 0.0f && dot(dir, -ns) <= 0.0f) {
            SdfHit hb = sceneFull(U, p + dir * (2.0f * offsetTol(D, p)));
            if (hb.insideW > -1.0e6f && hb.a < D->nmats) w |= PT_HR_BEYOND | (hb.a << 16);   // inside something (-1e7: nothing, air)
        }
        st4(D->hitRec, slot, ns.x, ns.y, ns.z, __uint_as_float(w));
    }
}

// ---------------------------------------------------------------------------------------------------
// SHADE (core/wavefront.js:1244-1442): escape to the environment weighted by the last vertex's MIS; from the hit
// record, the facing and material; an emitter's light weighted against face sampling; light sampling (faces, or the
// sky and sun) with the power heuristic; the BSDF's next direction; the bounce cap and Russian roulette; a point in
// the haze scatters by Henyey-Greenstein.

struct ShadeOut {
    bool requeue, shadow, store, photon;
    bool leaving, gathered, cut, guided, recorded;   // counted per wave by the caller
    V3 sp, spow, sl, sn;     // a photon to store: position, power, direction to its light, surface normal
    unsigned int ssub;       // ...and its sub-list
};

// The photon map's estimate at p for the lobes lb: sum over stored photons within r of f(l_p) P_p / (pi r^2) over the
// 2 x 2 x 2 cells of side 2r around p (core/sppm.js:487-520 sppmGather). A record counts only when walked from its
// own cell, so lists shared by hash collisions never count one twice. A cell holding more records than its share of
// the gather's budget is subsampled, not cut: m of its PT_SPPM_SUBS sub-lists walked (m from the count alone), each
// record found weighing PT_SPPM_SUBS / m -- every record is walked with probability m / PT_SPPM_SUBS, so the sum is
// exact in expectation. A hard bound on the records visited ends a gather early; such a gather is counted.
__device__ __forceinline__ V3 sppmGather(const PtDev* __restrict__ D, const Lobes& lb, V3 p, V3 nrm, bool& cut) {
    const float r = __uint_as_float(D->counters[PT_C_RADIUS]), inv = 1.0f / (2.0f * r);
    const float r2 = r * r;
    const unsigned int* heads = D->sppmHeads;
    const unsigned int* counts = heads + PT_SPPM_HEADS;
    const unsigned int* subs = heads + 2u * PT_SPPM_HEADS;
    const uint2* links = reinterpret_cast<const uint2*>(D->sppmNext);
    const float4* recs = reinterpret_cast<const float4*>(D->sppmRecs);
    int i0 = (int)floorf(p.x * inv - 0.5f), j0 = (int)floorf(p.y * inv - 0.5f), k0 = (int)floorf(p.z * inv - 0.5f);
    const unsigned int share = PT_SPPM_BUDGET / 8u;
    unsigned int visits = 0u;
    V3 sum = v3s(0.0f);
    for (int c = 0; c < 8 && !cut; ++c) {
        int ci = i0 + (c & 1), cj = j0 + ((c >> 1) & 1), ck = k0 + ((c >> 2) & 1);
        unsigned int h = sppmHash(ci, cj, ck);
        unsigned int cnt = counts[h];
        if (cnt == 0u) continue;
        unsigned int m = cnt <= share ? PT_SPPM_SUBS : max(1u, (PT_SPPM_SUBS * share) / cnt);
        float wgt = (float)PT_SPPM_SUBS / (float)m;
        for (unsigned int sl = 0; sl < (m == PT_SPPM_SUBS ? 1u : m) && !cut; ++sl) {
            bool whole = m == PT_SPPM_SUBS;
            unsigned int idx = whole ? heads[h] : subs[sppmHashSub(ci, cj, ck, sl)];
            while (idx != 0u) {
                if (++visits > PT_SPPM_BUDGET * PT_SPPM_HARD) { cut = true; break; }
                const float4 a = recs[2ull * (idx - 1u)], b = recs[2ull * (idx - 1u) + 1ull];
                uint2 ln = links[idx - 1u];
                unsigned int tag = __float_as_uint(b.w) & 31u;
                if (whole || tag == sl) {
                    V3 q = v3(a.x, a.y, a.z);
                    V3 d = q - p;
                    bool own = (int)floorf(q.x * inv) == ci && (int)floorf(q.y * inv) == cj && (int)floorf(q.z * inv) == ck;
                    if (own && dot(d, d) < r2 && dot(octDecode(__float_as_uint(b.w) >> 5, 13u), nrm) >= PT_SPPM_NORMAL_DOT) {
                        V3 f = lobesEval(lb, D->mf, octDecode(__float_as_uint(b.z)));
                        sum += f * v3(a.w, b.x, b.y) * wgt;
                    }
                }
                idx = whole ? ln.x : ln.y;
            }
        }
    }
    return sum * (1.0f / (PT_PI * r2));
}

// LIGHT TRACING (core/materials_pbr.js:554-700 ltConnect and camFootprint; core/haze.js:404-466 ltConnectMedium): a
// photon -- from the sun or a light face -- that came through a delta event (glass, a mirror) asks, where it first meets
// a surface or the haze, whether the camera sees that point, and adds its light to the pixel that sees it. The camera's
// paths leave exactly that light to the photons: a path whose first vertex left it by a non-delta lobe and then met the
// sun or a face through delta events only does not count it. Each path is counted by one side. Off for the thin lens and the illuminance view, and where
// the rays are not one a pixel.
__device__ __forceinline__ bool ltOn(const PtDev* __restrict__ D) {
    return D->u.lt != 0u && D->u.measureMode != 1u && !(D->u.cam.type == 0u && D->u.cam.lensR > 0.0f) &&
           D->tilesX == D->u.width && D->nrays == D->u.width * D->u.height;
}
// The camera's view of a point P: false where the camera cannot see it (behind it, off the picture); else the pixel,
// the way to the camera v and its length, and what one photon's light there is multiplied by -- the density, per unit
// area at P (surface: |n.v|) or per unit volume (in the haze: no normal), of the pixel's camera rays landing there,
// over the pixel's area on the image plane at distance 1. A pinhole: G = |n.v| dist / z^3 (a plane's area at distance
// 1 to solid angle, cos^3, to area at P, dist^2 / |n.v|); orthographic: G = |n.v|. A pixel is 4 fx fy / (w h) of the
// plane. The pixel is the one genrays' mapping sends a ray through P from.
__device__ __forceinline__ bool camSees(const PtDev* __restrict__ D, V3 P, V3 nrm, bool surface, unsigned int& pix, V3& v,
                                        float& dist, float& g) {
    const PtCamera& cam = D->u.cam;
    const V3 right = ld3(cam.right), up = ld3(cam.up), fwd = ld3(cam.fwd);
    const V3 rel = P - ld3(cam.pos);
    const float z = -dot(rel, fwd);                       // depth in front of the camera
    if (!(z > 1e-4f)) return false;
    float x, y, G;
    if (cam.type == 0u) {
        dist = length(rel);
        v = rel * (-1.0f / dist);
        x = dot(rel, right) / (z * cam.fx);
        y = dot(rel, up) / (z * cam.fy);
        G = (surface ? fabsf(dot(nrm, v)) : 1.0f) * dist / (z * z * z);
    } else {
        dist = z;
        v = fwd;
        x = dot(rel, right) / cam.fx;
        y = dot(rel, up) / cam.fy;
        G = surface ? fabsf(dot(nrm, fwd)) : 1.0f;
    }
    if (!(fabsf(x) < 1.0f && fabsf(y) < 1.0f)) return false;
    const float fw = (float)D->u.width, fh = (float)D->u.height;
    const unsigned int px = (unsigned int)fminf(fmaxf((x + 1.0f) * 0.5f * fw, 0.0f), fw - 1.0f);
    const unsigned int py = (unsigned int)fminf(fmaxf((y + 1.0f) * 0.5f * fh, 0.0f), fh - 1.0f);
    pix = px + py * D->u.width;
    g = G * fw * fh / (4.0f * cam.fx * cam.fy);
    return true;
}
// The connection, queued as the slot's shadow ray: from o towards the camera, stopping two offsets short of it; what it
// carries goes to its pixel if nothing is in the way (pt_occlude). shDir.w < 0 marks it a connection: no path's.
__device__ __forceinline__ void ltQueue(const PtDev* __restrict__ D, unsigned int slot, V3 o, V3 v, float len, V3 c,
                                        unsigned int pix) {
    st4(D->shOrig, slot, o.x, o.y, o.z, len - 2.0f * offsetTol(D, o));
    st4(D->shDir, slot, v.x, v.y, v.z, -1.0f);
    st4(D->shRad, slot, c.x, c.y, c.z, __uint_as_float(pix));
}

__device__ __forceinline__ ShadeOut shadeOne(const PtDev* __restrict__ D, const PtMaterial* __restrict__ smat,
                                             const unsigned int* seed, unsigned int slot) {
    ShadeOut out; out.requeue = false; out.shadow = false; out.store = false;
    out.leaving = out.gathered = out.cut = out.guided = out.recorded = false;
    float4 P = ld4(D->pos, slot), Dd = ld4(D->dir, slot), B = ld4(D->beta, slot), R = ld4(D->rad, slot), A = ld4(D->aux, slot);
    uint4 M = ld4u(D->meta, slot);
    V3 p = v3(P.x, P.y, P.z), dir = v3(Dd.x, Dd.y, Dd.z), beta = v3(B.x, B.y, B.z), rad = v3(R.x, R.y, R.z);
    float ltot = P.w, n = Dd.w, prevPdf = B.w, vtxLtot = R.w, sunmis = A.x, collide = A.y, lambda = A.z;
    // the absorbing medium the path travels in (materials.js attenuationDistance): its material + 1, 0 none
    const unsigned int medIn = __float_as_uint(A.w);
    unsigned int medOut = medIn;
    unsigned int pix = M.x, counts = M.y, z = M.z, tag = M.w;
    unsigned int bounces = mzBounces(z), state = mzState(z), flags = mzFlags(z);
    const PtEnv& env = D->u.env;
    const bool photon = (flags & PT_F_PHOTON) != 0u;
    out.photon = photon;
    // guide training: a recorded path's previous vertex learns S, the radiance the path had when this event came
    unsigned int gr = 0u;
    const bool grec = !photon && guideRecorded(D, slot, gr);
    const bool spend = grec && (flags & PT_F_GPEND);
    V3 skStart = v3s(0.0f);
    const V3 radStart = rad;
    bool emitted = false;
    if (spend) {
        const float4 nee = reinterpret_cast<const float4*>(D->gPath)[2 * gr + 1];
        skStart = rad + v3(nee.x, nee.y, nee.z);
        flags &= ~PT_F_GPEND;
    }
    // photons draw from their own stream, so the map and the camera paths of one sample are independent
    Rng rng = rngAt(seed, slot, bounces, (0x5AADEu ^ ((counts & 0xFFFFu) << 12)) ^ (photon ? 0x80000000u : 0u));
    bool terminate = false;
    bool direct = false;
    // a gathering path's light that reached its gather vertex through delta events only is the photon map's; the sun's
    // that reached the camera's first vertex so is the light tracer's
    const bool mapped = (flags & (PT_F_CAUSTIC | PT_F_SDELTA)) == (PT_F_CAUSTIC | PT_F_SDELTA);
    const bool traced = (flags & (PT_F_LTCLS | PT_F_SDELTA)) == (PT_F_LTCLS | PT_F_SDELTA);
    const bool lt = ltOn(D);

    if (state == PT_ESCAPED) {
        // the illuminance view: a camera ray that met nothing has no surface to measure on, and shows 0
        if (!photon && !(D->u.measureMode == 1u && bounces == 0u)) {
            // the sun through its aperture emits photons, so a mapped path leaves its light to the map; the sky
            // emits none and stays path traced
            V3 Le = envBackground(env, dir) + (((mapped || traced) && sunFlux(env) > 0.0f) ? v3s(0.0f) : envSun(env, p, dir));
            rad += beta * Le * sunmis;
        }
        direct = bounces == 0u;
        terminate = true;
    } else if (state == PT_MAXITER) {
        terminate = true;
    } else if (state == PT_SCATTER && photon) {
        // a photon ends where it scatters in the haze; a sun photon that came through a delta event sends its light
        // there to the camera through the phase function, the albedo its collision's weight (haze.js ltConnectMedium)
        if (lt && (flags & PT_F_HADDELTA)) {
            unsigned int cp; V3 v; float dist, g;
            if (camSees(D, p, v3s(0.0f), false, cp, v, dist, g)) {
                const V3 c = beta * (D->u.haze.albedo * hgEval(D->u.haze.g, dot(dir, v)) * g);
                if (max3(c) > 0.0f) { ltQueue(D, slot, p, v, dist, c, cp); out.shadow = true; }
            }
        }
        terminate = true;
    } else if (state == PT_SCATTER) {
        // A POINT IN THE HAZE (core/haze.js:292-600 mediumVertex): albedo, light sampling with the phase function,
        // a new direction from it.
        const PtHaze& hz = D->u.haze;
        bounces += 1u;
        beta *= hz.albedo;
        float fs = faceShareAt(D, bounces);
        V4 u0 = rngNext(rng), u1 = rngNext(rng);
        if (D->u.nee != 0u) {
            V3 L = v3s(0.0f), le = v3s(0.0f), orig = p;
            float pS = 0.0f, tmax = 3.0e38f;
            if (u0.x < fs) {
                const PtLightFace& f = D->faces[facePick(D, u1.w * D->faceTotal)];
                V3 y = ld3(f.centre);
                if (f.shape == 1u) {
                    float rr = f.halfx * sqrtf(u0.y), sp, cp; __sincosf(2.0f * PT_PI * u0.z, &sp, &cp);
                    y += (ld3(f.u) * cp + ld3(f.v) * sp) * rr;
                } else {
                    y += ld3(f.u) * ((2.0f * u0.y - 1.0f) * f.halfx) + ld3(f.v) * ((2.0f * u0.z - 1.0f) * f.halfy);
                }
                V3 dy = y - p; float dist = length(dy); L = dy * (1.0f / fmaxf(dist, 1e-20f));
                float cosY = -dot(ld3(f.n), L);
                if (dist > 0.0f && fabsf(cosY) > 1e-6f && (cosY > 0.0f || f.twoSided)) {
                    pS = faceDensity(D, fs, f, dist, cosY); le = ld3(f.le);
                    tmax = dist - 2.0f * offsetTol(D, y);
                }
            } else if (envLit(env)) {
                float pick = (u0.x - fs) / fmaxf(1.0f - fs, 1e-6f);
                L = pick < sunPick(env) ? sampleSunCone(env, u0.y, u0.z) : sampleCosine(v3(0.0f, 0.0f, 1.0f), u0.y, u0.z);
                float pe = envPdf(env, L, v3(0.0f, 0.0f, 1.0f));
                pS = (1.0f - fs) * pe;
                le = envBackground(env, L) + envSun(env, p, L);
            }
            if (pS > 0.0f) {
                float ph = hgEval(hz.g, dot(dir, L));
                V3 contrib = beta * le * (ph * misWeight(pS, ph) / pS);
                if (max3(contrib) > 1e-7f) {
                    st4(D->shOrig, slot, orig.x, orig.y, orig.z, tmax);
                    st4(D->shDir, slot, L.x, L.y, L.z, n);
                    st4(D->shRad, slot, contrib.x, contrib.y, contrib.z, __uint_as_float(pix));
                    out.shadow = true;
                }
            }
        }
        V3 nd = hgSample(hz.g, dir, u1.x, u1.y);
        float ph = hgEval(hz.g, dot(dir, nd));
        float pLout = (1.0f - fs) * (envLit(env) ? envPdf(env, nd, v3(0.0f, 0.0f, 1.0f)) : 0.0f);
        sunmis = (D->u.nee != 0u && pLout > 0.0f) ? misWeight(ph, pLout) : 1.0f;
        prevPdf = D->u.nee != 0u ? ph : -1.0f;
        vtxLtot = ltot;
        dir = nd;
        flags &= ~(PT_F_DELTAONLY | PT_F_CAUSTIC | PT_F_SDELTA | PT_F_LTCLS);
        // the camera's first vertex in the haze: the sun it meets past here through delta events is the light tracer's
        if (lt && bounces == 1u) flags |= PT_F_LTCLS;
        if (bounces > D->u.maxbounce) terminate = true;
        else {
            collide = ltot + hazeDraw(hz, p, dir, u1.z) * n;
            out.requeue = true;
        }
    } else {
        // A SURFACE HIT: what the surface kernel found there
        const float4 HR = ld4(D->hitRec, slot);
        const unsigned int hw = __float_as_uint(HR.w);
        V3 ns = v3(HR.x, HR.y, HR.z);
        bool inside = (hw & PT_HR_INSIDE) != 0u;
        V3 nrm = inside ? -ns : ns;
        if (dot(dir, nrm) > 0.0f) {
            // hit from behind (core/scene_api.js:127-136 pathFacing): moved on, no bounce
            float off = offsetTol(D, p);
            p = p + dir * off; ltot += off * n;
            out.requeue = true;
            out.leaving = true;
        } else {
            bounces += 1u;
            // ABSORPTION ALONG THE WAY (tracebasics.js absorbAlong): inside an absorbing medium the path loses
            // base^(d / attenuation) per channel over the distance d since its last vertex
            if (medIn != 0u) beta *= mediumTr(smat[medIn - 1u], fmaxf(ltot - vtxLtot, 0.0f) / fmaxf(n, 1e-4f));
            // the illuminance view (measureMode 1): the camera's first hit is the white Lambertian kept after the table
#if PT_SCENE_MATCODE
            // the material's own code runs on the hit's material (not on the illuminance view's white)
            const bool whiteView = D->u.measureMode == 1u && bounces == 1u && !photon;
            PtMaterial m = smat[whiteView ? PT_MAX_MATERIALS : (hw & 0xFFu)];
            if (!whiteView) sceneMaterial((const PT_CONST float*)D->params, hw & 0xFFu, p, ns, m);
#else
            const PtMaterial& m = smat[(D->u.measureMode == 1u && bounces == 1u && !photon) ? PT_MAX_MATERIALS : (hw & 0xFFu)];
#endif
            V3 Le = ld3(m.emissive);
            if (max3(Le) > 0.0f && (photon || ((mapped || traced) && m.face >= 0 && faceFlux(D->faces[m.face]) > 0.0f))) {
                // a photon ends at a light; a mapped path's light is the map's, a traced one's the connections' (the
                // face's photons')
                terminate = true;
                emitted = true;
            } else if (max3(Le) > 0.0f) {
                // AN EMITTER HIT: weighted against the light sampling the previous vertex ran
                float w = 1.0f;
                if (m.face >= 0) {
                    const PtLightFace& f = D->faces[m.face];
                    float cosY = -dot(ld3(f.n), dir);
                    if (!(cosY > 0.0f || f.twoSided)) Le = v3s(0.0f);
                    else if (prevPdf >= 0.0f) {
                        float dist = fmaxf(ltot - vtxLtot, 0.0f) / fmaxf(n, 1e-4f);
                        w = misWeight(prevPdf, faceDensity(D, faceShareAt(D, bounces - 1u), f, dist, cosY));
                    }
                }
                rad += beta * Le * w;
                direct = bounces == 1u;
                terminate = true;
                emitted = true;
            } else {
                // the medium beyond the surface: the material's own index going in, air (or the enclosing
                // material) going out
                float n0 = n;
                float n1 = fmaxf(m.ior, 1.0f);
                if (inside) {
                    n1 = 1.0f;
                    if (hw & PT_HR_BEYOND) n1 = fmaxf(smat[hw >> 16].ior, 1.0f);
                }
                // DISPERSION (core/spectral.js:121-135): a wavelength drawn at the first refraction into a
                // dispersive medium; its index n(lambda) by Cauchy through n_d and V_d, the ray's own n after
                V4 ud = rngNext(rng);
                bool draw = false;
                V3 specW = v3s(1.0f);
                float lambda0 = lambda;
                if (m.abbe > 0.0f && m.transmission > 0.0f && !inside) {
                    if (lambda <= 0.0f) {
                        draw = true;
                        lambda = 0.38f + 0.40f * ud.w;   // micrometres, uniform over the visible
                        specW = spectralWeight(lambda);
                    }
                    n1 = cauchyN(n1, m.abbe, lambda);
                }
                Lobes lb = makeLobes(nrm, -dir, m, n0 / n1, D->mf);
                float fs = faceShareAt(D, bounces);
                bool neeOn = D->u.nee != 0u && !photon;
                // PATH GUIDING: where it applies and the network is trained, directions come from its mixture with
                // probability gc (the network's own fraction, capped by guideFrac), from the BSDF otherwise; every
                // density below is the two strategies' mixture
                const bool gElig = !photon && guideMaterial(m);
                const bool guided = gElig && guideActive(D);
                // the mixture is read from the slot's record where it is used, not held across the shading
                const float gc = guided ? fminf(guideFracAt(D, slot), D->u.guideFrac) : 0.0f;
                bool gathered = false;
                // a surface with a lobe the map can be gathered through (any but a purely delta one)
                const bool evaluable = !lobesNone(m);
                if (photon && evaluable && (flags & PT_F_HADDELTA)) {
                    // a caustic photon lands where the material has a non-delta lobe: stored (the store itself is
                    // wave-aggregated by the caller), on one of the cell's sub-lists drawn uniformly; it goes on only
                    // through a delta lobe (below)
                    if (D->u.sppm != 0u) {
                        out.store = true;
                        out.sp = p; out.spow = beta; out.sl = -dir; out.sn = nrm;
                        out.ssub = min((unsigned int)(ud.z * (float)PT_SPPM_SUBS), PT_SPPM_SUBS - 1u);
                    }
                    // ...and a sun photon's light from here to the camera, by the non-delta lobes the camera's side
                    // leaves to it, seen from where the camera is (ltConnect); from air only
                    if (lt && fabsf(n0 - 1.0f) < 1e-3f) {
                        unsigned int cp; V3 v; float dist, g;
                        if (camSees(D, p, nrm, true, cp, v, dist, g) && dot(nrm, v) > 0.0f) {
                            const Lobes lc = makeLobes(nrm, v, m, n0 / n1, D->mf);
                            const V3 c = beta * lobesEval(lc, D->mf, -dir) * g;
                            if (max3(c) > 0.0f) {
                                const V3 o = p + nrm * (2.0f * offsetTol(D, p));
                                const V3 toCam = D->u.cam.type == 0u ? ld3(D->u.cam.pos) - o : v * dist;
                                const float len = length(toCam);
                                ltQueue(D, slot, o, toCam * (1.0f / len), len, c, cp);
                                out.shadow = true;
                            }
                        }
                    }
                } else if (!photon && D->u.sppm != 0u && (D->u.sppm >= 2u || (flags & PT_F_DELTAONLY)) && evaluable &&
                           !(lt && bounces == 1u)) {
                    // THE GATHER (core/sppm.js:487-494): the camera path's first non-delta vertex after delta events
                    // (sppm 1, the reference's); with sppm 2 every non-delta camera vertex, so that caustic light a
                    // path meets after a diffuse bounce is the map's too, not a rare BSDF hit on the sun through glass.
                    // With light tracing the first vertex is the photons' connections' instead
                    bool cut = false;
                    rad += beta * sppmGather(D, lb, p, nrm, cut);
                    gathered = true;
                    out.gathered = true;
                    out.cut = cut;
                }
                if (neeOn && !lobesNone(m)) {
                    V4 u0 = rngNext(rng), u1 = rngNext(rng);
                    V3 L = v3s(0.0f), le = v3s(0.0f);
                    float pS = 0.0f, tmax = 3.0e38f;
                    if (u0.x < fs) {
                        const PtLightFace& f = D->faces[facePick(D, u1.x * D->faceTotal)];
                        V3 y = ld3(f.centre);
                        if (f.shape == 1u) {
                            float rr = f.halfx * sqrtf(u0.y), sp, cp; __sincosf(2.0f * PT_PI * u0.z, &sp, &cp);
                            y += (ld3(f.u) * cp + ld3(f.v) * sp) * rr;
                        } else {
                            y += ld3(f.u) * ((2.0f * u0.y - 1.0f) * f.halfx) + ld3(f.v) * ((2.0f * u0.z - 1.0f) * f.halfy);
                        }
                        V3 dy = y - p; float dist = length(dy); L = dy * (1.0f / fmaxf(dist, 1e-20f));
                        float cosY = -dot(ld3(f.n), L);
                        if (dist > 0.0f && dot(nrm, L) > 0.0f && fabsf(cosY) > 1e-6f && (cosY > 0.0f || f.twoSided)) {
                            pS = faceDensity(D, fs, f, dist, cosY);
                            le = ld3(f.le) * (1.0f / fmaxf(n0 * n0, 1.0f));
                            V3 o = p + nrm * (2.0f * offsetTol(D, p));
                            tmax = length(y - o) - 2.0f * offsetTol(D, p);
                        }
                    } else if (envLit(env)) {
                        float pick = (u0.x - fs) / fmaxf(1.0f - fs, 1e-6f);
                        L = pick < sunPick(env) ? sampleSunCone(env, u0.y, u0.z) : sampleCosine(nrm, u0.y, u0.z);
                        if (dot(nrm, L) > 0.0f) {
                            pS = (1.0f - fs) * envPdf(env, L, nrm);
                            le = envBackground(env, L) + envSun(env, p, L);
                        }
                    }
                    if (pS > 0.0f && max3(le) > 0.0f) {
                        V3 f = lobesEval(lb, D->mf, L);
                        float pB = lobesPdf(lb, L);
                        if (guided) pB = gc * guidePdfAt(D, slot, L) + (1.0f - gc) * pB;
                        float NoL = dot(nrm, L);
                        V3 contrib = beta * f * le * (NoL * misWeight(pS, pB) / pS);
                        if (max3(contrib) > 1e-7f) {
                            V3 o = p + nrm * (2.0f * offsetTol(D, p));
                            if (tmax < 3.0e38f) L = normalize((o + L * (tmax + 2.0f * offsetTol(D, p))) - o);
                            st4(D->shOrig, slot, o.x, o.y, o.z, tmax);
                            st4(D->shDir, slot, L.x, L.y, L.z, n0);
                            st4(D->shRad, slot, contrib.x, contrib.y, contrib.z, __uint_as_float(pix));
                            out.shadow = true;
                        }
                    }
                }
                const V3 wi = dir;
                LobeSample s = lobesSample(lb, D->mf, rngNext(rng));
                float pTot = -1.0f;    // the sampled direction's density (non-delta): the BSDF's, or the mixture's
                if (guided) {
                    const V4 ug = rngNext(rng);
                    // a BSDF sample that failed (a microfacet direction below the horizon) stays a failed sample: its
                    // direction is a placeholder, not a draw of the BSDF's density
                    bool valid = max3(s.w) > 0.0f;
                    if (ug.w < gc) {
                        s.dir = guideSampleAt(D, slot, ug);
                        s.isDelta = false; s.transmitted = false; s.ev = PT_EV_EVAL;
                        valid = true;
                    }
                    if (!valid) {
                        s.w = v3s(0.0f);
                    } else if (s.isDelta) {
                        s.w = s.w * (1.0f / (1.0f - gc));
                    } else {
                        // the mixture's density and value over the whole sphere: a rough transmissive material's
                        // refraction is evaluated below the surface too, so a direction either strategy draws there
                        // carries what the BSDF passes (the plain tracer's refraction weight, reached by MIS)
                        lb.through = true;
                        pTot = gc * guidePdfAt(D, slot, s.dir) + (1.0f - gc) * lobesPdf(lb, s.dir);
                        const float cosO = dot(nrm, s.dir);
                        s.w = (pTot > 0.0f && cosO != 0.0f) ? lobesEval(lb, D->mf, s.dir) * (fabsf(cosO) / pTot) : v3s(0.0f);
                        s.transmitted = cosO < 0.0f;
                    }
                    out.guided = true;
                }
                if (photon && !s.isDelta) terminate = true;     // the map holds caustics: L S+ D only
                dir = s.dir;
                beta *= s.w;
                if (s.transmitted) {
                    n = n1;
                    if (draw) beta *= specW;
                    // through the surface: into its material, or out to the one beyond (absorbing ones only)
                    const unsigned int into = inside ? ((hw & PT_HR_BEYOND) ? (hw >> 16) : 0xFFFFFFFFu) : (hw & 0xFFu);
                    medOut = (into != 0xFFFFFFFFu && smat[into].attenuation > 0.0f) ? into + 1u : 0u;
                } else if (draw) {
                    lambda = lambda0;   // reflected: the wavelength is not the path's
                }
                if (!(max3(s.w) > 0.0f)) terminate = true;
                bool misOut = neeOn && !s.isDelta;
                float pOut = misOut ? (pTot >= 0.0f ? pTot : lobesPdf(lb, dir)) : 0.0f;
                bool above = dot(nrm, dir) > 0.0f && !s.isDelta;
                if (misOut) {
                    float pLout = above && envLit(env) ? (1.0f - fs) * envPdf(env, dir, nrm) : 0.0f;
                    sunmis = pLout > 0.0f ? misWeight(pOut, pLout) : 1.0f;
                    prevPdf = above ? pOut : -1.0f;
                } else {
                    sunmis = 1.0f;
                    prevPdf = -1.0f;
                }
                vtxLtot = ltot;
                if (s.ev == PT_EV_EVAL) counts += 0x10000u;
                if (!s.isDelta) flags &= ~PT_F_DELTAONLY;
                if (photon && s.isDelta) flags |= PT_F_HADDELTA;
                // THE CLASSES ANOTHER ESTIMATOR COUNTS: kept through delta events (SDELTA), closed at any other; opened
                // where a non-delta lobe leaves a vertex on its own side -- after a gather, the light met through
                // delta events only is the map's (it gathered it); at the camera's first vertex with light tracing,
                // the sun's so met is the photons'. Light that came through a rough transmission is neither's (the
                // map and the connections read the side the vertex is seen from), and a delta lobe leaves the path
                // as it was.
                if (s.isDelta) { if (flags & (PT_F_CAUSTIC | PT_F_LTCLS)) flags |= PT_F_SDELTA; }
                else flags &= ~(PT_F_CAUSTIC | PT_F_LTCLS | PT_F_SDELTA);
                if (!photon && !s.isDelta && !s.transmitted) {
                    if (gathered) flags |= PT_F_CAUSTIC;
                    else if (lt && bounces == 1u && evaluable) flags |= PT_F_LTCLS;
                }
                if (bounces > D->u.maxbounce) terminate = true;
                if (!terminate && D->u.diffuseCap != 0u && (counts >> 16) > D->u.diffuseCap) terminate = true;
                if (!terminate && D->u.rr != 0u && bounces >= D->u.rrStart) {
                    float pr = fminf(fmaxf(max3(beta), 0.05f), 0.95f);
                    if (ud.x > pr) terminate = true; else beta *= 1.0f / pr;
                }
                if (!terminate && grec && gElig && !s.isDelta && s.ev == PT_EV_EVAL && D->gVerts[gr] < D->gMaxV) {
                    // a training vertex: where, the view, the density it sampled w with, the normal, w, f cos at w,
                    // the throughput after it; S comes at the next event
                    const float pw = pTot >= 0.0f ? pTot : lobesPdf(lb, dir);
                    const V3 fc = lobesEval(lb, D->mf, dir) * fabsf(dot(nrm, dir));
                    const unsigned int k = D->gVerts[gr];
                    float4* rec = guideRec(D, gr, k);
                    rec[0] = make_float4(p.x, p.y, p.z, m.roughness);
                    rec[1] = make_float4(-wi.x, -wi.y, -wi.z, pw);
                    rec[2] = make_float4(nrm.x, nrm.y, nrm.z, 0.0f);
                    rec[3] = make_float4(dir.x, dir.y, dir.z, lobesPdf(lb, dir));
                    rec[4] = make_float4(fc.x, fc.y, fc.z, 0.0f);
                    rec[5] = make_float4(beta.x, beta.y, beta.z, 0.0f);
                    D->gVerts[gr] = k + 1u;
                    flags |= PT_F_GPEND;
                    out.recorded = true;
                }
                if (!terminate) {
                    float off = offsetTol(D, p);
                    p = p + dir * off; ltot += off * n;
                    collide = 3.0e38f;
                    if (hazeOn(D) && n <= 1.0f) collide = ltot + hazeDraw(D->u.haze, p, dir, ud.y) * n;
                    out.requeue = true;
                }
            }
        }
    }
    if (spend) {
        // the previous vertex's S: the path's radiance before this event -- and, with light sampling on, after the
        // light this event met on leaving the scene or at an emitter, which light sampling also reaches: the network
        // learns the light it must find itself, free of the MIS weights its own density sets
        V3 sk = skStart;
        if (D->u.nee != 0u && (state == PT_ESCAPED || emitted)) sk += rad - radStart;
        float4* rec = guideRec(D, gr, D->gVerts[gr] - 1u);
        rec[6] = make_float4(sk.x, sk.y, sk.z, 0.0f);
        rec[4].w = 1.0f;
    }
    if (terminate) {
        out.requeue = false;
        if (grec) reinterpret_cast<float4*>(D->gPath)[2 * gr] = make_float4(rad.x, rad.y, rad.z, 1.0f);
        if (!photon) {
            if (badDeposit(rad)) atomicAdd(&D->stats[PT_S_NONFINITE], 1ull);
            else depositRGB(D, pix, rad, direct);
            atomicAdd(reinterpret_cast<unsigned int*>(D->frame) + 4ull * pix + 3ull, 1u);
        }
    } else if (out.requeue) {
        st4(D->pos, slot, p.x, p.y, p.z, ltot);
        st4(D->dir, slot, dir.x, dir.y, dir.z, n);
        st4(D->beta, slot, beta.x, beta.y, beta.z, prevPdf);
        st4(D->rad, slot, rad.x, rad.y, rad.z, vtxLtot);
        st4u(D->meta, slot, make_uint4(pix, counts, mzPack(bounces, PT_ACTIVE, flags), tag));
        st4(D->aux, slot, sunmis, collide, lambda, __uint_as_float(medOut));
    }
    return out;
}

extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_shade(const PtDev* __restrict__ D, unsigned int qsel) {
    __shared__ PtMaterial smat[PT_MAX_MATERIALS + 1];
    {
        const unsigned int words = D->nmats * (unsigned int)(sizeof(PtMaterial) / 4);
        const unsigned int* src = reinterpret_cast<const unsigned int*>(D->mats);
        unsigned int* dst = reinterpret_cast<unsigned int*>(smat);
        for (unsigned int k = threadIdx.x; k < words; k += PT_BLOCK) dst[k] = src[k];
        // after the table: the illuminance view's white Lambertian
        if (threadIdx.x == 0) {
            PtMaterial& w = smat[PT_MAX_MATERIALS];
            w.base[0] = w.base[1] = w.base[2] = 1.0f;
            w.metallic = 0.0f; w.roughness = 1.0f; w.transmission = 0.0f; w.ior = 1.5f; w.specular = 0.0f;
            w.emissive[0] = w.emissive[1] = w.emissive[2] = 0.0f;
            w.abbe = 0.0f; w.face = -1; w.specularTint = 0.0f; w.pad[0] = w.pad[1] = 0.0f;
        }
    }
    __syncthreads();
    unsigned int* C = D->counters;
    // this round's rays: the queue march read (binned), every one of them now at its event
    const unsigned int count = C[qsel ? PT_C_MARCH0 : PT_C_MARCH1];
    unsigned int seed[4];
    sampleSeed(D, seed);
    unsigned int* qout = qsel ? D->qMarch1 : D->qMarch0;
    unsigned int* cout = &C[qsel ? PT_C_MARCH1 : PT_C_MARCH0];
    unsigned long long segs = 0ull, psegs = 0ull, stored = 0ull, dropped = 0ull;
    unsigned long long nLeave = 0ull, nGather = 0ull, nCut = 0ull, nGuided = 0ull, nRec = 0ull;
    for (unsigned int base = blockIdx.x * PT_BLOCK; base < count; base += gridDim.x * PT_BLOCK) {
        unsigned int i = base + threadIdx.x;
        ShadeOut o; o.requeue = false; o.shadow = false; o.store = false; o.photon = false;
        unsigned int slot = 0u;
        if (i < count) {
            slot = D->qSorted[i];
            o = shadeOne(D, smat, seed, slot);
            if (o.photon) psegs += 1ull; else segs += 1ull;
            nLeave += o.leaving; nGather += o.gathered; nCut += o.cut; nGuided += o.guided; nRec += o.recorded;
        }
        unsigned int qi = waveAppend(cout, o.requeue);
        if (o.requeue) qout[qi] = slot;
        unsigned int si = waveAppend(&C[PT_C_SHADOW], o.shadow);
        if (o.shadow) D->qShadow[si] = slot;
        // a photon into the map: its record, then the heads of its cell's list and of its sub-list swapped for it
        // and the cell's count raised
        unsigned int ri = waveAppend(&C[PT_C_PHOTON], o.store);
        if (o.store && ri < D->sppmCap) {
            float4* rec = reinterpret_cast<float4*>(D->sppmRecs) + 2ull * ri;
            rec[0] = make_float4(o.sp.x, o.sp.y, o.sp.z, o.spow.x);
            rec[1] = make_float4(o.spow.y, o.spow.z, __uint_as_float(octEncode(o.sl)),
                                 __uint_as_float((octEncode(o.sn, 13u) << 5) | o.ssub));
            const float inv = 1.0f / (2.0f * __uint_as_float(C[PT_C_RADIUS]));
            int ci = (int)floorf(o.sp.x * inv), cj = (int)floorf(o.sp.y * inv), ck = (int)floorf(o.sp.z * inv);
            unsigned int h = sppmHash(ci, cj, ck);
            unsigned int* heads = D->sppmHeads;
            uint2 ln;
            ln.x = atomicExch(&heads[h], ri + 1u);
            ln.y = atomicExch(&heads[2u * PT_SPPM_HEADS + sppmHashSub(ci, cj, ck, o.ssub)], ri + 1u);
            atomicAdd(&heads[PT_SPPM_HEADS + h], 1u);
            reinterpret_cast<uint2*>(D->sppmNext)[ri] = ln;
            stored += 1ull;
        } else if (o.store) {
            dropped += 1ull;
        }
    }
    statAdd(D, PT_S_SEGMENTS, segs);
    statAdd(D, PT_S_PHOTONSEGS, psegs);
    statAdd(D, PT_S_PHOTONS, stored);
    statAdd(D, PT_S_OVERFLOW, dropped);
    statAdd(D, PT_S_LEAVING, nLeave);
    statAdd(D, PT_S_GATHERS, nGather);
    statAdd(D, PT_S_GATHERCUT, nCut);
    statAdd(D, PT_S_GUIDED, nGuided);
    statAdd(D, PT_S_GRECORDS, nRec);
}

// ---------------------------------------------------------------------------------------------------
// OCCLUDE (core/wavefront.js:1481-1692): a shadow ray is blocked where |d| < marchTol, reaches its light when it
// leaves the scene (sky, sun) or when the sphere of free space around it covers what is left to the face; a
// reached ray deposits its contribution, dimmed by the haze it crossed. Persistent, like march.
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_occlude(const PtDev* __restrict__ D, unsigned int qsel) {
    unsigned int* C = D->counters;
    // the march queue this round read (shade has consumed it) and the march's fetch heads start empty for the round
    // after next
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) C[qsel ? PT_C_MARCH1 : PT_C_MARCH0] = 0u;
        if (threadIdx.x < PT_PARTS) C[PT_C_HEADS_MARCH + threadIdx.x] = 0u;
    }
    const unsigned int count = C[PT_C_SHADOW];
    const float setting = tolSetting(D);
    const float R2 = D->u.envRadius * D->u.envRadius;
    const float stepscale = D->u.stepscale;
    const unsigned int maxiter = D->u.maxiter;
    const unsigned int refill = D->refillOcclude;
    const float relax = D->u.relax > 1.0f ? D->u.relax : 1.0f;
    ParamPtr U = (const PT_CONST float*)D->params;
    int slot = -1;
    V3 p = v3s(0.0f), dir = v3s(0.0f), o = v3s(0.0f);
    float tmax = 0.0f, trav = 0.0f, dir4w = 0.0f;   // dir4w < 0: a light tracer's connection
    float om = 1.0f, rPrev = 0.0f, sPrev = 0.0f;    // over-relaxation, as the march's
    unsigned long long mask = ~0ull;
    unsigned int steps = 0u;
    WaveFetch f = fetchInit();
    unsigned long long nsteps = 0ull, nsegs = 0ull, nflops = 0ull, ncut = 0ull;
    for (;;) {
        unsigned long long idle = __ballot(slot < 0);
        if (!f.done && (unsigned int)__popcll(idle) >= refill) {
            unsigned int k = fetchNext(f, C + PT_C_HEADS_SHADOW, count, slot < 0);
            if (slot < 0 && k != 0xFFFFFFFFu) {
                slot = (int)ldu(D->qSorted, k);
                float4 Oo = ld4(D->shOrig, slot), Dd = ld4(D->shDir, slot);
                o = v3(Oo.x, Oo.y, Oo.z); p = o; tmax = Oo.w;
                dir = v3(Dd.x, Dd.y, Dd.z);
                dir4w = Dd.w;
                trav = 0.0f; steps = 0u;
                mask = ldu64(D->maskBuf, slot);
                nsegs += 1ull;
                om = relax;
                rPrev = sPrev = 0.0f;
            }
        }
        if (__ballot(slot >= 0) == 0ull) {
            if (f.done) break;
            continue;
        }
        if (slot >= 0) {
            int result = 0;   // 1 reached, 2 blocked
            #pragma unroll 1
            for (int k = 0; k < PT_MARCH_STEPS; ++k) {
                unsigned int w = PT_STEP_FLOPS;
                float d = sceneDistM(U, p, mask, w);
                nflops += w;
                float ea = fabsf(d);
                float r = ea * stepscale;
                if (om > 1.0f && r + rPrev <= sPrev) {
                    float back = sPrev - rPrev;
                    p = p - dir * back; trav -= back;
                    om = 1.0f;
                    steps += 1u; nsteps += 1ull;
                    continue;
                }
                if (ea < marchTol(setting, p)) { result = 2; break; }
                if (ea >= tmax - trav || (dot(p, dir) > 0.0f && dot(p, p) > R2)) { result = 1; break; }
                // cut short by maxiter: counted as reaching the light, as the reference does (blocked would bias dark)
                if (steps > maxiter) { result = 3; break; }
                float t = r * om;
                p = p + dir * t; trav += t;
                rPrev = r; sPrev = t;
                steps += 1u; nsteps += 1ull;
            }
            if (result != 0) {
                if (result == 3) { ncut += 1ull; result = 1; }
                if (result == 1) {
                    float4 Rr = ld4(D->shRad, slot);
                    V3 c = v3(Rr.x, Rr.y, Rr.z);
                    if (hazeOn(D)) c *= hazeTr(D->u.haze, o, dir, tmax < 3.0e38f ? tmax : 1.0e30f);
                    depositRGB(D, __float_as_uint(Rr.w), c, false);
                    unsigned int gr;
                    if (!(dir4w < 0.0f) && guideRecorded(D, (unsigned int)slot, gr)) {
                        // the path's light-sampling total, for its vertices' training radiance
                        float4* nee = reinterpret_cast<float4*>(D->gPath) + 2 * gr + 1;
                        *nee = make_float4(nee->x + c.x, nee->y + c.y, nee->z + c.z, 0.0f);
                    }
                }
                slot = -1;
            }
        }
    }
    statAdd(D, PT_S_SHADOWSTEPS, nsteps);
    statAdd(D, PT_S_SHADOWSEGS, nsegs);
    statAdd(D, PT_S_SHADOWFLOPS, nflops);
    statAdd(D, PT_S_SHADOWCUT, ncut);
}

// ---------------------------------------------------------------------------------------------------
// RESOLVE (core/wavefront.js:414-494; core/measure.js:36-57 stepMoments): the sample's mean per pixel into the
// running mean, and the exponentially weighted moments (V, S) of its luminance that the error bar reads,
// all in FP64; the frame cleared for the next sample.
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_resolve(const PtDev* __restrict__ D) {
    unsigned int i = blockIdx.x * PT_BLOCK + threadIdx.x;
    if (i == 0) D->counters[PT_C_SAMPLE] += 1u;
    if (i >= D->u.width * D->u.height) return;
    float4 fr = ld4(D->frame, i);
    unsigned int cnt = __float_as_uint(fr.w);
    st4(D->frame, i, 0.0f, 0.0f, 0.0f, 0.0f);
    if (cnt == 0u) return;
    double inv = 1.0 / (double)cnt;
    double s0 = (double)fr.x * inv, s1 = (double)fr.y * inv, s2 = (double)fr.z * inv;
    double* acc = D->accum + 4ull * i;
    double prevn = acc[3];
    if (D->u.maxaverages > 0u) prevn = fmin(prevn, (double)D->u.maxaverages);
    double c = prevn + 1.0;
    double m0 = acc[0], m1 = acc[1], m2 = acc[2];
    double y = 0.2126 * s0 + 0.7152 * s1 + 0.0722 * s2;
    double yPrev = 0.2126 * m0 + 0.7152 * m1 + 0.0722 * m2;
    acc[0] = (m0 * prevn + s0) / c;
    acc[1] = (m1 * prevn + s1) / c;
    acc[2] = (m2 * prevn + s2) / c;
    acc[3] = c;
    double w = 1.0 / c;
    double dlt = prevn > 0.0 ? y - yPrev : 0.0;
    double* mo = D->moments + 2ull * i;
    double V = mo[0], S = mo[1];
    mo[0] = (1.0 - w) * (V + w * dlt * dlt);
    mo[1] = (1.0 - w) * (1.0 - w) * S + w * w;
}

// ---------------------------------------------------------------------------------------------------
// FLUSH (core/wavefront.js:948-975, the last pass's flush): the camera paths still going when the sample's rounds ran
// out -- paths that met more hits from behind than the straggler rounds allow -- deposit the light they carry as they
// stand, with their count, so that every pixel closes every sample; a recorded training path is closed as at its end.
// qsel: the march queue the last round's shade filled.
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_flush(const PtDev* __restrict__ D, unsigned int qsel) {
    const unsigned int count = D->counters[qsel ? PT_C_MARCH1 : PT_C_MARCH0];
    const unsigned int* __restrict__ q = qsel ? D->qMarch1 : D->qMarch0;
    unsigned long long n = 0ull;
    for (unsigned int i = blockIdx.x * PT_BLOCK + threadIdx.x; i < count; i += gridDim.x * PT_BLOCK) {
        const unsigned int slot = q[i];
        const float4 R = ld4(D->rad, slot);
        const unsigned int pix = ld4u(D->meta, slot).x;
        const V3 rad = v3(R.x, R.y, R.z);
        if (badDeposit(rad)) atomicAdd(&D->stats[PT_S_NONFINITE], 1ull);
        else depositRGB(D, pix, rad, false);
        atomicAdd(reinterpret_cast<unsigned int*>(D->frame) + 4ull * pix + 3ull, 1u);
        unsigned int gr;
        if (guideRecorded(D, slot, gr)) reinterpret_cast<float4*>(D->gPath)[2 * gr] = make_float4(rad.x, rad.y, rad.z, 1.0f);
        n += 1ull;
    }
    statAdd(D, PT_S_FLUSHED, n);
}

// ---------------------------------------------------------------------------------------------------
// THE ALBEDO TABLES (core/microfacet.js:245-330 prog_mf_lut): the directional albedo Ec(mu; r, n) of the GGX
// dielectric reflection lobe, and E1(mu; r) with F = 1, by stratified visible-normal quadrature (32 x 32); then
// the cosine-weighted averages Ec_avg = 2 int Ec mu dmu.
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_mf_lut(float* T) {
    unsigned int t = blockIdx.x * PT_BLOCK + threadIdx.x;
    const unsigned int nEc = PT_MF_NIOR * PT_MF_NR * PT_MF_NMU;
    const unsigned int nE1 = PT_MF_NR * PT_MF_NMU;
    if (t >= nEc + nE1) return;
    bool e1 = t >= nEc;
    unsigned int k = e1 ? t - nEc : t;
    unsigned int im = k % PT_MF_NMU, ir = (k / PT_MF_NMU) % PT_MF_NR, ii = e1 ? 0u : k / (PT_MF_NMU * PT_MF_NR);
    float mu = ((float)im + 0.5f) / PT_MF_NMU;
    float r = (float)ir / (PT_MF_NR - 1);
    float nrel = 1.0f + 2.0f * (float)ii / (PT_MF_NIOR - 1);
    float a = fminf(fmaxf(r * r, PT_ALPHA_MIN), 1.0f);
    V3 o = v3(sqrtf(fmaxf(0.0f, 1.0f - mu * mu)), 0.0f, mu);
    // The albedo is the mean over the visible-normal unit square of F G2/G1 where the reflection stays above the
    // horizon. That integrand is smooth but for the cut where reflections go below the horizon, which a plain grid
    // resolves badly at low roughness (too high by up to 0.5 % at roughness 0.3). So: the midpoint rule on a 128 x 128
    // grid, and any cell whose corners do not all agree with its centre about the horizon taken again as 8 x 8
    // midpoints (within 2e-4 of a converged integration over roughness 0.1-0.8).
    const float ig1 = 1.0f / mfG1(o, a);
    auto at = [&](float u1, float u2, bool& above) {
        const V3 h = mfSampleVNDF(o, a, u1, u2);
        const V3 i = reflectLocal(o, h);
        above = i.z > 0.0f;
        if (!above) return 0.0f;
        const float F = e1 ? 1.0f : fresnelDielectric(dot(o, h), 1.0f / nrel);
        return F * mfG2(o, i, a) * ig1;
    };
    const int N = 128, S = 8;
    const float cell = 1.0f / N;
    float acc = 0.0f;
    for (int ca = 0; ca < N; ++ca) {
        float row = 0.0f;
        for (int cb = 0; cb < N; ++cb) {
            bool above, c;
            const float mid = at(((float)ca + 0.5f) * cell, ((float)cb + 0.5f) * cell, above);
            bool cut = false;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                at((float)(ca + (k & 1)) * cell, (float)(cb + (k >> 1)) * cell, c);
                cut |= c != above;
            }
            if (!cut) { row += mid; continue; }
            float sub = 0.0f;
            for (int sa = 0; sa < S; ++sa)
                for (int sb = 0; sb < S; ++sb)
                    sub += at(((float)ca + ((float)sa + 0.5f) / S) * cell, ((float)cb + ((float)sb + 0.5f) / S) * cell, c);
            row += sub * (1.0f / (S * S));
        }
        acc += row;
    }
    acc *= 1.0f / (float)(N * N);
    if (e1) T[PT_MF_E1 + ir * PT_MF_NMU + im] = acc;
    else T[PT_MF_EC + (ii * PT_MF_NR + ir) * PT_MF_NMU + im] = acc;
}
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_mf_lut_avg(float* T) {
    unsigned int t = blockIdx.x * PT_BLOCK + threadIdx.x;
    if (t >= PT_MF_NIOR * PT_MF_NR) return;
    float acc = 0.0f;
    for (int im = 0; im < PT_MF_NMU; ++im) {
        float mu = ((float)im + 0.5f) / PT_MF_NMU;
        acc += T[PT_MF_EC + t * PT_MF_NMU + im] * mu;
    }
    T[PT_MF_ECAVG + t] = 2.0f * acc / PT_MF_NMU;
}

// ---------------------------------------------------------------------------------------------------
// PICK (core/pick.js, R.pick): after a sample's first march, each camera ray's event, where it stopped, and the
// primitive and material of the surface it hit -- the decisions the oracle checks ray by ray.
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_pick(const PtDev* __restrict__ D, float* __restrict__ pos4, unsigned int* __restrict__ info) {
    unsigned int i = blockIdx.x * PT_BLOCK + threadIdx.x;
    if (i >= D->nrays) return;
    float4 P = ld4(D->pos, i);
    unsigned int state = mzState(D->meta[4u * i + 2u]);
    unsigned int prim = 0xFFFFFFFFu, mat = 0xFFFFFFu;
    if (state == PT_ACTIVE) {
        SdfHit h = sceneFull((const PT_CONST float*)D->params, v3(P.x, P.y, P.z));
        prim = h.prim; mat = h.m;
    }
    st4(pos4, i, P.x, P.y, P.z, P.w);
    info[2u * i] = state | (mat << 8);
    info[2u * i + 1u] = prim;
}

// ---------------------------------------------------------------------------------------------------
// THE PHOTON PASS (core/wavefront.js:542-626 genphotons, core/photons.js:31-76; sppm.js plan): the map cleared, then
// nphotons photons from the lights by their flux -- the sun's beam (flux E pi R^2, white; from the aperture's disc,
// directions over the sun's cone) and each light face (Lambertian: pi L A a side) -- each carrying the total flux
// over nphotons, then marched and shaded as any path (shade stores the caustic ones).
// r_k for the k-th sample since the clear (k = samples resolved + 1), in closed form:
// prod_{i=1}^{k-1} (i + a) / (i + 1) = Gamma(k + a) / (Gamma(1 + a) Gamma(k + 1)).
__device__ __forceinline__ float sppmRadiusAt(float r1, float alpha, unsigned int k) {
    double a = (double)alpha, kk = (double)k;
    double lg = lgamma(kk + a) - lgamma(1.0 + a) - lgamma(kk + 1.0);
    return (float)((double)r1 * sqrt(exp(lg)));
}
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_sppm_clear(const PtDev* __restrict__ D) {
    unsigned int i = blockIdx.x * PT_BLOCK + threadIdx.x;
    if (i == 0) {
        D->counters[PT_C_PHOTON] = 0u;
        float r = sppmRadiusAt(D->u.sppmRadius, D->u.sppmAlpha, D->counters[PT_C_SAMPLE] + 1u);
        D->counters[PT_C_RADIUS] = __float_as_uint(r);
        D->stats[PT_S_RADIUS] = (unsigned long long)__float_as_uint(r);   // reported, not accumulated
    }
    if (i < 3u * PT_SPPM_HEADS) D->sppmHeads[i] = 0u;
}
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_photon_gen(const PtDev* __restrict__ D) {
    unsigned int i = blockIdx.x * PT_BLOCK + threadIdx.x;
    unsigned int* C = D->counters;
    const PtEnv& env = D->u.env;
    float total = sunFlux(env);
    for (unsigned int k = 0; k < D->nfaces; ++k) total += faceFlux(D->faces[k]);
    // no light to emit photons from: an empty pass
    if (i == 0) { C[PT_C_MARCH0] = total > 0.0f ? D->nphotons : 0u; C[PT_C_MARCH1] = 0u; C[PT_C_SHADOW] = 0u; }
    if (i < 2u * PT_PARTS) C[PT_C_HEADS_MARCH + i] = 0u;
    if (i >= D->nphotons || !(total > 0.0f)) return;
    unsigned int seed[4];
    sampleSeed(D, seed);
    Rng r = rngAt(seed, i, 0u, 0x94070000u);
    V4 u0 = rngNext(r), u1 = rngNext(r);
    float pick = u0.x * total;
    float each = total / (float)D->nphotons;
    V3 o, d, power;
    float fsun = sunFlux(env);
    const bool sunPh = pick < fsun;
    if (sunPh) {
        V3 a = ld3(env.sunDir), t, b;
        frameOf(a, t, b);
        float rr = env.sunRadius * sqrtf(u0.y), sp, cp;
        __sincosf(2.0f * PT_PI * u0.z, &sp, &cp);
        o = ld3(env.sunPoint) + t * (rr * cp) + b * (rr * sp);
        d = -sampleSunCone(env, u1.x, u1.y);
        // directions uniform over the cone, positions over the disc: a photon's weight is L cos(theta) Omega A,
        // L = E / Omega, so E A cos(theta) -- each carries the flux share times cos(theta)
        power = v3s(each * fmaxf(-dot(d, a), 0.0f));
    } else {
        pick -= fsun;
        unsigned int k = 0u;
        for (; k + 1u < D->nfaces; ++k) { float ff = faceFlux(D->faces[k]); if (pick < ff) break; pick -= ff; }
        const PtLightFace& f = D->faces[k];
        V3 y = ld3(f.centre);
        if (f.shape == 1u) {
            float rr = f.halfx * sqrtf(u0.y), sp, cp; __sincosf(2.0f * PT_PI * u0.z, &sp, &cp);
            y += (ld3(f.u) * cp + ld3(f.v) * sp) * rr;
        } else {
            y += ld3(f.u) * ((2.0f * u0.y - 1.0f) * f.halfx) + ld3(f.v) * ((2.0f * u0.z - 1.0f) * f.halfy);
        }
        V3 side = (f.twoSided && u1.z < 0.5f) ? -ld3(f.n) : ld3(f.n);
        d = sampleCosine(side, u1.x, u1.y);
        o = y + side * (2.0f * offsetTol(D, y));
        V3 le = ld3(f.le);
        power = le * (each / fmaxf(luma(le), 1e-30f));
    }
    st4(D->pos, i, o.x, o.y, o.z, 0.0f);
    st4(D->dir, i, d.x, d.y, d.z, 1.0f);
    st4(D->beta, i, power.x, power.y, power.z, -1.0f);
    st4(D->rad, i, 0.0f, 0.0f, 0.0f, 0.0f);
    st4u(D->meta, i, make_uint4(0u, 0u, mzPack(0u, PT_ACTIVE, PT_F_PHOTON | (sunPh ? PT_F_SUNPH : 0u)), 0u));
    float collide = hazeOn(D) ? hazeDraw(D->u.haze, o, d, u1.w) : 3.0e38f;
    st4(D->aux, i, 1.0f, collide, 0.0f, 0.0f);
    D->qMarch0[i] = i;
}
