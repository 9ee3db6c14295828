// THE OPTICS TRACE over the loaded scene (optics/kernels.js:1-26, prog_opt_gen :504-531, prog_opt_shade :533-723;
// optics/index.js createOptics :92-825): rays leave the sources, walk the scene through this engine's own binning, march
// and surface kernels and its layered BSDF, and where their flux goes is tallied -- per material (incident, leaving,
// hits), escaped, lost, detected, absorbed in the haze -- into a far-field grid, and as detector hits.
//
// A batch: pt_opt_gen fills slots [0, rays) and the march queue; then rounds of bin -> march -> surface -> pt_opt_shade
// -> occlude (on an empty shadow queue, for its reset of the queue just read); pt_opt_flush counts what is still going
// after the last round as LOST.
//
// Tallies are fixed point in 64-bit integers (PT_OPT_FX per unit of ray weight): exact, and the same sum in any order,
// so a trace is deterministic. A block gathers its rays' tallies in LDS (64-bit LDS atomics) and adds them to the
// device's once. A ray's weight is the luminance of its throughput, which starts as the source's hue (luminance 1).
#define PT_F_OPTIC 0x200u     // the slot carries an optics-trace ray

__device__ __forceinline__ unsigned long long optFix(float w) {
    return (unsigned long long)((double)fminf(fmaxf(w, 0.0f), 4096.0f) * PT_OPT_FX + 0.5);
}

// the far-field bin of a direction: gamma from the axis, C about it counterclockwise from C0 seen from where the axis
// points (C90 = axis x C0); the gamma bins are centred on 0, 180 / (nG - 1), ... and the C bins on 0, 360 / nC, ...
__device__ __forceinline__ void optFarAdd(const PtOptDev* __restrict__ O, V3 d, float w) {
    const unsigned int nC = O->t.nC, nG = O->t.nG;
    if (nC == 0u || nG < 2u || !(w > 0.0f)) return;
    const V3 ax = ld3(O->t.farAxis), c0 = ld3(O->t.farC0), c90 = cross(ax, c0);
    const V3 dn = normalize(d);
    const float g = acosf(fminf(fmaxf(dot(dn, ax), -1.0f), 1.0f));
    float c = atan2f(dot(dn, c90), dot(dn, c0));
    if (c < 0.0f) c += 2.0f * PT_PI;
    if (!(c >= 0.0f && c <= 2.0f * PT_PI)) c = 0.0f;      // a direction that is not one (NaN): bin 0, in range
    const unsigned int ig = min((unsigned int)(g * (float)(nG - 1u) / PT_PI + 0.5f), nG - 1u);
    const unsigned int ic = (unsigned int)(c * (float)nC / (2.0f * PT_PI) + 0.5f) % nC;
    atomicAdd(&O->far[ig * nC + ic], optFix(w));
}

__device__ __forceinline__ bool optIsDetector(const PtOptDev* __restrict__ O, unsigned int mat) {
    return mat < 64u && ((O->t.detectors[mat >> 5] >> (mat & 31u)) & 1u);
}

// SOURCES (optics/sources.js; kernels.js SOURCE_RAY :270-420): which source by flux, a point on its shape, a direction
// about its axis -- collimated; uniform in solid angle within the cone (cos = 1 - u (1 - cos a)); cos^n within the cone
// (cos^(n+1) uniform on [cos^(n+1) a, 1]) -- and, for a sheet of the scene, a start two march tolerances off it.
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_opt_gen(const PtDev* __restrict__ D, const PtOptDev* __restrict__ O) {
    const unsigned int i = blockIdx.x * PT_BLOCK + threadIdx.x;
    unsigned int* C = D->counters;
    if (i == 0) { C[PT_C_MARCH0] = O->rays; C[PT_C_MARCH1] = 0u; C[PT_C_SHADOW] = 0u; }
    if (i < 2u * PT_PARTS) C[PT_C_HEADS_MARCH + i] = 0u;
    if (i >= O->rays) return;
    const unsigned int id = O->firstId + i;
    Rng r = rngAt(O->seed, id, 0u, 0x0971C500u);
    const V4 u0 = rngNext(r), u1 = rngNext(r);
    unsigned int k = 0u;
    while (k + 1u < O->nsrc && u0.x > O->cdf[k]) ++k;
    const PtOpticsSource& S = O->src[k];
    const V3 axis = ld3(S.axis), tu = ld3(S.u), tv = cross(axis, tu);
    V3 p = ld3(S.pos);
    if (S.shape == 1u) {
        const float rr = S.size[0] * sqrtf(u0.y);
        float sp, cp;
        __sincosf(2.0f * PT_PI * u0.z, &sp, &cp);
        p += (tu * cp + tv * sp) * rr;
    } else if (S.shape == 2u) {
        p += tu * ((2.0f * u0.y - 1.0f) * S.size[0]) + tv * ((2.0f * u0.z - 1.0f) * S.size[1]);
    }
    const V3 side = (S.twoSided && u0.w < 0.5f) ? -axis : axis;
    const float cosA = cosf(fminf(fmaxf(S.halfAngle, 0.0f), 90.0f) * (PT_PI / 180.0f));
    float c = 1.0f;
    if (S.distribution == 1u) {
        c = 1.0f - u1.x * (1.0f - cosA);
    } else if (S.distribution == 2u) {
        const float e = fmaxf(S.n, 0.0f) + 1.0f;
        const float lo = powf(fmaxf(cosA, 0.0f), e);
        c = powf(1.0f - u1.x * (1.0f - lo), 1.0f / e);
    }
    const float s = sqrtf(fmaxf(0.0f, 1.0f - c * c));
    float sp, cp;
    __sincosf(2.0f * PT_PI * u1.y, &sp, &cp);
    const V3 d = normalize(tu * (s * cp) + tv * (s * sp) + side * c);
    const V3 o = p + side * (S.onSurface ? 2.0f * marchTol(tolSetting(D), p) : 0.0f);
    const float n = fmaxf(S.emitN, 1.0f);
    st4(D->pos, i, o.x, o.y, o.z, 0.0f);
    st4(D->dir, i, d.x, d.y, d.z, n);
    st4(D->beta, i, S.colour[0], S.colour[1], S.colour[2], -1.0f);
    st4(D->rad, i, 0.0f, 0.0f, 0.0f, 0.0f);
    st4u(D->meta, i, make_uint4(id, 0u, mzPack(0u, PT_ACTIVE, PT_F_PHOTON | PT_F_OPTIC), 0u));
    const float collide = (hazeOn(D) && n <= 1.0f) ? hazeDraw(D->u.haze, o, d, u1.z) : 3.0e38f;
    st4(D->aux, i, 1.0f, collide, 0.0f, 0.0f);
    D->qMarch0[i] = i;
}

// One ray at its event: true when it goes on (requeued for the next round).
__device__ __forceinline__ bool opticsOne(const PtDev* __restrict__ D, const PtOptDev* __restrict__ O,
                                          unsigned long long* __restrict__ sT, unsigned int slot) {
    const float4 P = ld4(D->pos, slot), Dd = ld4(D->dir, slot), B = ld4(D->beta, slot), A = ld4(D->aux, slot);
    const uint4 M = ld4u(D->meta, slot);
    V3 p = v3(P.x, P.y, P.z), dir = v3(Dd.x, Dd.y, Dd.z), beta = v3(B.x, B.y, B.z);
    float ltot = P.w, n = Dd.w, collide = A.y, lambda = A.z;
    const unsigned int id = M.x, counts = M.y;
    unsigned int bounces = mzBounces(M.z), flags = mzFlags(M.z);
    const unsigned int state = mzState(M.z);
    float w = luma(beta);
    // the absorbing medium the ray travels in (material + 1, 0 none), and the optical length at its last vertex
    const unsigned int medIn = __float_as_uint(A.w);
    unsigned int medOut = medIn;
    float vtx = B.w;
    Rng rng = rngAt(O->seed, id, bounces, 0x0971C5ADu ^ ((counts & 0xFFFFu) << 12));
    if (state == PT_ESCAPED) {
        atomicAdd(&sT[PT_OT_ESCAPED], optFix(w));
        atomicAdd(&sT[PT_OT_CESC], 1ull);
        if (O->t.farMode == 1u) optFarAdd(O, dir, w);
        return false;
    }
    if (state == PT_MAXITER) {
        // march steps or optical length ran out: still going, counted LOST (kernels.js ONWARD)
        atomicAdd(&sT[PT_OT_LOST], optFix(w));
        atomicAdd(&sT[PT_OT_CLOST], 1ull);
        return false;
    }
    V4 u = rngNext(rng);
    if (state == PT_SCATTER) {
        // a collision in the haze: the albedo's complement is absorbed by the medium (VOLUME), the rest scatters by
        // Henyey-Greenstein
        const PtHaze& hz = D->u.haze;
        atomicAdd(&sT[PT_OT_VOLUME], optFix(w * (1.0f - hz.albedo)));
        beta *= hz.albedo;
        bounces += 1u;
        if (bounces > O->bounces) {
            atomicAdd(&sT[PT_OT_LOST], optFix(luma(beta)));
            atomicAdd(&sT[PT_OT_CLOST], 1ull);
            return false;
        }
        dir = hgSample(hz.g, dir, u.x, u.y);
        collide = ltot + hazeDraw(hz, p, dir, u.z) * n;
    } else {
        const float4 HR = ld4(D->hitRec, slot);
        const unsigned int hw = __float_as_uint(HR.w);
        const bool inside = (hw & PT_HR_INSIDE) != 0u;
        const V3 ns = v3(HR.x, HR.y, HR.z);
        const V3 nrm = inside ? -ns : ns;
        if (dot(dir, nrm) > 0.0f) {
            // met from behind (scene_api.js:127-136 pathFacing): moved on, nothing tallied
            const float off = offsetTol(D, p);
            p = p + dir * off; ltot += off * n;
            st4(D->pos, slot, p.x, p.y, p.z, ltot);
            stu(D->meta, 4u * slot + 2u, mzPack(bounces, PT_ACTIVE, flags));
            return true;
        }
        if (medIn != 0u) {
            // ABSORPTION ALONG THE WAY (tracebasics.js absorbAlong), since the last vertex: what the medium took is its
            // material's (counted in as incident, never as leaving)
            beta *= mediumTr(D->mats[medIn - 1u], fmaxf(ltot - vtx, 0.0f) / fmaxf(n, 1e-4f));
            const float wa = luma(beta);
            atomicAdd(&sT[PT_OT_MAT + 3u * (medIn - 1u)], optFix(w) - optFix(wa));
            w = wa;
        }
        if (O->t.maxDist > 0.0f && ltot > O->t.maxDist) {
            // a hit past the longest path traced: stopped at the limit
            atomicAdd(&sT[PT_OT_LOST], optFix(w));
            atomicAdd(&sT[PT_OT_CLOST], 1ull);
            return false;
        }
        const unsigned int mat = hw & 0xFFu;
#if PT_SCENE_MATCODE
        // the material's own code, at this hit
        PtMaterial m = D->mats[mat];
        sceneMaterial((const PT_CONST float*)D->params, mat, p, ns, m);
#else
        const PtMaterial& m = D->mats[mat];
#endif
        const unsigned long long fw = optFix(w);
        atomicAdd(&sT[PT_OT_MAT + 3u * mat], fw);
        atomicAdd(&sT[PT_OT_MAT + 3u * mat + 2u], 1ull);
        atomicAdd(&sT[PT_OT_CINTER], 1ull);
        if (!inside && optIsDetector(O, mat)) {
            // a detector: the ray stops, all of it caught, and the hit recorded while there is room
            atomicAdd(&sT[PT_OT_DETECTED], fw);
            atomicAdd(&sT[PT_OT_CDET], 1ull);
            const unsigned long long k = atomicAdd(&O->tally[PT_OT_HITS], 1ull);
            if (k < O->t.detCap) {
                float4* h = reinterpret_cast<float4*>(O->hits + k * PT_OPT_HIT_WORDS);
                h[0] = make_float4(p.x, p.y, p.z, w);
                h[1] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(mat));
            }
            if (O->t.farMode == 2u && mat == O->t.farMaterial) optFarAdd(O, dir, w);
            return false;
        }
        if (max3(ld3(m.emissive)) > 0.0f) {
            // an emitter ends the ray: a source, not a reflector -- everything absorbed
            atomicAdd(&sT[PT_OT_CABS], 1ull);
            return false;
        }
        bounces += 1u;
        // the medium beyond: the material's index going in, air (or the enclosing material) going out; a dispersive
        // glass draws the ray's wavelength at its first refraction (as the camera's paths do)
        const float n0 = n;
        float n1 = fmaxf(m.ior, 1.0f);
        if (inside) {
            n1 = 1.0f;
            if (hw & PT_HR_BEYOND) n1 = fmaxf(D->mats[hw >> 16].ior, 1.0f);
        }
        bool draw = false;
        V3 specW = v3s(1.0f);
        const float lambda0 = lambda;
        if (m.abbe > 0.0f && m.transmission > 0.0f && !inside) {
            if (lambda <= 0.0f) {
                draw = true;
                lambda = 0.38f + 0.40f * u.w;
                specW = spectralWeight(lambda);
            }
            n1 = cauchyN(n1, m.abbe, lambda);
        }
        const Lobes lb = makeLobes(nrm, -dir, m, n0 / n1, D->mf);
        const LobeSample s = lobesSample(lb, D->mf, rngNext(rng));
        beta *= s.w;
        if (s.transmitted) {
            n = n1;
            if (draw) beta *= specW;
            // through the surface: into its material, or out to the one beyond (absorbing ones only)
            const unsigned int into = inside ? ((hw & PT_HR_BEYOND) ? (hw >> 16) : 0xFFFFFFFFu) : mat;
            medOut = (into != 0xFFFFFFFFu && D->mats[into].attenuation > 0.0f) ? into + 1u : 0u;
        } else if (draw) {
            lambda = lambda0;
        }
        const float w2 = luma(beta);
        if (!(w2 > 0.0f)) {
            atomicAdd(&sT[PT_OT_CABS], 1ull);
            return false;
        }
        atomicAdd(&sT[PT_OT_MAT + 3u * mat + 1u], optFix(w2));
        if (bounces > O->bounces) {
            // the bounce cap: counted as leaving and as LOST, so the balance holds exactly
            atomicAdd(&sT[PT_OT_LOST], optFix(w2));
            atomicAdd(&sT[PT_OT_CLOST], 1ull);
            return false;
        }
        dir = s.dir;
        vtx = ltot;
        const float off = offsetTol(D, p);
        p = p + dir * off; ltot += off * n;
        collide = (hazeOn(D) && n <= 1.0f) ? ltot + hazeDraw(D->u.haze, p, dir, u.y) * n : 3.0e38f;
    }
    st4(D->pos, slot, p.x, p.y, p.z, ltot);
    st4(D->dir, slot, dir.x, dir.y, dir.z, n);
    st4(D->beta, slot, beta.x, beta.y, beta.z, vtx);
    st4u(D->meta, slot, make_uint4(id, counts, mzPack(bounces, PT_ACTIVE, flags), 0u));
    st4(D->aux, slot, 1.0f, collide, lambda, __uint_as_float(medOut));
    return true;
}

__device__ __forceinline__ void optTallyFlush(const PtOptDev* __restrict__ O, const unsigned long long* sT) {
    __syncthreads();
    for (unsigned int k = threadIdx.x; k < PT_OT_WORDS; k += PT_BLOCK)
        if (sT[k]) atomicAdd(&O->tally[k], sT[k]);
}

extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_opt_shade(const PtDev* __restrict__ D, const PtOptDev* __restrict__ O,
                                                                    unsigned int qsel) {
    __shared__ unsigned long long sT[PT_OT_WORDS];
    for (unsigned int k = threadIdx.x; k < PT_OT_WORDS; k += PT_BLOCK) sT[k] = 0ull;
    __syncthreads();
    unsigned int* C = D->counters;
    // this round's rays: the queue march read (binned), every one at its event
    const unsigned int count = C[qsel ? PT_C_MARCH0 : PT_C_MARCH1];
    unsigned int* qout = qsel ? D->qMarch1 : D->qMarch0;
    unsigned int* cout = &C[qsel ? PT_C_MARCH1 : PT_C_MARCH0];
    unsigned long long segs = 0ull;
    for (unsigned int base = blockIdx.x * PT_BLOCK; base < count; base += gridDim.x * PT_BLOCK) {
        const unsigned int i = base + threadIdx.x;
        bool requeue = false;
        unsigned int slot = 0u;
        if (i < count) {
            slot = D->qSorted[i];
            requeue = opticsOne(D, O, sT, slot);
            segs += 1ull;
        }
        const unsigned int qi = waveAppend(cout, requeue);
        if (requeue) qout[qi] = slot;
    }
    statAdd(D, PT_S_PHOTONSEGS, segs);
    optTallyFlush(O, sT);
}

// After the last round: the rays still going are LOST (kernels.js ONWARD: a run's last pass).
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_opt_flush(const PtDev* __restrict__ D, const PtOptDev* __restrict__ O,
                                                                    unsigned int qsel) {
    __shared__ unsigned long long sT[PT_OT_WORDS];
    for (unsigned int k = threadIdx.x; k < PT_OT_WORDS; k += PT_BLOCK) sT[k] = 0ull;
    __syncthreads();
    const unsigned int count = D->counters[qsel ? PT_C_MARCH1 : PT_C_MARCH0];
    const unsigned int* __restrict__ q = qsel ? D->qMarch1 : D->qMarch0;
    for (unsigned int i = blockIdx.x * PT_BLOCK + threadIdx.x; i < count; i += gridDim.x * PT_BLOCK) {
        const float4 B = ld4(D->beta, q[i]);
        atomicAdd(&sT[PT_OT_LOST], optFix(luma(v3(B.x, B.y, B.z))));
        atomicAdd(&sT[PT_OT_CLOST], 1ull);
    }
    optTallyFlush(O, sT);
}
