// PATH GUIDING: a small MLP proposes, at each evaluable path vertex, an 8-lobe von Mises-Fisher mixture over incident
// directions; shade samples it beside the BSDF under one-sample MIS (the mixture's density and the BSDF's, weighted by
// the guided fraction), so the estimator stays unbiased whatever the network says. The network is trained online, on
// the GPU, on the tracer's own radiance samples: a KL loss to the product of incident radiance and the BSDF, Adam on
// FP32 master weights, BF16 weights and activations on MFMA (v_mfma_f32_32x32x16_bf16) with FP32 accumulation.
//
// No reference code: the reference removed its guiding (lib/pathtracer/README.md:853-861). Everything here follows
// the mathematics stated with each function; the FP64 oracle (oracle/guide.py) states the same.
//
//   features (80): position in the scene box at 12 octaves -- sin, cos of 2^k pi x_i, k major, axis minor --, the view
//                  direction, the normal, the roughness, 1
//   network:       80 -> 64 -> 64 -> 64 -> 41, ReLU between; parameters flat W1[80][64], b1, W2[64][64], b2, W3, b3,
//                  W4[64][41], b4 (W[in][out])
//   head:          weights softmax(o[0:8]), means normalize(o[8 + 3j : 11 + 3j]), concentrations
//                  exp(clamp(o[32 + j], -3, 7)); vMF(w) = k / (2 pi (1 - e^{-2k})) exp(k (mu . w - 1)); the guided
//                  fraction c = sigmoid(o[40]) (capped by the uniforms' guideFrac)
//   losses:        the mixture learns the normalised product f cos L by its KL divergence -sum F log q(w) / sum F;
//                  the fraction learns how far to trust it by the combined density's -sum F log(c q + (1 - c) b) /
//                  sum F, b the BSDF's density at w, with q held fixed there (through that loss alone, a mixture the
//                  fraction had turned off would never learn again); the second is the loss reported
#pragma once

#define PT_G_IN 80
#define PT_G_H 64
#define PT_G_OUT 41
#define PT_G_LOBES 8
#define PT_G_NP 16169
#define PT_G_OFF_W1 0
#define PT_G_OFF_B1 5120
#define PT_G_OFF_W2 5184
#define PT_G_OFF_B2 9280
#define PT_G_OFF_W3 9344
#define PT_G_OFF_B3 13440
#define PT_G_OFF_W4 13504
#define PT_G_OFF_B4 16128
// the BF16 copy the MFMA reads, each layer's weights transposed (WT[out][in]): W1T 64 x 80, W2T, W3T 64 x 64, W4T
// 64 x 64 with rows 41..63 zero
#define PT_G_WT1 0
#define PT_G_WT2 5120
#define PT_G_WT3 9216
#define PT_G_WT4 13312
#define PT_G_WT_WORDS 17408
#define PT_G_BIAS 256          // b1, b2, b3 at 0, 64, 128; b4 at 192 (64, the last 23 zero)
#define PT_G_KLO -3.0f
#define PT_G_KHI 7.0f
#define PT_G_MIX_F4 11         // a slot's mixture: (kappa mu, folded exponent) x 8, the 8 weights, the guided fraction
#define PT_G_REC_F4 7          // a recorded vertex (below)
#define PT_G_SAMPLE_F4 4       // a training sample (below)
#define PT_G_TILE_LD 88        // the activation tile's row stride (bf16): 80 features and 8 of padding
#define PT_G_PARTS 256         // blocks of the gradient pass: each writes its own partial gradient row

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ __bf16 toBf16(float x) { return (__bf16)x; }
__device__ __forceinline__ float fromBf16(__bf16 x) { return (float)x; }

// ---------------------------------------------------------------------------------------------------
// The wave's GEMM on MFMA: C[32 MB][32 NB] = A[.][16 KS] B[16 KS][.], all blocks into registers before any store (so
// a layer may write its output over its input). Lane l supplies A[l % 32][8 (l / 32) + 0..7] and B[8 (l / 32) + 0..7]
// [l % 32] of each 16-deep step; element (8 (r / 4) + 4 (l / 32) + r % 4, l % 32) of a block comes back in its
// accumulator r.
template <int MB, int NB, int KS, class FA, class FB, class ST>
__device__ __forceinline__ void waveGemm(FA fa, FB fb, ST st) {
    const int l = (int)(threadIdx.x & 63u), lr = l & 31, lk = (l >> 5) * 8;
    f32x16 acc[MB][NB];
#pragma unroll
    for (int bm = 0; bm < MB; ++bm)
#pragma unroll
        for (int bn = 0; bn < NB; ++bn) {
            f32x16 a = {};
#pragma unroll
            for (int ks = 0; ks < KS; ++ks)
                a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa(bm * 32 + lr, ks * 16 + lk), fb(ks * 16 + lk, bn * 32 + lr), a, 0, 0, 0);
            acc[bm][bn] = a;
        }
#pragma unroll
    for (int bm = 0; bm < MB; ++bm)
#pragma unroll
        for (int bn = 0; bn < NB; ++bn)
#pragma unroll
            for (int r = 0; r < 16; ++r) st(bm * 32 + (r >> 2) * 8 + (l >> 5) * 4 + (r & 3), bn * 32 + lr, acc[bm][bn][r]);
}
// ...and the same block by block, each stored as soon as it is done (fewer registers; for outputs that are no input)
template <int MB, int NB, int KS, class FA, class FB, class ST>
__device__ __forceinline__ void waveGemmStream(FA fa, FB fb, ST st) {
    const int l = (int)(threadIdx.x & 63u), lr = l & 31, lk = (l >> 5) * 8;
    for (int bm = 0; bm < MB; ++bm)
        for (int bn = 0; bn < NB; ++bn) {
            f32x16 a = {};
#pragma unroll
            for (int ks = 0; ks < KS; ++ks)
                a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa(bm * 32 + lr, ks * 16 + lk), fb(ks * 16 + lk, bn * 32 + lr), a, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r) st(bm * 32 + (r >> 2) * 8 + (l >> 5) * 4 + (r & 3), bn * 32 + lr, a[r]);
        }
}
__device__ __forceinline__ bf16x8 ld8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }
// eight values down a column (stride apart): the transposed operands of the backward pass
__device__ __forceinline__ bf16x8 col8(const __bf16* p, int stride) {
    bf16x8 v;
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = p[i * stride];
    return v;
}

// The block's copy of the BF16 weights and the biases, from HBM into LDS.
struct GuideLds {
    __bf16 w[PT_G_WT_WORDS];
    float b[PT_G_BIAS];
};
__device__ __forceinline__ void guideLoadWeights(const PtDev* __restrict__ D, GuideLds& s) {
    const uint4* src = reinterpret_cast<const uint4*>(D->gW);
    uint4* dst = reinterpret_cast<uint4*>(s.w);
    for (unsigned int k = threadIdx.x; k < PT_G_WT_WORDS / 8u; k += blockDim.x) dst[k] = src[k];
    for (unsigned int k = threadIdx.x; k < PT_G_BIAS; k += blockDim.x) s.b[k] = D->gBias[k];
}

// The 80 features of a vertex into row `row` of a tile (bf16, stride PT_G_TILE_LD): position mapped into the box,
// 12 octaves of sin, cos (sincospi of 2^k x is exact range reduction), view direction, normal, roughness, 1.
__device__ __forceinline__ void guideFeatures(const PtDev* __restrict__ D, __bf16* tile, int row, V3 p, V3 wo, V3 n, float rough) {
    const PtGuideConfig& g = D->g;
    float x[3] = {(p.x - g.boxMin[0]) / fmaxf(g.boxMax[0] - g.boxMin[0], 1e-6f),
                  (p.y - g.boxMin[1]) / fmaxf(g.boxMax[1] - g.boxMin[1], 1e-6f),
                  (p.z - g.boxMin[2]) / fmaxf(g.boxMax[2] - g.boxMin[2], 1e-6f)};
    __bf16* r = tile + row * PT_G_TILE_LD;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float xa = fminf(fmaxf(x[a], 0.0f), 1.0f);
        float sc = 1.0f;
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            float s, c;
            sincospif(sc * xa, &s, &c);
            r[6 * k + 2 * a] = toBf16(s);
            r[6 * k + 2 * a + 1] = toBf16(c);
            sc *= 2.0f;
        }
    }
    r[72] = toBf16(wo.x); r[73] = toBf16(wo.y); r[74] = toBf16(wo.z);
    r[75] = toBf16(n.x); r[76] = toBf16(n.y); r[77] = toBf16(n.z);
    r[78] = toBf16(rough); r[79] = toBf16(1.0f);
}
__device__ __forceinline__ void guideZeroRow(__bf16* tile, int row) {
    for (int k = 0; k < PT_G_IN; ++k) tile[row * PT_G_TILE_LD + k] = toBf16(0.0f);
}

// The forward pass of one wave's 64-row tile: features in `tile` (rows x 80), the 40 head outputs written over it as
// FP32 rows of 40 (out). Activations between the layers stay in `tile` as BF16.
__device__ __forceinline__ void guideForward(const GuideLds& s, __bf16* tile, float* out) {
    __bf16* keep = nullptr;
    const __bf16* w = s.w;
    // layer 1: 80 -> 64
    waveGemm<2, 2, 5>([&](int i, int k) { return ld8(tile + i * PT_G_TILE_LD + k); },
                      [&](int k, int j) { return ld8(w + PT_G_WT1 + j * 80 + k); },
                      [&](int i, int j, float v) {
                          __bf16 h = toBf16(fmaxf(v + s.b[j], 0.0f));
                          tile[i * PT_G_TILE_LD + j] = h;
                          if (keep) keep[i * 64 + j] = h;
                      });
    // layers 2 and 3: 64 -> 64
#pragma unroll
    for (int L = 0; L < 2; ++L) {
        const __bf16* wt = w + (L == 0 ? PT_G_WT2 : PT_G_WT3);
        const float* b = s.b + 64 * (L + 1);
        __bf16* kp = keep ? keep + 4096 * (L + 1) : nullptr;
        waveGemm<2, 2, 4>([&](int i, int k) { return ld8(tile + i * PT_G_TILE_LD + k); },
                          [&](int k, int j) { return ld8(wt + j * 64 + k); },
                          [&](int i, int j, float v) {
                              __bf16 h = toBf16(fmaxf(v + b[j], 0.0f));
                              tile[i * PT_G_TILE_LD + j] = h;
                              if (kp) kp[i * 64 + j] = h;
                          });
    }
    // the head: 64 -> 40 (the W4T rows past 40 are zero)
    waveGemm<2, 2, 4>([&](int i, int k) { return ld8(tile + i * PT_G_TILE_LD + k); },
                      [&](int k, int j) { return ld8(w + PT_G_WT4 + j * 64 + k); },
                      [&](int i, int j, float v) { if (j < PT_G_OUT) out[i * PT_G_OUT + j] = v + s.b[192 + j]; });
}

// ---------------------------------------------------------------------------------------------------
// The mixture from 40 head outputs, and its density.
struct GuideMix {
    float4 lobe[PT_G_LOBES];   // mu, kappa
    float w[PT_G_LOBES];
    float c;                   // the guided fraction the network asks for
};
__device__ __forceinline__ float vmfLogNorm(float k) {
    // log(k / (2 pi (1 - e^{-2k})))
    return logf(k) - 1.8378770664f - log1pf(-expf(-2.0f * k));
}
__device__ __forceinline__ GuideMix guideMixOf(const float* o) {
    GuideMix m;
    float mx = o[0];
#pragma unroll
    for (int j = 1; j < PT_G_LOBES; ++j) mx = fmaxf(mx, o[j]);
    float sum = 0.0f;
#pragma unroll
    for (int j = 0; j < PT_G_LOBES; ++j) { m.w[j] = expf(o[j] - mx); sum += m.w[j]; }
#pragma unroll
    for (int j = 0; j < PT_G_LOBES; ++j) {
        m.w[j] /= sum;
        V3 v = v3(o[8 + 3 * j], o[9 + 3 * j], o[10 + 3 * j]);
        float l = fmaxf(length(v), 1e-6f);
        float k = expf(fminf(fmaxf(o[32 + j], PT_G_KLO), PT_G_KHI));
        m.lobe[j] = make_float4(v.x / l, v.y / l, v.z / l, k);
    }
    m.c = 1.0f / (1.0f + expf(-o[40]));
    return m;
}
// A direction from one lobe (mu, kappa): Jakob 2012's inversion w = 1 + log(u1 + (1 - u1) e^{-2k}) / k about its mean.
__device__ __forceinline__ V3 guideLobeSample(float4 L, V4 u) {
    float k = L.w;
    float wz = 1.0f + logf(u.y + (1.0f - u.y) * expf(-2.0f * k)) / k;
    wz = fminf(fmaxf(wz, -1.0f), 1.0f);
    float sr = sqrtf(fmaxf(0.0f, 1.0f - wz * wz)), sp, cp;
    __sincosf(2.0f * PT_PI * u.z, &sp, &cp);
    V3 mu = v3(L.x, L.y, L.z), t, b;
    frameOf(mu, t, b);
    return normalize(t * (sr * cp) + b * (sr * sp) + mu * wz);
}
// The lobe u falls in by the weights' running sum: compares unrolled, no index into a private array (which would put
// the weights in scratch memory)
__device__ __forceinline__ int guideLobePick(const float* w, float u) {
    int j = PT_G_LOBES - 1;
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < PT_G_LOBES - 1; ++k) {
        acc += w[k];
        if (j == PT_G_LOBES - 1 && u < acc) j = k;
    }
    return j;
}

// A vertex's mixture as shade reads it, straight from the slot's record where it is used (shade holds no copy of it).
// The record (PT_G_MIX_F4 float4s, written by the inference) folds each lobe's weight and normalisation into its
// exponent: lobe j as (k mu, ln(w_j C(k_j)) - k_j), C(k) = k / (2 pi (1 - e^{-2k})), so that the mixture's density at d
// is sum_j exp(v_j . d + b_j), one exponential a lobe; then the weights as they are (to pick a lobe); then the fraction.
__device__ __forceinline__ void guideStoreMix(const PtDev* __restrict__ D, unsigned int slot, const GuideMix& m) {
    float4* dst = reinterpret_cast<float4*>(D->gMix) + (size_t)PT_G_MIX_F4 * slot;
#pragma unroll
    for (int j = 0; j < PT_G_LOBES; ++j) {
        const float4 L = m.lobe[j];
        const float b = logf(fmaxf(m.w[j], 1e-38f)) + vmfLogNorm(L.w) - L.w;
        dst[j] = make_float4(L.x * L.w, L.y * L.w, L.z * L.w, b);
    }
    dst[8] = make_float4(m.w[0], m.w[1], m.w[2], m.w[3]);
    dst[9] = make_float4(m.w[4], m.w[5], m.w[6], m.w[7]);
    dst[10] = make_float4(m.c, 0.0f, 0.0f, 0.0f);
}
__device__ __forceinline__ const float4* guideMixAt(const PtDev* __restrict__ D, unsigned int slot) {
    return reinterpret_cast<const float4*>(D->gMix) + (size_t)PT_G_MIX_F4 * slot;
}
__device__ __forceinline__ float guideFracAt(const PtDev* __restrict__ D, unsigned int slot) { return guideMixAt(D, slot)[10].x; }
__device__ __forceinline__ float guidePdfAt(const PtDev* __restrict__ D, unsigned int slot, V3 d) {
    const float4* src = guideMixAt(D, slot);
    float q = 0.0f;
#pragma unroll
    for (int j = 0; j < PT_G_LOBES; ++j) {
        const float4 L = src[j];
        q += __expf(L.x * d.x + L.y * d.y + L.z * d.z + L.w);
    }
    return q;
}
__device__ __forceinline__ V3 guideSampleAt(const PtDev* __restrict__ D, unsigned int slot, V4 u) {
    const float4* src = guideMixAt(D, slot);
    const float4 wa = src[8], wb = src[9];
    const float w[PT_G_LOBES] = {wa.x, wa.y, wa.z, wa.w, wb.x, wb.y, wb.z, wb.w};
    const float4 V = src[guideLobePick(w, u.x)];
    const float k = sqrtf(V.x * V.x + V.y * V.y + V.z * V.z), inv = 1.0f / k;
    return guideLobeSample(make_float4(V.x * inv, V.y * inv, V.z * inv, k), u);
}

// Where guiding applies: a camera path's surface vertex whose material has a lobe to evaluate, reflects only (a
// transmissive lobe is sampled, not evaluated) and emits nothing.
// A vertex guiding can serve: a material with a lobe that has a density (a rough interface, refracting or not, or a
// diffuse base; not a mirror or clear glass, whose directions no proposal can share), and not an emitter.
__device__ __forceinline__ bool guideMaterial(const PtMaterial& m) {
    return !lobesNone(m) && max3(ld3(m.emissive)) <= 0.0f;
}
__device__ __forceinline__ bool guideActive(const PtDev* __restrict__ D) {
    return D->u.guide != 0u && D->u.guideFrac > 0.0f && D->counters[PT_C_GUIDE_T] >= D->g.warmup;
}
// the paths of this sample that record their vertices for training: every stride-th slot, from an offset that moves
// with the sample
__device__ __forceinline__ bool guideRecorded(const PtDev* __restrict__ D, unsigned int slot, unsigned int& r) {
    if (D->u.guide == 0u || D->gStride == 0u) return false;
    const unsigned int off = D->counters[PT_C_SAMPLE] % D->gStride;
    if (slot % D->gStride != off) return false;
    r = slot / D->gStride;
    return r < D->gPaths;
}

// ---------------------------------------------------------------------------------------------------
// INFERENCE: the round's surface vertices where guiding applies, 64 to a wave, features -> network -> mixture, kept by
// slot for shade. Runs only once training has passed its warm-up.
#define PT_G_WAVES 8
extern "C" __global__ __launch_bounds__(64 * PT_G_WAVES) void pt_guide_infer(const PtDev* __restrict__ D, unsigned int qsel) {
    if (!guideActive(D)) return;
    __shared__ GuideLds s;
    __shared__ __bf16 tiles[PT_G_WAVES][64 * PT_G_TILE_LD];
    guideLoadWeights(D, s);
    __syncthreads();
    const unsigned int count = D->counters[qsel ? PT_C_MARCH0 : PT_C_MARCH1];
    const unsigned int wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    __bf16* tile = tiles[wave];
    float* out = reinterpret_cast<float*>(tile);
    const unsigned int ntiles = (count + 63u) / 64u;
    unsigned long long nOn = 0ull, nTiles = 0ull;
    for (unsigned int t = blockIdx.x * PT_G_WAVES + wave; t < ntiles; t += gridDim.x * PT_G_WAVES) {
        const unsigned int i = t * 64u + lane;
        bool on = false;
        unsigned int slot = 0u;
        if (i < count) {
            slot = D->qSorted[i];
            const uint4 M = ld4u(D->meta, slot);
            if (mzState(M.z) == PT_ACTIVE && !(mzFlags(M.z) & PT_F_PHOTON)) {
                const float4 HR = ld4(D->hitRec, slot), P = ld4(D->pos, slot), Dd = ld4(D->dir, slot);
                const unsigned int hw = __float_as_uint(HR.w);
                V3 ns = v3(HR.x, HR.y, HR.z), dir = v3(Dd.x, Dd.y, Dd.z);
                V3 nrm = (hw & PT_HR_INSIDE) ? -ns : ns;
                const PtMaterial& m = D->mats[hw & 0xFFu];
                if (dot(dir, nrm) <= 0.0f && guideMaterial(m)) {
                    on = true;
                    guideFeatures(D, tile, (int)lane, v3(P.x, P.y, P.z), -dir, nrm, m.roughness);
                }
            }
        }
        if (!on) guideZeroRow(tile, (int)lane);
        guideForward(s, tile, out);
        if (on) guideStoreMix(D, slot, guideMixOf(out + lane * PT_G_OUT));
        nOn += on ? 1ull : 0ull;
        nTiles += lane == 0u ? 1ull : 0ull;
    }
    // the loop's bound is the wave's own, so every lane of the wave arrives here together
    statAdd(D, PT_S_GINFER, nOn);
    statAdd(D, PT_S_GTILES, nTiles);
}

// The forward pass on given vertices (the C ABI's evaluation for tests and tools): raw rows of 10 floats (position,
// view direction, normal, roughness) in, 41 head outputs a row out.
extern "C" __global__ __launch_bounds__(64) void pt_guide_eval(const PtDev* __restrict__ D, const float* __restrict__ raw,
                                                             unsigned int n, float* __restrict__ o41) {
    __shared__ GuideLds s;
    __shared__ __bf16 tile[64 * PT_G_TILE_LD];
    guideLoadWeights(D, s);
    __syncthreads();
    float* out = reinterpret_cast<float*>(tile);
    const unsigned int lane = threadIdx.x;
    for (unsigned int t = blockIdx.x; t * 64u < n; t += gridDim.x) {
        const unsigned int i = t * 64u + lane;
        if (i < n) {
            const float* r = raw + 10ull * i;
            guideFeatures(D, tile, (int)lane, v3(r[0], r[1], r[2]), v3(r[3], r[4], r[5]), v3(r[6], r[7], r[8]), r[9]);
        } else {
            guideZeroRow(tile, (int)lane);
        }
        guideForward(s, tile, out);
        if (i < n)
            for (int j = 0; j < PT_G_OUT; ++j) o41[(unsigned long long)PT_G_OUT * i + j] = out[lane * PT_G_OUT + j];
    }
}

// ---------------------------------------------------------------------------------------------------
// TRAINING DATA. A recorded path keeps, for each vertex it sampled at, 7 float4: position, roughness | view direction,
// the sampling density p(w) | normal | the sampled direction w | f cos at w (rgb), S-known | the throughput after the
// vertex (rgb) | S, the path's radiance so far (its own and its light sampling's) when the next event was met. At the
// sample's end the radiance that arrived along w is (total - S) / throughput, and the training weight
// F = luminance(f cos L) / p(w).
__device__ __forceinline__ float4* guideRec(const PtDev* __restrict__ D, unsigned int r, unsigned int k) {
    return reinterpret_cast<float4*>(D->gRec) + ((size_t)r * D->gMaxV + k) * PT_G_REC_F4;
}

// Collect: every recorded vertex with a known S becomes a training sample (position in the box, roughness | view
// direction, F | normal | w, the BSDF's density at w), appended; the sum of F over all of them kept for the steps'
// normalisation.
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_guide_collect(const PtDev* __restrict__ D) {
    const unsigned int idx = blockIdx.x * PT_BLOCK + threadIdx.x;
    const unsigned int r = idx / D->gMaxV, k = idx % D->gMaxV;
    bool ok = false;
    float4 a, b, c, d;
    float F = 0.0f;
    if (r < D->gPaths && k < D->gVerts[r]) {
        const float4* rec = guideRec(D, r, k);
        const float4 pr = rec[0], wp = rec[1], nn = rec[2], ww = rec[3], fc = rec[4], bt = rec[5], S = rec[6];
        const float4 tot = reinterpret_cast<const float4*>(D->gPath)[2 * r], nee = reinterpret_cast<const float4*>(D->gPath)[2 * r + 1];
        if (fc.w > 0.0f && tot.w > 0.0f) {
            V3 L = v3((tot.x + nee.x - S.x) / fmaxf(bt.x, 1e-20f), (tot.y + nee.y - S.y) / fmaxf(bt.y, 1e-20f),
                      (tot.z + nee.z - S.z) / fmaxf(bt.z, 1e-20f));
            F = luma(v3(fmaxf(fc.x * L.x, 0.0f), fmaxf(fc.y * L.y, 0.0f), fmaxf(fc.z * L.z, 0.0f))) / fmaxf(wp.w, 1e-20f);
            ok = isfinite(F) && F > 0.0f;
            const PtGuideConfig& g = D->g;
            a = make_float4((pr.x - g.boxMin[0]) / fmaxf(g.boxMax[0] - g.boxMin[0], 1e-6f),
                            (pr.y - g.boxMin[1]) / fmaxf(g.boxMax[1] - g.boxMin[1], 1e-6f),
                            (pr.z - g.boxMin[2]) / fmaxf(g.boxMax[2] - g.boxMin[2], 1e-6f), pr.w);
            b = make_float4(wp.x, wp.y, wp.z, F);
            c = nn; d = ww;
        }
    }
    const unsigned int at = waveAppend(&D->counters[PT_C_GUIDE_N], ok);
    if (ok && at < D->gBatchCap) {
        float4* dst = reinterpret_cast<float4*>(D->gBatch) + (size_t)PT_G_SAMPLE_F4 * at;
        dst[0] = a; dst[1] = b; dst[2] = c; dst[3] = d;
        atomicAdd(&D->gSumF[0], (double)F);
    }
}
// Step `step`'s batch: all N samples in order when they fit one batch (the exact batch gradient, F / sum F), else
// `batch` drawn uniformly with replacement from all N -- every image region in every step -- each weighing
// F N / (batch sum F), an unbiased estimate of the whole collection's gradient.
__device__ __forceinline__ unsigned int guidePick(const PtDev* __restrict__ D, unsigned int step, unsigned int i,
                                                  unsigned int N) {
    if (N <= D->g.batch) return i;
    Rng r;
    r.a = i; r.b = step * 0x9E3779B9u; r.c = D->counters[PT_C_GUIDE_T] ^ 0x6A09E667u; r.d = D->counters[PT_C_SAMPLE];
    pcg4d(r);
    return (unsigned int)(((unsigned long long)r.a * N) >> 32);
}

// ---------------------------------------------------------------------------------------------------
// THE GRADIENT of step `step`'s batch: one wave a block, tiles of 64 samples. Forward as inference (the hidden layers
// kept), the loss's gradient at the head from the mixture (d/do of -F log q / sum F, the oracle's dloss_do), then back
// through the layers on MFMA with the transposed operands gathered from LDS. Each block sums its tiles' weight
// gradients into its own row of the partial-gradient array (no atomics; Adam sums the rows).
struct GuideTrainLds {
    GuideLds s;
    __bf16 x[64 * PT_G_TILE_LD];   // features
    __bf16 h[3][64 * 64];          // the hidden layers, then their gradients in place
    float out[64 * PT_G_OUT];      // the head's outputs
    __bf16 dout[64 * 48];          // the head's gradient (K padded to 48)
};
// d/do at one sample (s = its share of the batch's weight) of -s log q(w) for the mixture's outputs and of
// -s log(c q(w) + (1 - c) b) for the fraction's; returns the combined density
__device__ __forceinline__ float guideHeadGrad(const float* o, V3 w, float bsdf, float s, float* g) {
    // responsibilities r_j = pi_j vMF_j(w) / q(w); the guide's share of the combined density rho = c q / p
    GuideMix m = guideMixOf(o);
    float comp[PT_G_LOBES], q = 0.0f;
#pragma unroll
    for (int j = 0; j < PT_G_LOBES; ++j) {
        const float4 L = m.lobe[j];
        comp[j] = m.w[j] * expf(vmfLogNorm(L.w) + L.w * (L.x * w.x + L.y * w.y + L.z * w.z - 1.0f));
        q += comp[j];
    }
    q = fmaxf(q, 1e-30f);
    const float p = fmaxf(m.c * q + (1.0f - m.c) * bsdf, 1e-30f);
    const float scale = s;
    g[40] = -s * (q - bsdf) * m.c * (1.0f - m.c) / p;
#pragma unroll
    for (int j = 0; j < PT_G_LOBES; ++j) {
        const float4 L = m.lobe[j];
        const float r = comp[j] / q, k = L.w;
        g[j] = -(r - m.w[j]) * scale;
        const float dt = L.x * w.x + L.y * w.y + L.z * w.z;
        V3 mraw = v3(o[8 + 3 * j], o[9 + 3 * j], o[10 + 3 * j]);
        const float ml = fmaxf(length(mraw), 1e-6f);
        const float cm = -r * k / ml * scale;
        g[8 + 3 * j] = cm * (w.x - dt * L.x);
        g[9 + 3 * j] = cm * (w.y - dt * L.y);
        g[10 + 3 * j] = cm * (w.z - dt * L.z);
        const float sraw = o[32 + j];
        const bool in = sraw > PT_G_KLO && sraw < PT_G_KHI;
        const float e2 = expf(-2.0f * k);
        const float dlogc = 1.0f / k - 2.0f * e2 / fmaxf(1.0f - e2, 1e-30f);
        g[32 + j] = in ? -r * (dlogc + (dt - 1.0f)) * k * scale : 0.0f;
    }
    return p;
}

extern "C" __global__ __launch_bounds__(64) void pt_guide_grad(const PtDev* __restrict__ D, unsigned int step) {
    __shared__ GuideTrainLds S;
    guideLoadWeights(D, S.s);
    const unsigned int lane = threadIdx.x;
    float* part = D->gPart + (size_t)blockIdx.x * PT_G_NP;
    for (unsigned int k = lane; k < PT_G_NP; k += 64u) part[k] = 0.0f;
    __syncthreads();
    const unsigned int N = D->counters[PT_C_GUIDE_N] < D->gBatchCap ? D->counters[PT_C_GUIDE_N] : D->gBatchCap;
    const unsigned int n = min(N, D->g.batch);
    const double sumF = D->gSumF[0];
    if (n == 0u || !(sumF > 0.0)) return;
    const float inv = (float)((double)N / ((double)n * sumF));
    const float meanF = (float)(sumF / (double)N);
    const __bf16* w = S.s.w;
    const float* b = S.s.b;
    float lossPart = 0.0f;
    for (unsigned int t = blockIdx.x; t * 64u < n; t += gridDim.x) {
        const unsigned int i = t * 64u + lane;
        const bool on = i < n;
        V3 wdir = v3(0.0f, 0.0f, 1.0f);
        float F = 0.0f, bsdf = 0.0f;
        if (on) {
            // the features, from the position already mapped into the box
            const float4* smp = reinterpret_cast<const float4*>(D->gBatch) + (size_t)PT_G_SAMPLE_F4 * guidePick(D, step, i, N);
            const float4 a = smp[0], bb = smp[1], c = smp[2], d = smp[3];
            __bf16* r = S.x + lane * PT_G_TILE_LD;
            const float xs[3] = {a.x, a.y, a.z};
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                float xa = fminf(fmaxf(xs[ax], 0.0f), 1.0f), sc = 1.0f;
#pragma unroll
                for (int k = 0; k < 12; ++k) {
                    float sn, cs;
                    sincospif(sc * xa, &sn, &cs);
                    r[6 * k + 2 * ax] = toBf16(sn);
                    r[6 * k + 2 * ax + 1] = toBf16(cs);
                    sc *= 2.0f;
                }
            }
            r[72] = toBf16(bb.x); r[73] = toBf16(bb.y); r[74] = toBf16(bb.z);
            r[75] = toBf16(c.x); r[76] = toBf16(c.y); r[77] = toBf16(c.z);
            r[78] = toBf16(a.w); r[79] = toBf16(1.0f);
            wdir = v3(d.x, d.y, d.z);
            bsdf = d.w;
            F = fminf(bb.w, meanF * 16.0f);   // a firefly's weight capped: it would steer a whole batch
        } else {
            guideZeroRow(S.x, (int)lane);
        }
        // forward, every layer kept: X -> H1 -> H2 -> H3 -> out
        waveGemmStream<2, 2, 5>([&](int ii, int k) { return ld8(S.x + ii * PT_G_TILE_LD + k); },
                                [&](int k, int j) { return ld8(w + PT_G_WT1 + j * 80 + k); },
                                [&](int ii, int j, float v) { S.h[0][ii * 64 + j] = toBf16(fmaxf(v + b[j], 0.0f)); });
        waveGemmStream<2, 2, 4>([&](int ii, int k) { return ld8(&S.h[0][ii * 64 + k]); },
                                [&](int k, int j) { return ld8(w + PT_G_WT2 + j * 64 + k); },
                                [&](int ii, int j, float v) { S.h[1][ii * 64 + j] = toBf16(fmaxf(v + b[64 + j], 0.0f)); });
        waveGemmStream<2, 2, 4>([&](int ii, int k) { return ld8(&S.h[1][ii * 64 + k]); },
                                [&](int k, int j) { return ld8(w + PT_G_WT3 + j * 64 + k); },
                                [&](int ii, int j, float v) { S.h[2][ii * 64 + j] = toBf16(fmaxf(v + b[128 + j], 0.0f)); });
        waveGemmStream<2, 2, 4>([&](int ii, int k) { return ld8(&S.h[2][ii * 64 + k]); },
                                [&](int k, int j) { return ld8(w + PT_G_WT4 + j * 64 + k); },
                                [&](int ii, int j, float v) { if (j < PT_G_OUT) S.out[ii * PT_G_OUT + j] = v + b[192 + j]; });
        // the loss's gradient at the head, this lane's row
        {
            float g[PT_G_OUT];
            if (on) {
                const float p = guideHeadGrad(S.out + lane * PT_G_OUT, wdir, bsdf, F * inv, g);
                lossPart += -F * inv * logf(p);
            } else {
#pragma unroll
                for (int j = 0; j < PT_G_OUT; ++j) g[j] = 0.0f;
            }
#pragma unroll
            for (int j = 0; j < 48; ++j) S.dout[lane * 48 + j] = toBf16(j < PT_G_OUT ? g[j] : 0.0f);
        }
        // b4, W4: gW4[h][o] = sum_rows H3[row][h] dO[row][o]
        if (lane < PT_G_OUT) {
            float sb = 0.0f;
            for (int row = 0; row < 64; ++row) sb += fromBf16(S.dout[row * 48 + lane]);
            part[PT_G_OFF_B4 + lane] += sb;
        }
        waveGemmStream<2, 2, 4>([&](int hh, int k) { return col8(&S.h[2][k * 64 + hh], 64); },
                                [&](int k, int o) { return col8(&S.dout[k * 48 + (o < 48 ? o : 0)], 48); },
                                [&](int hh, int o, float v) { if (o < PT_G_OUT) part[PT_G_OFF_W4 + hh * PT_G_OUT + o] += v; });
        // dZ3 = (dO W4^T) . [H3 > 0]   (B[k = o][n = h] = W4T[o][h])
        waveGemmStream<2, 2, 3>([&](int row, int k) { return ld8(&S.dout[row * 48 + k]); },
                                [&](int k, int hh) { return col8(w + PT_G_WT4 + k * 64 + hh, 64); },
                                [&](int row, int hh, float v) {
                                    __bf16* e = &S.h[2][row * 64 + hh];
                                    *e = toBf16(fromBf16(*e) > 0.0f ? v : 0.0f);
                                });
        // layers 3 and 2: gW[i][h] = sum_rows Hprev[row][i] dZ[row][h]; dZprev = (dZ W^T) . [Hprev > 0]
        for (int L = 2; L >= 1; --L) {
            __bf16* dz = S.h[L];
            __bf16* hp = S.h[L - 1];
            const int offW = L == 2 ? PT_G_OFF_W3 : PT_G_OFF_W2, offB = L == 2 ? PT_G_OFF_B3 : PT_G_OFF_B2;
            const __bf16* wt = w + (L == 2 ? PT_G_WT3 : PT_G_WT2);
            {
                float sb = 0.0f;
                for (int row = 0; row < 64; ++row) sb += fromBf16(dz[row * 64 + lane]);
                part[offB + lane] += sb;
            }
            waveGemmStream<2, 2, 4>([&](int ii, int k) { return col8(&hp[k * 64 + ii], 64); },
                                    [&](int k, int hh) { return col8(&dz[k * 64 + hh], 64); },
                                    [&](int ii, int hh, float v) { part[offW + ii * 64 + hh] += v; });
            // B[k = h][n = i] = W^T[h][i] = WT[h][i]
            waveGemmStream<2, 2, 4>([&](int row, int k) { return ld8(&dz[row * 64 + k]); },
                                    [&](int k, int ii) { return col8(wt + k * 64 + ii, 64); },
                                    [&](int row, int ii, float v) {
                                        __bf16* e = &hp[row * 64 + ii];
                                        *e = toBf16(fromBf16(*e) > 0.0f ? v : 0.0f);
                                    });
        }
        // layer 1: gW1[f][h] = sum_rows X[row][f] dZ1[row][h] (f padded to 96: rows of zeros past 80 are not stored)
        {
            float sb = 0.0f;
            for (int row = 0; row < 64; ++row) sb += fromBf16(S.h[0][row * 64 + lane]);
            part[PT_G_OFF_B1 + lane] += sb;
        }
        waveGemmStream<3, 2, 4>([&](int f, int k) {
                                    bf16x8 v;
#pragma unroll
                                    for (int q = 0; q < 8; ++q) v[q] = f < PT_G_IN ? S.x[(k + q) * PT_G_TILE_LD + f] : toBf16(0.0f);
                                    return v;
                                },
                                [&](int k, int hh) { return col8(&S.h[0][k * 64 + hh], 64); },
                                [&](int f, int hh, float v) { if (f < PT_G_IN) part[PT_G_OFF_W1 + f * 64 + hh] += v; });
    }
    lossPart = waveSum64f(lossPart);
    if (lane == 0 && lossPart != 0.0f) atomicAdd(&D->gLoss[step], (double)lossPart);
    // the rows this block ran: its tiles' in-range rows
    unsigned long long rows = 0ull;
    for (unsigned int t = blockIdx.x; t * 64u < n; t += gridDim.x) rows += (t * 64u + lane < n) ? 1ull : 0ull;
    statAdd(D, PT_S_GTRAIN, rows);
}

// ADAM (beta1 0.9, beta2 0.99, eps 1e-8) on the FP32 masters from the summed partial rows, the gradient cleared, the
// BF16 transposed copy and the biases refreshed. Step t = (steps so far) + 1; the commit kernel counts it afterwards.
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_guide_adam(const PtDev* __restrict__ D, unsigned int step) {
    const unsigned int p = blockIdx.x * PT_BLOCK + threadIdx.x;
    if (p >= PT_G_NP || !(D->gSumF[0] > 0.0)) return;
    float g = 0.0f;
    for (unsigned int b = 0; b < PT_G_PARTS; ++b) g += D->gPart[(size_t)b * PT_G_NP + p];
    const float t = (float)(D->counters[PT_C_GUIDE_T] + 1u);
    const float m = 0.9f * D->gM[p] + 0.1f * g;
    const float v = 0.99f * D->gV[p] + 0.01f * g * g;
    D->gM[p] = m;
    D->gV[p] = v;
    const float mh = m / (1.0f - powf(0.9f, t)), vh = v / (1.0f - powf(0.99f, t));
    const float th = D->gMaster[p] - D->g.lr * mh / (sqrtf(vh) + 1e-8f);
    D->gMaster[p] = th;
    if (D->gGradOut) D->gGradOut[p] = g;
    // the inference copy
    unsigned int idx;
    bool bias = false;
    if (p < PT_G_OFF_B1) { unsigned int i = p / 64u, o = p % 64u; idx = PT_G_WT1 + o * 80u + i; }
    else if (p < PT_G_OFF_W2) { bias = true; idx = p - PT_G_OFF_B1; }
    else if (p < PT_G_OFF_B2) { unsigned int q = p - PT_G_OFF_W2, i = q / 64u, o = q % 64u; idx = PT_G_WT2 + o * 64u + i; }
    else if (p < PT_G_OFF_W3) { bias = true; idx = 64u + p - PT_G_OFF_B2; }
    else if (p < PT_G_OFF_B3) { unsigned int q = p - PT_G_OFF_W3, i = q / 64u, o = q % 64u; idx = PT_G_WT3 + o * 64u + i; }
    else if (p < PT_G_OFF_W4) { bias = true; idx = 128u + p - PT_G_OFF_B3; }
    else if (p < PT_G_OFF_B4) { unsigned int q = p - PT_G_OFF_W4, i = q / PT_G_OUT, o = q % PT_G_OUT; idx = PT_G_WT4 + o * 64u + i; }
    else { bias = true; idx = 192u + p - PT_G_OFF_B4; }
    if (bias) D->gBias[idx] = th;
    else reinterpret_cast<__bf16*>(D->gW)[idx] = toBf16(th);
}
extern "C" __global__ void pt_guide_commit(const PtDev* __restrict__ D, unsigned int step) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && D->gSumF[0] > 0.0) {
        D->counters[PT_C_GUIDE_T] += 1u;
        atomicAdd(&D->stats[PT_S_GSTEPS], 1ull);
    }
}
