// THE DEVELOPED PICTURE: the FP64 accumulation (cd/m^2) exposed, tone mapped and encoded as 8-bit sRGB
// (core/rendertoscreen.js displayPixel :209-260, core/autoexposure.js :1-153, core/photometry.js :83-96).
//
// Exposure is photographic: at EV100 e, the luminance that clips is 1.2 2^e cd/m^2 (ISO 12232's saturation-based
// speed), so a pixel's value before the tone curve is L / (1.2 2^e). The meter follows ISO 2720's reflected-light
// calibration, e = log2(L_avg 100 / 12.5).
//
// AUTO-EXPOSURE, in two kernels and nothing read back on the way: pt_ae_hist bins every sampled pixel's log2 luminance
// into PT_AE_BINS bins over [PT_AE_LOG_MIN, PT_AE_LOG_MAX] (a block's histogram in LDS, then one atomic per bin);
// pt_ae_update, one thread, takes the average of log2 L over the pixels between the 10 % and 90 % quantiles (cut
// tails: black background below, the sun, speculars and fireflies above), meters it, and holds the 95 % quantile at
// most `headroom` stops above white (never brighter); the first call sets the exposure, later ones move it by
// `speed` of the way, so a progressive picture's exposure does not shimmer with its noise. Pixels darker than the
// histogram's floor do not vote: they are no scene (a void, a sky switched off), not a dark part of one.
#define PT_AE_BINS 128
#define PT_AE_LOG_MIN -10.0f
#define PT_AE_LOG_MAX 34.0f
#define PT_EV_CLIP 1.2f
#define PT_METER_K 12.5f

__device__ __forceinline__ V3 accumMean(const PtDev* __restrict__ D, unsigned int i, bool& sampled) {
    const double* a = D->accum + 4ull * i;
    sampled = a[3] > 0.0;
    return v3((float)a[0], (float)a[1], (float)a[2]);
}

extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_ae_hist(const PtDev* __restrict__ D, unsigned int* __restrict__ hist) {
    __shared__ unsigned int h[PT_AE_BINS];
    for (unsigned int k = threadIdx.x; k < PT_AE_BINS; k += PT_BLOCK) h[k] = 0u;
    __syncthreads();
    const unsigned int n = D->u.width * D->u.height;
    for (unsigned int i = blockIdx.x * PT_BLOCK + threadIdx.x; i < n; i += gridDim.x * PT_BLOCK) {
        bool sampled;
        const float Y = luma(accumMean(D, i, sampled));
        if (!sampled || !(Y >= exp2f(PT_AE_LOG_MIN))) continue;
        const float t = (log2f(Y) - PT_AE_LOG_MIN) / (PT_AE_LOG_MAX - PT_AE_LOG_MIN);
        const unsigned int bin = (unsigned int)fminf(fmaxf(t * (float)PT_AE_BINS, 0.0f), (float)(PT_AE_BINS - 1));
        atomicAdd(&h[bin], 1u);
    }
    __syncthreads();
    for (unsigned int k = threadIdx.x; k < PT_AE_BINS; k += PT_BLOCK)
        if (h[k]) atomicAdd(&hist[k], h[k]);
}

// state: [0] EV100 in force, [1] 1 once set, [2] the metered log2 average, [3] log2 of the protected quantile,
// [4] pixels counted. s: the developing settings. The histogram is cleared for the next call.
extern "C" __global__ void pt_ae_update(unsigned int* __restrict__ hist, float* __restrict__ state, const PtDevelop* __restrict__ s) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    float total = 0.0f;
    for (unsigned int k = 0; k < PT_AE_BINS; ++k) total += (float)hist[k];
    state[4] = total;
    if (total == 0.0f) return;                     // nothing sampled above the floor: the exposure stays
    const float lo = 0.10f * total, hi = 0.90f * total, q = 0.95f * total;
    const float w = (PT_AE_LOG_MAX - PT_AE_LOG_MIN) / (float)PT_AE_BINS;
    float below = 0.0f, kept = 0.0f, sum = 0.0f, logQ = PT_AE_LOG_MAX;
    bool haveQ = false;
    for (unsigned int k = 0; k < PT_AE_BINS; ++k) {
        const float c = (float)hist[k];
        hist[k] = 0u;
        // the part of this bin between the two cuts, at its centre
        const float a = fmaxf(below, lo), b = fminf(below + c, hi);
        if (b > a) { kept += b - a; sum += (b - a) * (PT_AE_LOG_MIN + ((float)k + 0.5f) * w); }
        // the protected quantile, placed within its bin by where the running count passes it
        if (!haveQ && c > 0.0f && below + c >= q) {
            logQ = PT_AE_LOG_MIN + ((float)k + fminf(fmaxf((q - below) / c, 0.0f), 1.0f)) * w;
            haveQ = true;
        }
        below += c;
    }
    const float logL = sum / fmaxf(kept, 1e-9f);
    const float evMeter = log2f(exp2f(logL) * 100.0f / PT_METER_K);
    const float evHigh = logQ - log2f(PT_EV_CLIP) - fmaxf(s->aeHeadroom, 0.0f);
    const float target = fmaxf(evMeter, evHigh);
    const float ev = state[1] != 0.0f ? state[0] + fminf(fmaxf(s->aeSpeed, 0.0f), 1.0f) * (target - state[0]) : target;
    state[0] = ev; state[1] = 1.0f; state[2] = logL; state[3] = logQ;
}

// ---- the tone curves, on linear light already divided by the clipping luminance (1 is white)
// Khronos PBR Neutral (2024): linear below the compression start but for a small toe offset, then the brightest channel
// compressed towards 1 with the hue kept, and a little desaturation where it compresses.
__device__ __forceinline__ V3 tonePbrNeutral(V3 c) {
    const float start = 0.76f, desat = 0.15f;
    const float m = fminf(c.x, fminf(c.y, c.z));
    const float off = m < 0.08f ? m - 6.25f * m * m : 0.04f;
    c = c - v3s(off);
    const float peak = fmaxf(c.x, fmaxf(c.y, c.z));
    if (peak < start) return c;
    const float d = 1.0f - start;
    const float np = 1.0f - d * d / (peak + d - start);
    c = c * (np / peak);
    const float g = 1.0f - 1.0f / (desat * (peak - np) + 1.0f);
    return c * (1.0f - g) + v3s(np * g);
}
// AgX (Sobotka), the widely used polynomial approximation (Wrensch 2023): into AgX's inset primaries, log2 over
// [-12.47393, 4.026069] stops about mid grey, a sixth-order fit of the default contrast curve, the outset back, and
// the 2.2 display power undone so that the sRGB encoding below applies. `punchy`: Blender's Punchy look in the
// encoded space (power 1.35, then saturation 1.4 about the Rec.709 luma from before the power).
__device__ __forceinline__ V3 mat3v(const float* M, V3 v) {
    return v3(M[0] * v.x + M[1] * v.y + M[2] * v.z, M[3] * v.x + M[4] * v.y + M[5] * v.z, M[6] * v.x + M[7] * v.y + M[8] * v.z);
}
__device__ __forceinline__ float agxCurve(float x) {
    const float x2 = x * x, x4 = x2 * x2;
    return 15.5f * x4 * x2 - 40.14f * x4 * x + 31.96f * x4 - 6.868f * x2 * x + 0.4298f * x2 + 0.1191f * x - 0.00232f;
}
__device__ __forceinline__ V3 toneAgX(V3 c, bool punchy) {
    // rows of the inset and outset matrices (each the transpose of the column-major listing the approximation gives)
    const float inset[9] = {0.842479062253094f, 0.0784335999999992f, 0.0792237451477643f,
                            0.0423282422610123f, 0.878468636469772f, 0.0791661274605434f,
                            0.0423756549057051f, 0.0784336f, 0.879142973793104f};
    const float outset[9] = {1.19687900512017f, -0.0980208811401368f, -0.0990297440797205f,
                             -0.0528968517574562f, 1.15190312990417f, -0.0989611768448433f,
                             -0.0529716355144438f, -0.0980434501171241f, 1.15107367264116f};
    const float lo = -12.47393f, hi = 4.026069f;
    V3 v = mat3v(inset, v3(fmaxf(c.x, 1e-10f), fmaxf(c.y, 1e-10f), fmaxf(c.z, 1e-10f)));
    v = v3(fminf(fmaxf(log2f(v.x), lo), hi), fminf(fmaxf(log2f(v.y), lo), hi), fminf(fmaxf(log2f(v.z), lo), hi));
    v = (v - v3s(lo)) * (1.0f / (hi - lo));
    v = v3(agxCurve(v.x), agxCurve(v.y), agxCurve(v.z));
    if (punchy) {
        // an ASC CDL power, then saturation about the luma the value had before it
        const float l = luma(v);
        v = v3(powf(fmaxf(v.x, 0.0f), 1.35f), powf(fmaxf(v.y, 0.0f), 1.35f), powf(fmaxf(v.z, 0.0f), 1.35f));
        v = v3s(l) + (v - v3s(l)) * 1.4f;
    }
    v = mat3v(outset, v);
    return v3(powf(fmaxf(v.x, 0.0f), 2.2f), powf(fmaxf(v.y, 0.0f), 2.2f), powf(fmaxf(v.z, 0.0f), 2.2f));
}
// ACES, Stephen Hill's fit of the reference rendering and sRGB output transforms: into the ACES working space, a
// rational fit of RRT + ODT per channel, back out. Rows of each matrix sum to 1, so grey stays grey.
__device__ __forceinline__ float acesFit(float v) {
    return (v * (v + 0.0245786f) - 0.000090537f) / (v * (0.983729f * v + 0.4329510f) + 0.238081f);
}
__device__ __forceinline__ V3 toneAces(V3 c) {
    const float in[9] = {0.59719f, 0.35458f, 0.04823f, 0.07600f, 0.90834f, 0.01566f, 0.02840f, 0.13383f, 0.83777f};
    const float out[9] = {1.60475f, -0.53108f, -0.07367f, -0.10208f, 1.10813f, -0.00605f, -0.00327f, -0.07276f, 1.07602f};
    V3 v = mat3v(in, c);
    v = v3(acesFit(v.x), acesFit(v.y), acesFit(v.z));
    v = mat3v(out, v);
    return v3(fminf(fmaxf(v.x, 0.0f), 1.0f), fminf(fmaxf(v.y, 0.0f), 1.0f), fminf(fmaxf(v.z, 0.0f), 1.0f));
}
// Reinhard on the luminance, L / (1 + L), the colour scaled with it: hue kept, nothing clips.
__device__ __forceinline__ V3 toneReinhard(V3 c) {
    const float L = luma(c);
    return L > 0.0f ? c * (1.0f / (1.0f + L)) : v3s(0.0f);
}
// Hable's Uncharted 2 filmic curve (A 0.15, B 0.50, C 0.10, D 0.20, E 0.02, F 0.30), exposure bias 2, white 11.2.
__device__ __forceinline__ float hableF(float x) {
    const float A = 0.15f, B = 0.50f, C = 0.10f, Dd = 0.20f, E = 0.02f, F = 0.30f;
    return (x * (A * x + C * B) + Dd * E) / (x * (A * x + B) + Dd * F) - E / F;
}
__device__ __forceinline__ V3 toneHable(V3 c) {
    const float w = 1.0f / hableF(11.2f);
    return v3(hableF(2.0f * c.x) * w, hableF(2.0f * c.y) * w, hableF(2.0f * c.z) * w);
}
// IEC 61966-2-1's sRGB encoding of a linear value in [0, 1]
__device__ __forceinline__ float srgbEncode(float v) {
    v = fminf(fmaxf(v, 0.0f), 1.0f);
    return v <= 0.0031308f ? 12.92f * v : 1.055f * powf(v, 1.0f / 2.4f) - 0.055f;
}

// One pixel per thread: the accumulation's mean, exposed, tone mapped, encoded, to 8 bits (RGBA, alpha 255; a pixel
// never sampled is black). Rows written top first when s->flipY (the accumulation's row 0 is the bottom).
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_develop(const PtDev* __restrict__ D, uchar4* __restrict__ out,
                                                                 const float* __restrict__ state, const PtDevelop* __restrict__ s) {
    const unsigned int W = D->u.width, H = D->u.height;
    const unsigned int i = blockIdx.x * PT_BLOCK + threadIdx.x;
    if (i >= W * H) return;
    bool sampled;
    V3 c = accumMean(D, i, sampled);
    const float ev = s->autoExposure ? state[0] : s->exposureEV;
    const float scale = 1.0f / (PT_EV_CLIP * exp2f(ev));
    c = v3(fmaxf(c.x, 0.0f), fmaxf(c.y, 0.0f), fmaxf(c.z, 0.0f)) * scale;
    switch (s->toneMap) {
        case 1u: c = tonePbrNeutral(c); break;
        case 2u: c = toneAgX(c, false); break;
        case 3u: c = toneAgX(c, true); break;
        case 4u: c = toneAces(c); break;
        case 5u: c = toneReinhard(c); break;
        case 6u: c = toneHable(c); break;
        default: break;                            // 0: linear, clipped at white
    }
    float r = srgbEncode(c.x) * 255.0f, g = srgbEncode(c.y) * 255.0f, b = srgbEncode(c.z) * 255.0f;
    float d = 0.5f;
    if (s->dither) {
        // +-0.5 of a step, uniform, from the pixel's hash: breaks 8-bit banding in smooth gradients
        Rng q; q.a = i; q.b = 0xD1u; q.c = 0x7E5u; q.d = 0x5EEDu;
        pcg4d(q);
        d = (float)(q.a >> 8) * (1.0f / 16777216.0f);
    }
    const unsigned int x = i % W, y = i / W;
    const unsigned int o = (s->flipY ? (H - 1u - y) : y) * W + x;
    out[o] = sampled ? make_uchar4((unsigned char)fminf(r + d, 255.0f), (unsigned char)fminf(g + d, 255.0f),
                                   (unsigned char)fminf(b + d, 255.0f), 255)
                     : make_uchar4(0, 0, 0, 255);
}
