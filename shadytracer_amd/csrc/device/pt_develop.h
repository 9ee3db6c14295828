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

// ---- THE DARKROOM (darkroom/grade.js, core/rendertoscreen.js developPixel :423-455): tools on the light between the
// exposure and the tone curve, tools on the encoded picture after it. Each runs only when its PT_LIGHT_* bit is on.

// Into the display's gamut (core/colorimetry.js gamutClip): a colour with a negative channel -- the exact colour of a
// saturated LED, which the accumulation keeps -- is desaturated towards the grey of its own luminance until that
// channel is 0; luminance kept, no light of its own: black.
__device__ __forceinline__ V3 gamutClip(V3 c) {
    const float lo = fminf(c.x, fminf(c.y, c.z));
    if (lo >= 0.0f) return c;
    const float Y = luma(c);
    if (!(Y > 0.0f)) return v3s(0.0f);
    const float t = Y / (Y - lo);
    return v3(fmaxf(Y + (c.x - Y) * t, 0.0f), fmaxf(Y + (c.y - Y) * t, 0.0f), fmaxf(Y + (c.z - Y) * t, 0.0f));
}
// Oklab (Ottosson 2020): linear sRGB to LMS, a cube root that keeps the sign (a colour outside the gamut goes through
// and back), then Lab; and its inverse
__device__ __forceinline__ float cbrtKeep(float x) { return copysignf(powf(fabsf(x), 1.0f / 3.0f), x); }
__device__ __forceinline__ V3 toOklab(V3 c) {
    const float l = cbrtKeep(0.4122214708f * c.x + 0.5363325363f * c.y + 0.0514459929f * c.z);
    const float m = cbrtKeep(0.2119034982f * c.x + 0.6806995451f * c.y + 0.1073969566f * c.z);
    const float s = cbrtKeep(0.0883024619f * c.x + 0.2817188376f * c.y + 0.6299787005f * c.z);
    return v3(0.2104542553f * l + 0.7936177850f * m - 0.0040720468f * s,
              1.9779984951f * l - 2.4285922050f * m + 0.4505937099f * s,
              0.0259040371f * l + 0.7827717662f * m - 0.8086757660f * s);
}
__device__ __forceinline__ V3 fromOklab(V3 q) {
    const float l = q.x + 0.3963377774f * q.y + 0.2158037573f * q.z;
    const float m = q.x - 0.1055613458f * q.y - 0.0638541728f * q.z;
    const float s = q.x - 0.0894841775f * q.y - 1.2914855480f * q.z;
    const float l3 = l * l * l, m3 = m * m * m, s3 = s * s * s;
    return v3(4.0767416621f * l3 - 3.3077115913f * m3 + 0.2309699292f * s3,
              -1.2684380046f * l3 + 2.6097574011f * m3 - 0.3413193965f * s3,
              -0.0041960863f * l3 - 0.7034186147f * m3 + 1.7076147010f * s3);
}
// Hable's rational curve with its shoulder and toe given (linear strength 0.50, angle 0.10, toe 0.02 / 0.30 his)
__device__ __forceinline__ float hableP(float x, float sh, float to) {
    return (x * (sh * x + 0.10f * 0.50f) + to * 0.02f) / (x * (sh * x + 0.50f) + to * 0.30f) - 0.02f / 0.30f;
}
// The light stages on an exposed colour (1.0 is white): white balance and gains, the tone equaliser (8 bands, a gain in
// stops interpolated in log2 of the luminance and held at the ends, the same on every channel), Oklab's saturation,
// vibrance (more where the chroma is low against the lightness: up to a quarter of it) and hue.
__device__ __forceinline__ V3 darkroomLight(V3 c, const PtDevelop* __restrict__ s, unsigned int on) {
    if (on & PT_LIGHT_WB) c = mat3v(s->wb, c);
    if (on & PT_LIGHT_EQ) {
        const float t = fminf(fmaxf(log2f(fmaxf(luma(c), 1e-30f)) + 7.0f, 0.0f), 7.0f);
        const unsigned int k = min((unsigned int)floorf(t), 6u);
        const float stops = s->eq[k] + (s->eq[k + 1] - s->eq[k]) * (t - (float)k);
        c = c * exp2f(stops);
    }
    if (on & PT_LIGHT_LAB) {
        const V3 q = toOklab(c);
        const float rel = sqrtf(q.y * q.y + q.z * q.z) / fmaxf(fabsf(q.x), 1e-6f);
        const float k = (1.0f + s->lab[0]) * (1.0f + s->lab[1] * (1.0f - fminf(rel / 0.25f, 1.0f)));
        float sn, cs;
        sincosf(s->lab[2], &sn, &cs);
        c = fromOklab(v3(q.x, (q.y * cs - q.z * sn) * k, (q.y * sn + q.z * cs) * k));
    }
    return c;
}
// PCG's output permutation on 32 bits (the hash of Jarzynski and Olano 2020): grain and dither from the picture pixel,
// the same at a pixel in every development
__device__ __forceinline__ unsigned int pcgHash(unsigned int n) {
    const unsigned int st = n * 747796405u + 2891336453u;
    const unsigned int w = ((st >> ((st >> 28u) + 4u)) ^ st) * 277803737u;
    return (w >> 22u) ^ w;
}
__device__ __forceinline__ float unit24(unsigned int h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }
__device__ __forceinline__ float grainCell(int ix, int iy) {
    return unit24(pcgHash((unsigned int)(ix + iy * 65536) ^ 0x632BE5ABu)) * 2.0f - 1.0f;
}
// value noise on cells `size` pixels wide, smoothstep-interpolated, -1..1
__device__ __forceinline__ float grainNoise(unsigned int px, unsigned int py, float size) {
    const float gx = ((float)px + 0.5f) / size, gy = ((float)py + 0.5f) / size;
    const float cx = floorf(gx), cy = floorf(gy);
    float fx = gx - cx, fy = gy - cy;
    fx = fx * fx * (3.0f - 2.0f * fx);
    fy = fy * fy * (3.0f - 2.0f * fy);
    const int ix = (int)cx, iy = (int)cy;
    const float a = grainCell(ix, iy), b = grainCell(ix + 1, iy), c = grainCell(ix, iy + 1), d = grainCell(ix + 1, iy + 1);
    const float top = a + (b - a) * fx;
    return top + ((c + (d - c) * fx) - top) * fy;
}
// triangular dither, -1..1 of one 8-bit step
__device__ __forceinline__ float ditherAt(unsigned int px, unsigned int py) {
    const unsigned int h1 = pcgHash((px + py * 65536u) ^ 0x1B873593u), h2 = pcgHash(h1);
    return (unit24(h1) + unit24(h2) - 1.0f) / 255.0f;
}
// a lens's vignette in stops: 0 at the centre, vignette[0] at the corners, from vignette[1] of the way out; vignette[2]
// 0 follows the picture's shape, 1 is a circle in pixels
__device__ __forceinline__ float vignetteStops(unsigned int px, unsigned int py, float W, float H, const float* v) {
    const float qx = (((float)px + 0.5f) / W - 0.5f) * 2.0f, qy = (((float)py + 0.5f) / H - 0.5f) * 2.0f;
    const float a = W / H;
    const float r0 = sqrtf(qx * qx + qy * qy) / sqrtf(2.0f);
    const float r1 = sqrtf(qx * a * qx * a + qy * qy) / sqrtf(a * a + 1.0f);
    const float r = r0 + (r1 - r0) * v[2];
    const float t = fminf(fmaxf((r - v[1]) / fmaxf(1e-6f, 1.0f - v[1]), 0.0f), 1.0f);
    return v[0] * t * t * (3.0f - 2.0f * t);
}
// the output transfer of a value in [0, 1]: 0 sRGB, 1 gamma 2.2, 2 gamma 2.4, 3 linear
__device__ __forceinline__ float encodeOut(float c, unsigned int transfer) {
    switch (transfer) {
        case 1u: return powf(c, 1.0f / 2.2f);
        case 2u: return powf(c, 1.0f / 2.4f);
        case 3u: return c;
        default: return srgbEncode(c);
    }
}
// one curve of the baked table at x: linear between its two nearest entries
__device__ __forceinline__ float lutAt(const float* __restrict__ lut, unsigned int ch, float x) {
    const float t = fminf(fmaxf(x, 0.0f), 1.0f) * (float)(PT_CURVE_LUT - 1);
    const unsigned int k = min((unsigned int)floorf(t), (unsigned int)(PT_CURVE_LUT - 2));
    const float a = lut[ch * PT_CURVE_LUT + k], b = lut[ch * PT_CURVE_LUT + k + 1];
    return a + (b - a) * (t - (float)k);
}
// Levels, the curves (master on every channel or on the luma alone, then red, green, blue) and the wheels as an ASC CDL
// (slope, offset, power), on the encoded picture
__device__ __forceinline__ V3 darkroomFinish(V3 c, const PtDevelop* __restrict__ s, unsigned int on) {
    if (on & PT_LIGHT_LEVELS) {
        const float span = fmaxf(1e-6f, s->levels[1] - s->levels[0]), ob = s->levels[3], ow = s->levels[4];
        const float g = s->levels[2];
        c = v3(ob + (ow - ob) * powf(fminf(fmaxf((c.x - s->levels[0]) / span, 0.0f), 1.0f), g),
               ob + (ow - ob) * powf(fminf(fmaxf((c.y - s->levels[0]) / span, 0.0f), 1.0f), g),
               ob + (ow - ob) * powf(fminf(fmaxf((c.z - s->levels[0]) / span, 0.0f), 1.0f), g));
    }
    if (on & PT_LIGHT_CURVES) {
        const float* lut = s->curveLut;
        if (on & PT_LIGHT_CURVE_LUMA) {
            const float y = luma(c), y2 = lutAt(lut, 0u, y);
            c = y > 1e-6f ? c * (y2 / y) : v3s(y2);
        } else {
            c = v3(lutAt(lut, 0u, c.x), lutAt(lut, 0u, c.y), lutAt(lut, 0u, c.z));
        }
        c = v3(lutAt(lut, 1u, c.x), lutAt(lut, 2u, c.y), lutAt(lut, 3u, c.z));
    }
    if (on & PT_LIGHT_CDL) {
        c = v3(powf(fmaxf(c.x * s->cdlSlope[0] + s->cdlOffset[0], 0.0f), s->cdlPower[0]),
               powf(fmaxf(c.y * s->cdlSlope[1] + s->cdlOffset[1], 0.0f), s->cdlPower[1]),
               powf(fmaxf(c.z * s->cdlSlope[2] + s->cdlOffset[2], 0.0f), s->cdlPower[2]));
    }
    return c;
}
// Google's Turbo colour map in its published polynomial fit, display-encoded already
__device__ __forceinline__ V3 turbo(float t) {
    const float x = fminf(fmaxf(t, 0.0f), 1.0f), x2 = x * x, x3 = x2 * x, x4 = x2 * x2, x5 = x4 * x;
    return v3(0.13572138f + 4.61539260f * x - 42.66032258f * x2 + 132.13108234f * x3 - 152.94239396f * x4 + 59.28637943f * x5,
              0.09140261f + 2.19418839f * x + 4.84296658f * x2 - 14.18503333f * x3 + 4.27729857f * x4 + 2.82956604f * x5,
              0.10667330f + 12.64194608f * x - 60.58204836f * x2 + 110.36276771f * x3 - 89.90310912f * x4 + 27.34824973f * x5);
}
// THE FALSE-COLOUR VIEWS (displayPixel's measure modes): the light as a number on a log scale from measureMin to
// measureMax, Turbo-coloured, stepped into measureBands bands when that is not 0; mode 1 reads pi Y as lux (the
// accumulation then holds E / pi), mode 2 Y in cd/m^2. The bottom 2.5 % of the picture is the scale itself. No light:
// black. (u, v) is the pixel centre in the picture, v = 0 at the bottom.
__device__ __forceinline__ V3 falseColour(V3 c, float u, float v, const PtDevelop* __restrict__ s) {
    const float Y = fmaxf(luma(c), 0.0f);
    const float val = s->measureMode == 1u ? 3.14159265f * Y : Y;
    const float lo = log2f(fmaxf(s->measureMin, 1e-6f));
    const float hi = fmaxf(log2f(fmaxf(s->measureMax, 1e-6f)), lo + 1e-3f);
    const bool strip = v < 0.025f;
    float t = strip ? u : (log2f(fmaxf(val, 1e-30f)) - lo) / (hi - lo);
    if (s->measureBands > 0u) {
        const float b = (float)s->measureBands;
        t = (floorf(fminf(fmaxf(t, 0.0f), 0.99999f) * b) + 0.5f) / b;
    }
    const V3 shown = (val <= 0.0f && !strip) ? v3s(0.0f) : turbo(t);
    return v3(fminf(fmaxf(shown.x, 0.0f), 1.0f), fminf(fmaxf(shown.y, 0.0f), 1.0f), fminf(fmaxf(shown.z, 0.0f), 1.0f));
}

// ONE PIXEL DEVELOPED: its mean light (cd/m^2) to display values in [0, 1], the reference's developPixel in its order --
// the light stages, the vignette, the gamut, the tone curve, the output transfer, the finish, the display's grading
// (contrast about 0.5, saturation about the luma, brightness), grain, dither. (px, py): the picture pixel, row 0 at
// the bottom.
__device__ __forceinline__ V3 developPixel(V3 light, float scale, unsigned int px, unsigned int py, unsigned int W, unsigned int H,
                                           const PtDevelop* __restrict__ s) {
    if (s->measureMode != 0u) return falseColour(light, ((float)px + 0.5f) / (float)W, ((float)py + 0.5f) / (float)H, s);
    const unsigned int on = s->lightOn;
    V3 c = light * scale;
    if (on) c = darkroomLight(c, s, on);
    if (on & PT_LIGHT_VIGNETTE) c = c * exp2f(vignetteStops(px, py, (float)W, (float)H, s->vignette));
    c = gamutClip(c);
    switch (s->toneMap) {
        case 1u: c = tonePbrNeutral(c); break;
        case 2u: c = toneAgX(c, false); break;
        case 3u: c = toneAgX(c, true); break;
        case 4u: c = toneAces(c); break;
        case 5u: c = toneReinhard(c); break;
        case 6u: c = toneHable(c); break;
        case 7u: {
            const float w = 1.0f / hableP(2.0f * s->tone7[0], s->tone7[2], s->tone7[1]);
            c = v3(hableP(2.0f * c.x, s->tone7[2], s->tone7[1]) * w, hableP(2.0f * c.y, s->tone7[2], s->tone7[1]) * w,
                   hableP(2.0f * c.z, s->tone7[2], s->tone7[1]) * w);
            break;
        }
        default: break;                            // 0: linear, clipped at white
    }
    c = v3(fminf(fmaxf(c.x, 0.0f), 1.0f), fminf(fmaxf(c.y, 0.0f), 1.0f), fminf(fmaxf(c.z, 0.0f), 1.0f));
    c = v3(encodeOut(c.x, s->transfer), encodeOut(c.y, s->transfer), encodeOut(c.z, s->transfer));
    if (on) c = darkroomFinish(c, s, on);
    // the display's grading; at 0 each step is exact (x (1 - 1) + y 1 = y), so an ungraded picture is untouched
    const float tc = s->contrast + 1.0f, ts = s->saturation + 1.0f;
    c = v3s(0.5f * (1.0f - tc)) + c * tc;
    const float l = luma(c);
    c = v3s(l * (1.0f - ts)) + c * ts;
    c = c + v3s(s->brightness);
    if (on & PT_LIGHT_GRAIN) {
        const float y = fminf(fmaxf(luma(c), 0.0f), 1.0f);
        c = c + v3s(grainNoise(px, py, s->grain[1]) * s->grain[0] * 0.08f * (0.3f + 2.8f * y * (1.0f - y)));
    }
    if (on & PT_LIGHT_DITHER) c = c + v3s(ditherAt(px, py));
    return v3(fminf(fmaxf(c.x, 0.0f), 1.0f), fminf(fmaxf(c.y, 0.0f), 1.0f), fminf(fmaxf(c.z, 0.0f), 1.0f));
}

__device__ __forceinline__ float developScale(const float* __restrict__ state, const PtDevelop* __restrict__ s) {
    const float ev = s->autoExposure ? state[0] : s->exposureEV;
    return 1.0f / (PT_EV_CLIP * exp2f(ev));
}

// One pixel per thread: developed, rounded to 8 bits (RGBA, alpha 255). Rows written top first when s->flipY (the
// accumulation's row 0 is the bottom).
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_develop(const PtDev* __restrict__ D, uchar4* __restrict__ out,
                                                                 const float* __restrict__ state, const PtDevelop* __restrict__ s) {
    const unsigned int W = D->u.width, H = D->u.height;
    const unsigned int i = blockIdx.x * PT_BLOCK + threadIdx.x;
    if (i >= W * H) return;
    bool sampled;
    const V3 light = accumMean(D, i, sampled);
    const unsigned int x = i % W, y = i / W;
    const V3 c = developPixel(light, developScale(state, s), x, y, W, H, s);
    const unsigned int o = (s->flipY ? (H - 1u - y) : y) * W + x;
    out[o] = make_uchar4((unsigned char)fminf(c.x * 255.0f + 0.5f, 255.0f), (unsigned char)fminf(c.y * 255.0f + 0.5f, 255.0f),
                         (unsigned char)fminf(c.z * 255.0f + 0.5f, 255.0f), 255);
}

// The same development to 16 bits a channel (RGBA, alpha 65535): round(v 65535), as a 16-bit PNG keeps it.
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_develop16(const PtDev* __restrict__ D, ushort4* __restrict__ out,
                                                                   const float* __restrict__ state, const PtDevelop* __restrict__ s) {
    const unsigned int W = D->u.width, H = D->u.height;
    const unsigned int i = blockIdx.x * PT_BLOCK + threadIdx.x;
    if (i >= W * H) return;
    bool sampled;
    const V3 light = accumMean(D, i, sampled);
    const unsigned int x = i % W, y = i / W;
    const V3 c = developPixel(light, developScale(state, s), x, y, W, H, s);
    const unsigned int o = (s->flipY ? (H - 1u - y) : y) * W + x;
    out[o] = make_ushort4((unsigned short)fminf(c.x * 65535.0f + 0.5f, 65535.0f), (unsigned short)fminf(c.y * 65535.0f + 0.5f, 65535.0f),
                          (unsigned short)fminf(c.z * 65535.0f + 0.5f, 65535.0f), 65535);
}

// A VIEW OF THE DEVELOPED PICTURE (rendertoscreen.js develop_prog's frag): an output of outW x outH pixels looking at
// the picture through view = (cx, cy, sx, sy) -- the picture point at the output's centre and the picture pixels one
// output pixel spans, both in picture pixels with y up. Each output pixel averages box x box taps spread over its
// footprint (a preview smaller than the picture does not stand one pixel for its neighbourhood; at a span of 1 or
// less, one tap, the nearest pixel); outside the picture it is the backdrop grey 0.09. clipWarn stripes diagonally
// (8 output pixels a period) a tap at white on any channel red and one black on all blue, in the picture view only.
// One output pixel per thread, rows top first when s->flipY.
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_develop_view(const PtDev* __restrict__ D, uchar4* __restrict__ out,
                                                                      const float* __restrict__ state, const PtDevelop* __restrict__ s,
                                                                      float4 view, unsigned int outW, unsigned int outH,
                                                                      unsigned int box, unsigned int clipWarn) {
    const unsigned int j = blockIdx.x * PT_BLOCK + threadIdx.x;
    if (j >= outW * outH) return;
    const unsigned int W = D->u.width, H = D->u.height;
    const unsigned int ox = j % outW, oy = j / outW;                     // oy counted from the bottom, as the picture
    const float fx = (float)ox + 0.5f, fy = (float)oy + 0.5f;
    const float cx = view.x + (fx - 0.5f * (float)outW) * view.z, cy = view.y + (fy - 0.5f * (float)outH) * view.w;
    V3 sum = v3s(0.0f);
    if (cx < 0.0f || cy < 0.0f || cx >= (float)W || cy >= (float)H) {
        sum = v3s(0.09f);
    } else {
        const float scale = developScale(state, s);
        const bool stripe = (fx + fy) / 8.0f + 0.0625f - floorf((fx + fy) / 8.0f + 0.0625f) < 0.5f;
        const float n = (float)box;
        for (unsigned int b = 0; b < box * box; ++b) {
            const float qx = cx - 0.5f * view.z + ((float)(b % box) + 0.5f) / n * view.z;
            const float qy = cy - 0.5f * view.w + ((float)(b / box) + 0.5f) / n * view.w;
            const unsigned int px = (unsigned int)fminf(fmaxf(qx, 0.0f), (float)W - 1.0f);
            const unsigned int py = (unsigned int)fminf(fmaxf(qy, 0.0f), (float)H - 1.0f);
            bool sampled;
            V3 c = developPixel(accumMean(D, py * W + px, sampled), scale, px, py, W, H, s);
            if (clipWarn && s->measureMode == 0u && stripe) {
                const float m = fmaxf(c.x, fmaxf(c.y, c.z));
                if (m >= 254.5f / 255.0f) c = v3(1.0f, 0.0f, 0.0f);
                else if (m <= 0.5f / 255.0f) c = v3(0.0f, 0.25f, 1.0f);
            }
            sum = sum + c;
        }
        sum = sum * (1.0f / (n * n));
    }
    const unsigned int o = (s->flipY ? (outH - 1u - oy) : oy) * outW + ox;
    out[o] = make_uchar4((unsigned char)fminf(sum.x * 255.0f + 0.5f, 255.0f), (unsigned char)fminf(sum.y * 255.0f + 0.5f, 255.0f),
                         (unsigned char)fminf(sum.z * 255.0f + 0.5f, 255.0f), 255);
}

// THE DEVELOPED PICTURE'S HISTOGRAM (rendertoscreen.js develop_histogram_prog): every `step`-th pixel each way developed
// as pt_develop does it and counted, 256 bins each of R, G, B and the Rec.709 luma of the display values (bin
// round(v 255)). A block's 1024 bins in LDS, one atomic a non-empty bin to `hist`; every thread reaches both barriers.
extern "C" __global__ __launch_bounds__(PT_BLOCK) void pt_develop_hist(const PtDev* __restrict__ D, unsigned int* __restrict__ hist,
                                                                      const float* __restrict__ state,
                                                                      const PtDevelop* __restrict__ s, unsigned int step) {
    __shared__ unsigned int h[4 * PT_HIST_BINS];
    for (unsigned int k = threadIdx.x; k < 4u * PT_HIST_BINS; k += PT_BLOCK) h[k] = 0u;
    __syncthreads();
    const unsigned int W = D->u.width, H = D->u.height;
    const unsigned int nx = (W + step - 1u) / step, ny = (H + step - 1u) / step;
    const float scale = developScale(state, s);
    for (unsigned int j = blockIdx.x * PT_BLOCK + threadIdx.x; j < nx * ny; j += gridDim.x * PT_BLOCK) {
        const unsigned int x = (j % nx) * step, y = (j / nx) * step;
        bool sampled;
        const V3 c = developPixel(accumMean(D, y * W + x, sampled), scale, x, y, W, H, s);
        const float bins = (float)(PT_HIST_BINS - 1);
        atomicAdd(&h[(unsigned int)(c.x * bins + 0.5f)], 1u);
        atomicAdd(&h[PT_HIST_BINS + (unsigned int)(c.y * bins + 0.5f)], 1u);
        atomicAdd(&h[2 * PT_HIST_BINS + (unsigned int)(c.z * bins + 0.5f)], 1u);
        atomicAdd(&h[3 * PT_HIST_BINS + (unsigned int)(fminf(fmaxf(luma(c), 0.0f), 1.0f) * bins + 0.5f)], 1u);
    }
    __syncthreads();
    for (unsigned int k = threadIdx.x; k < 4u * PT_HIST_BINS; k += PT_BLOCK)
        if (h[k]) atomicAdd(&hist[k], h[k]);
}
