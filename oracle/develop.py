"""The developed picture in FP64 numpy -- test infrastructure only (tests/ use it as the checker of pt_develop; the
product never imports it).

Restates lib/pathtracer/core/autoexposure.js (prog_ae_histogram :54-89, prog_ae_update :91-153: the log2-luminance
histogram, its 10 %-90 % trimmed average, the 95 % highlight guard, the easing), core/photometry.js :83-96 (the
exposure: white at 1.2 2^EV cd/m^2, ISO 12232; the meter log2(L 100 / 12.5), ISO 2720) and core/rendertoscreen.js
(the tone curves :35-150 and the sRGB encoding of displayPixel :209-260), from the published definitions each curve
names: Khronos PBR Neutral, AgX in Wrensch's polynomial approximation (with Blender's Punchy look), Hill's fitted ACES,
Reinhard on luminance, Hable's Uncharted 2 curve, IEC 61966-2-1.

The darkroom (core/rendertoscreen.js developPixel :423-455 and its stages :286-421; darkroom/grade.js lightStages
:279-290, finishEarly :296-310, finishLate :311-320, vignetteStops :160-168, grainNoise :218-226, ditherAt :230-233;
core/colorimetry.js gamutClip :110-117; the false-colour views of displayPixel :224-241) restated from the
definitions they name: Oklab (Ottosson 2020), Hable's curve with its toe and shoulder given, Levels, the ASC CDL,
PCG's output permutation (Jarzynski and Olano 2020), Google's Turbo in its published polynomial fit. `grade` is the
dict of PtDevelop's darkroom words the host makes (lightOn, wb rows, eq, lab, tone7, transfer, levels, cdl*,
vignette, grain, curveLut, contrast, saturation, brightness, measure*).
"""
import numpy as np

LUMA = np.array([0.2126, 0.7152, 0.0722])
AE_BINS, AE_LOG_MIN, AE_LOG_MAX = 128, -10.0, 34.0
EV_CLIP, METER_K = 1.2, 12.5


def ae_histogram(acc):
    """acc: [h, w, 4] FP64 (mean rgb, count). The bins of log2 Y of the sampled pixels at or above 2^-10."""
    Y = (acc[..., :3] @ LUMA)[acc[..., 3] > 0]
    Y = Y[Y >= 2.0 ** AE_LOG_MIN]
    t = (np.log2(Y) - AE_LOG_MIN) / (AE_LOG_MAX - AE_LOG_MIN)
    b = np.clip(t * AE_BINS, 0, AE_BINS - 1).astype(int)
    return np.bincount(b, minlength=AE_BINS).astype(float)


def ae_update(hist, headroom=2.0, speed=1.0, state=None):
    """(ev, log2 average, log2 of the 95 % quantile) from a histogram; state: the last (ev) or None for the first."""
    total = hist.sum()
    if total == 0:
        return (state, None, None)
    lo, hi, q = 0.10 * total, 0.90 * total, 0.95 * total
    w = (AE_LOG_MAX - AE_LOG_MIN) / AE_BINS
    below = kept = acc = 0.0
    logq = AE_LOG_MAX
    have = False
    for k, c in enumerate(hist):
        a, b = max(below, lo), min(below + c, hi)
        if b > a:
            kept += b - a
            acc += (b - a) * (AE_LOG_MIN + (k + 0.5) * w)
        if not have and c > 0 and below + c >= q:
            logq = AE_LOG_MIN + (k + min(max((q - below) / c, 0.0), 1.0)) * w
            have = True
        below += c
    logl = acc / max(kept, 1e-9)
    target = max(np.log2(2.0 ** logl * 100.0 / METER_K), logq - np.log2(EV_CLIP) - max(headroom, 0.0))
    ev = target if state is None else state + min(max(speed, 0.0), 1.0) * (target - state)
    return ev, logl, logq


def tone_neutral(c):
    c = np.array(c, float)
    m = c.min(axis=-1, keepdims=True)
    c = c - np.where(m < 0.08, m - 6.25 * m * m, 0.04)
    peak = c.max(axis=-1, keepdims=True)
    d = 1 - 0.76
    newp = 1 - d * d / (peak + d - 0.76)
    g = 1 - 1 / (0.15 * (peak - newp) + 1)
    comp = c * (newp / np.where(peak > 0, peak, 1)) * (1 - g) + newp * g
    return np.where(peak < 0.76, c, comp)


_AGX_IN = np.array([[0.842479062253094, 0.0784335999999992, 0.0792237451477643],
                    [0.0423282422610123, 0.878468636469772, 0.0791661274605434],
                    [0.0423756549057051, 0.0784336, 0.879142973793104]])
_AGX_OUT = np.array([[1.19687900512017, -0.0980208811401368, -0.0990297440797205],
                     [-0.0528968517574562, 1.15190312990417, -0.0989611768448433],
                     [-0.0529716355144438, -0.0980434501171241, 1.15107367264116]])


def tone_agx(c, punchy=False):
    v = np.maximum(np.array(c, float), 1e-10) @ _AGX_IN.T
    v = (np.clip(np.log2(v), -12.47393, 4.026069) + 12.47393) / (4.026069 + 12.47393)
    x2 = v * v
    x4 = x2 * x2
    v = 15.5 * x4 * x2 - 40.14 * x4 * v + 31.96 * x4 - 6.868 * x2 * v + 0.4298 * x2 + 0.1191 * v - 0.00232
    if punchy:
        l = (v @ LUMA)[..., None]
        v = np.maximum(v, 0.0) ** 1.35
        v = l + (v - l) * 1.4
    v = v @ _AGX_OUT.T
    return np.maximum(v, 0.0) ** 2.2


_ACES_IN = np.array([[0.59719, 0.35458, 0.04823], [0.07600, 0.90834, 0.01566], [0.02840, 0.13383, 0.83777]])
_ACES_OUT = np.array([[1.60475, -0.53108, -0.07367], [-0.10208, 1.10813, -0.00605], [-0.00327, -0.07276, 1.07602]])


def tone_aces(c):
    v = np.array(c, float) @ _ACES_IN.T
    v = (v * (v + 0.0245786) - 0.000090537) / (v * (0.983729 * v + 0.4329510) + 0.238081)
    return np.clip(v @ _ACES_OUT.T, 0.0, 1.0)


def tone_reinhard(c):
    c = np.array(c, float)
    L = (c @ LUMA)[..., None]
    return np.where(L > 0, c / (1 + L), 0.0)


def _hable(x):
    A, B, C, D, E, F = 0.15, 0.50, 0.10, 0.20, 0.02, 0.30
    return (x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F) - E / F


def tone_hable(c):
    return _hable(2.0 * np.array(c, float)) / _hable(11.2)


TONES = {0: lambda c: np.array(c, float), 1: tone_neutral, 2: tone_agx, 3: lambda c: tone_agx(c, True), 4: tone_aces,
         5: tone_reinhard, 6: tone_hable}


def srgb(v):
    v = np.clip(v, 0.0, 1.0)
    return np.where(v <= 0.0031308, 12.92 * v, 1.055 * v ** (1 / 2.4) - 0.055)


LIGHT = dict(ANY=1, WB=2, EQ=4, LAB=8, LEVELS=16, CURVES=32, CDL=64, VIGNETTE=128, GRAIN=256, DITHER=512, CURVE_LUMA=1024)
GRADE_DEFAULTS = dict(lightOn=0, wb=[1, 0, 0, 0, 1, 0, 0, 0, 1], eq=[0] * 8, lab=[0, 0, 0], tone7=[5.6, 0.2, 0.15], transfer=0,
                      levels=[0, 1, 1, 0, 1], cdlSlope=[1, 1, 1], cdlOffset=[0, 0, 0], cdlPower=[1, 1, 1], vignette=[0, 0.5, 0],
                      grain=[0, 1.5], curveLut=None, contrast=0.0, saturation=0.0, brightness=0.0, measureMode=0,
                      measureMin=1.0, measureMax=100000.0, measureBands=0)


def gamut_clip(c):
    """A negative channel desaturated to 0 towards the grey of the colour's own luminance; no luminance: black."""
    c = np.array(c, float)
    lo = c.min(axis=-1, keepdims=True)
    Y = (c @ LUMA)[..., None]
    t = Y / np.where(Y - lo > 0, Y - lo, 1.0)
    clipped = np.where(Y > 0, np.maximum(Y + (c - Y) * t, 0.0), 0.0)
    return np.where(lo >= 0, c, clipped)


_OK_M1 = np.array([[0.4122214708, 0.5363325363, 0.0514459929], [0.2119034982, 0.6806995451, 0.1073969566],
                   [0.0883024619, 0.2817188376, 0.6299787005]])
_OK_M2 = np.array([[0.2104542553, 0.7936177850, -0.0040720468], [1.9779984951, -2.4285922050, 0.4505937099],
                   [0.0259040371, 0.7827717662, -0.8086757660]])


def oklab(c):
    """Ottosson's Oklab of linear sRGB, the cube root keeping its sign."""
    return np.cbrt(np.array(c, float) @ _OK_M1.T) @ _OK_M2.T


def from_oklab(q):
    """Oklab's inverse, from the inverses of its two matrices (the published 10-digit inverses agree to 1e-9)."""
    return (np.array(q, float) @ np.linalg.inv(_OK_M2).T) ** 3 @ np.linalg.inv(_OK_M1).T


def eq_gain(bands, x):
    """The equaliser's gain in stops at x stops below white: the 8 bands at -7..0, interpolated, held at the ends."""
    t = np.clip(np.asarray(x, float) + 7.0, 0.0, 7.0)
    k = np.minimum(np.floor(t), 6).astype(int)
    b = np.asarray(bands, float)
    return b[k] + (b[k + 1] - b[k]) * (t - k)


def lab_adjust(c, sat, vib, hue):
    q = oklab(c)
    rel = np.hypot(q[..., 1], q[..., 2]) / np.maximum(np.abs(q[..., 0]), 1e-6)
    k = (1 + sat) * (1 + vib * (1 - np.minimum(rel / 0.25, 1.0)))
    cs, sn = np.cos(hue), np.sin(hue)
    a = (q[..., 1] * cs - q[..., 2] * sn) * k
    b = (q[..., 1] * sn + q[..., 2] * cs) * k
    return from_oklab(np.stack([q[..., 0], a, b], axis=-1))


def hable_p(x, sh, to):
    """Hable's rational curve with shoulder strength sh and toe strength to (linear 0.50, angle 0.10, toe 0.02/0.30)."""
    return (x * (sh * x + 0.10 * 0.50) + to * 0.02) / (x * (sh * x + 0.50) + to * 0.30) - 0.02 / 0.30


def tone7(c, white, toe, shoulder):
    return hable_p(2.0 * np.asarray(c, float), shoulder, toe) / hable_p(2.0 * white, shoulder, toe)


def vignette_stops(px, py, w, h, amount, mid, rnd):
    qx, qy = ((px + 0.5) / w - 0.5) * 2, ((py + 0.5) / h - 0.5) * 2
    a = w / h
    r0 = np.hypot(qx, qy) / np.sqrt(2.0)
    r1 = np.hypot(qx * a, qy) / np.hypot(a, 1.0)
    r = r0 + (r1 - r0) * rnd
    t = np.clip((r - mid) / max(1e-6, 1 - mid), 0.0, 1.0)
    return amount * t * t * (3 - 2 * t)


def encode_out(x, transfer):
    c = np.clip(x, 0.0, 1.0)
    if transfer == 1:
        return c ** (1 / 2.2)
    if transfer == 2:
        return c ** (1 / 2.4)
    if transfer == 3:
        return c
    return srgb(c)


def levels(x, in_black, in_white, inv_gamma, out_black, out_white):
    t = np.clip((x - in_black) / max(1e-6, in_white - in_black), 0.0, 1.0)
    return out_black + (out_white - out_black) * t ** inv_gamma


def lut_at(lut, ch, x):
    n = len(lut) // 4
    t = np.clip(x, 0.0, 1.0) * (n - 1)
    k = np.minimum(np.floor(t), n - 2).astype(int)
    a, b = lut[ch * n + k], lut[ch * n + k + 1]
    return a + (b - a) * (t - k)


def cdl(x, slope, offset, power):
    return np.maximum(x * slope + offset, 0.0) ** power


def pcg_hash(n):
    """PCG's output permutation on 32 bits (uint64 arrays keep the products exact before the mask)."""
    M = np.uint64(0xFFFFFFFF)
    st = (np.asarray(n, np.uint64) * np.uint64(747796405) + np.uint64(2891336453)) & M
    w = (((st >> ((st >> np.uint64(28)) + np.uint64(4))) ^ st) * np.uint64(277803737)) & M
    return (w >> np.uint64(22)) ^ w


def unit24(h):
    return (np.asarray(h, np.uint64) >> np.uint64(8)).astype(float) / 16777216.0


def grain_noise(px, py, size):
    gx, gy = (px + 0.5) / size, (py + 0.5) / size
    ix, iy = np.floor(gx), np.floor(gy)
    fx, fy = gx - ix, gy - iy
    fx, fy = fx * fx * (3 - 2 * fx), fy * fy * (3 - 2 * fy)

    def cell(i, j):
        key = (np.asarray(i + j * 65536, np.int64) & 0xFFFFFFFF).astype(np.uint64) ^ np.uint64(0x632BE5AB)
        return unit24(pcg_hash(key)) * 2 - 1
    a, b, c, d = cell(ix, iy), cell(ix + 1, iy), cell(ix, iy + 1), cell(ix + 1, iy + 1)
    top = a + (b - a) * fx
    return top + ((c + (d - c) * fx) - top) * fy


def dither_at(px, py):
    key = (np.asarray(px + py * 65536, np.int64) & 0xFFFFFFFF).astype(np.uint64) ^ np.uint64(0x1B873593)
    h1 = pcg_hash(key)
    h2 = pcg_hash(h1)
    return (unit24(h1) + unit24(h2) - 1) / 255


_TURBO = np.array([[0.13572138, 4.61539260, -42.66032258, 132.13108234, -152.94239396, 59.28637943],
                   [0.09140261, 2.19418839, 4.84296658, -14.18503333, 4.27729857, 2.82956604],
                   [0.10667330, 12.64194608, -60.58204836, 110.36276771, -89.90310912, 27.34824973]])


def turbo(t):
    x = np.clip(np.asarray(t, float), 0.0, 1.0)
    return np.stack([x ** k for k in range(6)], axis=-1) @ _TURBO.T


def false_colour(light, u, v, g):
    Y = np.maximum(light @ LUMA, 0.0)
    val = np.pi * Y if g['measureMode'] == 1 else Y
    lo = np.log2(max(g['measureMin'], 1e-6))
    hi = max(np.log2(max(g['measureMax'], 1e-6)), lo + 1e-3)
    strip = v < 0.025
    t = np.where(strip, u, (np.log2(np.maximum(val, 1e-30)) - lo) / (hi - lo))
    if g['measureBands'] > 0:
        t = (np.floor(np.clip(t, 0.0, 0.99999) * g['measureBands']) + 0.5) / g['measureBands']
    shown = np.where(((val <= 0) & ~strip)[..., None], 0.0, turbo(t))
    return np.clip(shown, 0.0, 1.0)


def develop_pixels(light, scale, px, py, W, H, tone, grade=None):
    """The reference's developPixel on light [..., 3] (cd/m^2) at picture pixels (px, py), row 0 at the bottom: display
    values in [0, 1]."""
    g = dict(GRADE_DEFAULTS, **(grade or {}))
    if g['measureMode']:
        return false_colour(np.asarray(light, float), (px + 0.5) / W, (py + 0.5) / H, g)
    on = int(g['lightOn'])
    c = np.asarray(light, float) * scale
    if on & LIGHT['WB']:
        c = c @ np.asarray(g['wb'], float).reshape(3, 3).T
    if on & LIGHT['EQ']:
        c = c * 2.0 ** eq_gain(g['eq'], np.log2(np.maximum(c @ LUMA, 1e-30)))[..., None]
    if on & LIGHT['LAB']:
        c = lab_adjust(c, *g['lab'])
    if on & LIGHT['VIGNETTE']:
        c = c * 2.0 ** vignette_stops(px, py, W, H, *g['vignette'])[..., None]
    c = gamut_clip(c)
    c = tone7(c, *g['tone7']) if tone == 7 else TONES[tone](c)
    c = encode_out(np.clip(c, 0.0, 1.0), g['transfer'])
    if on & LIGHT['LEVELS']:
        c = levels(c, *g['levels'])
    if on & LIGHT['CURVES']:
        lut = np.asarray(g['curveLut'], float)
        if on & LIGHT['CURVE_LUMA']:
            y = c @ LUMA
            y2 = lut_at(lut, 0, y)
            c = np.where((y > 1e-6)[..., None], c * (y2 / np.where(y > 1e-6, y, 1.0))[..., None], y2[..., None])
        else:
            c = lut_at(lut, 0, c)
        c = np.stack([lut_at(lut, k + 1, c[..., k]) for k in range(3)], axis=-1)
    if on & LIGHT['CDL']:
        c = cdl(c, np.asarray(g['cdlSlope']), np.asarray(g['cdlOffset']), np.asarray(g['cdlPower']))
    c = 0.5 + (c - 0.5) * (g['contrast'] + 1)
    y = (c @ LUMA)[..., None]
    c = y + (c - y) * (g['saturation'] + 1)
    c = c + g['brightness']
    if on & LIGHT['GRAIN']:
        y = np.clip(c @ LUMA, 0.0, 1.0)
        c = c + (grain_noise(px, py, g['grain'][1]) * g['grain'][0] * 0.08 * (0.3 + 2.8 * y * (1 - y)))[..., None]
    if on & LIGHT['DITHER']:
        c = c + dither_at(px, py)[..., None]
    return np.clip(c, 0.0, 1.0)


def develop(acc, tone=1, ev=None, headroom=2.0, grade=None):
    """acc [h, w, 4] FP64 accumulation (row 0 at the bottom): (8-bit RGB [h, w, 3] top row first, the display values
    times 255 before rounding, top row first, the EV used). ev None: metered from the picture."""
    if ev is None:
        ev = ae_update(ae_histogram(acc), headroom)[0]
    H, W = acc.shape[:2]
    py, px = np.mgrid[0:H, 0:W].astype(float)
    v = develop_pixels(acc[..., :3], 1.0 / (EV_CLIP * 2.0 ** ev), px, py, W, H, tone, grade) * 255.0
    return np.floor(np.minimum(v + 0.5, 255.0))[::-1].astype(np.uint8), v[::-1], ev


def histogram(acc, tone=1, ev=None, headroom=2.0, grade=None, step=1):
    """What develop() writes, counted over every `step`-th pixel each way: [4, 256] (R, G, B, luma), bin round(v 255)."""
    if ev is None:
        ev = ae_update(ae_histogram(acc), headroom)[0]
    sub = acc[::step, ::step]
    H, W = acc.shape[:2]
    py, px = np.mgrid[0:H:step, 0:W:step].astype(float)
    c = develop_pixels(sub[..., :3], 1.0 / (EV_CLIP * 2.0 ** ev), px, py, W, H, tone, grade)
    chans = [c[..., 0], c[..., 1], c[..., 2], np.clip(c @ LUMA, 0.0, 1.0)]
    return np.stack([np.bincount((np.clip(x, 0, 1) * 255 + 0.5).astype(int).ravel(), minlength=256) for x in chans])


def develop_view(acc, view, out_w, out_h, tone=1, ev=None, headroom=2.0, grade=None, clip_warn=False):
    """rendertoscreen.js develop_prog's frag as pt_develop_view runs it: an out_h x out_w view (8-bit RGB, the top row
    first, and the values x 255 before rounding) of the picture through view = (cx, cy, sx, sy), picture pixels, y up;
    each output pixel the mean of box x box taps over its footprint (box = ceil(max(sx, sy)), 1..8), 0.09 outside the
    picture, clip stripes (red at white, blue at black) on the diagonal bands when clip_warn."""
    if ev is None:
        ev = ae_update(ae_histogram(acc), headroom)[0]
    g = dict(GRADE_DEFAULTS, **(grade or {}))
    H, W = acc.shape[:2]
    py, px = np.mgrid[0:H, 0:W].astype(float)
    dev = develop_pixels(acc[..., :3], 1.0 / (EV_CLIP * 2.0 ** ev), px, py, W, H, tone, g)
    cx0, cy0, sx, sy = [float(v) for v in view]
    box = int(min(8, max(1, np.ceil(max(sx, sy) - 1e-9))))
    oy, ox = np.mgrid[0:out_h, 0:out_w].astype(float)
    fx, fy = ox + 0.5, oy + 0.5
    cx, cy = cx0 + (fx - 0.5 * out_w) * sx, cy0 + (fy - 0.5 * out_h) * sy
    inside = (cx >= 0) & (cy >= 0) & (cx < W) & (cy < H)
    stripe = np.mod((fx + fy) / 8.0 + 0.0625, 1.0) < 0.5
    total = np.zeros((out_h, out_w, 3))
    for b in range(box * box):
        qx = cx - 0.5 * sx + ((b % box) + 0.5) / box * sx
        qy = cy - 0.5 * sy + ((b // box) + 0.5) / box * sy
        ix = np.clip(qx, 0, W - 1).astype(int)
        iy = np.clip(qy, 0, H - 1).astype(int)
        c = dev[iy, ix]
        if clip_warn and not g['measureMode']:
            m = c.max(axis=-1)
            c = np.where((stripe & (m >= 254.5 / 255))[..., None], [1.0, 0.0, 0.0], c)
            c = np.where((stripe & (m < 254.5 / 255) & (m <= 0.5 / 255))[..., None], [0.0, 0.25, 1.0], c)
        total += c
    v = np.where(inside[..., None], total / (box * box), 0.09) * 255.0
    return np.floor(np.minimum(v + 0.5, 255.0))[::-1].astype(np.uint8), v[::-1]
