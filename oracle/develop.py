"""The developed picture in FP64 numpy -- test infrastructure only (tests/ use it as the checker of pt_develop; the
product never imports it).

Restates lib/pathtracer/core/autoexposure.js (prog_ae_histogram :54-89, prog_ae_update :91-153: the log2-luminance
histogram, its 10 %-90 % trimmed average, the 95 % highlight guard, the easing), core/photometry.js :83-96 (the
exposure: white at 1.2 2^EV cd/m^2, ISO 12232; the meter log2(L 100 / 12.5), ISO 2720) and core/rendertoscreen.js
(the tone curves :35-150 and the sRGB encoding of displayPixel :209-260), from the published definitions each curve
names: Khronos PBR Neutral, AgX in Wrensch's polynomial approximation (with Blender's Punchy look), Hill's fitted ACES,
Reinhard on luminance, Hable's Uncharted 2 curve, IEC 61966-2-1.
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


def develop(acc, tone=1, ev=None, headroom=2.0):
    """acc [h, w, 4] FP64 accumulation (row 0 at the bottom): (8-bit RGB [h, w, 3] top row first, before rounding as
    floats too, the EV used). ev None: metered from the picture."""
    if ev is None:
        ev = ae_update(ae_histogram(acc), headroom)[0]
    c = np.maximum(acc[..., :3], 0.0) / (EV_CLIP * 2.0 ** ev)
    v = srgb(TONES[tone](c)) * 255.0
    v = np.where((acc[..., 3] > 0)[..., None], v, 0.0)
    return np.floor(np.minimum(v + 0.5, 255.0))[::-1].astype(np.uint8), v[::-1], ev
