'use strict'
/*
 * The light files: the accumulated render written as numbers (linear Rec.709, cd/m^2 or exposed), the reference's
 * darkroom "light" formats (lib/pathtracer/darkroom/export/index.js FORMATS:16, pfm.js, hdr.js, exr.js) -- PFM, Radiance
 * HDR and OpenEXR -- with the same writer interface: make*({ w, h, ... }), addRows(rgb, k) with k rows of w x 3 floats
 * TOP FIRST, finish() -> a Buffer (the reference returns a Blob). The formats are the published ones:
 *   PFM       "PF\n<w> <h>\n-1.0\n", little-endian float32 rows, bottom row first (the format's order).
 *   HDR       Radiance RGBE: one shared exponent a pixel; rows 8..32767 wide in the 1991 run-length form (each
 *             component's runs, 128 + n then the byte, and copies, n then n bytes), other widths flat; "-Y h +X w".
 *   OpenEXR   single part, scanlines, channels B G R (sorted), float32 or half, compression NONE (a row a block) or ZIP
 *             (16 rows a block: the rows' bytes split even/odd, each byte as its difference from the last + 128, then
 *             zlib; raw where zlib does not shrink them), Rec.709 / D65 chromaticities and whiteLuminance.
 * lightRows() turns the engine's FP64 accumulation (bottom row first, r g b n per pixel) into these rows.
 */
const zlib = require('zlib')

const RADIANCE_EFFICACY = 179          // Radiance's units: W/sr/m^2 at 179 lm/W
const REC709_D65 = [0.64, 0.33, 0.30, 0.60, 0.15, 0.06, 0.3127, 0.3290]

// ---- PFM
function makePFM({ w, h }) {
    const rows = []
    return {
        addRows(rgb, k) {
            if (rows.length + k > h) throw new Error(`[pfm] ${rows.length + k} rows for a picture of ${h}`)
            for (let r = 0; r < k; r++) {
                const b = Buffer.alloc(w * 12)
                for (let i = 0; i < w * 3; i++) b.writeFloatLE(rgb[r * w * 3 + i], 4 * i)
                rows.push(b)
            }
        },
        finish() {
            if (rows.length !== h) throw new Error(`[pfm] ${rows.length} rows written of ${h}`)
            return Buffer.concat([Buffer.from(`PF\n${w} ${h}\n-1.0\n`, 'latin1'), ...rows.slice().reverse()])
        },
    }
}

// ---- Radiance HDR
/** One pixel's RGBE: the largest channel's binary exponent shared, each mantissa a byte (truncated); 0 below 1e-32.
 *  The exponent is frexp's (the largest channel's mantissa in [0.5, 1)) even where log2 rounds across an integer. */
function rgbe(r, g, b, out = Buffer.alloc(4), o = 0) {
    const v = Math.max(r, g, b)
    if (!(v >= 1e-32)) { out[o] = out[o + 1] = out[o + 2] = out[o + 3] = 0; return out }
    let e = Math.floor(Math.log2(v)) + 1
    if (v / 2 ** e >= 1) e++                       // log2 rounded up at an exact power of two
    else if (v / 2 ** e < 0.5) e--                 // ...or down just below one
    const s = 256 / 2 ** e
    const q = x => Math.min(255, Math.max(0, Math.floor(x * s)))
    out[o] = q(r); out[o + 1] = q(g); out[o + 2] = q(b); out[o + 3] = e + 128
    return out
}
// one component of a row in the run-length form: a run where 4 or more bytes repeat, literal copies of up to 128
// bytes between runs
function rleComponent(data, out) {
    const w = data.length
    let i = 0
    while (i < w) {
        // the next run of 4 or more from i on
        let runAt = i, runLen = 0
        while (runAt < w) {
            runLen = 1
            while (runAt + runLen < w && runLen < 127 && data[runAt + runLen] === data[runAt]) runLen++
            if (runLen >= 4) break
            runAt += runLen
        }
        if (runAt >= w) runLen = 0
        while (i < runAt) {                        // the literals before it
            const n = Math.min(128, runAt - i)
            out.push(n)
            for (let k = 0; k < n; k++) out.push(data[i + k])
            i += n
        }
        if (runLen >= 4) { out.push(128 + runLen, data[runAt]); i = runAt + runLen }
    }
}
function makeHDR({ w, h, lines = [] }) {
    const parts = [Buffer.from(`#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n${lines.map(l => l + '\n').join('')}\n-Y ${h} +X ${w}\n`, 'latin1')]
    const flat = w < 8 || w > 0x7fff
    let rows = 0
    return {
        addRows(rgb, k) {
            if (rows + k > h) throw new Error(`[hdr] ${rows + k} rows for a picture of ${h}`)
            const px = Buffer.alloc(w * 4)
            for (let r = 0; r < k; r++) {
                for (let x = 0; x < w; x++) { const s = (r * w + x) * 3; rgbe(rgb[s], rgb[s + 1], rgb[s + 2], px, 4 * x) }
                if (flat) { parts.push(Buffer.from(px)); continue }
                const out = [2, 2, (w >> 8) & 0xff, w & 0xff]
                const comp = Buffer.alloc(w)
                for (let c = 0; c < 4; c++) { for (let x = 0; x < w; x++) comp[x] = px[4 * x + c]; rleComponent(comp, out) }
                parts.push(Buffer.from(out))
            }
            rows += k
        },
        finish() {
            if (rows !== h) throw new Error(`[hdr] ${rows} rows written of ${h}`)
            return Buffer.concat(parts)
        },
    }
}

// ---- OpenEXR
const f32a = new Float32Array(1), u32a = new Uint32Array(f32a.buffer)
/** float32 to half bits, to the nearest even; past 65504 infinity, under the smallest subnormal 0, NaN kept. */
function toHalf(v) {
    f32a[0] = v
    const x = u32a[0], sign = (x >>> 16) & 0x8000, e = (x >>> 23) & 0xff, m = x & 0x7fffff
    if (e === 0xff) return sign | 0x7c00 | (m ? 0x200 : 0)
    const E = e - 112                               // the half's biased exponent
    if (E >= 0x1f) return sign | 0x7c00
    if (E <= 0) {                                   // a subnormal half, or 0
        if (E < -10) return sign
        const mm = m | 0x800000, shift = 14 - E
        let hm = mm >>> shift
        const rest = mm & ((1 << shift) - 1), half = 1 << (shift - 1)
        if (rest > half || (rest === half && (hm & 1))) hm++
        return sign | hm
    }
    let hv = (E << 10) | (m >>> 13)
    const rest = m & 0x1fff
    if (rest > 0x1000 || (rest === 0x1000 && (hv & 1))) hv++      // a carry into the exponent is still right
    return sign | hv
}
/** ZIP's pre-pass: the bytes split into the even-indexed then the odd-indexed, then each as (its value minus the last
 *  one's + 128) mod 256, the first kept. */
function zipPrepass(raw) {
    const n = raw.length, t = Buffer.alloc(n), half = (n + 1) >> 1
    for (let i = 0, a = 0, b = half; i < n; i += 2) { t[a++] = raw[i]; if (i + 1 < n) t[b++] = raw[i + 1] }
    for (let i = n - 1; i > 0; i--) t[i] = (t[i] - t[i - 1] + 128) & 0xff
    return t
}
function exrHeader({ w, h, half, compression, attributes }) {
    const parts = []
    const attr = (name, type, value) => {
        const size = Buffer.alloc(4); size.writeInt32LE(value.length)
        parts.push(Buffer.from(`${name}\0${type}\0`, 'latin1'), size, value)
    }
    const i32 = xs => { const b = Buffer.alloc(4 * xs.length); xs.forEach((x, i) => b.writeInt32LE(x, 4 * i)); return b }
    const f32 = xs => { const b = Buffer.alloc(4 * xs.length); xs.forEach((x, i) => b.writeFloatLE(x, 4 * i)); return b }
    const ch = []
    for (const c of ['B', 'G', 'R']) ch.push(Buffer.from(c + '\0', 'latin1'), i32([half ? 1 : 2]), Buffer.from([0, 0, 0, 0]), i32([1, 1]))
    ch.push(Buffer.from([0]))
    attr('channels', 'chlist', Buffer.concat(ch))
    attr('compression', 'compression', Buffer.from([compression === 'zip' ? 3 : 0]))
    attr('dataWindow', 'box2i', i32([0, 0, w - 1, h - 1]))
    attr('displayWindow', 'box2i', i32([0, 0, w - 1, h - 1]))
    attr('lineOrder', 'lineOrder', Buffer.from([0]))
    attr('pixelAspectRatio', 'float', f32([1]))
    attr('screenWindowCenter', 'v2f', f32([0, 0]))
    attr('screenWindowWidth', 'float', f32([1]))
    attr('chromaticities', 'chromaticities', f32(REC709_D65))
    attr('whiteLuminance', 'float', f32([attributes.whiteLuminance !== undefined ? attributes.whiteLuminance : 1]))
    for (const [k, v] of Object.entries(attributes.strings || {})) attr(k, 'string', Buffer.from(String(v), 'utf8'))
    parts.push(Buffer.from([0]))
    const magic = Buffer.alloc(8); magic.writeInt32LE(20000630, 0); magic.writeInt32LE(2, 4)
    return Buffer.concat([magic, ...parts])
}
function makeEXR({ w, h, half = false, compression = 'zip', attributes = {} }) {
    if (compression !== 'zip' && compression !== 'none') throw new Error(`[exr] no compression '${compression}'`)
    const per = compression === 'zip' ? 16 : 1, bpv = half ? 2 : 4
    const blocks = [], pending = []
    let rows = 0
    const flush = () => {
        const k = pending.length
        const raw = Buffer.alloc(k * w * 3 * bpv)
        let o = 0
        for (const row of pending) {
            for (const c of [2, 1, 0]) {                    // B, G, R: the channels in name order
                for (let x = 0; x < w; x++, o += bpv) {
                    const v = row[3 * x + c]
                    if (half) raw.writeUInt16LE(toHalf(v), o); else raw.writeFloatLE(v, o)
                }
            }
        }
        let data = raw
        if (compression === 'zip') { const z = zlib.deflateSync(zipPrepass(raw)); if (z.length < raw.length) data = z }
        const head = Buffer.alloc(8); head.writeInt32LE(rows - k, 0); head.writeInt32LE(data.length, 4)
        blocks.push(Buffer.concat([head, data]))
        pending.length = 0
    }
    return {
        addRows(rgb, k) {
            if (rows + k > h) throw new Error(`[exr] ${rows + k} rows for a picture of ${h}`)
            for (let r = 0; r < k; r++) {
                pending.push(rgb.slice(r * w * 3, (r + 1) * w * 3)); rows++
                if (pending.length === per) flush()
            }
        },
        finish() {
            if (pending.length) flush()
            if (rows !== h) throw new Error(`[exr] ${rows} rows written of ${h}`)
            const head = exrHeader({ w, h, half, compression, attributes })
            const table = Buffer.alloc(8 * blocks.length)
            let at = head.length + table.length
            blocks.forEach((b, i) => { table.writeBigUInt64LE(BigInt(at), 8 * i); at += b.length })
            return Buffer.concat([head, table, ...blocks])
        },
    }
}

/** The engine's accumulation (Float64, r g b n a pixel, bottom row first) as rows of w x 3 Float32, top row first,
 *  times k (1 for cd/m^2, 1 / 179 for Radiance's units, the exposure's scale for exposed values); a pixel no sample
 *  reached is 0. clip: negatives (colours outside Rec.709) to 0, as RGBE must. */
function lightRows(acc, w, h, { k = 1, clip = false } = {}) {
    const out = new Float32Array(w * h * 3)
    for (let y = 0; y < h; y++) {
        const src = (h - 1 - y) * w
        for (let x = 0; x < w; x++) {
            const i = 4 * (src + x), o = 3 * (y * w + x)
            for (let c = 0; c < 3; c++) { const v = acc[i + 3] > 0 ? acc[i + c] * k : 0; out[o + c] = clip ? Math.max(v, 0) : v }
        }
    }
    return out
}

/** One call: the accumulation written as format 'pfm' | 'hdr' | 'exr' (absolute values; HDR in Radiance's units). */
function writeLight(format, acc, w, h, opts = {}) {
    if (format === 'pfm') { const P = makePFM({ w, h }); P.addRows(lightRows(acc, w, h, opts), h); return P.finish() }
    if (format === 'hdr') {
        const R = makeHDR({ w, h, lines: ['PRIMARIES=0.640 0.330 0.300 0.600 0.150 0.060 0.3127 0.3290', ...(opts.lines || [])] })
        R.addRows(lightRows(acc, w, h, { k: (opts.k || 1) / RADIANCE_EFFICACY, clip: true }), h)
        return R.finish()
    }
    if (format === 'exr') {
        const E = makeEXR({ w, h, half: !!opts.half, compression: opts.compression || 'zip', attributes: { whiteLuminance: 1 / (opts.k || 1), strings: opts.strings || {} } })
        E.addRows(lightRows(acc, w, h, opts), h)
        return E.finish()
    }
    throw new Error(`[export] no light format '${format}'`)
}

module.exports = { makePFM, makeHDR, makeEXR, rgbe, toHalf, zipPrepass, lightRows, writeLight, RADIANCE_EFFICACY }
