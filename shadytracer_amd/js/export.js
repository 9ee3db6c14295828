'use strict'
/*
 * Light files: the engine's FP64 accumulation written as numbers (linear Rec.709, cd/m^2 or scaled), in the three
 * formats the reference's darkroom offers for light (lib/pathtracer/darkroom/export/index.js FORMATS:16): PFM, Radiance
 * HDR and OpenEXR. Each writer takes the picture in slices of rows, top row first, and ends in one Buffer:
 *     const W = makePFM({ w, h }); W.addRows(rgb, k); ...; const bytes = W.finish()
 * (the reference's writer interface; its finish returns a Blob). Written from the formats' own descriptions:
 *   PFM      Poskanzer's portable float map: "PF", the size, a negative scale for little-endian, then float32 RGB rows
 *            from the bottom of the picture up.
 *   HDR      Greg Ward's Radiance picture: RGBE pixels (three 8-bit mantissas sharing one exponent byte), each row of
 *            8..32767 pixels in the adaptive run-length form -- a 4-byte row mark, then each of the four byte planes
 *            as runs (count + 128, then the byte) and literal stretches (count, then the bytes) -- other widths flat.
 *   OpenEXR  one part of scanlines; channels B, G, R (the format stores them in name order) as float32 or half; blocks
 *            of one row uncompressed or sixteen rows ZIP-compressed (the format's predictor: even bytes before odd,
 *            each stored as its difference from the one before plus 128, then zlib; a block that does not shrink is
 *            kept raw), Rec.709/D65 chromaticities and a whiteLuminance.
 * lightRows() turns the accumulation (bottom row first, r g b n a pixel) into such rows.
 */
const zlib = require('zlib')

const RADIANCE_EFFICACY = 179          // Radiance's radiometric unit: one W/sr/m^2 is 179 cd/m^2
const REC709_D65 = [0.64, 0.33, 0.30, 0.60, 0.15, 0.06, 0.3127, 0.3290]

// Counts the rows a writer is given and refuses more than the picture holds, or fewer at the end.
function rowCounter(format, h) {
    let seen = 0
    return {
        take(k) {
            if (!(k >= 0) || seen + k > h) throw new Error(`export ${format}: rows ${seen + 1}..${seen + k} given, the picture is ${h} tall`)
            const first = seen
            seen += k
            return first
        },
        done() { if (seen !== h) throw new Error(`export ${format}: finished after ${seen} of ${h} rows`) },
    }
}

// ---- PFM: float32 little-endian, bottom row first
function makePFM({ w, h }) {
    const rows = rowCounter('pfm', h)
    const body = new Float32Array(w * h * 3)
    return {
        addRows(rgb, k) {
            const top = rows.take(k)
            for (let r = 0; r < k; r++) body.set(rgb.subarray ? rgb.subarray(r * w * 3, (r + 1) * w * 3) : rgb.slice(r * w * 3, (r + 1) * w * 3), (h - 1 - (top + r)) * w * 3)
        },
        finish() {
            rows.done()
            const le = new Uint8Array(body.buffer)
            const head = Buffer.from(`PF\n${w} ${h}\n-1.0\n`, 'latin1')
            // the platform's Float32Array is little-endian on every host node runs on here; written byte by byte otherwise
            if (new Uint8Array(new Float32Array([1]).buffer)[3] === 0x3f) return Buffer.concat([head, Buffer.from(le)])
            const out = Buffer.alloc(body.length * 4)
            body.forEach((v, i) => out.writeFloatLE(v, 4 * i))
            return Buffer.concat([head, out])
        },
    }
}

// ---- Radiance HDR
// frexp by the float64's own exponent field: v = m 2^e with m in [0.5, 1) for a normal positive v
const f64 = new Float64Array(1), f64w = new Uint32Array(f64.buffer)
function binaryExponent(v) {
    f64[0] = v
    return ((f64w[1] >>> 20) & 0x7ff) - 1022
}
/** One pixel as RGBE: e the largest channel's binary exponent, each channel's mantissa byte floor(c 256 / 2^e); black
 *  below 1e-32 (the format's own floor). Written into out at o (a fresh 4-byte Buffer by default). */
function rgbe(r, g, b, out = Buffer.alloc(4), o = 0) {
    const top = Math.max(r, g, b)
    if (!(top >= 1e-32)) { out.fill(0, o, o + 4); return out }
    const e = binaryExponent(top)
    const scale = 2 ** (8 - e)
    const byte = c => c > 0 ? Math.min(255, Math.floor(c * scale)) : 0
    out[o] = byte(r); out[o + 1] = byte(g); out[o + 2] = byte(b); out[o + 3] = e + 128
    return out
}
// One byte plane of a row in the run-length form. A stretch of equal bytes worth a run (4 or more, at most 127 a run)
// is sent as a run; what lies between runs goes as literal packets of at most 128 bytes.
function encodePlane(plane, out) {
    const n = plane.length
    let lit = 0                     // where the pending literals start
    const literals = upTo => {
        for (let s = lit; s < upTo; s += 128) {
            const m = Math.min(128, upTo - s)
            out.push(m)
            for (let j = 0; j < m; j++) out.push(plane[s + j])
        }
    }
    let i = 0
    while (i < n) {
        let j = i + 1
        while (j < n && j - i < 127 && plane[j] === plane[i]) j++
        if (j - i >= 4) {
            literals(i)
            out.push(128 + (j - i), plane[i])
            lit = j
        }
        i = j
    }
    literals(n)
}
function makeHDR({ w, h, lines = [] }) {
    const rows = rowCounter('hdr', h)
    const head = ['#?RADIANCE', 'FORMAT=32-bit_rle_rgbe', ...lines, '', `-Y ${h} +X ${w}`, ''].join('\n')
    const chunks = [Buffer.from(head, 'latin1')]
    const encoded = w >= 8 && w <= 0x7fff
    const px = Buffer.alloc(4 * w), plane = Buffer.alloc(w)
    return {
        addRows(rgb, k) {
            rows.take(k)
            for (let r = 0; r < k; r++) {
                for (let x = 0; x < w; x++) { const s = 3 * (r * w + x); rgbe(rgb[s], rgb[s + 1], rgb[s + 2], px, 4 * x) }
                if (!encoded) { chunks.push(Buffer.from(px)); continue }
                const out = [2, 2, w >> 8, w & 0xff]
                for (let c = 0; c < 4; c++) {
                    for (let x = 0; x < w; x++) plane[x] = px[4 * x + c]
                    encodePlane(plane, out)
                }
                chunks.push(Buffer.from(out))
            }
        },
        finish() { rows.done(); return Buffer.concat(chunks) },
    }
}

// ---- OpenEXR
// half from float32, round to nearest even: the value scaled so that its leading bit sits at 2^10 (or, below the
// smallest normal half, in units of the smallest subnormal 2^-24) and rounded; a carry past 2^11 moves the exponent.
const f32 = new Float32Array(1)
function roundEven(x) {
    const f = Math.floor(x), r = x - f
    return r > 0.5 || (r === 0.5 && f % 2 === 1) ? f + 1 : f
}
function toHalf(v) {
    if (Number.isNaN(v)) return 0x7e00
    f32[0] = v
    const x = f32[0]
    const sign = x < 0 || Object.is(x, -0) ? 0x8000 : 0
    const a = Math.abs(x)
    if (a === Infinity) return sign | 0x7c00
    if (a === 0) return sign
    const e = binaryExponent(a) - 1                // a = 1.f 2^e
    if (e < -14) return sign | roundEven(a * 2 ** 24)   // subnormal (a round up to 1024 is the smallest normal's bits)
    let m = roundEven(a * 2 ** (10 - e)), ex = e
    if (m === 2048) { m = 1024; ex++ }
    if (ex > 15) return sign | 0x7c00
    return sign | ((ex + 15) << 10) | (m - 1024)
}
/** The ZIP block predictor: the block's even-indexed bytes, then its odd-indexed ones; then, after the first, each byte
 *  as its difference from the previous one plus 128, mod 256. */
function zipPrepass(raw) {
    const n = raw.length, evens = (n + 1) >> 1, t = Buffer.alloc(n)
    for (let i = 0; i < n; i++) t[(i & 1) ? evens + (i >> 1) : i >> 1] = raw[i]
    let prev = n ? t[0] : 0
    for (let i = 1; i < n; i++) { const cur = t[i]; t[i] = (cur - prev + 128) & 0xff; prev = cur }
    return t
}
// The header: magic and version, then attributes (name, type, size, value), then an empty name.
function exrHeader({ w, h, half, compression, attributes }) {
    const ints = xs => { const b = Buffer.alloc(4 * xs.length); xs.forEach((v, i) => b.writeInt32LE(v, 4 * i)); return b }
    const floats = xs => { const b = Buffer.alloc(4 * xs.length); xs.forEach((v, i) => b.writeFloatLE(v, 4 * i)); return b }
    const list = [
        ['channels', 'chlist', Buffer.concat([...['B', 'G', 'R'].map(c => Buffer.concat([Buffer.from(c + '\0', 'latin1'), ints([half ? 1 : 2, 0, 1, 1])])), Buffer.from([0])])],
        ['compression', 'compression', Buffer.from([compression === 'zip' ? 3 : 0])],
        ['dataWindow', 'box2i', ints([0, 0, w - 1, h - 1])],
        ['displayWindow', 'box2i', ints([0, 0, w - 1, h - 1])],
        ['lineOrder', 'lineOrder', Buffer.from([0])],
        ['pixelAspectRatio', 'float', floats([1])],
        ['screenWindowCenter', 'v2f', floats([0, 0])],
        ['screenWindowWidth', 'float', floats([1])],
        ['chromaticities', 'chromaticities', floats(REC709_D65)],
        ['whiteLuminance', 'float', floats([attributes.whiteLuminance !== undefined ? attributes.whiteLuminance : 1])],
        ...Object.entries(attributes.strings || {}).map(([k, v]) => [k, 'string', Buffer.from(String(v), 'utf8')]),
    ]
    const parts = [ints([20000630, 2])]
    for (const [name, type, value] of list) parts.push(Buffer.from(`${name}\0${type}\0`, 'latin1'), ints([value.length]), value)
    parts.push(Buffer.from([0]))
    return Buffer.concat(parts)
}
function makeEXR({ w, h, half = false, compression = 'zip', attributes = {} }) {
    if (compression !== 'zip' && compression !== 'none') throw new Error(`export exr: compression must be 'zip' or 'none', not '${compression}'`)
    const rows = rowCounter('exr', h)
    const perBlock = compression === 'zip' ? 16 : 1, bytes = half ? 2 : 4
    const blocks = []
    let block = [], blockY = 0
    const seal = () => {
        if (!block.length) return
        const raw = Buffer.alloc(block.length * w * 3 * bytes)
        let o = 0
        for (const row of block) for (const c of [2, 1, 0]) for (let x = 0; x < w; x++, o += bytes) {
            if (half) raw.writeUInt16LE(toHalf(row[3 * x + c]), o); else raw.writeFloatLE(row[3 * x + c], o)
        }
        let data = raw
        if (compression === 'zip') { const z = zlib.deflateSync(zipPrepass(raw)); if (z.length < raw.length) data = z }
        const head = Buffer.alloc(8)
        head.writeInt32LE(blockY, 0)
        head.writeInt32LE(data.length, 4)
        blocks.push(head, data)
        blockY += block.length
        block = []
    }
    return {
        addRows(rgb, k) {
            rows.take(k)
            for (let r = 0; r < k; r++) {
                block.push(rgb.slice(r * w * 3, (r + 1) * w * 3))
                if (block.length === perBlock) seal()
            }
        },
        finish() {
            seal()
            rows.done()
            const head = exrHeader({ w, h, half, compression, attributes })
            const nBlocks = blocks.length / 2, offsets = Buffer.alloc(8 * nBlocks)
            let at = head.length + offsets.length
            for (let b = 0; b < nBlocks; b++) {
                offsets.writeBigUInt64LE(BigInt(at), 8 * b)
                at += blocks[2 * b].length + blocks[2 * b + 1].length
            }
            return Buffer.concat([head, offsets, ...blocks])
        },
    }
}

/** The engine's accumulation (Float64, r g b n a pixel, bottom row first) as rows of w x 3 Float32, top row first,
 *  times k (1 for cd/m^2, 1 / 179 for Radiance's units, an exposure's scale for exposed values); a pixel no sample
 *  reached is 0. clip: negatives (colours outside Rec.709) to 0, as RGBE must. */
function lightRows(acc, w, h, { k = 1, clip = false } = {}) {
    const out = new Float32Array(w * h * 3)
    for (let y = 0; y < h; y++) {
        const src = (h - 1 - y) * w
        for (let x = 0; x < w; x++) {
            const i = 4 * (src + x), o = 3 * (y * w + x), seen = acc[i + 3] > 0
            for (let c = 0; c < 3; c++) {
                const v = seen ? acc[i + c] * k : 0
                out[o + c] = clip && v < 0 ? 0 : v
            }
        }
    }
    return out
}

/** One call: the accumulation written as 'pfm' | 'hdr' | 'exr' (absolute values; HDR in Radiance's units). */
function writeLight(format, acc, w, h, opts = {}) {
    const k = opts.k || 1
    let W
    if (format === 'pfm') W = makePFM({ w, h })
    else if (format === 'hdr') W = makeHDR({ w, h, lines: ['PRIMARIES=0.640 0.330 0.300 0.600 0.150 0.060 0.3127 0.3290', ...(opts.lines || [])] })
    else if (format === 'exr') W = makeEXR({ w, h, half: !!opts.half, compression: opts.compression || 'zip', attributes: { whiteLuminance: 1 / k, strings: opts.strings || {} } })
    else throw new Error(`export: '${format}' is not a light format (pfm, hdr, exr)`)
    W.addRows(format === 'hdr' ? lightRows(acc, w, h, { k: k / RADIANCE_EFFICACY, clip: true }) : lightRows(acc, w, h, opts), h)
    return W.finish()
}

module.exports = { makePFM, makeHDR, makeEXR, rgbe, toHalf, zipPrepass, lightRows, writeLight, RADIANCE_EFFICACY }
