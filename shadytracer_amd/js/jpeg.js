'use strict'
/*
 * JPEG files of a developed picture (the darkroom's 'jpeg' format, darkroom/export/index.js FORMATS; the reference
 * hands its canvas to the browser's encoder, which Node does not have). A baseline sequential JFIF file written from
 * the standard (ITU-T T.81, JFIF 1.02): RGB to full-range YCbCr (T.871), chroma averaged 2 x 2 (4:2:0, what browsers
 * write), 8 x 8 blocks level-shifted by 128, the orthonormal 2-D DCT (separable, from its definition), quantised by the
 * example tables of T.81 Annex K.1 scaled to the quality as the IJG library scales them (quality q in 0..1 is
 * round(100 q): below 50 the tables times 50 / Q, above it times 2 - Q / 50), and Huffman-coded with the example
 * tables of Annex K.3 (DC differences by size category, AC run/size pairs with ZRL and EOB), 0xFF bytes stuffed.
 *
 *   const buf = encodeJPEG(rgba, width, height, quality = 0.92)     rgba: Uint8Array, width x height x 4, top row first
 */

// T.81 Annex K.1: the example quantisation tables, natural (row-major) order
const LUMA_Q = [16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55, 14, 13, 16, 24, 40, 57, 69, 56,
    14, 17, 22, 29, 51, 87, 80, 62, 18, 22, 37, 56, 68, 109, 103, 77, 24, 35, 55, 64, 81, 104, 113, 92,
    49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99]
const CHROMA_Q = [17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99, 24, 26, 56, 99, 99, 99, 99, 99,
    47, 66, 99, 99, 99, 99, 99, 99].concat(new Array(32).fill(99))

// the zig-zag scan: ZIGZAG[k] is the natural index of the k-th coefficient, walking the anti-diagonals alternately
const ZIGZAG = (() => {
    const z = []
    for (let s = 0; s < 15; s++) {
        const cells = []
        for (let i = 0; i <= s; i++) { const r = i, c = s - i; if (r < 8 && c < 8) cells.push(r * 8 + c) }
        // even diagonals run from bottom-left to top-right (row falling), odd ones the other way
        z.push(...(s % 2 === 0 ? cells.reverse() : cells))
    }
    return z
})()

// T.81 Annex K.3: the example Huffman tables as (codes of each length 1..16, symbols in code order)
const DC_LUMA = [[0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0], [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11]]
const DC_CHROMA = [[0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0], [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11]]
const AC_LUMA = [[0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d], [
    0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61, 0x07,
    0x22, 0x71, 0x14, 0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52, 0xd1, 0xf0,
    0x24, 0x33, 0x62, 0x72, 0x82, 0x09, 0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25, 0x26, 0x27, 0x28,
    0x29, 0x2a, 0x34, 0x35, 0x36, 0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49,
    0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59, 0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69,
    0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89,
    0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7,
    0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3, 0xc4, 0xc5,
    0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2,
    0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8,
    0xf9, 0xfa]]
const AC_CHROMA = [[0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77], [
    0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07, 0x61, 0x71,
    0x13, 0x22, 0x32, 0x81, 0x08, 0x14, 0x42, 0x91, 0xa1, 0xb1, 0xc1, 0x09, 0x23, 0x33, 0x52, 0xf0,
    0x15, 0x62, 0x72, 0xd1, 0x0a, 0x16, 0x24, 0x34, 0xe1, 0x25, 0xf1, 0x17, 0x18, 0x19, 0x1a, 0x26,
    0x27, 0x28, 0x29, 0x2a, 0x35, 0x36, 0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48,
    0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59, 0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68,
    0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x82, 0x83, 0x84, 0x85, 0x86, 0x87,
    0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3, 0xa4, 0xa5,
    0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3,
    0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda,
    0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8,
    0xf9, 0xfa]]

// the canonical codes of a table (T.81 Annex C): lengths ascending, each code the last one plus 1, shifted left a bit
// at every new length; symbol -> [code, length]
function huffmanCodes([counts, symbols]) {
    const out = new Map()
    let code = 0, k = 0
    for (let len = 1; len <= 16; len++) {
        for (let i = 0; i < counts[len - 1]; i++) out.set(symbols[k++], [code++, len])
        code <<= 1
    }
    return out
}

// IJG's quality scaling of a table: Q in 1..100 (50 the table itself), each entry rounded and held in 1..255
function scaledTable(base, Q) {
    const s = Q < 50 ? Math.floor(5000 / Q) : 200 - 2 * Q
    return base.map(v => Math.min(255, Math.max(1, Math.floor((v * s + 50) / 100))))
}

// the DCT's basis: C[u][x] = c(u) / 2 cos((2x + 1) u pi / 16), c(0) = 1 / sqrt 2, else 1 -- orthonormal in 2-D
const BASIS = (() => {
    const b = new Float64Array(64)
    for (let u = 0; u < 8; u++) for (let x = 0; x < 8; x++) b[u * 8 + x] = (u ? 0.5 : 0.5 / Math.SQRT2) * Math.cos((2 * x + 1) * u * Math.PI / 16)
    return b
})()
function fdct(block, out, tmp) {
    for (let y = 0; y < 8; y++) for (let u = 0; u < 8; u++) {          // rows
        let s = 0
        for (let x = 0; x < 8; x++) s += BASIS[u * 8 + x] * block[y * 8 + x]
        tmp[y * 8 + u] = s
    }
    for (let u = 0; u < 8; u++) for (let v = 0; v < 8; v++) {          // columns
        let s = 0
        for (let y = 0; y < 8; y++) s += BASIS[v * 8 + y] * tmp[y * 8 + u]
        out[v * 8 + u] = s
    }
}

function bitWriter() {
    const bytes = []
    let acc = 0, n = 0
    return {
        put(code, len) {
            for (let i = len - 1; i >= 0; i--) {
                acc = (acc << 1) | ((code >>> i) & 1)
                if (++n === 8) { bytes.push(acc); if (acc === 0xFF) bytes.push(0); acc = 0; n = 0 }
            }
        },
        // the last byte padded with 1 bits (T.81 F.1.2.3)
        finish() { while (n) this.put(1, 1); return bytes },
    }
}
// a value's size category and its bits (negative values as the value - 1 in that many bits, T.81 F.1.2.1)
function category(v) {
    const a = Math.abs(v)
    let s = 0
    while ((a >> s) > 0) s++
    return [s, v < 0 ? (v - 1) & ((1 << s) - 1) : v]
}

function encodeJPEG(rgba, width, height, quality = 0.92) {
    if (rgba.length !== width * height * 4) throw new Error(`jpeg: ${rgba.length} bytes for a ${width} x ${height} RGBA picture`)
    if (!(width > 0 && height > 0 && width < 65536 && height < 65536)) throw new Error(`jpeg: ${width} x ${height} is no JPEG size`)
    const Q = Math.min(100, Math.max(1, Math.round(100 * Math.min(1, Math.max(0, +quality || 0.92)))))
    const qt = [scaledTable(LUMA_Q, Q), scaledTable(CHROMA_Q, Q)]
    const dc = [huffmanCodes(DC_LUMA), huffmanCodes(DC_CHROMA)], ac = [huffmanCodes(AC_LUMA), huffmanCodes(AC_CHROMA)]
    // the planes, full-range YCbCr (T.871); chroma then averaged over 2 x 2, the edges repeated to whole MCUs
    const W16 = Math.ceil(width / 16) * 16, H16 = Math.ceil(height / 16) * 16
    const Y = new Float64Array(W16 * H16), Cb = new Float64Array(W16 * H16), Cr = new Float64Array(W16 * H16)
    for (let y = 0; y < H16; y++) for (let x = 0; x < W16; x++) {
        const i = 4 * (Math.min(y, height - 1) * width + Math.min(x, width - 1))
        const r = rgba[i], g = rgba[i + 1], b = rgba[i + 2], o = y * W16 + x
        Y[o] = 0.299 * r + 0.587 * g + 0.114 * b
        Cb[o] = -0.168736 * r - 0.331264 * g + 0.5 * b + 128
        Cr[o] = 0.5 * r - 0.418688 * g - 0.081312 * b + 128
    }
    const bw = bitWriter()
    const pred = [0, 0, 0]
    const block = new Float64Array(64), coef = new Float64Array(64), tmp = new Float64Array(64)
    const code = (comp, t) => {
        fdct(block, coef, tmp)
        const q = qt[t]
        const zz = new Int32Array(64)
        for (let k = 0; k < 64; k++) zz[k] = Math.round(coef[ZIGZAG[k]] / q[ZIGZAG[k]])
        const [ds, dbits] = category(zz[0] - pred[comp])
        pred[comp] = zz[0]
        const d = dc[t].get(ds)
        bw.put(d[0], d[1])
        if (ds) bw.put(dbits, ds)
        let run = 0
        for (let k = 1; k < 64; k++) {
            if (zz[k] === 0) { run++; continue }
            while (run > 15) { const z = ac[t].get(0xF0); bw.put(z[0], z[1]); run -= 16 }
            const [s, bits] = category(zz[k])
            const a = ac[t].get((run << 4) | s)
            bw.put(a[0], a[1])
            bw.put(bits, s)
            run = 0
        }
        if (run) { const e = ac[t].get(0x00); bw.put(e[0], e[1]) }
    }
    for (let my = 0; my < H16; my += 16) for (let mx = 0; mx < W16; mx += 16) {
        for (const [ox, oy] of [[0, 0], [8, 0], [0, 8], [8, 8]]) {
            for (let y = 0; y < 8; y++) for (let x = 0; x < 8; x++) block[y * 8 + x] = Y[(my + oy + y) * W16 + mx + ox + x] - 128
            code(0, 0)
        }
        for (const [plane, comp] of [[Cb, 1], [Cr, 2]]) {
            for (let y = 0; y < 8; y++) for (let x = 0; x < 8; x++) {
                const o = (my + 2 * y) * W16 + mx + 2 * x
                block[y * 8 + x] = (plane[o] + plane[o + 1] + plane[o + W16] + plane[o + W16 + 1]) / 4 - 128
            }
            code(comp, 1)
        }
    }
    const data = bw.finish()
    // the markers around it
    const seg = (marker, body) => [0xFF, marker, (body.length + 2) >> 8, (body.length + 2) & 0xFF, ...body]
    const dqt = t => [t, ...ZIGZAG.map(k => qt[t][k])]
    const dht = (cls, id, [counts, symbols]) => [(cls << 4) | id, ...counts, ...symbols]
    const head = [0xFF, 0xD8,
        ...seg(0xE0, [0x4A, 0x46, 0x49, 0x46, 0, 1, 2, 0, 0, 1, 0, 1, 0, 0]),          // JFIF 1.02, aspect 1:1, no thumbnail
        ...seg(0xDB, [...dqt(0), ...dqt(1)]),
        ...seg(0xC0, [8, height >> 8, height & 0xFF, width >> 8, width & 0xFF, 3, 1, 0x22, 0, 2, 0x11, 1, 3, 0x11, 1]),
        ...seg(0xC4, [...dht(0, 0, DC_LUMA), ...dht(1, 0, AC_LUMA), ...dht(0, 1, DC_CHROMA), ...dht(1, 1, AC_CHROMA)]),
        ...seg(0xDA, [3, 1, 0x00, 2, 0x11, 3, 0x11, 0, 63, 0])]
    return Buffer.concat([Buffer.from(head), Buffer.from(data), Buffer.from([0xFF, 0xD9])])
}

module.exports = { encodeJPEG, scaledTable, ZIGZAG, LUMA_Q, CHROMA_Q, AC_LUMA, AC_CHROMA, DC_LUMA, DC_CHROMA }
