'use strict'
/*
 * PNG files of a developed picture (darkroom/export/png.js's role), from the format's specification (RFC 2083): the
 * signature, IHDR (8 bits a channel, RGBA), an sRGB chunk, the rows each behind a filter byte and deflated (Node's zlib)
 * into IDAT, IEND; every chunk closed by the CRC-32 of its type and data (the table-driven CRC of the specification's
 * annex, polynomial 0xEDB88320). Each row takes the filter whose output has the least sum of absolute (signed) values.
 *
 *   const buf = encodePNG(rgba, width, height)      rgba: Uint8Array of width x height x 4, row 0 at the top
 *   const b16 = encodePNG16(rgb, width, height)     rgb: Uint16Array of width x height x 3 (16 bits a channel, RGB)
 *   insertText(buf, [[keyword, value], ...])         the same PNG with iTXt chunks (UTF-8 text) right after IHDR
 *
 * Both writers take { text } too: iTXt chunks (keyword, uncompressed, no language, UTF-8) after IHDR, which is how
 * the reference's saves say how they were made (darkroom/export/png.js itxt, insertText).
 */
const zlib = require('zlib')

const CRC = (() => {
    const t = new Uint32Array(256)
    for (let n = 0; n < 256; n++) {
        let c = n
        for (let k = 0; k < 8; k++) c = c & 1 ? 0xEDB88320 ^ (c >>> 1) : c >>> 1
        t[n] = c >>> 0
    }
    return t
})()
function crc32(bufs) {
    let c = 0xFFFFFFFF
    for (const b of bufs) for (let i = 0; i < b.length; i++) c = CRC[(c ^ b[i]) & 0xFF] ^ (c >>> 8)
    return (c ^ 0xFFFFFFFF) >>> 0
}
function chunk(type, data) {
    const head = Buffer.alloc(8)
    head.writeUInt32BE(data.length, 0)
    head.write(type, 4, 'ascii')
    const tail = Buffer.alloc(4)
    tail.writeUInt32BE(crc32([head.subarray(4), data]), 0)
    return Buffer.concat([head, data, tail])
}
function paeth(a, b, c) {
    const p = a + b - c, pa = Math.abs(p - a), pb = Math.abs(p - b), pc = Math.abs(p - c)
    return pa <= pb && pa <= pc ? a : pb <= pc ? b : c
}
function filterRow(row, prev, bpp, kind, out) {
    for (let i = 0; i < row.length; i++) {
        const a = i >= bpp ? row[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0
        const pred = kind === 0 ? 0 : kind === 1 ? a : kind === 2 ? b : kind === 3 ? (a + b) >> 1 : paeth(a, b, c)
        out[i] = (row[i] - pred) & 0xFF
    }
    return out
}

// an iTXt chunk: keyword (1-79 Latin-1 bytes) and a NUL, compression flag 0, method 0, empty language tag and
// translated keyword (each NUL-ended), then the text in UTF-8
function itxt(keyword, text) {
    const k = Buffer.from(String(keyword).slice(0, 79), 'latin1')
    return chunk('iTXt', Buffer.concat([k, Buffer.alloc(5), Buffer.from(String(text), 'utf8')]))
}
function insertText(png, text = []) {
    const at = 8 + 25                            // the signature and IHDR, always first
    return Buffer.concat([Buffer.from(png.subarray(0, at)), ...text.map(([k, v]) => itxt(k, v)), Buffer.from(png.subarray(at))])
}

// rows of `bytes` (height rows of `stride` bytes, top first) filtered and deflated into a PNG of the given IHDR depth
// and colour type; bpp: bytes a pixel, what the filters look back by
function encodeRows(bytes, width, height, stride, bpp, depth, colourType, { level = 6, text = [] } = {}) {
    const raw = Buffer.alloc(height * (stride + 1))
    let prev = new Uint8Array(stride)
    const cand = new Uint8Array(stride), best = new Uint8Array(stride)
    for (let y = 0; y < height; y++) {
        const row = bytes.subarray(y * stride, (y + 1) * stride)
        let bestKind = 0, bestCost = Infinity
        for (let k = 0; k < 5; k++) {
            filterRow(row, prev, bpp, k, cand)
            let cost = 0
            for (let i = 0; i < stride; i++) cost += cand[i] < 128 ? cand[i] : 256 - cand[i]
            if (cost < bestCost) { bestCost = cost; bestKind = k; best.set(cand) }
        }
        raw[y * (stride + 1)] = bestKind
        raw.set(best, y * (stride + 1) + 1)
        prev = row
    }
    const ihdr = Buffer.alloc(13)
    ihdr.writeUInt32BE(width, 0)
    ihdr.writeUInt32BE(height, 4)
    ihdr[8] = depth; ihdr[9] = colourType; ihdr[10] = 0; ihdr[11] = 0; ihdr[12] = 0
    return Buffer.concat([Buffer.from([0x89, 0x50, 0x4E, 0x47, 0x0D, 0x0A, 0x1A, 0x0A]), chunk('IHDR', ihdr),
        ...text.map(([k, v]) => itxt(k, v)), chunk('sRGB', Buffer.from([0])), chunk('IDAT', zlib.deflateSync(raw, { level })),
        chunk('IEND', Buffer.alloc(0))])
}

function encodePNG(rgba, width, height, opts = {}) {
    if (rgba.length !== width * height * 4) throw new Error(`png: ${rgba.length} bytes for a ${width} x ${height} RGBA picture`)
    return encodeRows(rgba, width, height, width * 4, 4, 8, 6, opts)
}

// 16 bits a channel, RGB (colour type 2): each sample big-endian, as PNG stores it
function encodePNG16(rgb, width, height, opts = {}) {
    if (rgb.length !== width * height * 3) throw new Error(`png16: ${rgb.length} samples for a ${width} x ${height} RGB picture`)
    const bytes = new Uint8Array(rgb.length * 2)
    for (let i = 0; i < rgb.length; i++) { bytes[2 * i] = rgb[i] >>> 8; bytes[2 * i + 1] = rgb[i] & 0xFF }
    return encodeRows(bytes, width, height, width * 6, 6, 16, 2, opts)
}

module.exports = { encodePNG, encodePNG16, insertText, itxt, crc32 }
