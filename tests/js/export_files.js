'use strict'
// Writes the light files of a seeded synthetic accumulation (bottom row first, r g b n a pixel: a sun of 1.6e9, values
// past half's range, negatives, pixels no sample reached) for tests/test_export.py to decode, and prints the writers'
// known answers as JSON.
const path = require('path')
const fs = require('fs')
const X = require(path.join(__dirname, '..', '..', 'shadytracer_amd', 'js', 'export'))
const [dir, W, H] = [process.argv[2], +process.argv[3], +process.argv[4]]
let a = 12345
const rnd = () => { a = (Math.imul(a, 1103515245) + 12345) >>> 0; return a / 4294967296 }
const acc = new Float64Array(W * H * 4)
for (let i = 0; i < W * H; i++) {
    const big = rnd() < 0.02 ? 1.6e9 : 1
    for (let c = 0; c < 3; c++) acc[4 * i + c] = big * (rnd() < 0.05 ? -0.3 * rnd() : 2000 * rnd() ** 3)
    acc[4 * i + 3] = rnd() < 0.01 ? 0 : 64
}
// a flat run for the run-length form
for (let x = 0; x < W; x++) for (let c = 0; c < 4; c++) acc[4 * (3 * W + x) + c] = c < 3 ? 100 : 64
fs.writeFileSync(path.join(dir, 'acc.bin'), Buffer.from(acc.buffer))
const files = {
    'light.pfm': X.writeLight('pfm', acc, W, H),
    'light.hdr': X.writeLight('hdr', acc, W, H, { lines: ['EXPOSURE=2.0'] }),
    'light.exr': X.writeLight('exr', acc, W, H, { strings: { software: 'shadytracer_amd test' } }),
    'half.exr': X.writeLight('exr', acc, W, H, { half: true, compression: 'none', k: 1 / 1000 }),
    'narrow.hdr': X.writeLight('hdr', acc.slice(0, 4 * 5 * 2), 5, 2),
}
for (const [f, b] of Object.entries(files)) fs.writeFileSync(path.join(dir, f), b)
const flat = Buffer.alloc(64, 7)
process.stdout.write(JSON.stringify({
    rgbe: [X.rgbe(1, 0.5, 0.25), X.rgbe(0, 0, 0), X.rgbe(1e-33, 0, 0), X.rgbe(0.999999, 0.5, 0)].map(b => [...b]),
    half: [1, 65504, 65520, -2, 0.5, 6.103515625e-05, 5.960464477539063e-08, 2.98e-08, 1e-9, 1 + 2 ** -11, 1 + 3 * 2 ** -11, Infinity, NaN].map(X.toHalf),
    prepassFlat: [...X.zipPrepass(flat)],
}))
