'use strict'
// The developed-picture writers without a GPU: a synthetic RGBA picture (argv[2]: its raw bytes are written there) as
// JPEG at qualities 0.92 and 0.5, as 8-bit PNG with text, and as 16-bit PNG with text; files into the folder argv[3].
const path = require('path')
const fs = require('fs')
const js = path.join(__dirname, '..', '..', 'shadytracer_amd', 'js')
const { encodeJPEG } = require(path.join(js, 'jpeg'))
const { encodePNG, encodePNG16, insertText } = require(path.join(js, 'png'))

const w = 83, h = 45
const px = new Uint8Array(w * h * 4), px16 = new Uint16Array(w * h * 3)
for (let y = 0; y < h; y++) for (let x = 0; x < w; x++) {
    const i = 4 * (y * w + x)
    px[i] = Math.round(255 * x / (w - 1))
    px[i + 1] = Math.round(127.5 + 100 * Math.sin(x / 7) * Math.cos(y / 5))
    px[i + 2] = Math.round(255 * y / (h - 1))
    px[i + 3] = 255
    const j = 3 * (y * w + x)
    px16[j] = Math.round(65535 * x / (w - 1)); px16[j + 1] = (x * 977 + y * 131) & 0xFFFF; px16[j + 2] = Math.round(65535 * y / (h - 1))
}
const dir = process.argv[3]
fs.writeFileSync(process.argv[2], Buffer.from(px.buffer))
fs.writeFileSync(path.join(dir, 'p16.raw'), Buffer.from(px16.buffer))
fs.writeFileSync(path.join(dir, 'q92.jpg'), encodeJPEG(px, w, h, 0.92))
fs.writeFileSync(path.join(dir, 'q50.jpg'), encodeJPEG(px, w, h, 0.5))
fs.writeFileSync(path.join(dir, 'q100.jpg'), encodeJPEG(px, w, h, 1))
const text = [['Software', 'shadytracer_amd'], ['Comment', JSON.stringify({ unit: 'cd/m²', ev: 9.5 })]]
fs.writeFileSync(path.join(dir, 'p8.png'), encodePNG(px, w, h, { text }))
fs.writeFileSync(path.join(dir, 'p8late.png'), insertText(encodePNG(px, w, h), text))
fs.writeFileSync(path.join(dir, 'p16.png'), encodePNG16(px16, w, h, { text }))
process.stdout.write(JSON.stringify({ w, h }))
