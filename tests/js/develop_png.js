'use strict'
// The developed picture from JavaScript: createPathTracer -> the C1 floor, a few samples -> RT.develop / RT.savePicture
// ('png') -> the PNG file at argv[2] and the FP64 accumulation (row 0 at the bottom) at argv[3], optionally a 16-bit PNG
// at argv[4] and a JPEG at argv[5]; prints JSON.
const path = require('path')
const fs = require('fs')
const js = path.join(__dirname, '..', '..', 'shadytracer_amd', 'js')
const { createPathTracer } = require(path.join(js, 'index'))
const { CONFIGS } = require(path.join(js, 'configs'))

async function main() {
    const made = CONFIGS['plane-c1']({ width: 48, height: 32 })
    const RT = createPathTracer({ width: 48, height: 32, CAMERA: made.camera })
    await RT.init()
    await RT.setScene(made.doc)
    RT.set({ maxbounce: 4 })
    await RT.step(4)
    const pic = RT.develop({ toneMap: 'neutral' })
    // the PNG from the same picture: the exposure held where the first development set it
    fs.writeFileSync(process.argv[2], RT.savePicture('png', { toneMap: 'neutral', aeSpeed: 0 }))
    fs.writeFileSync(process.argv[3], Buffer.from(RT.readAccum().acc.buffer))
    // and the same picture at 16 bits (with text) and as a JPEG, when asked for
    if (process.argv[4]) fs.writeFileSync(process.argv[4], RT.savePicture('png16', { toneMap: 'neutral', aeSpeed: 0 }))
    if (process.argv[5]) fs.writeFileSync(process.argv[5], RT.savePicture('jpeg', { toneMap: 'neutral', aeSpeed: 0, quality: 0.95 }))
    RT.dispose()
    process.stdout.write(JSON.stringify({ width: pic.width, height: pic.height, ev: pic.ev }))
}
main().catch(e => { console.error(e); process.exit(1) })
