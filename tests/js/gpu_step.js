'use strict'
// The drop-in path end to end on a GPU: createPathTracer -> init -> setScene -> step(n) -> imageStats / readRegion,
// through the N-API addon. Prints JSON for tests/test_gpu.py.
const path = require('path')
const js = path.join(__dirname, '..', '..', 'shadytracer_amd', 'js')
const { createPathTracer } = require(path.join(js, 'index'))
const { CONFIGS } = require(path.join(js, 'configs'))

async function main() {
    const made = CONFIGS['plane-c1']({ width: 64, height: 64 })
    const RT = createPathTracer({ width: 64, height: 64, CAMERA: made.camera })
    await RT.init()
    const ms = RT.setScene(made.doc)
    RT.set({ maxbounce: 4 })
    await RT.step(16)
    const st = RT.imageStats()
    const floor = RT.readRegion(0, 0, 63, 15)
    const log = RT.wavefrontLog()
    if (process.argv[2]) {
        require('fs').writeFileSync(path.join(process.argv[2], 'c1.exr'), RT.saveLight('exr'))
        require('fs').writeFileSync(path.join(process.argv[2], 'c1.pfm'), RT.saveLight('pfm'))
    }
    RT.dispose()
    process.stdout.write(JSON.stringify({ ms, stats: st, floor, log }))
}
main().catch(e => { process.stderr.write(String(e && e.stack || e) + '\n'); process.exit(1) })
