'use strict'
// The drop-in path end to end on a GPU: createPathTracer -> init -> setScene -> step(n) -> imageStats / readRegion,
// through the N-API addon; a compile the event loop keeps running through (a timer ticks while hipRTC works on a
// worker thread), keepPicture, a trace refused during a build, and the reference's lighting settings through set().
// Prints JSON for tests/test_gpu.py.
const path = require('path')
const js = path.join(__dirname, '..', '..', 'shadytracer_amd', 'js')
const { createPathTracer } = require(path.join(js, 'index'))
const { CONFIGS } = require(path.join(js, 'configs'))
const { generate, environmentOf } = require(path.join(js, 'hipgen'))

// a scene no code-object cache holds yet: the same program under a comment of its own, so the build compiles
function uncached(doc, tag) {
    const sc = generate(doc)
    sc.hip = `// build ${tag} ${process.pid} ${Date.now()}\n` + sc.hip
    return sc
}

async function main() {
    const made = CONFIGS['plane-c1']({ width: 64, height: 64 })
    const RT = createPathTracer({ width: 64, height: 64, CAMERA: made.camera })
    await RT.init()
    const ms = await RT.setScene(made.doc)
    RT.set({ maxbounce: 4 })
    await RT.step(16)
    const st = RT.imageStats()
    const floor = RT.readRegion(0, 0, 63, 15)
    const log = RT.wavefrontLog()
    if (process.argv[2]) {
        require('fs').writeFileSync(path.join(process.argv[2], 'c1.exr'), RT.saveLight('exr'))
        require('fs').writeFileSync(path.join(process.argv[2], 'c1.pfm'), RT.saveLight('pfm'))
    }
    // the compile on a worker thread: the event loop's timer goes on ticking meanwhile; the picture kept
    let ticks = 0
    const timer = setInterval(() => { ticks++ }, 10)
    const compileMs = await RT.setScene(uncached(made.doc, 'a'), { keepPicture: true })
    clearInterval(timer)
    const kept = RT.imageStats().samples
    // nothing else while a build runs: a step is refused, and the build goes on to the end
    const building = RT.setScene(uncached(made.doc, 'b'))
    let refused = null
    try { await RT.step(1) } catch (e) { refused = e.message }
    await building
    // the reference's lighting settings steer the environment (uniforms.schema.js Lighting): the photometry bench,
    // the sun alone at elevation 20 (tests.js photometry, 'sun only, elevation 20')
    const bench = CONFIGS['sunsky-bench']({ width: 32, height: 32 })
    RT.setSize(32, 32)
    RT.set({ CAMERA: bench.camera })
    await RT.setScene(bench.doc)
    const unknown = RT.set({ uSunElevation: 20, envSky: 0, maxbounce: 10, noSuchSetting: 1 })
    await RT.step(256)
    const sun = RT.imageStats()
    const env = environmentOf(Object.assign({}, bench.doc.environment, { uSunElevation: 20, envSky: 0 }), bench.doc)
    const want = 0.8 * env.sunE * Math.sin(20 * Math.PI / 180) / Math.PI
    RT.dispose()
    process.stdout.write(JSON.stringify({ ms, stats: st, floor, log, ticks, compileMs, kept, refused, unknown,
        sunOnly: { mean: sun.meanLuma, relSE: sun.relSE, want } }))
}
main().catch(e => { process.stderr.write(String(e && e.stack || e) + '\n'); process.exit(1) })
