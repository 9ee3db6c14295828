'use strict'
// The optics trace over a scene from JavaScript: createPathTracer -> createOptics(RT) -> addon.opticsRun ->
// pt_optics_run. A Lambertian disc shining up over the optics floor, its far field in 5-degree bins; prints JSON (for
// tests/test_optics_trace_gpu.py). A timer ticks while the trace runs: the run resolves a Promise from a worker thread.
const path = require('path')
const js = path.join(__dirname, '..', '..', 'shadytracer_amd', 'js')
const { createPathTracer } = require(path.join(js, 'index'))
const { createOptics } = require(path.join(js, 'optics'))
const farfield = require(path.join(js, 'farfield'))
const { CONFIGS } = require(path.join(js, 'configs'))

async function main() {
    const RT = createPathTracer({ width: 16, height: 16 })
    await RT.init()
    await RT.setScene(CONFIGS['optics-floor']({ albedo: 0.5 }).doc)
    RT.set({ eps: 1e-5 })
    const O = createOptics(RT)
    O.setSource({ type: 'disc', pos: [0, 0, 1], axis: [0, 0, 1], u: [1, 0, 0], size: [0.1, 0.1], distribution: 'cosine',
                  halfAngle: 90, n: 1, flux: 1000 })
    O.setFar({ mode: 'escaped', nC: 24, nG: 37, axis: [0, 0, 1], c0: [1, 0, 0] })
    let ticks = 0
    const timer = setInterval(() => { ticks++ }, 1)
    const res = await O.run({ rays: 1 << 20, batches: 4 })
    clearInterval(timer)
    const I = farfield.intensity(O.far, { symmetry: 'rotational' })
    let farSum = 0
    for (const f of O.far.flux) farSum += f
    // the floor as a detector, lit from above: its hits
    O.settings.detectors = [O.materialId('floor')]
    O.settings.detCap = 1000
    O.setFar({ mode: 'off' })
    O.setSource({ type: 'disc', pos: [0, 0, 1], axis: [0, 0, -1], size: [0.1, 0.1], distribution: 'collimated', flux: 500 })
    const det = await O.run({ rays: 1 << 16, batches: 1 })
    const hits = await O.readHits()
    process.stdout.write(JSON.stringify({ emitted: res.emitted, escaped: res.escaped, seEscaped: res.se.escaped, balance: res.balance,
        batches: res.batches, rays: res.rays, farSum, beam: farfield.width(I, 0, 0.5), field: farfield.width(I, 0, 0.1), ticks,
        detected: det.detected, hitCount: hits.count, hitsKept: hits.hits.length, firstHit: hits.hits[0], seed: O.seed }))
    RT.dispose()
}
main().catch(e => { console.error(e); process.exit(1) })
