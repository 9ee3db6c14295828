'use strict'
/*
 * The optics bench from JavaScript: createOptics's run and results (lib/pathtracer/optics/index.js) over libpt's
 * pt_optics_trace through the N-API addon. A scene is an object with PtOpticsScene's field names
 * (include/pt_types.h); each result an object with PtOpticsResult's, beam and field as [C0-C180, C90-C270] pairs.
 *
 *   const O = createOptics({ device: 0 })
 *   const { results, far, lux, walkMs } = O.run(scenes, { far: true, lux: true })
 *     far: Float64Array n x 181 x 360 (lm a bin, gamma rows), lux: Float64Array n x 256 x 256 (lx at z = 1 m)
 */
const { pack, sizeOf, loadAddon } = require('./raytrace_main')

const SCENE = [['ledHalf', 'f', 1], ['flux', 'f', 1], ['xyz', 'f', 3], ['lensZ', 'f', 1], ['lensT', 'f', 1], ['lensD', 'f', 1],
    ['c1', 'f', 1], ['k1', 'f', 1], ['c2', 'f', 1], ['k2', 'f', 1], ['ior', 'f', 1], ['reflF', 'f', 1], ['reflR0', 'f', 1],
    ['reflZ1', 'f', 1], ['rho', 'f', 3], ['boardR', 'f', 1], ['photons', 'u', 1], ['seed', 'u', 1]]
const RESULT = ['emitted', 'escaped', 'absorbed', 'efficiency', 'Imax', 'gammaMax', 'cMax', 'beam0', 'beam1', 'field0', 'field1',
    'X', 'Y', 'Z', 'x', 'y', 'cct', 'duv', 'luxMax', 'luxCentre', 'photons', 'events', 'flops']
const NG = 181, NC = 360, LUX = 256
const STRUCT_SCENE = 9, STRUCT_RESULT = 10   // pt_struct_size indices

function createOptics({ device = 0 } = {}) {
    const addon = loadAddon()
    if (addon.structSize(STRUCT_SCENE) !== sizeOf(SCENE) || addon.structSize(STRUCT_RESULT) !== RESULT.length * 8)
        throw new Error('optics: the addon and this module disagree on PtOpticsScene / PtOpticsResult')
    return {
        run(scenes, { far = false, lux = false } = {}) {
            const n = scenes.length
            const buf = new ArrayBuffer(n * sizeOf(SCENE))
            const view = new DataView(buf)
            let off = 0
            for (const s of scenes) off = pack(SCENE, s, view, off)
            const raw = new Float64Array(n * RESULT.length)
            const F = far ? new Float64Array(n * NG * NC) : null
            const L = lux ? new Float64Array(n * LUX * LUX) : null
            const [walkMs, numbersMs] = addon.opticsTrace(device, buf, n, raw, F, L)
            const results = []
            for (let i = 0; i < n; i++) {
                const r = {}
                RESULT.forEach((k, j) => { r[k] = raw[i * RESULT.length + j] })
                r.beam = [r.beam0, r.beam1]
                r.field = [r.field0, r.field1]
                for (const k of ['beam0', 'beam1', 'field0', 'field1']) delete r[k]
                results.push(r)
            }
            return { results, far: F, lux: L, walkMs, numbersMs }
        },
    }
}

module.exports = { createOptics, SCENE, RESULT, NG, NC, LUX }
