'use strict'
/*
 * THE OPTICS TRACE over the loaded scene, as lib/pathtracer/optics/index.js createOptics(RT, WG) offers it, over libpt's
 * pt_optics_run through the N-API addon: rays leave the sources, walk the scene through the engine's own march, surface
 * kernel and materials, and where their flux goes is tallied per material, into the far field and at detectors.
 *
 *   const O = createOptics(RT)                 RT: createPathTracer(...) or RAYTRACER(...), a scene loaded
 *   O.setSources([{ type: 'disc', pos, axis, u, size, distribution: 'cosine', halfAngle, n, flux, ... }])
 *   O.setFar({ mode: 'escaped', nC: 72, nG: 181, axis: [0, 0, 1], c0: [1, 0, 0] })
 *   O.settings.detectors = [materialId, ...]   surfaces that stop and record rays (engine material ids; O.materialId)
 *   const results = await O.run({ rays: 100000, batches: 4 })    lm, each with its standard error over the batches
 *   O.far                                       { mode, nC, nG, axis, c0, flux: Float64Array (gamma-major), se, emitted }
 *   const { hits, count, missed } = await O.readHits()   the last batch's detector hits
 *
 * Units are the engine's, metres and lumens. A source is optics/sources.js's description (defaultSource below): a shape
 * (point, disc, rectangle), an angular distribution about its axis (collimated, uniform within a cone, cos^n within a
 * cone), a flux, the index it emits into, one face or two, a colour (its hue; the flux is its luminance).
 *
 * THE OPTICS BENCH (64 luminaires in one persistent launch, pt_optics_trace) is createOpticsBench({ device }).run(scenes).
 */
const { pack, sizeOf, loadAddon } = require('./raytrace_main')

// PtOpticsSource, PtOpticsTrace (include/pt_types.h)
const SOURCE = [['pos', 'f', 3], ['shape', 'u', 1], ['axis', 'f', 3], ['distribution', 'u', 1], ['u', 'f', 3], ['halfAngle', 'f', 1],
    ['size', 'f', 2], ['n', 'f', 1], ['flux', 'f', 1], ['colour', 'f', 3], ['emitN', 'f', 1], ['twoSided', 'u', 1], ['onSurface', 'u', 1],
    ['pad', 'u', 2]]
const TRACE = [['rays', 'u', 1], ['bounces', 'u', 1], ['seed', 'u', 1], ['detCap', 'u', 1], ['maxDist', 'f', 1], ['farMode', 'u', 1],
    ['farMaterial', 'u', 1], ['nC', 'u', 1], ['nG', 'u', 1], ['farAxis', 'f', 3], ['farC0', 'f', 3], ['detectors', 'u', 2],
    ['marchSteps', 'u', 1], ['pad', 'u', 1]]
const SHAPES = { point: 0, disc: 1, rectangle: 2 }
const DISTRIBUTIONS = { collimated: 0, uniform: 1, cosine: 2 }
const FAR_MODES = { off: 0, escaped: 1, detector: 2 }
const MAX_MATERIALS = 64, HIT_WORDS = 8
const STRUCT_SOURCE = 12, STRUCT_TRACE = 13, STRUCT_TALLY = 14     // pt_struct_size indices
const TALLY_BYTES = 6 * 8 + 8 * 8 + MAX_MATERIALS * 24

/** A source for a script: a 1 x 1 Lambertian square 1 up, facing down, 1000 lm, into air, one face (sources.js
 *  defaultSource). */
const defaultSource = () => ({ type: 'rectangle', pos: [0, 0, 1], axis: [0, 0, -1], u: [1, 0, 0], size: [0.5, 0.5],
    distribution: 'cosine', halfAngle: 90, n: 1, flux: 1000, emitN: 1, twoSided: false, onSurface: false, colour: [1, 1, 1] })

/** PtOpticsTally from its bytes: totals, counts, and per material incident / leaving / hits. */
function readTally(buf) {
    const dv = new DataView(buf)
    const f = k => dv.getFloat64(8 * k, true)
    const u = k => Number(dv.getBigUint64(48 + 8 * k, true))
    const t = { emitted: f(0), escaped: f(1), lost: f(2), detected: f(3), volume: f(4), absorbed: f(5),
        rays: u(0), counts: { escaped: u(1), lost: u(2), interactions: u(3), absorbed: u(4), detected: u(5) }, detectorHits: u(6) }
    const inc = 112, lea = inc + 8 * MAX_MATERIALS, hit = lea + 8 * MAX_MATERIALS
    t.materials = []
    for (let m = 0; m < MAX_MATERIALS; m++) {
        const incident = dv.getFloat64(inc + 8 * m, true), leaving = dv.getFloat64(lea + 8 * m, true)
        t.materials.push({ id: m, incident, leaving, absorbed: incident - leaving, hits: Number(dv.getBigUint64(hit + 8 * m, true)) })
    }
    return t
}

/** The mean of the batches and its standard error (their spread over sqrt(n)): the reference's estimate of a run. */
function estimate(batches) {
    const n = batches.length
    const stat = get => {
        const v = batches.map(get)
        const mean = v.reduce((a, b) => a + b, 0) / n
        const se = n > 1 ? Math.sqrt(v.reduce((a, b) => a + (b - mean) ** 2, 0) / (n - 1) / n) : null
        return { mean, se }
    }
    const out = { unit: 'lm', batches: n, rays: batches.reduce((a, b) => a + b.rays, 0), se: {} }
    for (const k of ['emitted', 'escaped', 'lost', 'detected', 'volume', 'absorbed']) {
        const s = stat(b => b[k])
        out[k] = s.mean
        out.se[k] = s.se
    }
    out.counts = {}
    for (const k of Object.keys(batches[0].counts)) out.counts[k] = batches.reduce((a, b) => a + b.counts[k], 0)
    out.materials = batches[0].materials.map((m, i) => {
        const inc = stat(b => b.materials[i].incident), lea = stat(b => b.materials[i].leaving)
        return { id: m.id, incident: inc.mean, leaving: lea.mean, absorbed: inc.mean - lea.mean, hits: batches.reduce((a, b) => a + b.materials[i].hits, 0),
                 se: { incident: inc.se, leaving: lea.se } }
    }).filter(m => m.hits > 0)
    // emitted = escaped + lost + volume + what the materials took: exact per batch, so exact on the mean
    out.balance = out.emitted - (out.escaped + out.lost + out.volume + out.materials.reduce((a, m) => a + m.absorbed, 0))
    return out
}

function createOptics(RT, WG) {
    if (!RT || typeof RT.opticsRun !== 'function') throw new Error('[optics] createOptics(RT): RT is a path tracer with a scene')
    const A = loadAddon()
    if (A.structSize(STRUCT_SOURCE) !== sizeOf(SOURCE) || A.structSize(STRUCT_TRACE) !== sizeOf(TRACE) || A.structSize(STRUCT_TALLY) !== TALLY_BYTES)
        throw new Error('[optics] the addon and this module disagree on PtOpticsSource / PtOpticsTrace / PtOpticsTally')
    let busy = false
    let last = null                 // the last batch's hits
    const O = {
        sources: [defaultSource()],
        /* bounces: interactions a ray is followed through; marchSteps: sphere-tracing steps over its whole path; detCap:
           detector hits kept a batch; maxDist: the longest optical path traced, metres (0: none); detectors: material
           ids whose surfaces stop and record rays; far: the far field (optics/index.js O.settings) */
        settings: { bounces: 64, marchSteps: 2000, detCap: 262144, maxDist: 0, detectors: [],
                    far: { mode: 'off', material: 0, nC: 72, nG: 181, axis: [0, 0, 1], c0: [1, 0, 0] } },
        results: null, far: null, rays: 0, seed: 1,
    }
    O.reseed = () => { O.seed = ((O.seed * 1664525 + 1013904223) >>> 0) || 1 }
    O.setSources = list => { O.sources = (list || []).map(s => Object.assign(defaultSource(), s)) }
    O.setSource = src => O.setSources([src])
    O.sourceFlux = () => {
        const live = O.sources.filter(s => s.flux > 0)
        return { flux: live.reduce((a, s) => a + s.flux, 0), count: live.length }
    }
    O.setFar = far => { O.settings.far = Object.assign({}, O.settings.far, far) }
    /** The engine's material id of the scene's material `id` (the document's; 0 is the emitter's default). */
    O.materialId = id => {
        const sc = RT.scene && RT.scene()
        const k = sc ? sc.materials.findIndex(m => m.id === id) : -1
        if (k < 0) throw new Error(`[optics] no material '${id}' in the scene`)
        return k
    }
    function packed() {
        const list = O.sources
        const sbuf = new ArrayBuffer(Math.max(1, list.length) * sizeOf(SOURCE))
        const sv = new DataView(sbuf)
        list.forEach((s, i) => {
            const size = Array.isArray(s.size) ? s.size : [s.size, s.size]
            pack(SOURCE, { pos: s.pos, shape: SHAPES[s.type], axis: s.axis, distribution: DISTRIBUTIONS[s.distribution], u: s.u || [1, 0, 0],
                halfAngle: s.halfAngle, size: [size[0] || 0, size[1] !== undefined ? size[1] : size[0] || 0], n: s.n, flux: s.flux,
                colour: s.colour || [1, 1, 1], emitN: s.emitN || 1, twoSided: s.twoSided ? 1 : 0, onSurface: s.onSurface ? 1 : 0 }, sv, i * sizeOf(SOURCE))
            if (SHAPES[s.type] === undefined || DISTRIBUTIONS[s.distribution] === undefined)
                throw new Error(`[optics] source ${i}: type '${s.type}' or distribution '${s.distribution}' unknown`)
        })
        return { sbuf, n: list.length }
    }
    function traceBuf(rays) {
        const S = O.settings, F = S.far
        const det = [0, 0]
        for (const m of S.detectors || []) det[m >> 5] |= (1 << (m & 31)) >>> 0
        const buf = new ArrayBuffer(sizeOf(TRACE))
        pack(TRACE, { rays, bounces: S.bounces, seed: O.seed, detCap: S.detCap, maxDist: S.maxDist || 0, farMode: FAR_MODES[F.mode] || 0,
            farMaterial: F.material || 0, nC: F.nC, nG: F.nG, farAxis: F.axis, farC0: F.c0, detectors: det, marchSteps: S.marchSteps || 0 },
            new DataView(buf), 0)
        return buf
    }
    /** Trace `rays` rays to the end, `batches` times, each batch with its own seed; the reading averages the batches
     *  (optics/index.js O.run :434-477). onBatch(results, b) after each. The run's seed is given back at the end, so
     *  the same run traces the same rays. */
    O.run = async ({ rays = 100000, batches = 4, onBatch } = {}) => {
        if (busy) throw new Error('[optics] a trace is already running')
        busy = true
        const seed = O.seed
        const F = O.settings.far
        const bins = FAR_MODES[F.mode] ? F.nC * F.nG : 0
        try {
            const { sbuf, n } = packed()
            const got = [], fars = []
            for (let b = 0; b < batches; b++) {
                O.reseed()
                const r = await RT.opticsRun(sbuf, n, traceBuf(rays), bins, O.settings.detCap)
                got.push(readTally(r.tally))
                if (r.far) fars.push(r.far)
                last = { hits: r.hits, count: got[got.length - 1].detectorHits }
                O.results = Object.assign(estimate(got), { mode: 'run' })
                if (fars.length) {
                    const flux = new Float64Array(bins), se = new Float64Array(bins)
                    for (let i = 0; i < bins; i++) {
                        let s = 0, s2 = 0
                        for (const f of fars) { s += f[i]; s2 += f[i] * f[i] }
                        const m = s / fars.length
                        flux[i] = m
                        se[i] = fars.length > 1 ? Math.sqrt(Math.max(0, s2 / fars.length - m * m) * fars.length / (fars.length - 1) / fars.length) : 0
                    }
                    O.far = Object.assign({}, F, { flux, se, emitted: O.results.emitted, unit: 'lm' })
                }
                if (onBatch) await onBatch(O.results, b + 1)
            }
            O.rays = rays
            return O.results
        } finally { O.seed = seed; busy = false }
    }
    /** The last batch's detector hits: position (m), flux (lm), the direction it arrived along, the material. */
    O.readHits = async () => {
        if (!last) return { hits: [], count: 0, missed: 0, unit: 'lm' }
        const h = last.hits, out = []
        const mat = new Uint32Array(h.buffer, h.byteOffset, h.length)
        for (let i = 0; i + HIT_WORDS <= h.length; i += HIT_WORDS)
            out.push({ pos: [h[i], h[i + 1], h[i + 2]], flux: h[i + 3], dir: [h[i + 4], h[i + 5], h[i + 6]], material: mat[i + 7] })
        return { hits: out, count: last.count, missed: last.count - out.length, unit: 'lm' }
    }
    return O
}

// ---- the optics bench: 64 luminaire scenes in one persistent launch (pt_optics_trace)
const SCENE = [['ledHalf', 'f', 1], ['flux', 'f', 1], ['xyz', 'f', 3], ['lensZ', 'f', 1], ['lensT', 'f', 1], ['lensD', 'f', 1],
    ['c1', 'f', 1], ['k1', 'f', 1], ['c2', 'f', 1], ['k2', 'f', 1], ['ior', 'f', 1], ['reflF', 'f', 1], ['reflR0', 'f', 1],
    ['reflZ1', 'f', 1], ['rho', 'f', 3], ['boardR', 'f', 1], ['photons', 'u', 1], ['seed', 'u', 1]]
const RESULT = ['emitted', 'escaped', 'absorbed', 'efficiency', 'Imax', 'gammaMax', 'cMax', 'beam0', 'beam1', 'field0', 'field1',
    'X', 'Y', 'Z', 'x', 'y', 'cct', 'duv', 'luxMax', 'luxCentre', 'photons', 'events', 'flops']
const NG = 181, NC = 360, LUX = 256
const STRUCT_SCENE = 9, STRUCT_RESULT = 10   // pt_struct_size indices

/*
 *   const B = createOpticsBench({ device: 0 })
 *   const { results, far, lux, walkMs } = B.run(scenes, { far: true, lux: true })
 *     scenes: objects with PtOpticsScene's fields; results: PtOpticsResult's, beam and field as [C0-C180, C90-C270];
 *     far: Float64Array n x 181 x 360 (lm a bin, gamma rows), lux: Float64Array n x 256 x 256 (lx at z = 1 m)
 */
function createOpticsBench({ device = 0 } = {}) {
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

module.exports = { createOptics, createOpticsBench, defaultSource, readTally, SOURCE, TRACE, SCENE, RESULT, NG, NC, LUX, TALLY_BYTES }
