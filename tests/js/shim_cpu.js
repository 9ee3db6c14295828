'use strict'
// The JavaScript shim without a GPU: the C structures' packing against libpt's own sizes, the facade's uniform bag
// (unknown names returned, as lib/pathtracer/index.js set does), and the emitter's refusals. Prints JSON.
const path = require('path')
const js = path.join(__dirname, '..', '..', 'shadytracer_amd', 'js')
const m = require(path.join(js, 'raytrace_main'))
const { generate } = require(path.join(js, 'hipgen'))
const { CONFIGS } = require(path.join(js, 'configs'))
const { createPathTracer } = require(path.join(js, 'index'))

const out = {}
const a = m.loadAddon()
const optics = require(path.join(js, 'optics'))
optics.createOpticsBench({ device: 0 })   // checks the bench's layouts against the addon's structure sizes; no GPU call
out.traceSizes = [m.sizeOf(optics.SOURCE), a.structSize(12), m.sizeOf(optics.TRACE), a.structSize(13), optics.TALLY_BYTES, a.structSize(14)]
out.opticsSizes = [m.sizeOf(optics.SCENE), a.structSize(9), optics.RESULT.length * 8, a.structSize(10)]
out.developSize = [m.sizeOf(m.DEVELOP), a.structSize(15)]
out.sizes = [m.sizeOf(m.UNIFORMS), a.structSize(0), m.sizeOf(m.MATERIAL), a.structSize(1), m.sizeOf(m.FACE), a.structSize(2)]
out.addonKeys = Object.keys(a).sort()
// packing: a known value lands at its C offset (width 0, height 4, seed 60, cam.pos 64, env.type 128)
const buf = new ArrayBuffer(m.sizeOf(m.UNIFORMS))
m.pack(m.UNIFORMS, { width: 640, height: 480, seed: 77, cam: { pos: [1, 2, 3] }, env: { type: 2, L: 1000 } }, new DataView(buf), 0)
const dv = new DataView(buf)
out.packed = [dv.getUint32(0, true), dv.getUint32(4, true), dv.getUint32(60, true), dv.getFloat32(64, true), dv.getFloat32(72, true),
    dv.getUint32(144, true), dv.getFloat32(148, true)]
const RT = createPathTracer({ width: 8, height: 8 })
out.unknown = RT.set({ maxbounce: 3, notAUniform: 1, stepscale: 0.8 })
out.maxbounce = RT.get('maxbounce')
// the reference's lighting settings (uniforms.schema.js Lighting) are known names, kept for the scene's environment
out.lighting = RT.set({ uSunElevation: 20, envSky: 0, sunOutside: 128000, envMode: 0 })
out.elevation = RT.get('uSunElevation')
try {
    generate({ root: { children: [{ id: 'u1', kind: 'object', type: 'code', block: 'b' }] }, blocks: [{ id: 'b', code: 'C = add_new(C, myShape(p), MATERIAL);' }] })
    out.refused = null
} catch (e) { out.refused = String(e.message) }
const c2 = generate(CONFIGS['room-c2']().doc)
out.c2 = { prims: c2.program.filter(o => o.op === 'prim').length, materials: c2.materials.length, faces: c2.faces.length, flops: c2.flops.dist }
process.stdout.write(JSON.stringify(out))
