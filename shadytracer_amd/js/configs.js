'use strict'
/*
 * The benchmark and parity scenes as builder documents (format 'raytracer-scene', version 31), built from a seed.
 * World Z-up, metres. Each returns { doc, camera, settings }: the document for hipgen, the pinhole camera as the
 * engine's uniforms take it (genrays.js:61-92: right, up, fwd columns of invView, fx, fy of invProjection), and
 * the render settings of the configuration.
 *
 *   plane-c1     infinite SDF plane z = 0, Lambertian albedo 0.8, uniform sky of 1000 cd/m^2 over the upper
 *                hemisphere, camera at (0, -3, 1.5) aimed at the origin, 60 deg vertical, 256 x 256, 64 spp, depth 4:
 *                the floor reads 0.8 x 1000 = 800 cd/m^2 exactly.
 *   room-c2      a 5 x 4 x 3 m room (walls albedo 0.7) holding 32 primitives (spheres, boxes, tori, rounded
 *                cylinders) placed from the seed, sizes U(0.1, 0.5) m, 8 smooth-union pairs; GGX materials with
 *                roughness U(0.05, 0.9), metallic 0 or 1 (p = 0.3), base U(0.2, 0.9) per channel; a 0.5 x 0.5 m
 *                ceiling panel of 5000 cd/m^2; 1920 x 1080, 256 spp, depth 8, NEE on.
 *   caustic-c3   the room with a 2 x 2 x 2 m haze box (sigma_s 0.05 /m, g 0.6, albedo 1), a dispersive glass sphere
 *                (r 0.4 m, n = 1.5 + 0.0042 / lambda^2 um) and a 100 klx sun through a 0.3 m aperture in the ceiling;
 *                1280 x 720; the photon map at every camera vertex past the first, light tracing at the first.
 *   sppm-slab    a collimated beam through a glass slab onto a floor: the photon map's closed form.
 *   furnace-*    closed-form checks: a white furnace (albedo-1 Lambertian plane under a uniform sky, and a haze box
 *                under a uniform sphere of sky).
 *   studio-floor, lens-prism, reflector-condenser, turned-parts, gallery, joins
 *                documents in the builder's own form (millimetres, the studio environment, lights by CCT and flux,
 *                an orbit view), the others with the lens, prism, reflector, aspheric lens, outline and the
 *                remaining solid builtins.
 */

const SEED = 20261203

// mulberry32: a small seeded generator for scene layout (not the renderer's)
function rng(seed) {
    let a = seed >>> 0
    return () => {
        a = (a + 0x6D2B79F5) >>> 0
        let t = a
        t = Math.imul(t ^ (t >>> 15), t | 1)
        t ^= t + Math.imul(t ^ (t >>> 7), t | 61)
        return ((t ^ (t >>> 14)) >>> 0) / 4294967296
    }
}

function camera(pos, target, vfovDeg, width, height) {
    const v = [target[0] - pos[0], target[1] - pos[1], target[2] - pos[2]]
    const lv = Math.hypot(v[0], v[1], v[2])
    const f = v.map(x => x / lv)
    let r = [f[1], -f[0], 0]                      // cross(f, z)
    const lr = Math.hypot(r[0], r[1], r[2]) || 1
    r = r.map(x => x / lr)
    const u = [r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]]
    const fy = Math.tan(0.5 * vfovDeg * Math.PI / 180)
    return { pos, right: r, up: u, fwd: f.map(x => -x), fx: fy * width / height, fy, type: 0 }
}

const xf = (x, y, z, rx = 0, ry = 0, rz = 0, s = 1) => ({ pos: { x, y, z }, rot: { x: rx, y: ry, z: rz }, scale: s })
const BLOCKS = [
    { id: 'b-plane', label: 'plane', kind: 'object', builtin: 'plane@1', params: {} },
    { id: 'b-box', label: 'box', kind: 'object', builtin: 'box@1', params: { size: { x: 0.5, y: 0.5, z: 0.5 }, rounding: 0.02 }, lengths: ['size', 'rounding'] },
    { id: 'b-sphere', label: 'sphere', kind: 'object', builtin: 'sphere@1', params: { radius: 0.5 }, lengths: ['radius'] },
    { id: 'b-cylinder', label: 'cylinder', kind: 'object', builtin: 'cylinder@1', params: { radius: 0.5, height: 1 }, lengths: ['radius', 'height'] },
    { id: 'b-torus', label: 'torus', kind: 'object', builtin: 'torus@1', params: { ring: 0.4, tube: 0.1 }, lengths: ['ring', 'tube'] },
]
function material(id, base, roughness, metallic, extra) {
    return { id, label: id, params: Object.assign({ baseColor: { r: base[0], g: base[1], b: base[2] }, metallic, roughness, specular: 0.5, transmission: 0, ior: 1.5,
        emissive: { r: 1, g: 1, b: 1 }, emissiveLuminance: 0, abbe: 0 }, extra || {}) }
}
function doc(title, materials, children, environment, envRadius) {
    return { format: 'raytracer-scene', version: 31, title, unit: 'm', envRadius: envRadius || 400, materials, blocks: BLOCKS,
        root: { id: 'root', kind: 'group', label: title, enabled: true, blend: 0, children }, environment: environment || {} }
}
const obj = (id, block, material, transform, params, extra) => Object.assign({ id, kind: 'object', type: 'code', label: id, enabled: true, block, material, transform, params: params || {} }, extra || {})

function planeC1(opts = {}) {
    const albedo = opts.albedo !== undefined ? opts.albedo : 0.8
    const mats = [material('floor', [albedo, albedo, albedo], 1, 0, { specular: 0 })]
    // envRadius 20 m: a ray leaving the floor at a grazing angle reaches the sky sphere within maxiter steps
    // down to 0.7 degrees (1.1 at the 400 m default), so the march's step cap costs the floor under 0.03 %
    const d = doc('plane-c1', mats, [obj('floor', 'b-plane', 'floor', xf(0, 0, 0))], { sky: 'hemisphere', uniformSky: opts.sky || 1000 }, 20)
    const w = opts.width || 256, h = opts.height || 256
    return { doc: d, camera: camera([0, -3, 1.5], [0, 0, 0], 60, w, h), settings: { width: w, height: h, spp: 64, maxbounce: 4, nee: 1 } }
}

// The room: a box shell (outer minus inner) of albedo-0.7 walls, inner 5 x 4 x 3 m, floor at z = 0.
function roomChildren(extraCut) {
    const shell = { id: 'room', kind: 'group', label: 'room', enabled: true, blend: 0, transform: xf(0, 0, 1.5), children: [
        obj('room-outer', 'b-box', 'walls', xf(0, 0, 0), { size: { x: 2.7, y: 2.2, z: 1.7 }, rounding: 0 }),
        obj('room-inner', 'b-box', 'walls', xf(0, 0, 0), { size: { x: 2.5, y: 2.0, z: 1.5 }, rounding: 0 }, { join: 'subtract' }),
    ] }
    if (extraCut) shell.children.push(extraCut)
    return shell
}

function roomObjects(r, mats) {
    const kinds = ['sphere', 'box', 'torus', 'cylinder']
    const out = []
    const place = (size) => [(-2.5 + size + r() * (5 - 2 * size)) * 0.9, (-2.0 + size + r() * (4 - 2 * size)) * 0.9, size + r() * (2.2 - 2 * size)]
    const prim = (id, kind, size, at, matId) => {
        const rot = [r() * 360, r() * 360, r() * 360]
        const t = xf(at[0], at[1], at[2], rot[0], rot[1], rot[2])
        if (kind === 'sphere') return obj(id, 'b-sphere', matId, t, { radius: size })
        if (kind === 'box') return obj(id, 'b-box', matId, t, { size: { x: size, y: size * (0.5 + 0.5 * r()), z: size * (0.5 + 0.5 * r()) }, rounding: 0.02 })
        if (kind === 'torus') return obj(id, 'b-torus', matId, t, { ring: size * 0.7, tube: size * 0.3 })
        return obj(id, 'b-cylinder', matId, t, { radius: size * 0.6, height: size * 1.6, rounding: 0.03 })
    }
    const newMat = (i) => {
        const base = [0.2 + 0.7 * r(), 0.2 + 0.7 * r(), 0.2 + 0.7 * r()]
        const m = material('m' + i, base, 0.05 + 0.85 * r(), r() < 0.3 ? 1 : 0)
        mats.push(m)
        return m.id
    }
    let k = 0
    for (let pair = 0; pair < 8; pair++) {
        const s1 = 0.1 + 0.4 * r(), s2 = 0.1 + 0.4 * r()
        const at = place(Math.max(s1, s2))
        const dir = [r() - 0.5, r() - 0.5, r() - 0.5]
        const l = Math.hypot(dir[0], dir[1], dir[2]) || 1
        const at2 = at.map((v, i) => v + dir[i] / l * 0.8 * (s1 + s2) * 0.5)
        at2[2] = Math.max(at2[2], s2)
        const a = prim('p' + k, kinds[k % 4], s1, at, newMat(k)); k++
        const b = prim('p' + k, kinds[k % 4], s2, at2, newMat(k)); k++
        out.push({ id: 'pair' + pair, kind: 'group', label: 'smooth pair', enabled: true, blend: 0.05, transform: xf(0, 0, 0), children: [a, b] })
    }
    while (k < 32) {
        const s = 0.1 + 0.4 * r()
        out.push(prim('p' + k, kinds[k % 4], s, place(s), newMat(k)))
        k++
    }
    return out
}

function panel() {
    return { id: 'panel', kind: 'object', type: 'light', label: 'ceiling panel', enabled: true, material: null,
        transform: xf(0, 0, 2.99, 180, 0, 0), params: { width: 0.5, height: 0.5 }, light: { shape: 'rectangle', luminance: 5000 } }
}

function roomC2(opts = {}) {
    const r = rng(opts.seed || SEED)
    const mats = [material('walls', [0.7, 0.7, 0.7], 1, 0, { specular: 0 })]
    const objs = roomObjects(r, mats)
    const d = doc('room-c2', mats, [roomChildren(), ...objs, panel()], { sky: 'none' }, 10)
    const w = opts.width || 1920, h = opts.height || 1080
    return { doc: d, camera: camera([0, -1.95, 1.6], [0, 0.5, 0.8], 70, w, h), settings: { width: w, height: h, spp: 256, maxbounce: 8, nee: 1 } }
}

function causticC3(opts = {}) {
    const r = rng(opts.seed || SEED)
    const mats = [material('walls', [0.7, 0.7, 0.7], 1, 0, { specular: 0 }),
        // n = 1.5 + 0.0042 / lambda^2 (um) exactly: n_d = 1.5 + 0.0042 / 0.5876^2, V_d = (n_d - 1) / (n_F - n_C)
        material('glass', [1, 1, 1], 0, 0, { transmission: 1, ior: 1.5121643, abbe: 63.83197, specular: 0.5 })]
    const objs = roomObjects(r, mats)
    // the sun (azimuth 26.57, elevation 70) through a 0.3 m beam aimed at the glass: the beam's axis meets the
    // ceiling's mid-plane (z 3.1) where the line from the sphere's centre towards the sun does; the hole about it is
    // 0.4 m across, so the slanted beam clears the 0.2 m slab
    const az = 26.57 * Math.PI / 180, el = 70 * Math.PI / 180
    const sd = [Math.cos(el) * Math.cos(az), Math.cos(el) * Math.sin(az), Math.sin(el)]
    const sphere = [0.3, 0.15, 0.9], t = (3.1 - sphere[2]) / sd[2]
    const hx = sphere[0] + sd[0] * t, hy = sphere[1] + sd[1] * t
    const hole = obj('aperture', 'b-cylinder', 'walls', xf(hx, hy, 1.6), { radius: 0.2, height: 0.6 }, { join: 'subtract' })
    const glass = obj('glass-sphere', 'b-sphere', 'glass', xf(sphere[0], sphere[1], sphere[2]), { radius: 0.4 })
    const hazeBox = { id: 'haze', kind: 'object', type: 'haze', label: 'haze', enabled: true, material: null, transform: xf(0, 0, 1.2),
        params: { size: { x: 2, y: 2, z: 2 } }, haze: { mfp: 20, g: 0.6, albedo: { r: 1, g: 1, b: 1 } } }
    // keep the glass and the beam clear of the random objects: drop those near them
    const clear = objs.filter(o => {
        const c = o.kind === 'group' ? o.children[0].transform.pos : o.transform.pos
        return Math.hypot(c.x - 0.3, c.y - 0.15) > 0.9
    })
    const env = { sky: 'none', sunIlluminance: 100000, uSunAzimuth: 26.57, uSunElevation: 70, sunDiameter: 0.53,
        sunAperture: { point: { x: hx, y: hy, z: 3.1 }, radius: 0.15 } }
    const d = doc('caustic-c3', mats, [roomChildren(hole), ...clear, glass, hazeBox, panel()], env, 10)
    const w = opts.width || 1280, h = opts.height || 720
    return { doc: d, camera: camera([0, -1.95, 1.6], [0.2, 0.5, 0.5], 70, w, h),
        settings: { width: w, height: h, spp: 10, maxbounce: 8, nee: 1, sppm: 2, sppmIterations: 10, sppmPhotons: 1 << 22, sppmRadius: 0.02, sppmAlpha: 0.7,
            lt: 1 } }
}

// A collimated beam through a glass slab onto a floor (the closed form of core/test/sppm.test.mjs's collimated case):
// the sun straight down (100 klx, 0.3 m radius beam about (0, 0, 2)), a 2 x 2 x 0.1 m slab of n = 1.5 at z = 1, a
// Lambertian floor of albedo 0.5 at z = 0 seen from under the slab. At normal incidence each face reflects
// R = ((n - 1) / (n + 1))^2 = 0.04, the slab passes (1 - R) / (1 + R) of the beam with every internal reflection
// counted, and the floor in the beam reads 0.5 x 100000 x 0.96 / 1.04 / pi = 14691.3 cd/m^2 -- all of it caustic,
// the photon map's alone (the shadow rays towards the sun meet the glass).
// Options (for the map-against-path-tracing check): floorRoughness / floorSpecular (a glossy floor), sunDiameter
// (degrees), beam (the beam's radius).
function sppmSlab(opts = {}) {
    const fr = opts.floorRoughness !== undefined ? opts.floorRoughness : 1
    const fsp = opts.floorSpecular !== undefined ? opts.floorSpecular : 0
    const fa = opts.floorAlbedo !== undefined ? opts.floorAlbedo : 0.5
    const mats = [material('floor', [fa, fa, fa], fr, 0, { specular: fsp }),
        material('glass', [1, 1, 1], 0, 0, { transmission: 1, ior: 1.5, specular: 0.5 })]
    const env = { sky: 'none', sunIlluminance: 100000, uSunAzimuth: 0, uSunElevation: 90, sunDiameter: opts.sunDiameter || 0.53,
        sunAperture: { point: { x: 0, y: 0, z: 2 }, radius: opts.beam || 0.3 } }
    const children = [obj('floor', 'b-plane', 'floor', xf(0, 0, 0)),
        obj('slab', 'b-box', 'glass', xf(0, 0, 1), { size: { x: 1, y: 1, z: 0.05 }, rounding: 0 })]
    // with opts.mfp: a haze box under the slab, the beam's track through it (light tracing's connections from the dust)
    if (opts.mfp) children.push({ id: 'haze', kind: 'object', type: 'haze', label: 'haze', enabled: true, material: null,
        transform: xf(0, 0, 0.5), params: { size: { x: 1.2, y: 1.2, z: 0.8 } }, haze: { mfp: opts.mfp, g: opts.g || 0, albedo: { r: 1, g: 1, b: 1 } } })
    const d = doc('sppm-slab', mats, children, env, 20)
    const w = opts.width || 64, h = opts.height || 48
    return { doc: d, camera: camera([0, -0.8, 0.5], [0, 0, 0], 40, w, h),
        settings: { width: w, height: h, spp: 16, maxbounce: 8, nee: 1, sppm: 1, sppmPhotons: 1 << 20, sppmRadius: 0.02, sppmAlpha: 0.7,
            lt: opts.lt || 0 } }
}

function furnacePlane(opts = {}) {
    const mats = [material('white', [1, 1, 1], opts.roughness !== undefined ? opts.roughness : 1, opts.metallic || 0, { specular: opts.specular !== undefined ? opts.specular : 0 })]
    const d = doc('furnace-plane', mats, [obj('floor', 'b-plane', 'white', xf(0, 0, 0))], { sky: 'hemisphere', uniformSky: 1000 }, 20)
    const w = opts.width || 64, h = opts.height || 64
    return { doc: d, camera: camera([0, -3, 1.5], [0, 0, 0], 60, w, h), settings: { width: w, height: h, spp: 64, maxbounce: 16, nee: 1 } }
}

// furnace-sphere: the reference's white furnace (lib/pathtracer/tests.js furnace :1555, testScene 'sphere'): a unit
// sphere of one material in a uniform sphere of sky, 1000 cd/m^2, seen from 3 m with a 40-degree vertical field of view
// (FURNACE_CAMERA), maxbounce 40. A material that neither makes nor loses light vanishes: the picture reads 1000.
function furnaceSphere(opts = {}) {
    const num = (v, d) => v !== undefined ? v : d
    const b = num(opts.albedo, 1)
    const mats = [material('white', [b, b, b], num(opts.roughness, 1), num(opts.metallic, 0),
        { specular: num(opts.specular, 0), transmission: num(opts.transmission, 0), ior: num(opts.ior, 1.5), abbe: num(opts.abbe, 0) })]
    const d = doc('furnace-sphere', mats, [obj('ball', 'b-sphere', 'white', xf(0, 0, 0), { radius: 1 })], { sky: 'sphere', uniformSky: 1000 }, 20)
    const w = opts.width || 64, h = opts.height || 64
    const el = 30 * Math.PI / 180, az = 30 * Math.PI / 180
    const pos = [3 * Math.cos(el) * Math.cos(az), 3 * Math.cos(el) * Math.sin(az), 3 * Math.sin(el)]
    // the surface tolerance the reference traces this scene at: photometry.js autoEps, 1e-4 of its extent (tests.js EXTENT
    // sphere: 2 sqrt 3 m); with maxiter 400 over a path, what a path trapped in rough glass loses depends on it
    const epsAuto = 1e-4 * 2 * Math.sqrt(3)
    return { doc: d, camera: camera(pos, [0, 0, 0], 40, w, h), settings: { width: w, height: h, spp: 64, maxbounce: 40, nee: 1, maxiter: 400, epsAuto } }
}

// sky-code: a floor and a sphere under the studio written as the scene's own sky code (builder/sky.js: background(dir)
// with its parameters), the sphere's code calling a function of the document's functions block. With { native: true }
// the same scene under the builtin studio (envMode 1, 1000 lx, top 0.5), so the two can be held to one another.
function skyCode(opts = {}) {
    const { cameraOf } = require('./hipgen')
    const top = 0.5, E = 1000
    const L0 = E / (Math.PI * (1 + 2 * top / 3))
    const mats = [material('floor', [0.7, 0.7, 0.7], 1, 0, { specular: 0 }), material('ball', [0.8, 0.3, 0.2], 0.4, 0)]
    const blocks = BLOCKS.concat([{ id: 'u-ball', label: 'ball', kind: 'object', code: 'C = add_new(C, ball(p, radius), MATERIAL);', params: { radius: 40 } }])
    const d = studioDoc('sky-code', mats, [obj('floor', 'b-plane', 'floor', xf(0, 0, 0)), obj('ball', 'u-ball', 'ball', xf(0, 0, 40), {})], blocks,
        { mode: 'orbit', camtype: 0, fov: 40, target: [0, 0, 30], azimuth: -90, elevation: 70, distance: 400, up: [0, 0, 1], aperture: 0 })
    d.environment = { envMode: 1, studioIlluminance: E, studioTop: top, studioColor: { r: 1, g: 1, b: 1 } }
    if (!opts.native) {
        d.functions = `// a ball of radius r: its outward normal and distance
fn ball(q: vec3f, r: f32) -> vec4f {
    let l = length(q);
    return vec4f(select(vec3f(0.0, 0.0, 1.0), q / l, l > 0.0), l - r);
}`
        d.sky = `// the studio: brighter above
const BASE: f32 = 1.0;
fn background(dir: VEC3) -> VEC3 {
    return VEC3(L0 * max(BASE + top * dir.z, 0.0));
}`
        d.skyParams = [{ name: 'L0', value: L0 }, { name: 'top', value: top }]
    } else {
        d.blocks = BLOCKS.concat([{ id: 'u-ball', label: 'ball', kind: 'object', builtin: 'sphere@1', params: { radius: 40 } }])
        d.root.children[1].params = { radius: 40 }
    }
    const w = opts.width || 64, h = opts.height || 48
    return { doc: d, camera: cameraOf(d, w, h), settings: { width: w, height: h, spp: 64, maxbounce: 6, nee: 1 } }
}

// absorb-slab: a smooth glass slab (n 1.5) 20 mm thick in a uniform sphere of sky, its glass filtering light to a
// colour of (0.5, 0.7, 0.9) at 10 mm (materials.js attenuationDistance), seen head-on: the centre reads
// L (R + (1 - R)^2 (R tau^2 + tau) / (1 - R^2 tau^2)) per channel, R = 0.04 the face's Fresnel reflectance head-on and
// tau = colour^(20 / 10) the inside's transmittance (the series of passes back and forth)
function absorbSlab(opts = {}) {
    const att = opts.attenuation !== undefined ? opts.attenuation : 10
    const mats = [material('glass', [0.5, 0.7, 0.9], 0, 0, { specular: 0.5, transmission: 1, ior: 1.5, attenuationDistance: att })]
    const d = doc('absorb-slab', mats, [obj('slab', 'b-box', 'glass', xf(0, 0, 0), { size: { x: 200, y: 10, z: 200 }, rounding: 0 })],
        { sky: 'sphere', uniformSky: 1000 }, 20000)
    d.unit = 'mm'
    const w = opts.width || 32, h = opts.height || 32
    // a fine tolerance: the march stops within it of a face, and the path counted inside differs from the slab's by
    // up to it at each face (2 um against 20 mm here)
    return { doc: d, camera: camera([0, -0.5, 0], [0, 0, 0], 4, w, h), settings: { width: w, height: h, spp: 64, maxbounce: 40, nee: 1, eps: 2e-6 } }
}

function furnaceHaze(opts = {}) {
    // a haze box of albedo 1 in a uniform sphere of sky: every pixel reads the sky's luminance (energy conservation)
    const g = opts.g !== undefined ? opts.g : 0.6
    const hazeBox = { id: 'haze', kind: 'object', type: 'haze', label: 'haze', enabled: true, material: null, transform: xf(0, 0, 0),
        params: { size: { x: 2, y: 2, z: 2 } }, haze: { mfp: opts.mfp || 0.5, g, albedo: { r: 1, g: 1, b: 1 } } }
    // a far sphere marks nothing: the scene needs a surface, so a tiny sphere far outside the view
    const mats = [material('none', [0, 0, 0], 1, 0, { specular: 0 })]
    const d = doc('furnace-haze', mats, [obj('speck', 'b-sphere', 'none', xf(0, 0, -50), { radius: 0.01 }), hazeBox], { sky: 'sphere', uniformSky: 1000 }, 20)
    const w = opts.width || 64, h = opts.height || 64
    return { doc: d, camera: camera([0, -4, 0], [0, 0, 0], 30, w, h), settings: { width: w, height: h, spp: 64, maxbounce: 250, nee: 1 } }
}

// Documents as the builder writes them (millimetres, the studio environment -- envMode 1, no `sky` of ours -- lights by
// CCT and flux, the camera as an orbit view), for the emitter's reading of the reference's own scene format.
// studio-floor: a Lambertian floor of albedo 0.6 under a studio of 500 lx (top 0.5): the floor reads 0.6 x 500 / pi
// cd/m^2 exactly, whatever the top light; the sky above the horizon L0 (1 + 0.5 d.z).
function studioDoc(title, materials, children, blocks, view) {
    return { format: 'raytracer-scene', version: 31, title, unit: 'mm', envRadius: 20000, materials, blocks,
        root: { id: 'root', kind: 'group', label: title, enabled: true, blend: 0, children },
        environment: { envMode: 1, studioIlluminance: 500, studioTop: 0.5, studioColor: { r: 1, g: 1, b: 1 } }, view: { camera: view } }
}
function studioFloor(opts = {}) {
    const { cameraOf } = require('./hipgen')
    const albedo = opts.albedo !== undefined ? opts.albedo : 0.6
    const d = studioDoc('studio-floor', [material('floor', [albedo, albedo, albedo], 1, 0, { specular: 0 })],
        [obj('floor', 'b-plane', 'floor', xf(0, 0, 0))], BLOCKS,
        { mode: 'orbit', camtype: 0, fov: 50, target: [0, 0, 0], azimuth: -90, elevation: 70, distance: 2000, up: [0, 0, 1], aperture: 0 })
    const w = opts.width || 64, h = opts.height || 48
    return { doc: d, camera: cameraOf(d, w, h), settings: { width: w, height: h, spp: 64, maxbounce: 4, nee: 1 } }
}
// lens-prism: a plano-convex glass lens (60 mm across, 12 mm thick, r1 62 mm, flat back) standing with its axis along
// -y in front of a glass prism (triangle of 30 mm corner radius, 80 mm long) lying on a grey floor; a 40 x 40 mm
// panel of 3000 K, 200 lm behind and above them, and the studio of 500 lx. The lens and the prism are the builder's
// builtins (lens@1, prism@1).
function lensPrism(opts = {}) {
    const { cameraOf } = require('./hipgen')
    const blocks = BLOCKS.concat([
        { id: 'b-lens', label: 'lens', kind: 'object', builtin: 'lens@1', params: { diameter: 1, thickness: 0.2, r1: 2, r2: 2 } },
        { id: 'b-prism', label: 'prism', kind: 'object', builtin: 'prism@1', params: { sides: 3, radius: 0.5, height: 1 } }])
    const glass = material('glass', [1, 1, 1], 0, 0, { transmission: 1, ior: 1.5168, abbe: 64.17, specular: 0.5 })
    const d = studioDoc('lens-prism', [material('floor', [0.5, 0.5, 0.5], 0.8, 0, { specular: 0.5 }), glass], [
        obj('floor', 'b-plane', 'floor', xf(0, 0, 0)),
        obj('lens', 'b-lens', 'glass', xf(-40, -30, 45, 90, 0, 0), { diameter: 60, thickness: 12, r1: 62, r2: 0 }),
        obj('prism', 'b-prism', 'glass', xf(45, 20, 0.0, 0, 90, 0), { sides: 3, radius: 30, height: 80 }),
        { id: 'panel', kind: 'object', type: 'light', label: 'panel', enabled: true, material: null, transform: xf(0, 150, 160, 120, 0, 0),
            params: { width: 40, height: 40 }, light: { shape: 'rectangle', colourBy: 'cct', cct: 3000, duv: 0, flux: 200 } },
    ], blocks, { mode: 'orbit', camtype: 0, fov: 40, target: [0, 0, 30], azimuth: -100, elevation: 62, distance: 260, up: [0, 0, 1], aperture: 0 })
    const w = opts.width || 96, h = opts.height || 64
    return { doc: d, camera: cameraOf(d, w, h), settings: { width: w, height: h, spp: 64, maxbounce: 8, nee: 1 } }
}

// reflector-condenser: a mirror reflector (focal 8 mm, aperture 25 mm, wall 1 mm) on a post, opening up, with a 3 mm
// disc of 200 lm at its focus facing into the bowl; and the aspheric condenser builtin at its defaults, 30 times (a
// 30 mm condenser of k -2.25 over a flat back) standing on a grey floor; the studio of 500 lx.
function reflectorCondenser(opts = {}) {
    const { cameraOf } = require('./hipgen')
    const blocks = BLOCKS.concat([
        { id: 'b-refl', label: 'reflector', kind: 'object', builtin: 'reflector@1', params: { focal: 0.125, aperture: 0.5, wall: 0.01, solid: 0, focusOffset: 0 } },
        { id: 'b-asph', label: 'aspheric lens', kind: 'object', builtin: 'aspheric lens@1', params: { diameter: 1, thickness: 0.3, r1: 0.5, k1: -2.25,
            a4_1: 0, a6_1: 0, a8_1: 0, a10_1: 0, r2: 0, k2: 0, a4_2: 0, a6_2: 0, a8_2: 0, a10_2: 0 } }])
    const mirror = material('mirror', [0.95, 0.95, 0.95], 0.02, 1)
    const glass = material('glass', [1, 1, 1], 0, 0, { transmission: 1, ior: 1.5, specular: 0.5 })
    const d = studioDoc('reflector-condenser', [material('floor', [0.5, 0.5, 0.5], 0.8, 0, { specular: 0.5 }), mirror, glass], [
        obj('floor', 'b-plane', 'floor', xf(0, 0, 0)),
        obj('post', 'b-cylinder', 'floor', xf(-30, 0, 14), { radius: 4, height: 28 }),
        obj('reflector', 'b-refl', 'mirror', xf(-30, 0, 37), { focal: 8, aperture: 25, wall: 1 }),
        { id: 'lamp', kind: 'object', type: 'light', label: 'lamp', enabled: true, material: null, transform: xf(-30, 0, 37, 180, 0, 0),
            params: { radius: 1.5 }, light: { shape: 'disc', colourBy: 'cct', cct: 4000, flux: 200 } },
        obj('condenser', 'b-asph', 'glass', xf(35, 0, 15, 90, 0, 0, 30), {}),
    ], blocks, { mode: 'orbit', camtype: 0, fov: 40, target: [0, 0, 25], azimuth: -90, elevation: 62, distance: 150, up: [0, 0, 1], aperture: 0 })
    const w = opts.width || 96, h = opts.height || 64
    return { doc: d, camera: cameraOf(d, w, h), settings: { width: w, height: h, spp: 64, maxbounce: 8, nee: 1 } }
}

// turned-parts: outlines -- a Fresnel lens (six facet rings over a flat back, 40 mm across, glass) held up on a post, a
// knob turned from a filleted outline and a notched plate extruded from one, on a grey floor under a 3000 K panel
// and the studio.
function turnedParts(opts = {}) {
    const { cameraOf } = require('./hipgen')
    const fresnel = [[0, 0], [0.5, 0], [0.5, 0.025], [0.4167, 0.0918], [0.4167, 0.025], [0.3333, 0.0746], [0.3333, 0.025], [0.25, 0.0611],
        [0.25, 0.025], [0.1667, 0.0497], [0.1667, 0.025], [0.0833, 0.0394], [0.0833, 0.025], [0, 0.0297]]
    const knob = [[0, -0.42], [0.5, -0.42, 0.08], [0.5, 0.08], [0.21, 0.25, 0.125], [0.33, 0.5, 0.08], [0, 0.5]]
    const plate = [[-0.5, -0.35, 0.08], [0.5, -0.35, 0.08], [0.5, 0.35, 0.08], [0.125, 0.35, 0.035], [0, 0.125, 0.05], [-0.125, 0.35, 0.035], [-0.5, 0.35, 0.08]]
    const blocks = BLOCKS.concat([
        { id: 'b-fresnel', label: 'fresnel lens', kind: 'object', builtin: 'fresnel lens@1', params: { outline: fresnel } },
        { id: 'b-turn', label: 'revolution', kind: 'object', builtin: 'revolution@1', params: { outline: knob } },
        { id: 'b-extr', label: 'extrusion', kind: 'object', builtin: 'extrusion@1', params: { outline: plate, height: 0.2 } }])
    const glass = material('glass', [1, 1, 1], 0, 0, { transmission: 1, ior: 1.49, specular: 0.5 })
    const d = studioDoc('turned-parts', [material('floor', [0.5, 0.5, 0.5], 0.8, 0, { specular: 0.5 }), glass,
        material('brass', [0.9, 0.7, 0.35], 0.3, 1), material('paint', [0.2, 0.35, 0.7], 0.5, 0)], [
        obj('floor', 'b-plane', 'floor', xf(0, 0, 0)),
        obj('post', 'b-cylinder', 'paint', xf(-35, 10, 15, 0, 0, 0), { radius: 3, height: 30 }),
        obj('fresnel', 'b-fresnel', 'glass', xf(-35, 10, 30, 0, 0, 0, 40), {}),
        obj('knob', 'b-turn', 'brass', xf(5, 15, 17, 0, 0, 0, 40), {}),
        obj('plate', 'b-extr', 'paint', xf(40, -5, 4, 0, 0, 25, 40), { height: 0.2 }),
        { id: 'panel', kind: 'object', type: 'light', label: 'panel', enabled: true, material: null, transform: xf(0, 60, 120, 150, 0, 0),
            params: { width: 40, height: 40 }, light: { shape: 'rectangle', colourBy: 'cct', cct: 3000, flux: 300 } },
    ], blocks, { mode: 'orbit', camtype: 0, fov: 40, target: [0, 5, 15], azimuth: -95, elevation: 62, distance: 170, up: [0, 0, 1], aperture: 0 })
    const w = opts.width || 96, h = opts.height || 64
    return { doc: d, camera: cameraOf(d, w, h), settings: { width: w, height: h, spp: 64, maxbounce: 8, nee: 1 } }
}

// gallery: the remaining solid builtins at their defaults, 40 times (millimetres) -- a cone, a round cone, an
// octahedron, a capped torus (270 degrees), a chain link, a tube, a dome and a pyramid -- in a row on a grey floor,
// a 3000 K panel and the studio.
function gallery(opts = {}) {
    const { cameraOf } = require('./hipgen')
    const kinds = ['cone', 'round cone', 'octahedron', 'capped torus', 'link', 'tube', 'dome', 'pyramid']
    const blocks = BLOCKS.concat(kinds.map(k => ({ id: `b-${k}`, label: k, kind: 'object', builtin: `${k}@1`, params: {} })))
    const mats = [material('floor', [0.5, 0.5, 0.5], 0.8, 0, { specular: 0.5 }), material('brass', [0.9, 0.7, 0.35], 0.3, 1),
        material('paint', [0.2, 0.35, 0.7], 0.5, 0)]
    // each sits on the floor: its local z extent at scale 40 (the octahedron on a corner, the torus and link on edge)
    const place = { cone: [20, 0], 'round cone': [23, 0], octahedron: [20, 0], 'capped torus': [15.1, 90], link: [20, 90], tube: [20, 0], dome: [0, 0], pyramid: [20, 0] }
    const children = [obj('floor', 'b-plane', 'floor', xf(0, 0, 0))].concat(kinds.map((k, i) =>
        obj(k.replace(' ', '-'), `b-${k}`, i % 2 ? 'brass' : 'paint', xf(-147 + 42 * i, 8 * (i % 2), place[k][0], place[k][1], 0, 0, 40), {})))
    children.push({ id: 'panel', kind: 'object', type: 'light', label: 'panel', enabled: true, material: null, transform: xf(0, 80, 140, 150, 0, 0),
        params: { width: 60, height: 40 }, light: { shape: 'rectangle', colourBy: 'cct', cct: 3000, flux: 400 } })
    const d = studioDoc(opts.sun ? 'sunlit' : 'gallery', mats, children, blocks,
        { mode: 'orbit', camtype: 0, fov: 40, target: [0, 0, 15], azimuth: -90, elevation: 65, distance: 230, up: [0, 0, 1], aperture: 0 })
    if (opts.sun) d.environment = { envMode: 0, uSunAzimuth: -140, uSunElevation: 35, sunDiameter: 0.53, sunOutside: 128000, uSkyTurbidity: 0.7, sunHint: 1 }
    const w = opts.width || 160, h = opts.height || 64
    return { doc: d, camera: cameraOf(d, w, h), settings: { width: w, height: h, spp: 64, maxbounce: 8, nee: 1 } }
}

// modifiers: the builtin modifier groups (builder/modifiers.js) in millimetres, laid out as the reference's
// scenes/modifiers.json lays its own: a twisted bar, a bent plank, a repeated ball, a mirrored torus, a rounded cube,
// a shelled ball with a box cut out of it, and two circular arrays (balls, and boxes overlapping into one ring); the
// studio's light and a CCT panel. Options: `only` a modifier's name (that group alone over the floor), and `rate` to
// set the twist's (0: the untwisted bar).
function modifiers(opts = {}) {
    const { cameraOf } = require('./hipgen')
    const mods = ['twist', 'bend', 'repeat', 'mirror', 'round', 'shell', 'circular']
    const blocks = BLOCKS.concat(mods.map(k => ({ id: `b-${k}`, label: k, kind: 'group', builtin: `${k}@1`, params: {} })))
    const mats = [material('floor', [0.5, 0.5, 0.5], 0.8, 0, { specular: 0.5 }), material('brass', [0.9, 0.7, 0.35], 0.3, 1),
        material('paint', [0.2, 0.35, 0.7], 0.5, 0), material('clay', [0.75, 0.4, 0.3], 0.9, 0, { specular: 0.2 })]
    const group = (id, kind, at, params, children) => ({ id, kind: 'group', label: kind, enabled: true, blend: 0, block: `b-${kind}`,
        transform: at, params, children })
    const rate = opts.rate !== undefined ? opts.rate : 0.039269908
    const G = {
        twist: group('twisted', 'twist', xf(-135, 50, 0), { rate, extent: 22 },
            [obj('bar', 'b-box', 'paint', xf(0, 0, 40), { size: { x: 14, y: 14, z: 40 }, rounding: 1 })]),
        bend: group('bent', 'bend', xf(-45, 50, 25), { amount: 0.01, extent: 40 },
            [obj('plank', 'b-box', 'brass', xf(0, 0, 0), { size: { x: 40, y: 8, z: 4 }, rounding: 1 })]),
        repeat: group('repeated', 'repeat', xf(45, 50, 7), { spacing: { x: 20, y: 20, z: 20 }, count: { x: 1, y: 1, z: 0 } },
            [obj('ball', 'b-sphere', 'clay', xf(0, 0, 0), { radius: 7 })]),
        mirror: group('mirrored', 'mirror', xf(135, 50, 0), {},
            [obj('ring', 'b-torus', 'brass', xf(16, 0, 25, 90, 0, 30), { ring: 14, tube: 5 })]),
        round: group('rounded', 'round', xf(-135, -50, 0), { radius: 5 },
            [obj('cube', 'b-box', 'brass', xf(0, 0, 22), { size: { x: 17, y: 17, z: 17 }, rounding: 0 })]),
        shell: { id: 'cut', kind: 'group', label: 'cut shell', enabled: true, blend: 0, transform: xf(-45, -50, 0), children: [
            group('shelled', 'shell', xf(0, 0, 0), { thickness: 2 }, [obj('globe', 'b-sphere', 'clay', xf(0, 0, 25), { radius: 25 })]),
            obj('window', 'b-box', 'clay', xf(-18, -18, 45), { size: { x: 20, y: 20, z: 20 }, rounding: 0 }, { join: 'subtract' })] },
        circular: group('bolts', 'circular', xf(45, -50, 0), { count: 8, reach: 32 },
            [obj('bolt', 'b-sphere', 'paint', xf(24, 0, 7), { radius: 7 })]),
    }
    const ring = group('ring', 'circular', xf(135, -50, 0), { count: 12, reach: 32 },
        [obj('tooth', 'b-box', 'brass', xf(6, 0, 8), { size: { x: 18, y: 6, z: 8 }, rounding: 1 })])
    const chosen = opts.only ? [G[opts.only]] : mods.map(k => G[k]).concat([ring])
    const children = [obj('floor', 'b-plane', 'floor', xf(0, 0, 0))].concat(chosen)
    children.push({ id: 'panel', kind: 'object', type: 'light', label: 'panel', enabled: true, material: null, transform: xf(0, 100, 160, 150, 0, 0),
        params: { width: 80, height: 50 }, light: { shape: 'rectangle', colourBy: 'cct', cct: 3500, flux: 600 } })
    const d = studioDoc('modifiers', mats, children, blocks,
        { mode: 'orbit', camtype: 0, fov: 45, target: [0, 0, 20], azimuth: -90, elevation: 55, distance: 420, up: [0, 0, 1], aperture: 0 })
    const w = opts.width || 160, h = opts.height || 96
    return { doc: d, camera: cameraOf(d, w, h), settings: { width: w, height: h, spp: 64, maxbounce: 8, nee: 1 } }
}

// usercode: the modifiers scene with every modifier written as the document's own group code (WGSL the emitter
// translates, not the native modifiers), and two objects of their own code: a cone from its closed form and a Menger
// sponge of three levels. With { native: true } the same groups use the builtin modifiers, so the two documents
// describe one scene: what the translation draws can be held to what the native code draws.
const USER_CODE = {
    twist: `// turned about z by rate radians per unit up
let a = rate * p.z;
let c = cos(a);
let s = sin(a);
p = vec3<f32>(c * p.x - s * p.y, s * p.x + c * p.y, p.z);
warp(1.0 / sqrt(1.0 + rate * rate * extent * extent));
C = children(C, p);`,
    bend: `let a = amount * p.x;
p = VEC3(cos(a) * p.x - sin(a) * p.z, p.y, sin(a) * p.x + cos(a) * p.z);
warp(1.0 / sqrt(1.0 + amount * amount * extent * extent));
C = children(C, p);`,
    repeat: `p = p - spacing * clamp(round(p / spacing), -count, count);
C = children(C, p);`,
    mirror: `p.x = abs(p.x);
warp(1.0);
C = children(C, p);`,
    round: `var H = children(emptyHit(r), p);
H.s.w -= radius * XF_S;
C = fold(C, H.s, H.m);`,
    shell: `var H = children(emptyHit(r), p);
H.s = VEC4(H.s.xyz * sign(H.s.w), abs(H.s.w) - thickness * XF_S);
C = fold(C, H.s, H.m);`,
    circular: `/* every copy, turned by i / n of a turn */
let n = u32(clamp(round(count), 1.0, 64.0));
let slice = 6.2831853 / f32(n);
let xfm = XF_M;
for (var i = 0u; i < n; i++) {
    let t = f32(i) * slice;
    let c = cos(t); let s = sin(t);
    XF_M = xfm * mat3x3<f32>(c, s, 0.0, -s, c, 0.0, 0.0, 0.0, 1.0);
    C = children(C, VEC3(c * p.x + s * p.y, c * p.y - s * p.x, p.z));
}
XF_M = xfm;`,
    cone: `// a cone's side between z = -height/2 and height/2, radius base below and top above, and its caps
let k = (top - base) / height;
let c = inverseSqrt(1.0 + k * k);
let rho = length(p.xy);
let side = (rho - (0.5 * (base + top) + k * p.z)) * c;
let caps = abs(p.z) - 0.5 * height;
let radial = select(vec2f(1.0, 0.0), p.xy / rho, rho > 1e-12);
var f = VEC4(VEC3(radial * c, -k * c), side);
if (caps > side) { f = VEC4(0.0, 0.0, sign(p.z), caps); }
C = add_new(C, f, MATERIAL);`,
    menger: `let s = size * 0.5;
C = add_new(C, sdMenger(p / s, iterations) * s, MATERIAL);`,
}
// the materials' own code: a checker of the floor in the world, grain in the paint and stones in the clay in their
// objects' frames (value and gradient noise, fBm, crossed Worley cells)
const MATERIAL_CODE = {
    floor: `let q = p / mm(1.0);
let w = cos(PI * q / square);
m.baseColor = mix(m.baseColor, second, step(0.0, w.x * w.y));`,
    paint: `let q = p_local * S_local / mm(1.0);
let bend = gnoise3(q / vec3f(6.0, 6.0, 24.0), 7u).w;
let ring = fract(length(q.xy) / 2.5 + 0.6 * bend);
let late = smoothstep(0.7, 1.0, ring);
let fib = 0.5 + 0.5 * vnoise3(q / vec3f(0.4, 0.4, 8.0), 8u).w;
m.baseColor = mix(m.baseColor, latewood, late) * (0.8 + 0.4 * fib);
m.roughness = min(1.0, m.roughness + 0.2 * late);`,
    clay: `let q = p_local * S_local / mm(1.0);
let w = worley3x(q / 4.0, 12u);
let f = 0.5 + 0.5 * fbm3(q / 10.0, 3u, 4u, 2.0, 0.5).w;
m.baseColor = m.baseColor * (0.6 + 0.4 * f) * (1.0 - 0.5 * (1.0 - smoothstep(0.0, 0.3, w.y - w.x)));`,
}
function usercode(opts = {}) {
    const made = modifiers(Object.assign({}, opts, { width: opts.width || 160, height: opts.height || 96 }))
    const d = made.doc
    if (!opts.noMaterialCode) {
        const extra = { floor: { square: 20, second: { r: 0.15, g: 0.15, b: 0.15 } }, paint: { latewood: { r: 0.3, g: 0.15, b: 0.05 } }, clay: {} }
        for (const m of d.materials) {
            if (!MATERIAL_CODE[m.id]) continue
            m.code = MATERIAL_CODE[m.id]
            Object.assign(m.params, extra[m.id])
        }
    }
    if (!opts.native) {
        // each modifier group's block: its code in place of the builtin, the same parameters
        d.blocks = d.blocks.map(b => {
            const k = b.builtin && b.builtin.split('@')[0]
            return USER_CODE[k] && b.kind === 'group' ? { id: b.id, label: b.label, kind: 'group', code: USER_CODE[k], params: b.params } : b
        })
    }
    if (!opts.only) {
        // two tiles wearing reliefs: ripples of a closed form, and fBm plaster
        d.materials.push(Object.assign(material('tile', [0.6, 0.55, 0.5], 0.4, 0), { relief: { code: `let k = 6.2831853 / size;
h = amplitude * sin(k * q.x) * cos(k * q.y);
slope = amplitude * k;`, params: { amplitude: 0.5, size: 8 } } }),
        Object.assign(material('plaster', [0.8, 0.78, 0.74], 0.9, 0), { relief: { code: `let peak = nzFbmLip(4u, 1.0, 0.5, 1.0);
let n = fbm3(q / size, 1u, 4u, 2.0, 0.5);
h = amplitude * n.w / peak;
slope = amplitude / size * nzFbmLip(4u, 2.0, 0.5, 4.3) / peak;`, params: { amplitude: 1.0, size: 6 } } }))
        d.root.children.push(obj('tile', 'b-box', 'tile', xf(-60, -125, 3), { size: { x: 25, y: 25, z: 3 }, rounding: 0 }),
            obj('plaster', 'b-box', 'plaster', xf(60, -125, 3), { size: { x: 25, y: 25, z: 3 }, rounding: 0 }))
        d.blocks = d.blocks.concat([
            { id: 'u-cone', label: 'cone', kind: 'object', code: USER_CODE.cone, params: { base: 10, top: 4, height: 30 } },
            { id: 'u-menger', label: 'Menger sponge', kind: 'object', code: USER_CODE.menger, params: { size: 30, iterations: 3 },
                constants: { iterations: 'u32' } },
        ])
        d.root.children.splice(1, 0, obj('cone', 'u-cone', 'paint', xf(0, 0, 15), {}), obj('sponge', 'u-menger', 'clay', xf(0, 120, 15, 0, 0, 20), {}))
    }
    return made
}

// sunlit: the gallery under the builder's default environment, the sun & sky (envMode 0): the sun 35 degrees up at
// azimuth -140, 128 klx above the air, turbidity 0.7.
const sunlit = (opts = {}) => gallery(Object.assign({}, opts, { sun: true }))
// sunsky-floor: a Lambertian floor of albedo 0.6 under the sun & sky, the sun 40 degrees up: every pixel that sees the
// floor reads 0.6 / pi (E_sun sin 40 + E_sky), the sun's light head-on times the sine of its elevation plus the sky's
// on the level.
function sunskyFloor(opts = {}) {
    const { cameraOf } = require('./hipgen')
    const d = studioDoc('sunsky-floor', [material('floor', [0.6, 0.6, 0.6], 1, 0, { specular: 0 })],
        [obj('floor', 'b-plane', 'floor', xf(0, 0, 0))], BLOCKS,
        { mode: 'orbit', camtype: 0, fov: 50, target: [0, 0, 0], azimuth: -90, elevation: 50, distance: 2000, up: [0, 0, 1], aperture: 0 })
    d.environment = { envMode: 0, uSunAzimuth: 30, uSunElevation: 40, sunOutside: 128000, uSkyTurbidity: 0.7 }
    const w = opts.width || 64, h = opts.height || 48
    return { doc: d, camera: cameraOf(d, w, h), settings: { width: w, height: h, spp: 64, maxbounce: 4, nee: 1 } }
}

// sunsky-bench: the reference's photometry bench (lib/pathtracer/tests.js photometry :447, illuminance :553, cosine-law
// :613, bsdf-parity :1868): a plane through the origin under the sun & sky -- the sun at azimuth 45, elevation 45, 128
// klx above the air, turbidity 0.7 -- seen by FLOOR_CAMERA (straight down from 5 units, fov 40) or, with oblique,
// OBLIQUE_CAMERA (70 degrees off vertical); the floor Lambertian of albedo 0.8 (floorScene) unless the options give its
// material, tilted towards the sun's bearing by `tilt` degrees (testScene 'tilted'); in metres or millimetres (unit).
// Options: albedo, metallic, roughness, specular, transmission, tilt, unit, oblique, and the environment's envSun,
// envSky, uSunElevation, uSkyTurbidity, sunDiameter.
function sunskyBench(opts = {}) {
    const { cameraOf } = require('./hipgen')
    const num = (v, d) => v !== undefined ? v : d
    const b = num(opts.albedo, 0.8)
    const mat = material('floor', [b, b, b], num(opts.roughness, 1), num(opts.metallic, 0),
        { specular: num(opts.specular, 0), transmission: num(opts.transmission, 0), ior: 1.5 })
    const tilt = num(opts.tilt, 0), az = 45
    // the plane's normal (+z of its frame) tilted by `tilt` towards the sun's bearing (sin az, cos az, 0): Rz(90 - az) Ry(tilt)
    const floor = obj('floor', 'b-plane', 'floor', xf(0, 0, 0, 0, tilt, tilt ? 90 - az : 0))
    const unit = opts.unit || 'm'
    const d = { format: 'raytracer-scene', version: 31, title: 'sunsky-bench', unit, envRadius: 400 / UNIT[unit], materials: [mat], blocks: BLOCKS,
        root: { id: 'root', kind: 'group', label: 'bench', enabled: true, blend: 0, children: [floor] },
        environment: { envMode: 0, uSunAzimuth: az, uSunElevation: num(opts.uSunElevation, 45), uSkyTurbidity: num(opts.uSkyTurbidity, 0.7),
            sunOutside: 128000, sunDiameter: num(opts.sunDiameter, 0.53), envSun: num(opts.envSun, 1), envSky: num(opts.envSky, 1) },
        view: { camera: { mode: 'orbit', camtype: 0, target: [0, 0, 0], azimuth: 30, elevation: opts.oblique ? 70 : 8, distance: 5, fov: 40 } } }
    // the studio instead (envMode 1): opts.studio = { lux, top, tint }
    if (opts.studio) {
        const t = opts.studio.tint || [1, 1, 1]
        d.environment = { envMode: 1, studioIlluminance: opts.studio.lux, studioTop: opts.studio.top, studioColor: { r: t[0], g: t[1], b: t[2] } }
    }
    const w = opts.width || 32, h = opts.height || 32
    // the reference's bench plane says it is 10 units across (tests.js EXTENT plane): its tolerance is autoEps of that
    const epsAuto = Math.min(Math.max(1e-4 * 10 * UNIT[unit], 1e-8), 1e-3)
    return { doc: d, camera: cameraOf(d, w, h), settings: { width: w, height: h, spp: 32, maxbounce: 10, nee: 1, maxiter: 400, epsAuto } }
}
const UNIT = { m: 1, mm: 0.001 }

// joins: the builder's shaped joins (joinShape, model version 28) on millimetre parts -- a box intersected with a
// ball by a round, a box with a cylinder cut from it by a chamfer, two boxes unioned by a round (a concave fillet), and
// a chamfered union -- on a grey floor under the studio and a 4000 K panel.
function joins(opts = {}) {
    const { cameraOf } = require('./hipgen')
    const mats = [material('floor', [0.5, 0.5, 0.5], 0.8, 0, { specular: 0.5 }), material('brass', [0.9, 0.7, 0.35], 0.3, 1),
        material('paint', [0.2, 0.35, 0.7], 0.5, 0)]
    const box = (id, sx, sy, sz, t, extra) => Object.assign(obj(id, 'b-box', null, t, { size: { x: sx, y: sy, z: sz }, rounding: 0 }), extra || {})
    const part = (id, mat, t, kids) => ({ id, kind: 'group', label: id, enabled: true, blend: 0, transform: t, children: kids.map(k => Object.assign(k, { material: mat })) })
    const children = [obj('floor', 'b-plane', 'floor', xf(0, 0, 0)),
        part('rounded', 'paint', xf(-75, 0, 20), [box('r-box', 18, 18, 18, xf(0, 0, 0)),
            Object.assign(obj('r-ball', 'b-sphere', null, xf(0, 0, 0), { radius: 24 }), { join: 'intersect', joinBlend: 5, joinShape: 'round' })]),
        part('chamfered', 'brass', xf(-25, 0, 20), [box('c-box', 18, 18, 18, xf(0, 0, 0)),
            Object.assign(obj('c-hole', 'b-cylinder', null, xf(0, 0, 0), { radius: 9, height: 60 }), { join: 'subtract', joinBlend: 4, joinShape: 'chamfer' })]),
        part('filleted', 'paint', xf(25, 0, 14), [box('f-base', 18, 18, 6, xf(0, 0, -8)),
            box('f-post', 7, 7, 14, xf(0, 0, 4), { join: 'union', joinBlend: 6, joinShape: 'round' })]),
        part('bevelled', 'brass', xf(75, 0, 14), [box('b-base', 18, 18, 6, xf(0, 0, -8)),
            box('b-post', 7, 7, 14, xf(0, 0, 4), { join: 'union', joinBlend: 5, joinShape: 'chamfer' })]),
    ]
    children.push({ id: 'panel', kind: 'object', type: 'light', label: 'panel', enabled: true, material: null, transform: xf(0, 80, 140, 150, 0, 0),
        params: { width: 60, height: 40 }, light: { shape: 'rectangle', colourBy: 'cct', cct: 4000, flux: 400 } })
    const d = studioDoc('joins', mats, children, BLOCKS,
        { mode: 'orbit', camtype: 0, fov: 40, target: [0, 0, 12], azimuth: -70, elevation: 60, distance: 170, up: [0, 0, 1], aperture: 0 })
    const w = opts.width || 160, h = opts.height || 64
    return { doc: d, camera: cameraOf(d, w, h), settings: { width: w, height: h, spp: 64, maxbounce: 8, nee: 1 } }
}

// THE OPTICS TRACE'S TEST SCENES (lib/pathtracer/tests.js optics-balance :1060-1100, optics-farfield :1101-1137,
// optics-refraction :1138-1200): a floor, a closed shell round a point source, a glass slab over a floor, a glass ball.
// Material 0 is the one the rows name (the floor, the wall, the glass); the slab scene's floor is material 0 (black, a
// detector in its rows) and its glass material 1. No environment: what leaves the scene has escaped.
function opticsFloor(opts = {}) {
    const num = (v, d) => v !== undefined ? v : d
    const b = num(opts.albedo, 0.5)
    const mats = [material('floor', [b, b, b], num(opts.roughness, 1), num(opts.metallic, 0), { specular: num(opts.specular, 0) })]
    const d = doc('optics-floor', mats, [obj('floor', 'b-plane', 'floor', xf(0, 0, 0))], { sky: 'none' }, 20)
    return { doc: d, camera: camera([0, -3, 1.5], [0, 0, 0], 60, 32, 32), settings: { width: 32, height: 32, spp: 1, maxbounce: 8, nee: 1, eps: 1e-5 } }
}
function opticsShell(opts = {}) {
    const b = opts.albedo !== undefined ? opts.albedo : 0.8
    const mats = [material('wall', [b, b, b], 1, 0, { specular: 0 })]
    const shell = { id: 'shell', kind: 'group', label: 'shell', enabled: true, blend: 0, transform: xf(0, 0, 0), children: [
        obj('outer', 'b-sphere', 'wall', xf(0, 0, 0), { radius: 1.05 }),
        obj('inner', 'b-sphere', 'wall', xf(0, 0, 0), { radius: 1.0 }, { join: 'subtract' }),
    ] }
    const d = doc('optics-shell', mats, [shell], { sky: 'none' }, 20)
    return { doc: d, camera: camera([0, -0.5, 0.2], [0, 0, 0], 60, 32, 32), settings: { width: 32, height: 32, spp: 1, maxbounce: 8, nee: 1, eps: 1e-5 } }
}
function opticsSlab(opts = {}) {
    const t = opts.thickness !== undefined ? opts.thickness : 0.1
    // opts.attenuation: the glass absorbs inside, to `colour` (default white) at that depth (metres)
    const att = opts.attenuation || 0
    const mats = [material('floor', [0, 0, 0], 1, 0, { specular: 0 }),
        material('glass', att > 0 ? (opts.colour || [0.5, 0.5, 0.5]) : [1, 1, 1], 0, 0,
            { transmission: 1, ior: opts.ior || 1.5, specular: 0.5, attenuationDistance: att })]
    const d = doc('optics-slab', mats, [obj('floor', 'b-plane', 'floor', xf(0, 0, 0)),
        obj('slab', 'b-box', 'glass', xf(0, 0, opts.z !== undefined ? opts.z : 0.5), { size: { x: 4, y: 4, z: t / 2 }, rounding: 0 })], { sky: 'none' }, 20)
    return { doc: d, camera: camera([0, -3, 1.5], [0, 0, 0], 60, 32, 32), settings: { width: 32, height: 32, spp: 1, maxbounce: 8, nee: 1, eps: 1e-5 } }
}
function opticsBall(opts = {}) {
    const mats = [material('glass', [1, 1, 1], 0, 0, { transmission: 1, ior: opts.ior || 1.5, specular: 0.5 })]
    const d = doc('optics-ball', mats, [obj('ball', 'b-sphere', 'glass', xf(0, 0, 0), { radius: opts.radius || 0.5 })], { sky: 'none' }, 20)
    return { doc: d, camera: camera([0, -3, 1.5], [0, 0, 0], 60, 32, 32), settings: { width: 32, height: 32, spp: 1, maxbounce: 8, nee: 1, eps: 1e-5 } }
}

const CONFIGS = { 'plane-c1': planeC1, 'optics-floor': opticsFloor, 'optics-shell': opticsShell, 'optics-slab': opticsSlab,
    'optics-ball': opticsBall, 'room-c2': roomC2, 'caustic-c3': causticC3, 'furnace-plane': furnacePlane, 'furnace-sphere': furnaceSphere,
    'furnace-haze': furnaceHaze, 'absorb-slab': absorbSlab, 'sky-code': skyCode,
    'sppm-slab': sppmSlab, 'studio-floor': studioFloor, 'lens-prism': lensPrism, 'reflector-condenser': reflectorCondenser,
    'turned-parts': turnedParts, gallery, sunlit, modifiers, usercode, 'sunsky-floor': sunskyFloor, 'sunsky-bench': sunskyBench, joins }

module.exports = { CONFIGS, camera, rng, SEED }
