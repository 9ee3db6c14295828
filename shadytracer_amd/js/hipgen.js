'use strict'
/*
 * HIPGEN -- the scene builder's document becomes HIP C++ for the MI355X engine, beside the reference's WGSL
 * emitter (builder/codegen.js generate :974, emitScene :1238, emitMaterials :648, emitLightFaces :861).
 *
 * Input: a builder document (format 'raytracer-scene', model version 31: builder/model.js:183), the same JSON the
 * reference's scenes/*.json hold. Output: the scene contract the engine loads (lib/pathtracer/README.md:126-145),
 * with `hip` in place of `wgsl`:
 *   hip        sceneDist(U, p) (distance only, the march's) and sceneFull(U, p) (normal, distance, closest and
 *              enclosing material), straight-line code over the document's tree
 *   params     the parameter block U both read: every node's world transform and its own parameters, in metres
 *   layout     where each node's parameters sit in U (for updates without a rebuild)
 *   materials  the material table; lights' faces; the haze box; the environment; envRadius
 *   program    the same tree as a postfix list, for the CPU oracle
 *   flops      FP32 operations of one sceneDist call (the march's roofline count)
 *
 * Scope: the builder's builtin shapes (plane, box, sphere, cylinder, capsule, torus, rectangle, disc), groups with
 * their joins (union, subtract, intersect; smooth blends), transforms (rotation Rz*Ry*Rx in degrees, uniform scale:
 * builder/transform.js:9-38), lights (rectangle and disc faces), one haze box and the studio environment. A code
 * block that is not a builtin -- user WGSL -- is refused with its node named.
 */

const UNIT_METRES = { m: 1, cm: 0.01, mm: 0.001, um: 1e-6, in: 0.0254 }
const LUMA = [0.2126, 0.7152, 0.0722]

// FP32 operations per distance-only primitive evaluation in its frame (an FMA counts 2; the transform in and out
// counted apart), as the emitted code does them. Counted by the kernels per evaluated node: the march's roofline.
const PRIM_FLOPS = { plane: 1, sphere: 7, box: 27, cylinder: 24, capsule: 14, torus: 14, rectangle: 14, disc: 13 }
const XF_FLOPS = 22      // p - pos, three dot products, times 1/s; the distance times s
const FOLD_FLOPS = 2
const JOIN_FLOPS = 9
const CULL_FLOPS = 11    // |p - c|^2 against (r + |best|)^2, widened a hair (boundNear)

// The builtin shapes: their parameters in U (after the 13 of the transform), and the HIP calls.
const SHAPES = {
    plane: { params: () => [], dist: () => 'Dplane(q)', full: () => 'Nplane(q)' },
    sphere: { params: P => [P.radius], dist: o => `Dsphere(q, U[${o}])`, full: o => `Nsphere(q, U[${o}])` },
    box: {
        params: P => [P.size.x, P.size.y, P.size.z, P.rounding || 0],
        dist: o => `DboxRounded(q, v3(U[${o}], U[${o + 1}], U[${o + 2}]), U[${o + 3}])`,
        full: o => `NboxRounded(q, v3(U[${o}], U[${o + 1}], U[${o + 2}]), U[${o + 3}])`,
    },
    cylinder: {
        // half height, radius, edge rounding (0: the builder's sharp cylinder)
        params: P => [0.5 * P.height, P.radius, P.rounding || 0],
        dist: o => `DroundCylinder(q, U[${o}], U[${o + 1}], U[${o + 2}])`,
        full: o => `NroundCylinder(q, U[${o}], U[${o + 1}], U[${o + 2}])`,
    },
    capsule: { params: P => [0.5 * P.height, P.radius], dist: o => `Dcapsule(q, U[${o}], U[${o + 1}])`, full: o => `Ncapsule(q, U[${o}], U[${o + 1}])` },
    torus: { params: P => [P.ring, P.tube], dist: o => `Dtorus(q, U[${o}], U[${o + 1}])`, full: o => `Ntorus(q, U[${o}], U[${o + 1}])` },
    rectangle: {
        params: P => [0.5 * P.width, 0.5 * P.height],
        dist: o => `Drectangle(q, U[${o}], U[${o + 1}])`, full: o => `Nrectangle(q, U[${o}], U[${o + 1}])`,
    },
    disc: { params: P => [P.radius], dist: o => `Ddisc(q, U[${o}])`, full: o => `Ndisc(q, U[${o}])` },
}
// the length-valued parameters of each shape (scaled by the unit); the rest are plain numbers
const LENGTHS = { sphere: ['radius'], box: ['size', 'rounding'], cylinder: ['radius', 'height', 'rounding'], capsule: ['radius', 'height'],
    torus: ['ring', 'tube'], rectangle: ['width', 'height'], disc: ['radius'], plane: [] }

// ---- transforms: world = parent * local; R = Rz Ry Rx (columns: local axes in the parent)
function rotationOf(t) {
    const r = Math.PI / 180
    const rot = (t && t.rot) || { x: 0, y: 0, z: 0 }
    const cx = Math.cos(rot.x * r), sx = Math.sin(rot.x * r)
    const cy = Math.cos(rot.y * r), sy = Math.sin(rot.y * r)
    const cz = Math.cos(rot.z * r), sz = Math.sin(rot.z * r)
    return [
        [cz * cy, cz * sy * sx - sz * cx, cz * sy * cx + sz * sx],
        [sz * cy, sz * sy * sx + cz * cx, sz * sy * cx - cz * sx],
        [-sy, cy * sx, cy * cx],
    ]
}
const IDENT = { R: [[1, 0, 0], [0, 1, 0], [0, 0, 1]], t: [0, 0, 0], s: 1 }
function mul3(A, B) {
    const C = [[0, 0, 0], [0, 0, 0], [0, 0, 0]]
    for (let i = 0; i < 3; i++) for (let j = 0; j < 3; j++) for (let k = 0; k < 3; k++) C[i][j] += A[i][k] * B[k][j]
    return C
}
function apply3(A, v) { return [0, 1, 2].map(i => A[i][0] * v[0] + A[i][1] * v[1] + A[i][2] * v[2]) }
function compose(parent, t, unit) {
    if (!t) return parent
    const pos = t.pos || { x: 0, y: 0, z: 0 }
    const local = [pos.x * unit, pos.y * unit, pos.z * unit]
    const R = rotationOf(t)
    const s = t.scale === undefined ? 1 : t.scale
    if (!(s > 0)) throw new Error('hipgen: a transform scale must be > 0')
    const off = apply3(parent.R, local).map(v => v * parent.s)
    return { R: mul3(parent.R, R), t: [0, 1, 2].map(i => parent.t[i] + off[i]), s: parent.s * s }
}
// 14 floats: position, the rotation's three columns, scale, and its inverse (the march multiplies, never divides)
function xfParams(X) {
    const R = X.R
    return [X.t[0], X.t[1], X.t[2], R[0][0], R[1][0], R[2][0], R[0][1], R[1][1], R[2][1], R[0][2], R[1][2], R[2][2], X.s, 1 / X.s]
}
// A bounding sphere of a shape in its own frame (a radius about its origin); Infinity for the unbounded plane.
function localRadius(key, P) {
    switch (key) {
        case 'plane': return Infinity
        case 'sphere': return P.radius
        case 'box': return Math.hypot(P.size.x, P.size.y, P.size.z)
        case 'cylinder': return Math.hypot(P.radius, 0.5 * P.height)
        case 'capsule': return 0.5 * P.height + P.radius
        case 'torus': return P.ring + P.tube
        case 'rectangle': return 0.5 * Math.hypot(P.width, P.height)
        case 'disc': return P.radius
    }
    return Infinity
}
// The smallest sphere found by enclosing two spheres.
function encloseSpheres(a, b) {
    if (!isFinite(a.r) || !isFinite(b.r)) return { c: a.c, r: Infinity }
    const d = Math.hypot(b.c[0] - a.c[0], b.c[1] - a.c[1], b.c[2] - a.c[2])
    if (d + b.r <= a.r) return a
    if (d + a.r <= b.r) return b
    const r = 0.5 * (d + a.r + b.r)
    const t = d > 0 ? (r - a.r) / d : 0
    return { c: [0, 1, 2].map(i => a.c[i] + (b.c[i] - a.c[i]) * t), r }
}

function fmt(x) { return Number.isFinite(x) ? x : 0 }

function colourOf(c, d) { return c ? [fmt(c.r), fmt(c.g), fmt(c.b)] : d }

// ---- the document's materials as the engine's table
function materialOf(m) {
    const P = (m && m.params) || {}
    const base = colourOf(P.baseColor, [0.8, 0.8, 0.8])
    const ecol = colourOf(P.emissive, [1, 1, 1])
    const el = fmt(P.emissiveLuminance || 0)
    const y = LUMA[0] * ecol[0] + LUMA[1] * ecol[1] + LUMA[2] * ecol[2]
    const emissive = el > 0 && y > 0 ? ecol.map(c => c / y * el) : [0, 0, 0]
    return {
        id: m ? m.id : null,
        base,
        metallic: fmt(P.metallic || 0),
        roughness: P.roughness === undefined ? 0.5 : fmt(P.roughness),
        transmission: fmt(P.transmission || 0),
        ior: P.ior === undefined ? 1.5 : fmt(P.ior),
        specular: P.specular === undefined ? 0.5 : fmt(P.specular),
        emissive,
        abbe: fmt(P.abbe || 0),
        face: -1,
        specularTint: fmt(P.specularTint || 0),
    }
}

function generate(doc, opts) {
    opts = opts || {}
    if (!doc || !doc.root) throw new Error('hipgen: not a builder document (no root)')
    const unit = UNIT_METRES[doc.unit || 'm']
    if (!unit) throw new Error(`hipgen: unknown unit '${doc.unit}'`)
    const blocks = {}
    for (const b of doc.blocks || []) blocks[b.id] = b
    const materials = [materialOf(null)]           // 0: the default
    const matIndex = {}
    for (const m of doc.materials || []) { matIndex[m.id] = materials.length; materials.push(materialOf(m)) }
    if (materials.length > 64) throw new Error('hipgen: more than 64 materials')

    const params = []
    const layout = {}
    const program = []
    const faces = []
    let haze = null
    let flops = 0
    const dist = [], full = [], maskCode = []
    let tmp = 0, maskBits = 0

    const push = vals => { const o = params.length; for (const v of vals) params.push(fmt(v)); return o }

    function shapeOf(n) {
        if (n.type === 'light') return n.light && n.light.shape === 'disc' ? 'disc' : 'rectangle'
        const b = blocks[n.block]
        if (!b) throw new Error(`hipgen: node '${n.id}' has no block '${n.block}'`)
        const key = b.builtin ? String(b.builtin).split('@')[0] : null
        if (!key || !SHAPES[key]) throw new Error(`hipgen: node '${n.id}' (${n.label || ''}) is user code, not a builtin shape: WGSL translation is not supported`)
        return key
    }
    function paramsOf(n, key) {
        const b = blocks[n.block] || {}
        const P = Object.assign({}, b.params || {}, n.params || {})
        const L = LENGTHS[key]
        const out = {}
        for (const k of Object.keys(P)) {
            const v = P[k]
            const isLen = L.indexOf(k) >= 0
            if (v && typeof v === 'object') out[k] = { x: v.x * (isLen ? unit : 1), y: v.y * (isLen ? unit : 1), z: v.z * (isLen ? unit : 1) }
            else out[k] = isLen ? v * unit : v
        }
        return out
    }

    // one shape: its world transform and parameters into U; the expressions of its (local) value
    function shape(n, X) {
        const key = shapeOf(n)
        let P = paramsOf(n, key)
        if (n.type === 'light') {
            P = n.light.shape === 'disc' ? { radius: (n.params.radius || 0.5) * unit } : { width: (n.params.width || 1) * unit, height: (n.params.height || 1) * unit }
        }
        const xo = push(xfParams(X))
        const po = push(SHAPES[key].params(P))
        layout[n.id] = { xf: xo, params: po, shape: key }
        return { key, xo, po, P, bs: { c: X.t.slice(), r: localRadius(key, P) * X.s }, flops: PRIM_FLOPS[key] + XF_FLOPS }
    }
    function matOf(n) {
        if (n.type === 'light') return lightMaterial(n)
        if (n.material && matIndex[n.material] !== undefined) return matIndex[n.material]
        return 0
    }
    function lightMaterial(n) {
        // a light: an emissive face (one-sided, front along +z of its frame) and an entry in the face table
        const L = n.light || {}
        const X = n.__world
        const shape = L.shape === 'disc' ? 1 : 2
        const half = shape === 1 ? [(n.params.radius || 0.5) * unit * X.s, 0] : [0.5 * (n.params.width || 1) * unit * X.s, 0.5 * (n.params.height || 1) * unit * X.s]
        const area = shape === 1 ? Math.PI * half[0] * half[0] : 4 * half[0] * half[1]
        const col = colourOf(L.color, [1, 1, 1])
        const y = LUMA[0] * col[0] + LUMA[1] * col[1] + LUMA[2] * col[2]
        // luminance given, or from the flux of a one-sided Lambertian face: L = flux / (pi A)
        const lum = L.luminance !== undefined ? L.luminance : (L.flux || 0) / (Math.PI * area * (L.twoSided ? 2 : 1))
        const le = col.map(c => c / (y || 1) * lum)
        const R = X.R
        faces.push({
            centre: X.t.slice(), halfx: half[0], u: [R[0][0], R[1][0], R[2][0]], halfy: half[1], v: [R[0][1], R[1][1], R[2][1]], shape,
            n: [R[0][2], R[1][2], R[2][2]], twoSided: L.twoSided ? 1 : 0, le, weight: 1,
        })
        const m = materialOf(null)
        m.base = [0, 0, 0]
        m.emissive = le
        m.face = faces.length - 1
        m.id = 'light:' + n.id
        materials.push(m)
        return materials.length - 1
    }

    const isCsg = g => (g.blend > 0) || (g.children || []).some((c, i) => i > 0 && ((c.join && c.join !== 'union') || c.joinBlend > 0))
    const drawn = c => c.enabled !== false && c.kind !== undefined && c.type !== 'haze'

    // a CSG group's children joined in order into one value; returns the names of its expressions
    function csgGroup(g, X) {
        const kids = (g.children || []).filter(drawn)
        if (!kids.length) return null
        const blendG = (g.blend || 0) * unit * X.s
        let acc = null
        for (let i = 0; i < kids.length; i++) {
            const c = kids[i]
            const v = valueOf(c, X)
            if (!v) continue
            if (!acc) { acc = Object.assign({}, v); continue }
            const join = i === 0 ? 'union' : (c.join || 'union')
            const k = c.joinBlend !== undefined && c.joinBlend !== null ? c.joinBlend * unit * X.s : blendG
            // a union's value is at least the smaller of its sides less k/4; an intersection's or a cut's at least the
            // first side's: the bound that keeps culling exact
            if (join === 'union') { acc.bs = encloseSpheres(acc.bs, v.bs); acc.bs = { c: acc.bs.c, r: acc.bs.r + 0.25 * Math.max(k, 0) } }
            acc.flops += v.flops + JOIN_FLOPS
            const ko = push([k, k > 0 ? 1 / k : 0])
            const fn = { union: ['dUnion', 'csgUnion'], subtract: ['dSubtract', 'csgSubtract'], intersect: ['dIntersect', 'csgIntersect'] }[join]
            if (!fn) throw new Error(`hipgen: node '${c.id}': unknown join '${join}'`)
            const wins = { union: `${v.f}.w < ${acc.f}.w`, subtract: `-${v.f}.w > ${acc.f}.w`, intersect: `${v.f}.w > ${acc.f}.w` }[join]
            dist.push(`        ${acc.d} = ${fn[0]}(${acc.d}, ${v.d}, U[${ko}], U[${ko + 1}]);`)
            full.push(`        if (${wins}) { ${acc.m} = ${v.m}; ${acc.k} = ${v.k}; }`)
            full.push(`        ${acc.f} = ${fn[1]}(${acc.f}, ${v.f}, U[${ko}], U[${ko + 1}]);`)
            program.push({ op: 'join', join, k: ko })
        }
        return acc
    }
    // a node's value as expressions (for use inside a CSG group)
    function valueOf(n, parentX) {
        const X = compose(parentX, n.transform, unit)
        if (n.kind === 'group') {
            if (isCsg(n)) return csgGroup(n, X)
            // a plain group inside a CSG group: its children unioned
            return csgGroup(Object.assign({}, n, { blend: 0, children: (n.children || []).map(c => Object.assign({}, c, { join: 'union', joinBlend: undefined })) }), X)
        }
        n.__world = X
        const s = shape(n, X)
        const mat = matOf(n)
        const i = tmp++
        const d = `d${i}`, f = `f${i}`, m = `m${i}`, k = `k${i}`
        const prim = program.filter(o => o.op === 'prim').length
        dist.push(`        float ${d} = ${SHAPES[s.key].dist(s.po)} * U[${s.xo + 12}];`.replace(/\bq\b/, `xfIn(U + ${s.xo}, p)`))
        full.push(`        V4 ${f} = xfOut(U + ${s.xo}, ${SHAPES[s.key].full(s.po)});`.replace(/\bq\b/, `xfIn(U + ${s.xo}, p)`))
        full.push(`        unsigned int ${m} = ${mat}u, ${k} = ${prim}u;`)
        program.push({ op: 'prim', shape: s.key, xf: s.xo, p: s.po, mat, prim })
        return { d, f, m, k, bs: s.bs, flops: s.flops }
    }
    // the closest-surface fold over a plain group's descendants
    function foldNode(n, parentX) {
        if (n.enabled === false) return
        if (n.type === 'haze') { hazeOf(n, compose(parentX, n.transform, unit)); return }
        if (n.kind === 'group' && !isCsg(n)) {
            const X = compose(parentX, n.transform, unit)
            for (const c of n.children || []) foldNode(c, X)
            return
        }
        // the node's code, guarded by its bounding sphere: skipped where |p - c| - r cannot beat the closest |d| so
        // far, which leaves the fold's answer exactly as it was (a skipped node is never the closest, nor negative).
        // And by the ray's node mask (sceneMask): a node whose sphere, widened by the hit tolerance, the ray's segment
        // never meets can neither be hit nor stop the ray, so its march leaves it out.
        const start = program.length
        program.push({ op: 'node' })
        const d0 = dist.length, f0 = full.length
        const v = valueOf(n, parentX)
        if (!v) { program.splice(start); return }
        const body = dist.splice(d0), fbody = full.splice(f0)
        const nodeFlops = v.flops + FOLD_FLOPS
        if (isFinite(v.bs.r)) {
            const co = push([v.bs.c[0], v.bs.c[1], v.bs.c[2], v.bs.r])
            const near = best => `boundNear(p, U[${co}], U[${co + 1}], U[${co + 2}], U[${co + 3}], ${best})`
            const bit = maskBits < 64 ? maskBits++ : -1
            const inMask = bit >= 0 ? `((mask >> ${bit}) & 1ull) && ` : ''
            dist.push(`    if (${inMask}${near('best')}) {`, ...body, `        foldDist(best, ${v.d});`, `        w += ${nodeFlops}u;`, '    }')
            dist.push(`    w += ${CULL_FLOPS}u;`)
            full.push(`    if (${near('h.d')}) {`, ...fbody, `        foldFull(h, ${v.f}, ${v.m}, ${v.k});`, '    }')
            if (bit >= 0) {
                maskCode.push(`    { V3 oc = v3(U[${co}], U[${co + 1}], U[${co + 2}]) - o; float t = fminf(fmaxf(dot(oc, d), 0.0f), tmax);`,
                    `      V3 q = oc - d * t; float rr = U[${co + 3}] + margin; if (dot(q, q) > rr * rr) m &= ~(1ull << ${bit}); }`)
            }
            program[start] = { op: 'node', bs: co, bit, end: -1 }
            flops += nodeFlops + CULL_FLOPS
        } else {
            dist.push('    {', ...body, `        foldDist(best, ${v.d});`, `        w += ${nodeFlops}u;`, '    }')
            full.push('    {', ...fbody, `        foldFull(h, ${v.f}, ${v.m}, ${v.k});`, '    }')
            program[start] = { op: 'node', bs: -1, bit: -1, end: -1 }
            flops += nodeFlops
        }
        program.push({ op: 'fold' })
        program[start].end = program.length - 1
    }
    function hazeOf(n, X) {
        if (haze) throw new Error('hipgen: one haze box a scene')
        const H = Object.assign({ mfp: 1, albedo: { r: 0.9, g: 0.9, b: 0.9 }, g: 0.7 }, n.haze || {})
        const size = (n.params && n.params.size) || { x: 1, y: 1, z: 1 }
        const rotated = Math.abs(X.R[0][0] - 1) + Math.abs(X.R[1][1] - 1) + Math.abs(X.R[2][2] - 1) > 1e-6
        if (rotated) throw new Error(`hipgen: haze box '${n.id}' is turned; boxes are axis-aligned here`)
        const alb = colourOf(H.albedo, [0.9, 0.9, 0.9])
        haze = {
            centre: X.t.slice(), half: [size.x, size.y, size.z].map(v => 0.5 * v * unit * X.s),
            sigmaS: 1 / H.mfp, g: Math.max(-0.95, Math.min(0.95, H.g)),
            albedo: (alb[0] + alb[1] + alb[2]) / 3,
        }
    }

    for (const c of doc.root.children || []) foldNode(c, compose(IDENT, doc.root.transform, unit))

    const envRadius = (doc.envRadius !== undefined ? doc.envRadius * unit : 400)
    const env = environmentOf(doc.environment || {}, doc)

    const hip = [
        `// emitted by shadytracer_amd/js/hipgen.js from '${(doc.title || 'scene').replace(/[\n\r]/g, ' ')}'`,
        `#define PT_SCENE_MATERIALS ${materials.length}`,
        `#define PT_SCENE_MASK_BITS ${maskBits}`,
        '// the nodes a ray from o along d can reach within tmax: those whose bounding sphere, widened by margin, its',
        '// segment meets (bit k of the mask: the k-th bounded top-level node); unbounded nodes are always marched',
        '__device__ __forceinline__ unsigned long long sceneMask(ParamPtr U, V3 o, V3 d, float tmax, float margin) {',
        '    unsigned long long m = ~0ull;',
        ...maskCode,
        '    return m;',
        '}',
        '// the closest signed distance over the nodes in mask; w counts the FP32 operations spent',
        '__device__ __forceinline__ float sceneDistM(ParamPtr U, V3 p, unsigned long long mask, unsigned int& w) {',
        '    float best = 1.0e7f;',
        ...dist,
        '    return best;',
        '}',
        '__device__ __forceinline__ float sceneDist(ParamPtr U, V3 p) { unsigned int w = 0u; return sceneDistM(U, p, ~0ull, w); }',
        '__device__ __forceinline__ SdfHit sceneFull(ParamPtr U, V3 p) {',
        '    SdfHit h = sdfEmpty();',
        ...full,
        '    return h;',
        '}',
        '',
    ].join('\n')
    return {
        id: doc.title || 'scene',
        hip,
        params: Float32Array.from(params),
        layout,
        materials,
        faces,
        haze,
        env,
        envRadius,
        unit,
        program,
        flops: { dist: flops, maskNode: 14 },   // every node evaluated: the bound a culled step stays under
    }
}

// The studio environment (the builder's envMode 0, studioIlluminance lux on a horizontal surface): here a uniform
// sky over the upper hemisphere whose luminance gives that illuminance (E = pi L); `uniformSky` (cd/m^2) and
// `sky: 'sphere' | 'hemisphere' | 'none'` name it directly. An optional sun: direction from azimuth/elevation
// (sdf.js:72-76, Z-up), illuminance in lux, angular diameter in degrees, and an aperture (a beam radius about a point).
function environmentOf(E, doc) {
    const env = { type: 0, L: 0, col: [1, 1, 1], sunDir: [0, 0, 1], sunCos: 1, sunE: 0, sunRadius: 0, sunPoint: [0, 0, 0] }
    const kind = E.sky || (E.envMode === 0 || E.envMode === undefined ? 'hemisphere' : 'none')
    let L = E.uniformSky !== undefined ? E.uniformSky : (E.studioIlluminance !== undefined ? E.studioIlluminance / Math.PI : 0)
    if (kind === 'none') L = 0
    env.type = L > 0 ? (kind === 'sphere' ? 1 : 2) : 0
    env.L = L
    if (E.studioColor) {
        const c = colourOf(E.studioColor, [1, 1, 1])
        const y = LUMA[0] * c[0] + LUMA[1] * c[1] + LUMA[2] * c[2]
        env.col = c.map(v => v / (y || 1))
    }
    if (E.sunIlluminance > 0) {
        const r = Math.PI / 180
        const az = (E.uSunAzimuth || 0) * r, el = (E.uSunElevation || 45) * r
        env.sunDir = [Math.cos(el) * Math.cos(az), Math.cos(el) * Math.sin(az), Math.sin(el)]
        env.sunCos = Math.cos(0.5 * (E.sunDiameter || 0.53) * r)
        env.sunE = E.sunIlluminance
        const unit = UNIT_METRES[doc.unit || 'm']
        if (E.sunAperture) {
            env.sunRadius = E.sunAperture.radius * unit
            env.sunPoint = [E.sunAperture.point.x * unit, E.sunAperture.point.y * unit, E.sunAperture.point.z * unit]
        }
    }
    return env
}

module.exports = { generate, environmentOf, rotationOf, UNIT_METRES, PRIM_FLOPS }
