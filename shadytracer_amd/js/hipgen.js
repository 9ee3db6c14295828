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
 * Scope: the builder's builtin shapes (plane, box, sphere, cylinder, capsule, torus, rectangle, disc, lens, prism),
 * groups with their joins (union, subtract, intersect; smooth blends, rounds and chamfers), transforms (rotation
 * Rz*Ry*Rx in degrees, uniform scale: builder/transform.js:9-38), the builtin modifier groups (twist, bend, repeat,
 * mirror, circular array, round, shell: builder/modifiers.js), lights (rectangle and disc faces), one haze box and the
 * studio environment. An object or a group whose code is not one of those builtins -- the author's own WGSL, or a
 * builtin this file does not draw natively -- is translated (usercode.js, wgsl.js) and runs as written, in the scene's
 * unit; code that uses what the translation does not know is refused with its node, line and column named.
 */

const sky = require('./sky')
const usercode = require('./usercode')
const UNIT_METRES = { m: 1, cm: 0.01, mm: 0.001, um: 1e-6, in: 0.0254 }
const LUMA = [0.2126, 0.7152, 0.0722]

// FP32 operations per distance-only primitive evaluation in its frame (an FMA counts 2; the transform in and out
// counted apart), as the emitted code does them. Counted by the kernels per evaluated node: the march's roofline.
// The library calls (the prism's atan2f and sincospif, the reflector's cbrtf or cosf and atan2f) are counted as 25 each
// (their polynomials and the division), an estimate.
const PRIM_FLOPS = { plane: 1, sphere: 7, box: 27, cylinder: 24, capsule: 14, torus: 14, rectangle: 14, disc: 13, lens: 25, prism: 75,
    reflector: 80, 'aspheric lens': 52, revolution: 0, 'fresnel lens': 0, extrusion: 0, cone: 30, 'round cone': 20, octahedron: 24,
    'capped torus': 22, link: 18, tube: 16, dome: 24, pyramid: 90 }
const XF_FLOPS = 22      // p - pos, three dot products, times 1/s; the distance times s
// an object's or a group's own code: not counted operation by operation (an estimate of a small shape's cost)
const USER_FLOPS = 40
const FOLD_FLOPS = 2
// a join per shape: its device functions (distance only, with normals), FP32 operations, and how far below the smaller
// side a union's value can go per unit of its radius (the smooth blend's k/4; the round's and chamfer's allowance)
const JOIN_FNS = {
    smooth: { union: ['dUnion', 'csgUnion'], subtract: ['dSubtract', 'csgSubtract'], intersect: ['dIntersect', 'csgIntersect'] },
    round: { union: ['dRoundU', 'csgRoundU'], subtract: ['dRoundS', 'csgRoundS'], intersect: ['dRoundI', 'csgRoundI'] },
    chamfer: { union: ['dChamferU', 'csgChamferU'], subtract: ['dChamferS', 'csgChamferS'], intersect: ['dChamferI', 'csgChamferI'] },
}
const JOIN_FLOPS = { smooth: 9, round: 19, chamfer: 14 }
const JOIN_REACH = { smooth: 0.25, round: 0.41421356, chamfer: 0.70710678 }
const CULL_FLOPS = 11    // |p - c|^2 against (r + |best|)^2, widened a hair (boundNear)

// THE BUILTIN MODIFIER GROUPS (builder/modifiers.js MODIFIERS): a group whose block is one of these changes the point
// its children are evaluated at (space: twist, bend, repeat, mirror, the circular array) or their distance (shape:
// round, shell). Each is drawn by a device function of its own (pt_prelude.h m*), its parameters in U after the
// group's transform: lengths in metres, per-length rates per metre. A warp that stretches space scales the children's
// distances by its Lipschitz bound k (k = 1 / sqrt(1 + (rate extent)^2)) and leaves the normal to differences (a zero
// normal: pt_surface takes it from the distance). `bound` is the modifier's bounding sphere in the group's frame from
// its children's (c, r) and its parameters. Displacement and noise reliefs (noise functions) are not drawn.
const MODIFIERS = {
    twist: {
        lengths: ['extent'], perLength: ['rate'], warp: true,
        params: P => [P.rate, 1 / Math.sqrt(1 + (P.rate * P.extent) ** 2)],
        point: (o, q) => `mTwist(${q}, U[${o}])`, flops: 30,
        // a turn about z keeps |p|: the children's sphere's points are within |c| + r of the axis' origin
        bound: (c, r) => ({ c: [0, 0, 0], r: Math.hypot(...c) + r }),
    },
    bend: {
        lengths: ['extent'], perLength: ['amount'], warp: true,
        params: P => [P.amount, 1 / Math.sqrt(1 + (P.amount * P.extent) ** 2)],
        point: (o, q) => `mBend(${q}, U[${o}])`, flops: 30,
        bound: (c, r) => ({ c: [0, 0, 0], r: Math.hypot(...c) + r }),
    },
    repeat: {
        lengths: ['spacing'], perLength: [],
        params: P => [P.spacing.x, P.spacing.y, P.spacing.z, P.count.x, P.count.y, P.count.z],
        point: (o, q) => `mRepeat(${q}, v3(${us(o, 3)}), v3(${us(o + 3, 3)}))`, flops: 24,
        bound: (c, r, P) => ({ c, r: r + Math.hypot(P.spacing.x * P.count.x, P.spacing.y * P.count.y, P.spacing.z * P.count.z) }),
    },
    mirror: {
        lengths: [], perLength: [], warp: true, unscaled: true,
        params: () => [1],
        point: (o, q) => `mMirror(${q})`, flops: 1,
        bound: (c, r) => ({ c: [0, c[1], c[2]], r: Math.abs(c[0]) + r }),
    },
    circular: {
        lengths: [], perLength: [], copies: true,
        // copies 1..64 (rounded) and the angle between them
        params: P => { const n = Math.min(Math.max(Math.round(P.count), 1), 64); return [n, 2 * Math.PI / n] },
        flops: 12,
        bound: (c, r) => ({ c: [0, 0, c[2]], r: Math.hypot(c[0], c[1]) + r }),
    },
    round: {
        lengths: ['radius'], perLength: [],
        params: P => [P.radius], after: (o, d) => `(${d} - U[${o}])`, afterFull: (o, f) => `mGrow(${f}, U[${o}])`, flops: 1,
        bound: (c, r, P) => ({ c, r: r + Math.max(P.radius, 0) }),
    },
    shell: {
        lengths: ['thickness'], perLength: [],
        params: P => [P.thickness], after: (o, d) => `(fabsf(${d}) - U[${o}])`, afterFull: (o, f) => `mShell(${f}, U[${o}])`, flops: 2,
        bound: (c, r, P) => ({ c, r: r + Math.max(P.thickness, 0) }),
    },
}
const MODIFIER_DEFAULTS = {
    twist: { rate: 1, extent: 1 }, bend: { amount: 0.3, extent: 2 }, repeat: { spacing: { x: 2, y: 2, z: 2 }, count: { x: 2, y: 2, z: 0 } },
    mirror: {}, circular: { count: 6, reach: 32 }, round: { radius: 0.1 }, shell: { thickness: 0.05 },
}

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
    // rim radius, vertex half thickness, the faces' signed radii (0 flat), the planes closing hollow faces at the rim
    lens: {
        params: (P, unit) => lensParams(P, unit).U,
        dist: o => `Dlens(q, ${[0, 1, 2, 3, 4, 5].map(i => `U[${o + i}]`).join(', ')})`,
        full: o => `Nlens(q, ${[0, 1, 2, 3, 4, 5].map(i => `U[${o + i}]`).join(', ')})`,
    },
    // sides, 1 / sides, apothem, half a side, half height; the raw sides and corner radius after them (the oracle's)
    prism: {
        params: P => prismParams(P).U,
        dist: o => `Dprism(q, ${[0, 1, 2, 3, 4].map(i => `U[${o + i}]`).join(', ')})`,
        full: o => `Nprism(q, ${[0, 1, 2, 3, 4].map(i => `U[${o + i}]`).join(', ')})`,
    },
    // focal length, aperture radius, wall, solid (> 0.5), the focus ring's radius held at most the aperture
    reflector: {
        params: P => [P.focal, P.aperture, P.wall, P.solid, Math.min(P.focusOffset, Math.max(P.aperture, 0))],
        dist: o => `Dparabolic(q, ${[0, 1, 2, 3, 4].map(i => `U[${o + i}]`).join(', ')})`,
        full: o => `Nparabolic(q, ${[0, 1, 2, 3, 4].map(i => `U[${o + i}]`).join(', ')})`,
    },
    // rim radius, vertex half thickness, then per face c, k, a4..a10 (metres), 1 / sqrt(1 + m^2); the radii after them
    'aspheric lens': { params: (P, unit) => asphereParams(P, unit).U, dist: o => `Dasphere(q, U + ${o})`, full: o => `Nasphere(q, U + ${o})` },
    // an outline turned about z, or extruded along z: U holds the mode, half the height, the count of pieces and 9
    // numbers a piece (outlinePieces); the node's own function, written by the emitter, walks them (profileCode)
    revolution: { profile: 'revolve', params: (P, unit) => profileParams(P, unit, 'revolve').U, dist: o => `Dprofile${o}(U, q)`, full: o => `Nprofile${o}(U, q)` },
    'fresnel lens': { profile: 'revolve', params: (P, unit) => profileParams(P, unit, 'revolve').U, dist: o => `Dprofile${o}(U, q)`, full: o => `Nprofile${o}(U, q)` },
    extrusion: { profile: 'extrude', params: (P, unit) => profileParams(P, unit, 'extrude').U, dist: o => `Dprofile${o}(U, q)`, full: o => `Nprofile${o}(U, q)` },
    // the builder's other primitives (builder/primitives.js:139-236): their parameters as the shape functions take them
    cone: { params: P => [0.5 * P.height, P.base, P.top], dist: o => `Dcone(q, ${us(o, 3)})`, full: o => `Ncone(q, ${us(o, 3)})` },
    'round cone': { params: P => [P.bottom, P.top, P.height], dist: o => `NroundCone(q, ${us(o, 3)}).w`, full: o => `NroundCone(q, ${us(o, 3)})` },
    octahedron: { params: P => [P.size], dist: o => `Noctahedron(q, U[${o}]).w`, full: o => `Noctahedron(q, U[${o}])` },
    // ring, tube, and the sine and cosine of half the angle (the sine held above 1e-6, as shapes.js NcappedTorus does)
    'capped torus': {
        params: P => { const h = Math.min(Math.max(P.angle, 0), 360) * Math.PI / 360; return [P.ring, P.tube, Math.max(Math.sin(h), 1e-6), Math.cos(h)] },
        dist: o => `NcappedTorus(q, ${us(o, 4)}).w`, full: o => `NcappedTorus(q, ${us(o, 4)})`,
    },
    link: { params: P => [P.span, P.ring, P.tube], dist: o => `Nlink(q, ${us(o, 3)}).w`, full: o => `Nlink(q, ${us(o, 3)})` },
    tube: { params: P => [P.outer, P.inner, P.height], dist: o => `Ntube(q, ${us(o, 3)}).w`, full: o => `Ntube(q, ${us(o, 3)})` },
    dome: { params: P => [P.radius, P.height], dist: o => `Ndome(q, ${us(o, 2)}).w`, full: o => `Ndome(q, ${us(o, 2)})` },
    pyramid: { params: P => [P.base.x, P.base.y, P.top.x, P.top.y, P.height], dist: o => `Nfrustum(q, ${us(o, 5)}).w`, full: o => `Nfrustum(q, ${us(o, 5)})` },
}
function us(o, n) { return [...Array(n).keys()].map(i => `U[${o + i}]`).join(', ') }
// the builtins' own defaults for parameters a document leaves out (builder/primitives.js:172 prism, :237 lens, :255
// aspheric lens, :294 reflector)
const DEFAULTS = { lens: { diameter: 1.0, thickness: 0.2, r1: 2.0, r2: 2.0 }, prism: { sides: 3, radius: 0.5, height: 1.0 },
    reflector: { focal: 0.125, aperture: 0.5, wall: 0.01, solid: 0, focusOffset: 0 },
    'aspheric lens': { diameter: 1.0, thickness: 0.3, r1: 0.5, k1: -2.25, a4_1: 0, a6_1: 0, a8_1: 0, a10_1: 0, r2: 0, k2: 0, a4_2: 0, a6_2: 0,
        a8_2: 0, a10_2: 0 },
    cone: { base: 0.5, top: 0.0, height: 1.0 }, 'round cone': { bottom: 0.3, top: 0.15, height: 0.55 }, octahedron: { size: 0.5 },
    'capped torus': { ring: 0.42, tube: 0.08, angle: 270 }, link: { span: 0.36, ring: 0.25, tube: 0.07 }, tube: { outer: 0.5, inner: 0.4, height: 1.0 },
    dome: { radius: 0.5, height: 0.3 }, pyramid: { base: { x: 1, y: 1 }, top: { x: 0, y: 0 }, height: 1 } }

// ---- outlines (the builder's revolution, Fresnel lens and extrusion shapes): a closed polygon given as entries
// [x, y] or [x, y, r], r the radius of a fillet that rounds that corner, compiled into the pieces a profile function
// walks: straight sides, and an arc at each rounded corner, in counter-clockwise order. A side lying on x = 0 is
// marked as the axis (turned about the axis, it is no surface). What cannot make a sound outline is refused with the
// entries named by their place in the list as written.
const sub2 = (a, b) => [a[0] - b[0], a[1] - b[1]]
const cross2 = (a, b) => a[0] * b[1] - a[1] * b[0]
const dot2 = (a, b) => a[0] * b[0] + a[1] * b[1]
const unit2 = v => { const l = Math.hypot(v[0], v[1]); return [v[0] / l, v[1] / l] }
// Whether the closed segments ab and cd share a point: a + t (b - a) = c + s (d - c) solved by Cramer's rule for
// t, s in [0, 1]; parallel sides share a point only when they lie on one line and their spans along it overlap.
function sidesTouch(a, b, c, d) {
    const r = sub2(b, a), q = sub2(d, c), w = sub2(c, a)
    const den = cross2(r, q)
    if (den !== 0) {
        const t = cross2(w, q) / den, s = cross2(w, r) / den
        return t >= 0 && t <= 1 && s >= 0 && s <= 1
    }
    if (cross2(w, r) !== 0) return false
    const rr = dot2(r, r), t0 = dot2(w, r) / rr, t1 = dot2(sub2(d, a), r) / rr
    return Math.max(t0, t1) >= 0 && Math.min(t0, t1) <= 1
}
// A corner's fillet: the circle of radius r tangent to both sides at p. With u and v the unit directions from p to its
// neighbours and c = u.v the cosine of the corner's angle th, the tangent points lie r / tan(th/2) = r sqrt((1+c)/(1-c))
// along each side and the centre r / sin(th/2) = r sqrt(2/(1-c)) along the bisector. A straight run has nothing to round.
function fillet(prev, p, next, r) {
    const u = unit2(sub2(prev, p)), v = unit2(sub2(next, p)), c = Math.max(-1, Math.min(1, dot2(u, v)))
    if (!(r > 0) || 1 + c < 1e-15) return null
    const along = r * Math.sqrt((1 + c) / (1 - c)), out = r * Math.sqrt(2 / (1 - c)), bis = unit2([u[0] + v[0], u[1] + v[1]])
    return { along, r, inAt: [p[0] + u[0] * along, p[1] + u[1] * along], outAt: [p[0] + v[0] * along, p[1] + v[1] * along],
        centre: [p[0] + bis[0] * out, p[1] + bis[1] * out], convex: cross2(sub2(p, prev), sub2(next, p)) > 0 }
}
function outlinePieces(points, where) {
    const refuse = why => { throw new Error(`hipgen: ${where}: ${why}`) }
    const n = Array.isArray(points) ? points.length : 0
    if (n < 3) refuse(`only ${n} entries: an outline is a polygon of three or more`)
    for (let i = 0; i < n; i++) {
        const e = points[i]
        if (!Array.isArray(e) || e.length < 2 || e.length > 3 || !e.every(Number.isFinite)) refuse(`entry ${i + 1} is not [x, y] or [x, y, fillet] in numbers`)
        if (e.length === 3 && e[2] < 0) refuse(`entry ${i + 1} has a fillet radius below zero`)
    }
    const next = i => (i + 1) % n
    for (let i = 0; i < n; i++) if (points[i][0] === points[next(i)][0] && points[i][1] === points[next(i)][1]) refuse(`entries ${i + 1} and ${next(i) + 1} coincide`)
    // the shoelace area: its sign is the winding, its size whether there is a shape at all
    let twice = 0
    for (let i = 0; i < n; i++) twice += cross2(points[i], points[next(i)])
    const scale = points.reduce((m, e) => Math.max(m, Math.abs(e[0]), Math.abs(e[1])), 1e-30)
    if (!(Math.abs(twice) > 2e-12 * scale * scale)) refuse('its area is zero')
    for (let i = 0; i < n; i++) for (let j = i + 2; j < n; j++) {
        if (next(j) === i) continue
        if (sidesTouch(points[i], points[next(i)], points[j], points[next(j)])) refuse(`sides ${i + 1}-${next(i) + 1} and ${j + 1}-${next(j) + 1} intersect`)
    }
    // counter-clockwise, each entry keeping its place in the list for messages
    const ord = points.map((_, i) => i)
    if (twice < 0) ord.reverse()
    const V = ord.map(i => points[i])
    const F = V.map((p, i) => fillet(V[(i + n - 1) % n], p, V[next(i)], p.length === 3 ? p[2] : 0))
    for (let i = 0; i < n; i++) {
        const used = (F[i] ? F[i].along : 0) + (F[next(i)] ? F[next(i)].along : 0)
        if (used > Math.hypot(V[next(i)][0] - V[i][0], V[next(i)][1] - V[i][1]) * (1 + 1e-9))
            refuse(`side ${ord[i] + 1}-${ord[next(i)] + 1} is too short for the fillets at its ends`)
    }
    const pieces = []
    const onAxis = q => Math.abs(q[0]) <= 1e-12 * scale
    for (let i = 0; i < n; i++) {
        const f = F[i], g = F[next(i)]
        if (f) pieces.push({ kind: 'arc', c: f.centre, r: f.r, a: f.inAt, b: f.outAt, convex: f.convex })
        const from = f ? f.outAt : [V[i][0], V[i][1]], to = g ? g.inAt : [V[next(i)][0], V[next(i)][1]]
        if (Math.hypot(to[0] - from[0], to[1] - from[1]) > 1e-12 * scale) pieces.push({ kind: 'segment', a: from, b: to, axis: onAxis(from) && onAxis(to) })
    }
    return pieces
}
// U for an outline node: mode (0 revolve, 1 extrude), half the height, the count of pieces, then 9 numbers a piece:
// kind (0 segment, 1 arc), a segment's a, b and whether it is a surface, or an arc's centre, radius, a, b and
// whether it is convex. Lengths in metres. reach: the outline's farthest point from its origin (bounding sphere).
function profileParams(P, unit, mode) {
    const pts = (P.outline || []).map(p => p.map(v => v * unit))
    const pieces = outlinePieces(pts, 'an outline')
    const U = [mode === 'extrude' ? 1 : 0, mode === 'extrude' ? Math.max(0.5 * (P.height || 0), 0) : 0, pieces.length]
    for (const k of pieces) {
        if (k.kind === 'segment') U.push(0, k.a[0], k.a[1], k.b[0], k.b[1], k.axis && mode === 'revolve' ? 0 : 1, 0, 0, 0)
        else U.push(1, k.c[0], k.c[1], k.r, k.a[0], k.a[1], k.b[0], k.b[1], k.convex ? 1 : 0)
    }
    const reach = Math.max(...pts.map(p => Math.hypot(p[0], p[1]))) * 1.001
    return { U, pieces, reach, mode }
}
// The node's profile function (straight-line code, a call a piece: the kinds are the outline's shape, the numbers read
// from U) and its solid: revolved (q = (|p.xy|, z), the axis edges no surface) or extruded.
function profileCode(o, pr) {
    const u = i => `U[${o + i}]`
    const lines = [`__device__ __forceinline__ Mer profile${o}(ParamPtr U, float qx, float qy) {`, '    PT_SCENE_PARAMS(U);', '    Prof A = pfStart();']
    pr.pieces.forEach((k, i) => {
        const b = 3 + 9 * i
        if (k.kind === 'segment') lines.push(`    pfSegment(A, qx, qy, ${u(b + 1)}, ${u(b + 2)}, ${u(b + 3)}, ${u(b + 4)}, ${k.axis && pr.mode === 'revolve' ? 'false' : 'true'});`)
        else lines.push(`    pfArc(A, qx, qy, ${u(b + 1)}, ${u(b + 2)}, ${u(b + 3)}, ${u(b + 4)}, ${u(b + 5)}, ${u(b + 6)}, ${u(b + 7)}, ${k.convex ? 'true' : 'false'});`)
    })
    lines.push('    return pfEnd(A);', '}')
    if (pr.mode === 'revolve') {
        lines.push(`__device__ __forceinline__ V4 Nprofile${o}(ParamPtr U, V3 p) { return merTo3(p, profile${o}(U, fsqrt(p.x * p.x + p.y * p.y), p.z)); }`)
    } else {
        lines.push(`__device__ __forceinline__ V4 Nprofile${o}(ParamPtr U, V3 p) { PT_SCENE_PARAMS(U); return merExtrude(p, profile${o}(U, p.x, p.y), ${u(1)}); }`)
    }
    lines.push(`__device__ __forceinline__ float Dprofile${o}(ParamPtr U, V3 p) { return Nprofile${o}(U, p).w; }`)
    return lines.join('\n')
}

// An aspheric lens (shapes.js Nasphere:658, asphSag, asphSlopeMax): a radius under 1e-9 scene units is a flat face;
// the a's go from the scene's unit to metres as length^(1 - 2i) (a4 a unit^-3, ...); m, a face's steepest slope
// within the rim, is the conic's at the rim plus every polynomial term's size there. reach: how far past its vertex
// a face's surface goes (a hollow face), for the bounding sphere.
function asphereParams(P, unit) {
    const rho = Math.max(0.5 * P.diameter, 1e-6), ht = 0.5 * P.thickness
    const face = i => {
        const R = P['r' + i], c = Math.abs(R) < 1e-9 * unit ? 0 : 1 / R, k = P['k' + i]
        const a = [4, 6, 8, 10].map(e => P[`a${e}_${i}`] * Math.pow(unit, 1 - e))
        const sag = r => {
            const r2 = r * r, u = Math.sqrt(Math.max(1 - (1 + k) * c * c * r2, 1e-4))
            return c * r2 / (1 + u) + r2 * r2 * (a[0] + r2 * (a[1] + r2 * (a[2] + r2 * a[3])))
        }
        const r2 = rho * rho, u = Math.sqrt(Math.max(1 - (1 + k) * c * c * r2, 1e-4))
        const m = Math.abs(c) * rho / u + rho * r2 * (4 * Math.abs(a[0]) + r2 * (6 * Math.abs(a[1]) + r2 * (8 * Math.abs(a[2]) + r2 * 10 * Math.abs(a[3]))))
        let lowest = 0
        for (let j = 0; j <= 1024; j++) lowest = Math.min(lowest, sag(rho * j / 1024))
        return { U: [c, k, ...a, 1 / Math.sqrt(1 + m * m)], R: c === 0 ? 0 : R, reach: -lowest }
    }
    const f1 = face(1), f2 = face(2)
    return { U: [rho, ht, ...f1.U, ...f2.U, f1.R, f2.R], rho, reach: ht + 1.01 * Math.max(f1.reach, f2.reach) }
}

// A lens (sdf.js Nlens:463): a radius under 1e-9 scene units is a flat face; a hollow face (negative radius) is closed
// by the plane through its rim, its sag |R| - sqrt(R^2 - rim^2) beyond the vertex.
function lensParams(P, unit) {
    const rim = 0.5 * P.diameter, vz = 0.5 * P.thickness
    const R = r => Math.abs(r) < 1e-9 * unit ? 0 : r
    const r1 = R(P.r1), r2 = R(P.r2)
    const sag = r => r < 0 ? Math.abs(r) - Math.sqrt(Math.max(r * r - rim * rim, 0)) : 0
    return { U: [rim, vz, r1, r2, vz + sag(r1), vz + sag(r2)], rim, reach: vz + Math.max(sag(r1), sag(r2)) }
}
// A prism (shapes.js Nprism:285): sides rounded (half to even, as WGSL's round) and held to 3..64, the corner radius
// at least 0.
function roundHalfEven(x) { const f = Math.floor(x), r = x - f; return r > 0.5 || (r === 0.5 && f % 2 !== 0) ? f + 1 : f }
function prismParams(P) {
    const n = Math.min(Math.max(roundHalfEven(P.sides), 3), 64), R = Math.max(P.radius, 0)
    return { U: [n, 1 / n, R * Math.cos(Math.PI / n), R * Math.sin(Math.PI / n), Math.max(0.5 * P.height, 0), P.sides, P.radius], R }
}
// the length-valued parameters of each shape (scaled by the unit); the rest are plain numbers
const LENGTHS = { sphere: ['radius'], box: ['size', 'rounding'], cylinder: ['radius', 'height', 'rounding'], capsule: ['radius', 'height'],
    torus: ['ring', 'tube'], rectangle: ['width', 'height'], disc: ['radius'], plane: [], lens: ['diameter', 'thickness', 'r1', 'r2'],
    prism: ['radius', 'height'], reflector: ['focal', 'aperture', 'wall', 'focusOffset'], 'aspheric lens': ['diameter', 'thickness', 'r1', 'r2'],
    revolution: [], 'fresnel lens': [], extrusion: ['height'], cone: ['base', 'top', 'height'], 'round cone': ['bottom', 'top', 'height'],
    octahedron: ['size'], 'capped torus': ['ring', 'tube'], link: ['span', 'ring', 'tube'], tube: ['outer', 'inner', 'height'], dome: ['radius', 'height'],
    pyramid: ['base', 'top', 'height'] }

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
function localRadius(key, P, unit) {
    switch (key) {
        case 'plane': return Infinity
        case 'sphere': return P.radius
        case 'box': return Math.hypot(P.size.x, P.size.y, P.size.z)
        case 'cylinder': return Math.hypot(P.radius, 0.5 * P.height)
        case 'capsule': return 0.5 * P.height + P.radius
        case 'torus': return P.ring + P.tube
        case 'rectangle': return 0.5 * Math.hypot(P.width, P.height)
        case 'disc': return P.radius
        // the rim's cylinder from the farther face's reach to the other's: no face passes its vertex, or its rim's
        // plane for a hollow one
        case 'lens': { const L = lensParams(P, unit); return Math.hypot(L.rim, L.reach) }
        case 'prism': return Math.hypot(prismParams(P).R, 0.5 * Math.max(P.height, 0))
        // the rim's cylinder from the vertex less the wall to the rim's height (a focus ring widens the bowl)
        case 'reflector': {
            const f = Math.max(P.focal, 1e-6), A = Math.max(P.aperture, 0), w = Math.max(P.wall, 0), c = Math.abs(Math.min(P.focusOffset, A))
            return 1.001 * Math.hypot(A + w, Math.max(f + w, Math.abs((A + c) * (A + c) / (4 * f) - f) + w))
        }
        case 'aspheric lens': { const L = asphereParams(P, unit); return 1.001 * Math.hypot(L.rho, L.reach) }
        // an outline's farthest point from its origin; extruded, with half its height
        case 'revolution': case 'fresnel lens': return profileParams(P, unit, 'revolve').reach
        case 'extrusion': { const L = profileParams(P, unit, 'extrude'); return 1.001 * Math.hypot(L.reach, L.U[1]) }
        case 'cone': return Math.hypot(Math.max(P.base, P.top, 0), 0.5 * Math.abs(P.height))
        case 'round cone': return 0.5 * Math.abs(P.height) + Math.max(P.bottom, P.top, 0)
        case 'octahedron': return Math.max(P.size, 0)
        case 'capped torus': return P.ring + P.tube
        case 'link': return 0.5 * Math.max(P.span, 0) + P.ring + P.tube
        case 'tube': return Math.hypot(Math.max(P.outer, P.inner, 0), 0.5 * Math.abs(P.height))
        // the sphere's centre at c = height - R (held to +-R): within |c| + R of the origin
        case 'dome': { const R = Math.max(P.radius, 1e-6); return Math.abs(Math.min(Math.max(P.height - R, -R), R)) + R }
        case 'pyramid': { const h = 0.5 * Math.max(P.height, 1e-6); return Math.max(Math.hypot(0.5 * P.base.x, 0.5 * P.base.y, h), Math.hypot(0.5 * P.top.x, 0.5 * P.top.y, h)) }
    }
    return Infinity
}
// A shape's half extents in its own frame, for the scene's bounds (the builder's, model.js bounds): the common solids
// exactly, the rest as the cube round their bounding sphere; null for the unbounded plane.
function localHalf(key, P, unit) {
    switch (key) {
        case 'plane': return null
        case 'sphere': return [P.radius, P.radius, P.radius]
        case 'box': return [P.size.x, P.size.y, P.size.z]
        case 'cylinder': return [P.radius, P.radius, 0.5 * P.height]
        case 'capsule': return [P.radius, P.radius, 0.5 * P.height + P.radius]
        case 'torus': return [P.ring + P.tube, P.ring + P.tube, P.tube]
        case 'rectangle': return [0.5 * P.width, 0.5 * P.height, 0]
        case 'disc': return [P.radius, P.radius, 0]
    }
    const r = localRadius(key, P, unit)
    return isFinite(r) ? [r, r, r] : null
}
// THE SCENE'S TOLERANCE (photometry.js autoEps:117): 1e-4 of its size -- the diagonal of its bounds -- never under
// 1e-8 m nor over 1 mm; a scene with no bounded part keeps the 1 mm.
function autoEps(extentMetres) {
    return extentMetres > 0 && isFinite(extentMetres) ? Math.min(Math.max(1e-4 * extentMetres, 1e-8), 1e-3) : 1e-3
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
    const matIndex = {}, docMats = {}, checked = {}
    for (const m of doc.materials || []) {
        matIndex[m.id] = materials.length; docMats[m.id] = m
        const made = materialOf(m)
        // absorption inside (materials.js attenuationDistance): its depth in metres
        const ad = ((m.params || {}).attenuationDistance || 0)
        made.attenuation = ad > 0 ? ad * unit : 0
        materials.push(made)
    }
    // what the document asks for that the engine does not draw: refused, or with opts.approximate listed and left out
    const approximations = []
    const cannot = (where, what) => approximations.push(`${where}: ${what}`)
    // the document's functions block (builder/codegen.js:1366-1370): C++ ahead of every node's code, its names in scope
    // for every code of the document
    usercode.setShared({})
    let functionsCpp = ''
    if (usercode.isCode(doc.functions)) {
        try {
            const mod = usercode.moduleCode(doc.functions)
            functionsCpp = mod.text
            usercode.setShared(mod.own)
        } catch (e) {
            cannot('the functions block', e.message)
        }
    }
    if (materials.length > 64) throw new Error('hipgen: more than 64 materials')

    const params = []
    const layout = {}
    const bounds = [[Infinity, Infinity, Infinity], [-Infinity, -Infinity, -Infinity]]
    const program = []
    const faces = []
    let haze = null
    let flops = 0
    const dist = [], full = [], maskCode = [], helpers = []
    if (functionsCpp) helpers.push('// the document\'s functions block', functionsCpp)
    let tmp = 0, maskBits = 0, modDepth = 0
    // each primitive's object frame (its U offset): its node's own, or inside a modifier or coded group the outermost
    // such group's (the frame a material's code reads p_local in)
    const primFrames = []
    let outerFrame = null

    const push = vals => { const o = params.length; for (const v of vals) params.push(fmt(v)); return o }

    function shapeOf(n) {
        if (n.type === 'light') return n.light && n.light.shape === 'disc' ? 'disc' : 'rectangle'
        const b = blocks[n.block]
        if (!b) throw new Error(`hipgen: node '${n.id}' has no block '${n.block}'`)
        const key = b.builtin ? String(b.builtin).split('@')[0] : null
        if (key && SHAPES[key]) return key
        if (usercode.isCode(n.code) || usercode.isCode(b.code)) return 'code'
        throw new Error(`hipgen: node '${n.id}' (${n.label || ''}) has no code and is not a builtin shape`)
    }
    function paramsOf(n, key) {
        const b = blocks[n.block] || {}
        const P = Object.assign({}, DEFAULTS[key] || {}, b.params || {}, n.params || {})
        const L = LENGTHS[key]
        const out = {}
        for (const k of Object.keys(P)) {
            const v = P[k]
            const isLen = L.indexOf(k) >= 0
            if (Array.isArray(v)) out[k] = v                      // an outline: its own function scales it
            else if (v && typeof v === 'object') out[k] = { x: v.x * (isLen ? unit : 1), y: v.y * (isLen ? unit : 1), z: v.z * (isLen ? unit : 1) }
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
        let pr = null
        if (SHAPES[key].profile) {
            try { pr = profileParams(P, unit, SHAPES[key].profile) } catch (e) { throw new Error(`hipgen: node '${n.id}': ${e.message.replace(/^hipgen: /, '')}`) }
        }
        const xo = push(xfParams(X))
        const po = push(pr ? pr.U : SHAPES[key].params(P, unit))
        layout[n.id] = { xf: xo, params: po, shape: key }
        // the scene's bounds: the local box's corners in the world
        const half = localHalf(key, P, unit)
        // (inside a modifier group the frame is the group's, not the world's: the modifier adds its own bound)
        if (half && !modDepth) {
            for (let c = 0; c < 8; c++) {
                const l = [0, 1, 2].map(i => ((c >> i) & 1 ? 1 : -1) * half[i] * X.s)
                const w = [0, 1, 2].map(i => X.t[i] + X.R[i][0] * l[0] + X.R[i][1] * l[1] + X.R[i][2] * l[2])
                for (let i = 0; i < 3; i++) { bounds[0][i] = Math.min(bounds[0][i], w[i]); bounds[1][i] = Math.max(bounds[1][i], w[i]) }
            }
        }
        let primFlops = PRIM_FLOPS[key]
        if (pr) {
            helpers.push(profileCode(po, pr))
            // a segment's crossing test, projection and compare; an arc's; the revolve or extrude and the end
            primFlops = pr.pieces.reduce((s, k) => s + (k.kind === 'segment' ? 22 : 40), 0) + 12
        }
        return { key, xo, po, P, bs: { c: X.t.slice(), r: localRadius(key, P, unit) * X.s }, flops: primFlops + XF_FLOPS }
    }
    function matOf(n) {
        if (n.type === 'light') return lightMaterial(n)
        if (n.material && matIndex[n.material] !== undefined) {
            if (!checked[n.material]) { checked[n.material] = true; for (const why of materialLimits(docMats[n.material])) cannot(`material '${n.material}'`, why) }
            return matIndex[n.material]
        }
        return 0
    }
    function lightMaterial(n) {
        // a light: an emissive face (one-sided, front along +z of its frame) and an entry in the face table
        const L = n.light || {}
        const X = n.__world
        const shape = L.shape === 'disc' ? 1 : 2
        const half = shape === 1 ? [(n.params.radius || 0.5) * unit * X.s, 0] : [0.5 * (n.params.width || 1) * unit * X.s, 0.5 * (n.params.height || 1) * unit * X.s]
        const area = shape === 1 ? Math.PI * half[0] * half[0] : 4 * half[0] * half[1]
        for (const why of lightLimits(L)) cannot(`light '${n.id}'`, why)
        const col = lightRGB(L)
        // a flux (lm) rules over the luminance: the flux of a Lambertian face is pi L A a side (lights.js fluxFactor:201)
        const lum = L.flux !== undefined && L.flux !== null ? L.flux / (Math.PI * area * (L.twoSided ? 2 : 1))
            : (L.luminance !== undefined ? L.luminance : 1000)
        const le = col.map(c => c * lum)
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

    // a CSG group's children joined in order into one value; returns the names of its expressions. A child's join is
    // its op (union, subtract, intersect) in its shape: the smooth polynomial blend (no joinShape), or a round's or a
    // chamfer's of that radius (joinShape, model version 28)
    function csgGroup(g, X, pt = 'p') {
        const kids = (g.children || []).filter(drawn)
        if (!kids.length) return null
        const blendG = (g.blend || 0) * unit * X.s
        let acc = null
        for (let i = 0; i < kids.length; i++) {
            const c = kids[i]
            const v = valueOf(c, X, pt)
            if (!v) continue
            if (!acc) { acc = Object.assign({}, v); continue }
            const join = i === 0 ? 'union' : (c.join || 'union')
            const shape = c.joinShape === undefined || c.joinShape === null ? 'smooth' : c.joinShape
            const fn = (JOIN_FNS[shape] || {})[join]
            if (!JOIN_FNS[shape]) throw new Error(`hipgen: node '${c.id}': unknown joinShape '${shape}' (round, chamfer, or none for smooth)`)
            if (!fn) throw new Error(`hipgen: node '${c.id}': unknown join '${join}'`)
            const k = c.joinBlend !== undefined && c.joinBlend !== null ? c.joinBlend * unit * X.s : blendG
            // a union's value is at least the smaller of its sides less the most its blend takes off (k/4 smooth, the
            // round's and chamfer's corner allowance); an intersection's or a cut's at least the first side's: the bound
            // that keeps culling exact
            if (join === 'union') { acc.bs = encloseSpheres(acc.bs, v.bs); acc.bs = { c: acc.bs.c, r: acc.bs.r + JOIN_REACH[shape] * Math.max(k, 0) } }
            acc.flops += v.flops + JOIN_FLOPS[shape]
            const ko = push([k, k > 0 ? 1 / k : 0])
            const args = shape === 'smooth' ? `U[${ko}], U[${ko + 1}]` : `U[${ko}]`
            const wins = { union: `${v.f}.w < ${acc.f}.w`, subtract: `-${v.f}.w > ${acc.f}.w`, intersect: `${v.f}.w > ${acc.f}.w` }[join]
            dist.push(`        ${acc.d} = ${fn[0]}(${acc.d}, ${v.d}, ${args});`)
            full.push(`        if (${wins}) { ${acc.m} = ${v.m}; ${acc.k} = ${v.k}; }`)
            full.push(`        ${acc.f} = ${fn[1]}(${acc.f}, ${v.f}, ${args});`)
            program.push({ op: 'join', join, shape, k: ko })
        }
        return acc
    }
    // A group with code of its own -- a builtin modifier (twist, bend, repeat, circular, mirror, round, shell, displace,
    // relief: builder/modifiers.js) or the user's -- changes the point before its children or the distance after them
    // (builder/codegen.js group code). The engine runs no group code: such a group is refused, or with opts.approximate
    // drawn as a plain group of its children and listed.
    // A group block without a builtin and without statements only carries the group's parameters: a plain group.
    function modifierOf(g) {
        const b = g.block ? blocks[g.block] : null
        const key = b && b.builtin ? String(b.builtin).split('@')[0] : null
        return key && MODIFIERS[key] ? key : null
    }
    function modifierCheck(g) {
        if (modifierOf(g) || groupCodeOf(g)) return
        const b = g.block ? blocks[g.block] : null
        if (b && b.builtin) cannot(`group '${g.id}'`, `modifier '${b.builtin}' (group code)`)
    }
    // A group's own code (the group's, else its block's), when it has statements and is not a native modifier
    function groupCodeOf(g) {
        if (g.kind !== 'group' || modifierOf(g)) return null
        const b = g.block ? blocks[g.block] : null
        if (usercode.isCode(g.code)) return g.code
        if (b && usercode.isCode(b.code)) return b.code
        return null
    }
    // A material's relief (builder/codegen.js:1735-1760) on the object wearing it: the object's distance d (metres,
    // its parent's frame), then beyond two amplitudes d moved towards zero by one, within them d less the height at the
    // object's point (real millimetres in its own frame) and divided by one plus the height's slope bound; the normal by
    // differences. Returns the lines that change d and f, or null for a material without one.
    const reliefFns = {}
    function reliefOn(n, xo, pt, d, f) {
        const m = n.material && docMats[n.material]
        if (!m || !m.relief || !usercode.isCode(m.relief.code)) return null
        if (reliefFns[m.id] === undefined) {
            const R = m.relief, P = R.params || {}
            let fn
            try {
                fn = usercode.reliefFunction(reliefList.length, R.code, P, R.constants || {}, push)
            } catch (e) {
                cannot(`material '${m.id}'`, `its relief's code: ${e.message}`)
                reliefFns[m.id] = null
                return null
            }
            reliefFns[m.id] = { idx: reliefList.length, amp: push([0.001 * (+P.amplitude || 0)]) }
            reliefList.push({ material: m.id, host: fn.host })
            helpers.push(fn.device)
        }
        const R = reliefFns[m.id]
        if (!R) return null
        const i = tmp++
        const band = (v) => [
            `        { const float ra${i} = U[${R.amp}];`,
            `          if (fabsf(${v}) > 2.0f * ra${i}) ${v} -= copysignf(ra${i}, ${v});`,
            `          else { const V3 ql${i} = xfIn(U + ${xo}, ${pt}) * (U[${xo + 12}] * 1000.0f);`,
            `                 const wvec2 hs${i} = urelief${R.idx}(U, ofV3(ql${i}), ofV3(p * 1000.0f));`,
            `                 ${v} = (${v} - 0.001f * hs${i}.x) / (1.0f + hs${i}.y); } }`]
        program.push({ op: 'relief', index: R.idx, xf: xo, amp: R.amp })
        return { dist: band(d), full: [...band(`${f}.w`), `        ${f} = v4(v3s(0.0f), ${f}.w);`] }
    }
    const reliefList = []
    const codeParams = (n, b) => Object.assign({}, (b && b.params) || {}, n.params || {})
    const codeConstants = b => (b && b.constants) || {}
    const userFns = []
    // An object's code: a function of the node's local point (the emitter's frame, metres) giving what it folded
    function userObjectValue(n, X, pt) {
        const b = blocks[n.block] || {}
        const code = usercode.isCode(n.code) ? n.code : b.code
        const mat = matOf(n)
        const prim = program.filter(o => o.op === 'prim' || o.op === 'user').length
        const idx = userFns.length
        let fn
        try {
            fn = usercode.objectFunction(idx, code, codeParams(n, b), codeConstants(b), mat | (prim << 8), unit, push)
        } catch (e) {
            // code the translation does not take: refused, or with opts.approximate left out and listed
            cannot(`node '${n.id}' (${n.label || ''})`, `its code: ${e.message}`)
            return null
        }
        const xo = push(xfParams(X))
        primFrames[prim] = outerFrame !== null ? outerFrame : xo
        userFns.push({ kind: 'object', node: n.id, host: fn.host })
        // no bound is known for code, but the scene's extent (its tolerance) counts the node as a ball about its place of
        // its largest parameter (an estimate of its size, in the document's unit), never for culling
        if (!modDepth) {
            const sizes = []
            for (const v of Object.values(codeParams(n, b))) {
                if (typeof v === 'number') sizes.push(Math.abs(v))
                else if (v && typeof v === 'object' && !Array.isArray(v)) for (const c of ['x', 'y', 'z']) if (typeof v[c] === 'number') sizes.push(Math.abs(v[c]))
            }
            const r = (sizes.length ? Math.max(...sizes) : 0) * unit * X.s
            if (r > 0 && isFinite(r)) for (let a = 0; a < 3; a++) { bounds[0][a] = Math.min(bounds[0][a], X.t[a] - r); bounds[1][a] = Math.max(bounds[1][a], X.t[a] + r) }
        }
        helpers.push(fn.device)
        layout[n.id] = { xf: xo, shape: 'code' }
        const i = tmp++
        const d = `d${i}`, f = `f${i}`, m = `m${i}`, k = `k${i}`
        dist.push(`        unsigned int um${i}; float ${d} = ucode${idx}(U, xfIn(U + ${xo}, ${pt}), &um${i}).w * U[${xo + 12}];`)
        full.push(`        unsigned int um${i}; V4 ${f} = xfOut(U + ${xo}, ucode${idx}(U, xfIn(U + ${xo}, ${pt}), &um${i}));`)
        full.push(`        unsigned int ${m} = um${i} & 0xFFu, ${k} = um${i} >> 8;`)
        program.push({ op: 'user', index: idx, xf: xo, mat, prim })
        const rel = reliefOn(n, xo, pt, d, f)
        if (rel) { dist.push(...rel.dist); full.push(...rel.full) }
        // the code's cost is not counted (an estimate stands for it); no bound is known, so the node is never culled
        return { d, f, m, k, bs: { c: X.t.slice(), r: Infinity }, flops: USER_FLOPS + XF_FLOPS }
    }
    // A group's code around its children: children(C, q) evaluates them at q (the group's frame, scene units) and folds
    // each into C; the code's Hit, taken back to metres, is the group's value
    function userGroupValue(g, X, pt) {
        const b = g.block ? blocks[g.block] : null
        const code = groupCodeOf(g)
        for (const c of g.children || []) {
            if (c.type === 'light' || c.type === 'haze') throw new Error(`hipgen: node '${c.id}': a ${c.type} inside group '${g.id}' with code is not drawn`)
        }
        const idx = userFns.length
        let gc
        try {
            gc = usercode.groupCode(code, codeParams(g, b), codeConstants(b), unit, push)
        } catch (e) {
            // refused, or with opts.approximate drawn as a plain group of its children and listed
            cannot(`group '${g.id}' (${g.label || ''})`, `its code: ${e.message}`)
            return isCsg(g) ? csgGroup(g, X, pt) : csgGroup(unionOf(g), X, pt)
        }
        const xo = push(xfParams(X))
        const i = tmp++
        const q = `q${i}`
        const start = program.length
        program.push({ op: 'ugroup', index: idx, xf: xo, end: -1 })
        const d0 = dist.length, f0 = full.length
        const kids = []
        const frameWas = outerFrame
        if (outerFrame === null) outerFrame = xo
        modDepth++
        for (const c of (g.children || []).filter(drawn)) {
            const v = valueOf(c, IDENT, q)
            if (!v) continue
            program.push({ op: 'ufold' })
            kids.push(v)
        }
        modDepth--
        outerFrame = frameWas
        const body = dist.splice(d0), fbody = full.splice(f0)
        program.push({ op: 'ugroupend' })
        program[start].end = program.length - 1
        userFns.push({ kind: 'group', node: g.id, host: usercode.hostGroup(idx, gc) })
        const d = `d${i}`, f = `f${i}`, m = `m${i}`, k = `k${i}`
        const ind = l => '            ' + l
        const head = [`        WHit uh${i};`, `        const WF wunit_g${i} = WL(${unit.toExponential()});`, '        {', '            WXf wx;', ...gc.decls.map(ind)]
        const kidsD = kids.map(v => `                wCin = w_fold(wCin, wvec4(0.0f, 0.0f, 0.0f, ${v.d} / wunit * wx.S), 0u);`)
        const kidsF = kids.map(v => `                wCin = w_fold(wCin, mk_wvec4(wx.M * wvec3(${v.f}.x, ${v.f}.y, ${v.f}.z), ${v.f}.w / wunit * wx.S), ${v.m} | (${v.k} << 8));`)
        const tail = ['            wvec3 wp = ofV3(' + `xfIn(U + ${xo}, ${pt})` + ') / wunit;', '            WHit wC = w_emptyHit(0);',
            ...gc.lines.map(ind), `            uh${i} = wC;`, '        }']
        const lam = (lines, folds) => ['            auto wchildren = [&](WHit wCin, wvec3 wq) -> WHit {', `                const V3 ${q} = toV3(wq * wunit);`,
            ...lines.map(l => '        ' + l), ...folds, '                return wCin;', '            };']
        dist.push(...head, ...lam(body, kidsD), ...tail, `        float ${d} = uh${i}.s.w * wunit_g${i} * U[${xo + 12}];`)
        full.push(...head, ...lam(fbody, kidsF), ...tail,
            `        V4 ${f} = xfOut(U + ${xo}, v4(v3(uh${i}.s.x, uh${i}.s.y, uh${i}.s.z), uh${i}.s.w * wunit_g${i}));`,
            `        unsigned int ${m} = uh${i}.m & 0xFFu, ${k} = uh${i}.m >> 8;`)
        const flops = kids.reduce((a, v) => a + v.flops, 0) + USER_FLOPS + XF_FLOPS
        return { d, f, m, k, bs: { c: X.t.slice(), r: Infinity }, flops }
    }
    const unionOf = n => Object.assign({}, n, { blend: 0, children: (n.children || []).map(c => Object.assign({}, c, { join: 'union', joinBlend: undefined, joinShape: undefined })) })
    // A modifier group's value: the group's frame point changed by the modifier (or the copies' points), its children
    // evaluated there in the group's frame (their transforms relative to it) and joined as the group says, their distance
    // changed by the modifier, then scaled to the world by the group's scale; the normal turned to the world, or left to
    // differences after a warp.
    function modifierValue(g, X, pt) {
        const key = modifierOf(g)
        const M = MODIFIERS[key]
        const b = blocks[g.block] || {}
        const raw = Object.assign({}, MODIFIER_DEFAULTS[key], b.params || {}, g.params || {})
        const P = {}
        for (const k of Object.keys(raw)) {
            const v = raw[k], f = M.lengths.indexOf(k) >= 0 ? unit : M.perLength.indexOf(k) >= 0 ? 1 / unit : 1
            P[k] = v && typeof v === 'object' ? { x: v.x * f, y: v.y * f, z: v.z * f } : v * f
        }
        for (const c of g.children || []) {
            if (c.type === 'light' || c.type === 'haze') throw new Error(`hipgen: node '${c.id}': a ${c.type} inside modifier group '${g.id}' is not drawn`)
        }
        const xo = push(xfParams(X))
        const po = push(M.params(P))
        const i = tmp++
        const q = `q${i}`, d = `d${i}`, f = `f${i}`, m = `m${i}`, k = `k${i}`
        const start = program.length
        program.push({ op: 'mod', kind: key, xf: xo, p: po, end: -1 })
        const d0 = dist.length, f0 = full.length
        const frameWas = outerFrame
        if (outerFrame === null) outerFrame = xo
        modDepth++
        const inner = isCsg(g) ? csgGroup(g, IDENT, q) : csgGroup(unionOf(g), IDENT, q)
        modDepth--
        outerFrame = frameWas
        if (!inner) { program.splice(start); dist.splice(d0); full.splice(f0); return null }
        const body = dist.splice(d0), fbody = full.splice(f0)
        program.push({ op: 'modend' })
        program[start].end = program.length - 1
        const local = `xfIn(U + ${xo}, ${pt})`
        const scale = `U[${xo + 12}]`
        if (M.copies) {
            // the circular array: every copy, the point turned back by its angle, the least value; its normal turned
            const g0 = `g${i}`
            dist.push(`        const V3 ${g0} = ${local};`, `        float ${d} = 3.0e38f;`,
                `        for (int c${i} = 0; c${i} < (int)U[${po}]; ++c${i}) {`,
                `            float sn${i}, cs${i}; __sincosf((float)c${i} * U[${po + 1}], &sn${i}, &cs${i});`,
                `            const V3 ${q} = v3(cs${i} * ${g0}.x + sn${i} * ${g0}.y, cs${i} * ${g0}.y - sn${i} * ${g0}.x, ${g0}.z);`,
                ...body.map(l => '    ' + l), `            ${d} = fminf(${d}, ${inner.d});`, '        }', `        ${d} *= ${scale};`)
            full.push(`        const V3 ${g0} = ${local};`, `        V4 ${f} = v4(v3s(0.0f), 3.0e38f);`, `        unsigned int ${m} = 0u, ${k} = 0u;`,
                `        for (int c${i} = 0; c${i} < (int)U[${po}]; ++c${i}) {`,
                `            float sn${i}, cs${i}; __sincosf((float)c${i} * U[${po + 1}], &sn${i}, &cs${i});`,
                `            const V3 ${q} = v3(cs${i} * ${g0}.x + sn${i} * ${g0}.y, cs${i} * ${g0}.y - sn${i} * ${g0}.x, ${g0}.z);`,
                ...fbody.map(l => '    ' + l),
                `            if (${inner.f}.w < ${f}.w) { ${f} = v4(v3(cs${i} * ${inner.f}.x - sn${i} * ${inner.f}.y, sn${i} * ${inner.f}.x + cs${i} * ${inner.f}.y, ${inner.f}.z), ${inner.f}.w); ${m} = ${inner.m}; ${k} = ${inner.k}; }`,
                '        }', `        ${f} = xfOut(U + ${xo}, ${f});`)
        } else {
            const qv = M.point ? M.point(po, local) : local
            const after = v => M.after ? M.after(po, v) : v
            const kf = M.warp && !M.unscaled ? ` * U[${po + 1}]` : ''
            dist.push(`        const V3 ${q} = ${qv};`, ...body, `        float ${d} = ${after(inner.d)}${kf} * ${scale};`)
            const fl = M.warp ? `v4(v3s(0.0f), ${inner.f}.w${kf})` : M.afterFull ? M.afterFull(po, inner.f) : inner.f
            full.push(`        const V3 ${q} = ${qv};`, ...fbody, `        V4 ${f} = xfOut(U + ${xo}, ${fl});`,
                `        unsigned int ${m} = ${inner.m}, ${k} = ${inner.k};`)
        }
        // the modifier's bound in the group's frame, then in the world
        const lb = M.bound(inner.bs.c, inner.bs.r, P)
        const R = X.R
        const wc = [0, 1, 2].map(a => X.t[a] + X.s * (R[a][0] * lb.c[0] + R[a][1] * lb.c[1] + R[a][2] * lb.c[2]))
        const wr = lb.r * X.s
        if (isFinite(wr)) {
            for (let a = 0; a < 3; a++) { bounds[0][a] = Math.min(bounds[0][a], wc[a] - wr); bounds[1][a] = Math.max(bounds[1][a], wc[a] + wr) }
        }
        const copies = M.copies ? Math.min(Math.max(Math.round(P.count), 1), 64) : 1
        return { d, f, m, k, bs: { c: wc, r: wr }, flops: copies * (inner.flops + M.flops) + XF_FLOPS }
    }
    // a node's value as expressions (for use inside a CSG group); pt: the point it is evaluated at
    function valueOf(n, parentX, pt = 'p') {
        const X = compose(parentX, n.transform, unit)
        if (n.kind === 'group') {
            if (groupCodeOf(n)) return userGroupValue(n, X, pt)
            modifierCheck(n)
            if (modifierOf(n)) return modifierValue(n, X, pt)
            if (isCsg(n)) return csgGroup(n, X, pt)
            // a plain group inside a CSG group: its children unioned
            return csgGroup(unionOf(n), X, pt)
        }
        n.__world = X
        // a ray warp (its block the region where rays bend; builder/warp.js) is not a surface: refused, or left out
        if (n.warp) { cannot(`node '${n.id}' (${n.label || ''})`, 'a ray warp'); return null }
        if (n.type !== 'light' && shapeOf(n) === 'code') return userObjectValue(n, X, pt)
        const s = shape(n, X)
        const mat = matOf(n)
        const i = tmp++
        const d = `d${i}`, f = `f${i}`, m = `m${i}`, k = `k${i}`
        const prim = program.filter(o => o.op === 'prim' || o.op === 'user').length
        primFrames[prim] = outerFrame !== null ? outerFrame : s.xo
        dist.push(`        float ${d} = ${SHAPES[s.key].dist(s.po)} * U[${s.xo + 12}];`.replace(/\bq\b/, `xfIn(U + ${s.xo}, ${pt})`))
        full.push(`        V4 ${f} = xfOut(U + ${s.xo}, ${SHAPES[s.key].full(s.po)});`.replace(/\bq\b/, `xfIn(U + ${s.xo}, ${pt})`))
        full.push(`        unsigned int ${m} = ${mat}u, ${k} = ${prim}u;`)
        program.push({ op: 'prim', shape: s.key, xf: s.xo, p: s.po, mat, prim })
        const rel = n.type !== 'light' ? reliefOn(n, s.xo, pt, d, f) : null
        if (rel) {
            dist.push(...rel.dist); full.push(...rel.full)
            // the relief may lift the surface by its amplitude: the bound grows by it
            s.bs = { c: s.bs.c, r: s.bs.r + 2 * (docMats[n.material].relief.params.amplitude || 0) * 0.001 }
        }
        return { d, f, m, k, bs: s.bs, flops: s.flops }
    }
    // the closest-surface fold over a plain group's descendants
    function foldNode(n, parentX) {
        if (n.enabled === false) return
        if (n.type === 'haze') { hazeOf(n, compose(parentX, n.transform, unit)); return }
        if (n.kind === 'group' && !isCsg(n) && !modifierOf(n) && !groupCodeOf(n)) {
            modifierCheck(n)
            const X = compose(parentX, n.transform, unit)
            for (const c of n.children || []) foldNode(c, X)
            return
        }
        // the node's code, guarded by its bounding sphere: skipped where |p - c| - r cannot beat the closest |d| so
        // far, which leaves the fold's answer exactly as it was (a skipped node is never the closest, nor negative)
        // for an exact distance. For a bound (the lens's largest-of-pieces, a smooth join) the skipped value may be
        // smaller, but the node's true distance is not, so the step stays safe and its surfaces are where they were.
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

    // the materials' own code (those the drawn nodes use): one case each of the scene's material hook
    const matCases = []
    for (const id of Object.keys(checked)) {
        const m = docMats[id]
        if (!m || !usercode.isCode(m.code)) continue
        try {
            matCases.push(Object.assign({ mat: matIndex[id] }, usercode.materialCode(m.code, m.params || {}, unit, push)))
        } catch (e) {
            cannot(`material '${id}'`, `its code: ${e.message}`)
        }
    }
    const hook = matCases.length ? usercode.materialHook(matCases, primFrames.map(f => (f === undefined ? -1 : f)), unit) : null

    const envRadius = (doc.envRadius !== undefined ? doc.envRadius * unit : 400)
    let env = environmentOf(doc.environment || {}, doc, cannot)
    // the scene's own sky (builder/sky.js): its code lights the scene, whatever the environment's mode
    let skyMod = null
    if (usercode.isCode(doc.sky)) {
        try {
            skyMod = usercode.moduleCode(doc.sky, usercode.skyParamNames(doc.skyParams))
            if (!skyMod.own.background) throw new Error('it defines no background(dir)')
            env = Object.assign({}, env, { type: 5, L: 1, sunE: 0, sunL: [0, 0, 0] })
        } catch (e) {
            cannot('environment', `the sky's code: ${e.message}`)
            skyMod = null
        }
    } else if (((doc.environment || {}).envMode || 0) === 2) {
        cannot('environment', 'the scene\'s own code (envMode 2) with no sky code')
    }
    if (approximations.length && !opts.approximate) {
        throw new Error(`hipgen: the engine does not draw ${approximations.join('; ')} -- generate(doc, { approximate: true }) renders the scene without them`)
    }

    const hip = [
        `// emitted by shadytracer_amd/js/hipgen.js from '${(doc.title || 'scene').replace(/[\n\r]/g, ' ')}'`,
        `#define PT_SCENE_MATERIALS ${materials.length}`,
        `#define PT_SCENE_MASK_BITS ${maskBits}`,
        `#define PT_SCENE_MATCODE ${hook ? 1 : 0}`,
        `#define PT_SCENE_SKY ${skyMod ? 1 : 0}`,
        ...(skyMod ? ['// the scene\'s own sky', skyMod.text,
            '__device__ __forceinline__ V3 sceneSky(V3 d) { const wvec3 r = u_background(wvec3(d.x, d.y, d.z)); return v3(r.x, r.y, r.z); }'] : []),
        ...helpers,
        '// the nodes a ray from o along d can reach within tmax: those whose bounding sphere, widened by margin, its',
        '// segment meets (bit k of the mask: the k-th bounded top-level node); unbounded nodes are always marched',
        '__device__ __forceinline__ unsigned long long sceneMask(ParamPtr U, V3 o, V3 d, float tmax, float margin) {',
        '    PT_SCENE_PARAMS(U);',
        '    unsigned long long m = ~0ull;',
        ...maskCode,
        '    return m;',
        '}',
        '// the closest signed distance over the nodes in mask; w counts the FP32 operations spent',
        '__device__ __forceinline__ float sceneDistM(ParamPtr U, V3 p, unsigned long long mask, unsigned int& w) {',
        '    PT_SCENE_PARAMS(U);',
        '    float best = 1.0e7f;',
        ...dist,
        '    return best;',
        '}',
        '__device__ __forceinline__ float sceneDist(ParamPtr U, V3 p) { unsigned int w = 0u; return sceneDistM(U, p, ~0ull, w); }',
        '__device__ __forceinline__ SdfHit sceneFull(ParamPtr U, V3 p) {',
        '    PT_SCENE_PARAMS(U);',
        '    SdfHit h = sdfEmpty();',
        ...full,
        '    return h;',
        '}',
        ...(hook ? ['// the materials\' own code, run on a hit\'s material before it is used', hook.device] : []),
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
        // what the environment was made from: the document's settings and the little of the document it reads, so a
        // host's lighting settings (createPathTracer's set) can be laid over them later without the document
        envSettings: doc.environment || {},
        envDoc: { unit: doc.unit || 'm', sky: doc.sky },
        envRadius,
        unit,
        // the diagonal of the scene's bounds, metres (null when nothing in it is bounded), and the tolerance it implies
        extent: isFinite(bounds[0][0]) ? Math.hypot(...[0, 1, 2].map(i => bounds[1][i] - bounds[0][i])) : null,
        epsAuto: autoEps(isFinite(bounds[0][0]) ? Math.hypot(...[0, 1, 2].map(i => bounds[1][i] - bounds[0][i])) : null),
        program,
        // the objects' and groups' own code as the oracle compiles it (FP64 C++ over oracle/wgsl_host.h), by index
        user: userFns.map(u => ({ kind: u.kind, node: u.node, cpp: u.host })),
        reliefs: reliefList.map(r => ({ material: r.material, cpp: r.host })),
        materialCode: hook ? hook.host : null,
        functionsCode: functionsCpp || null,
        skyCode: skyMod ? [skyMod.text, 'extern "C" void usky(const double* d, double* out) {',
            '    const wvec3 r = u_background(wvec3(d[0], d[1], d[2]));', '    out[0] = r.x; out[1] = r.y; out[2] = r.z;', '}'].join('\n') : null,
        flops: { dist: flops, maskNode: 14 },   // every node evaluated: the bound a culled step stays under
        approximations,
    }
}

// ---- what the engine's material table, light faces and environments cannot carry (each a reason, or none)
// A code block with statements (not only comments) is user WGSL (builder/materials.js: the material's code runs per hit).
const hasStatements = code => typeof code === 'string' && code.split('\n').some(l => l.trim() && !l.trim().startsWith('//'))
function materialLimits(m) {
    const out = [], P = (m && m.params) || {}
    if (m.relief && !usercode.isCode(m.relief.code)) out.push('a relief without code')
    if (m.images && m.images.length) out.push('images')
    if (m.type && m.type !== 'emitter') out.push(`type '${m.type}'`)
    if (P.coating > 0) out.push('a thin-film coating')
    if (P.anisotropy > 0) out.push('anisotropy')
    if (P.haze > 0 || P.scatterDistance > 0) out.push('scattering inside')
    return out
}
// lights.js LIGHT_DEFAULTS:62-67: a face emitting cos^1 (Lambertian) over the hemisphere, into air, opaque
function lightLimits(L) {
    const out = []
    if (L.shape && L.shape !== 'rectangle' && L.shape !== 'disc') out.push(`a ${L.shape} light`)
    if ((L.distribution || 'cosine') !== 'cosine' || (L.n !== undefined && L.n !== 1)) out.push(`a ${L.distribution === 'cosine' ? 'cos^' + L.n : L.distribution} beam`)
    if (L.halfAngle !== undefined && L.halfAngle < 90) out.push('a beam limited to a cone')
    if (L.transparent) out.push('a transparent light')
    if (L.emitN !== undefined && L.emitN !== 1) out.push('emission into a medium (emitN)')
    if (['spectrum', 'file', 'wavelength'].includes(L.colourBy)) out.push(`colour by ${L.colourBy}`)
    return out
}

// ---- a light's colour, linear Rec.709 of luminance 1 (lights.js lightRGB:85-100 over core/colorimetry.js): given as
// RGB, as a chromaticity xy, or as a CCT and Duv -- the Planckian locus by Krystek's 1985 rational fit in CIE 1960
// uv, Duv along the locus's unit normal (positive towards green), then xy -> XYZ at Y = 1 -> RGB by the inverse of
// IEC 61966-2-1's four-decimal matrix (whose Y row is LUMA).
const RGB709_TO_XYZ = [[0.4124, 0.3576, 0.1805], LUMA, [0.0193, 0.1192, 0.9505]]
function inverse3(m) {
    const [[a, b, c], [d, e, f], [g, h, i]] = m
    const A = e * i - f * h, B = f * g - d * i, C = d * h - e * g, det = a * A + b * B + c * C
    return [[A / det, (c * h - b * i) / det, (b * f - c * e) / det], [B / det, (a * i - c * g) / det, (c * d - a * f) / det],
        [C / det, (b * g - a * h) / det, (a * e - b * d) / det]]
}
const XYZ_TO_RGB709 = inverse3(RGB709_TO_XYZ)
// Krystek's fit of the Planckian locus in CIE 1960 uv, each coordinate a ratio of quadratics in T (1000-15000 K),
// with the locus's direction dP/dT from the same ratios' derivatives, (N'D - N D') / D^2.
const KRYSTEK_U = [[0.860117757, 1.54118254e-4, 1.28641212e-7], [1, 8.42420234e-4, 7.08145163e-7]]
const KRYSTEK_V = [[0.317398726, 4.22806245e-5, 4.20481691e-8], [1, -2.89741816e-5, 1.61456053e-7]]
function ratioAndSlope([N, D], t) {
    const n = N[0] + t * (N[1] + t * N[2]), d = D[0] + t * (D[1] + t * D[2])
    const dn = N[1] + 2 * t * N[2], dd = D[1] + 2 * t * D[2]
    return [n / d, (dn * d - n * dd) / (d * d)]
}
// A CCT and Duv as a chromaticity: the locus's point at T moved duv along its unit normal, the tangent turned a quarter
// clockwise -- towards larger v, green, for duv > 0 -- then uv back to xy.
function cctDuvToXy(T, duv) {
    const t = Math.min(15000, Math.max(1000, T))
    const [u0, su] = ratioAndSlope(KRYSTEK_U, t), [v0, sv] = ratioAndSlope(KRYSTEK_V, t)
    const len = Math.hypot(su, sv) || 1
    const u = u0 + duv * sv / len, v = v0 - duv * su / len
    const den = 2 * u - 8 * v + 4
    return [3 * u / den, 2 * v / den]
}
function rgbFromXy(x, y) {
    const XYZ = y > 0 ? [x / y, 1, (1 - x - y) / y] : [0, 0, 0]
    return XYZ_TO_RGB709.map(r => r[0] * XYZ[0] + r[1] * XYZ[1] + r[2] * XYZ[2])
}
function lightRGB(L) {
    let rgb
    if (L.colourBy === 'cct') rgb = rgbFromXy(...cctDuvToXy(L.cct !== undefined ? L.cct : 4000, L.duv || 0))
    else if (L.colourBy === 'xy') { const c = L.xy || { x: 0.3127, y: 0.329 }; rgb = rgbFromXy(c.x, c.y) }
    else rgb = colourOf(L.colour || L.color, [1, 1, 1])
    const Y = LUMA[0] * rgb[0] + LUMA[1] * rgb[1] + LUMA[2] * rgb[2]
    return Y > 0 ? rgb.map(v => v / Y) : [1, 1, 1]
}

// ---- the document's own camera (its view: lib/camera.js syncOrbitPose:137 and frame:106): an orbit about target,
// azimuth about +z from +x, elevation the angle down from +z, the vertical field of view fov; metres
function cameraOf(doc, width, height, cannot) {
    const C = (doc.view && doc.view.camera) || {}
    const unit = UNIT_METRES[doc.unit || 'm']
    const r = Math.PI / 180
    if (C.mode && C.mode !== 'orbit') cannot('camera', `a ${C.mode} camera`)
    if (C.camtype) cannot('camera', 'an orthographic camera')
    if (C.roll) cannot('camera', 'roll')
    const t = (C.target || [0, 0, 0]).map(v => v * unit)
    const az = (C.azimuth !== undefined ? C.azimuth : 40) * r, el = (90 - (C.elevation !== undefined ? C.elevation : 70)) * r
    const dist = (C.distance !== undefined ? C.distance : 4) * unit
    const pos = [t[0] + dist * Math.cos(el) * Math.cos(az), t[1] + dist * Math.cos(el) * Math.sin(az), t[2] + dist * Math.sin(el)]
    const norm = v => { const l = Math.hypot(v[0], v[1], v[2]) || 1; return v.map(x => x / l) }
    const cross = (a, b) => [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]
    const z = norm([pos[0] - t[0], pos[1] - t[1], pos[2] - t[2]])
    let x = cross(C.up || [0, 0, 1], z)
    if (Math.hypot(x[0], x[1], x[2]) < 1e-6) x = cross(Math.abs(z[2]) < 0.9 ? [0, 0, 1] : [1, 0, 0], z)
    x = norm(x)
    const fy = Math.tan(0.5 * (C.fov || 60) * r)
    // the thin lens (lib/camera.js:336-338): its radius, half the aperture, and the distance in focus, the target's
    // for a focus of 0; both in the document's units
    const lensR = Math.max(+C.aperture || 0, 0) / 2 * unit
    const focus = (+C.focus > 0 ? +C.focus * unit : dist) || 1
    return { pos, right: x, up: cross(z, x), fwd: z, fx: fy * width / height, fy, type: 0, lensR, focus }
}

// The environment. A builder document's (uniforms.schema.js envMode:429): 1 the studio, L(d) = L0 (1 + top d.z) with
// L0 = E / (pi (1 + 2 top / 3)) so that a level surface facing up receives studioIlluminance lux (photometry.js
// studioLuminance:154, environment.js studioLight:165), its colour a tint of luminance 1, and no sun; 0 the sun & sky
// (js/sky.js); 2 the scene's own code, which the engine does not draw. The configurations here name theirs directly:
// `sky: 'sphere' | 'hemisphere' | 'none'` of `uniformSky` cd/m^2, and an optional sun -- direction from
// azimuth/elevation (sdf.js:72-76, Z-up), illuminance in lux, angular diameter in degrees, and an aperture (a beam
// radius about a point).
function environmentOf(E, doc, cannot) {
    cannot = cannot || (() => {})
    const env = { type: 0, L: 0, col: [1, 1, 1], sunDir: [0, 0, 1], sunCos: 1, sunE: 0, sunRadius: 0, sunPoint: [0, 0, 0], top: 0,
        sunL: [0, 0, 0], sunSin2: 0, sunOmega: 0, skyAerosol: 0, skyMs: 1, skyUp: 0 }
    const cone = d => { const c = sky.sunCone(d); env.sunCos = c.cos; env.sunSin2 = c.sin2; env.sunOmega = c.omega; return c }
    if (E.sky === undefined) {
        const mode = E.envMode || 0
        if (mode === 2) return env
        if (mode === 0) {
            // the sun & sky (environment.js SUNSKY_FN; photometry.js derive:183 puts the sun's cone in use while the
            // hint is on and its disc not wholly below the horizon)
            const num = (v, d) => v !== undefined && v !== null && isFinite(+v) ? +v : d
            const c = cone(num(E.sunDiameter, 0.53))
            const el = num(E.uSunElevation, 45), turb = num(E.uSkyTurbidity, 0.7)
            const s = sky.sunDirection(num(E.uSunAzimuth, 45), el)
            env.type = 4
            env.L = Math.max(num(E.sunOutside, 128000), 0)
            env.sunDir = s
            env.skyAerosol = sky.aerosolOf(turb)
            env.skyMs = sky.airMass(Math.max(s[2], 0))
            // envSky and envSun (uniforms.schema.js): the sky's and the sun's gains, 0 switching either off
            const skyGain = Math.max(num(E.envSky, 1), 0), sunGain = Math.max(num(E.envSun, 1), 0)
            env.skyUp = Math.min(Math.max((s[2] + 0.05) / 0.05, 0), 1)
            env.col = [skyGain, skyGain, skyGain]          // the sky's gain (a tint for the sun & sky)
            if (el * Math.PI / 180 + c.half > 0 && env.L > 0 && sunGain > 0) {
                env.sunL = sky.sunRadiance(s, turb, env.L, c.sin2).map(x => x * sunGain)
                env.sunE = (LUMA[0] * env.sunL[0] + LUMA[1] * env.sunL[1] + LUMA[2] * env.sunL[2]) * c.omega
            }
            if (env.L <= 0) env.type = 0
            return env
        }
        const top = Math.min(1, Math.max(0, E.studioTop !== undefined ? +E.studioTop : 0.5))
        const c = colourOf(E.studioColor, [1, 1, 1]).map(v => Math.max(v, 0))
        const y = LUMA[0] * c[0] + LUMA[1] * c[1] + LUMA[2] * c[2]
        env.L = Math.max(0, E.studioIlluminance !== undefined ? +E.studioIlluminance : 1000) / (Math.PI * (1 + 2 * top / 3))
        env.type = env.L > 0 ? 3 : 0
        env.col = c.map(v => v / Math.max(y, 1e-6))
        env.top = top
        return env
    }
    const kind = E.sky
    let L = E.uniformSky !== undefined ? E.uniformSky : 0
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
        cone(E.sunDiameter || 0.53)
        env.sunE = E.sunIlluminance
        const unit = UNIT_METRES[doc.unit || 'm']
        if (E.sunAperture) {
            env.sunRadius = E.sunAperture.radius * unit
            env.sunPoint = [E.sunAperture.point.x * unit, E.sunAperture.point.y * unit, E.sunAperture.point.z * unit]
        }
    }
    return env
}

module.exports = { generate, environmentOf, cameraOf, lightRGB, cctDuvToXy, rotationOf, autoEps, UNIT_METRES, PRIM_FLOPS }
