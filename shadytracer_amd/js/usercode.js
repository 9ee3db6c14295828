'use strict'
// Scene code written by the document's author -- an object's or a group's WGSL (builder/codegen.js:286-306,
// :609-648) -- as C++ for the scene module (hipRTC, FP32) and for the oracle's host compile (FP64), from one
// translation (wgsl.js). The code runs as the reference's does, in scene units: the point it is given is the node's
// local point in the document's unit, its parameters are the document's numbers, and what it folds into its Hit is
// taken back to metres by the emitter.
//
// In scope (the reference's names, core/scene_api.js and core/sdf.js, core/shapes.js, core/noise.js): p, C, r,
// MATERIAL, the node's parameters by name, add_new, fold, emptyHit, warp, children (a group's), XF_S, XF_M, XF_K,
// XF_SUB, LOCAL_CAPTURE, SURFACE_ONLY, READ_REACH, DIVIDED_ABOVE, mm, metres, PI, the builtins of WGSL, the shape
// functions and the noise below. Anything else is refused with its line and column.
const wgsl = require('./wgsl')

const BUILTINS = ['abs', 'sign', 'min', 'max', 'clamp', 'mix', 'step', 'smoothstep', 'sqrt', 'inverseSqrt', 'pow', 'exp',
    'exp2', 'log', 'log2', 'sin', 'cos', 'tan', 'asin', 'acos', 'atan', 'atan2', 'sinh', 'cosh', 'tanh', 'floor', 'ceil',
    'round', 'fract', 'trunc', 'saturate', 'radians', 'degrees', 'fma', 'length', 'distance', 'dot', 'cross', 'normalize',
    'reflect', 'refract', 'faceForward', 'select', 'any', 'all', 'transpose', 'determinant']
const SHAPE_FNS = ['sd', 'sdSphere', 'sdBox', 'sdRoundedBox', 'sdCylinder', 'sdCappedCone', 'sdMenger', 'Nsphere',
    'Nroundedbox', 'Nbox', 'NboxRounded', 'Ncylinder', 'Ncapsule', 'Ntorus', 'Nplane', 'Nrectangle', 'Ndisc', 'NroundCone',
    'Noctahedron', 'Nlink', 'Ntube', 'Ndome', 'safeDir', 'g2Extrude', 'g2Revolve']
const NOISE_FNS = ['vnoise3', 'gnoise3', 'fbm3', 'nzFbmLip', 'worley3', 'worley3x']

// the document's own functions (its `functions` block), callable from every code of the document: set per document
let shared = {}
function setShared(own) { shared = Object.assign({}, own || {}) }
function libraryNames() {
    const names = Object.assign({}, shared)
    for (const b of BUILTINS.concat(SHAPE_FNS, NOISE_FNS)) names[b] = `w_${b}`
    Object.assign(names, {
        NZ_GRAD_LIP_MEASURED: 'WL(4.3)', NZ_GRAD_VALUE_BOUND: 'WL(1.3856406460551018)', NZ_GRAD_SCALE: 'WL(1.6)',
        PI: 'WL(3.141592653589793)',
        add_new: 'wx.add_new', fold: 'w_fold', emptyHit: 'w_emptyHit', warp: 'wx.warp',
        XF_S: 'wx.S', XF_M: 'wx.M', XF_K: 'wx.K', XF_SUB: 'wx.SUB',
        LOCAL_CAPTURE: 'false', SURFACE_ONLY: 'false', READ_REACH: 'WL(1e30)', DIVIDED_ABOVE: 'false',
        r: '0', mm: 'w_mm', metres: 'w_metres',
    })
    return names
}
const HIT_FNS = new Set(['add_new', 'fold', 'emptyHit', 'children'])

/**
 * A module of the document's WGSL -- its `functions` block (builder/codegen.js:1366-1370, callable from every node's,
 * group's and material's code) or its sky (builder/sky.js, core/environment.js ENV_FN: `background(dir)` gives the
 * environment's radiance) -- as C++ functions and constants. `extra`: names in scope besides the library (the sky's
 * parameters). Returns { text (the C++), own (the module's names as C++) }.
 */
function moduleCode(src, extra) {
    const names = Object.assign(libraryNames(), extra || {})
    for (const k of ['add_new', 'fold', 'emptyHit', 'warp', 'r', 'XF_S', 'XF_M', 'XF_K', 'XF_SUB', 'children', 'mm', 'metres']) delete names[k]
    const t = wgsl.translate(src, { names, hitFns: new Set(), module: true })
    return { text: t.fns.join('\n'), own: t.own }
}
// The sky's parameters (doc.skyParams: [{name, value}], a number or three) as literals of the sky's code
function skyParamNames(params) {
    const out = {}
    for (const p of params || []) {
        if (!/^[A-Za-z_][A-Za-z0-9_]*$/.test(p.name || '')) throw new Error(`sky parameter name '${p.name}'`)
        const v = p.value
        out[p.name] = Array.isArray(v) ? `wvec3(WL(${(+v[0]).toExponential()}), WL(${(+v[1]).toExponential()}), WL(${(+v[2]).toExponential()}))` : `WL(${(+v).toExponential()})`
    }
    return out
}

// A parameter's C++ declaration: a number through U (live), a compile-time constant as a literal, a vector through U
function paramDecls(params, constants, push) {
    const decls = [], names = {}
    for (const k of Object.keys(params)) {
        const v = params[k]
        if (!/^[A-Za-z_][A-Za-z0-9_]*$/.test(k)) throw new Error(`parameter name '${k}'`)
        const cpp = `P_${k}`
        names[k] = cpp
        const ty = constants[k]
        if (ty) {
            const lit = ty === 'u32' ? `${Math.max(0, Math.round(+v))}u` : ty === 'i32' ? `${Math.round(+v)}` : `WL(${Number(v).toExponential()})`
            decls.push(`const ${ty === 'u32' ? 'unsigned int' : ty === 'i32' ? 'int' : 'WF'} ${cpp} = ${lit};`)
        } else if (typeof v === 'number' || typeof v === 'boolean') {
            decls.push(`const WF ${cpp} = (WF)U[${push([+v])}];`)
        } else if (v && typeof v === 'object' && !Array.isArray(v) && 'r' in v && 'g' in v && 'b' in v) {
            const o = push([v.r, v.g, v.b]); decls.push(`const wvec3 ${cpp} = wvec3((WF)U[${o}], (WF)U[${o + 1}], (WF)U[${o + 2}]);`)
        } else if (v && typeof v === 'object' && !Array.isArray(v) && 'x' in v && 'y' in v) {
            if ('w' in v) { const o = push([v.x, v.y, v.z, v.w]); decls.push(`const wvec4 ${cpp} = wvec4((WF)U[${o}], (WF)U[${o + 1}], (WF)U[${o + 2}], (WF)U[${o + 3}]);`) }
            else if ('z' in v) { const o = push([v.x, v.y, v.z]); decls.push(`const wvec3 ${cpp} = wvec3((WF)U[${o}], (WF)U[${o + 1}], (WF)U[${o + 2}]);`) }
            else { const o = push([v.x, v.y]); decls.push(`const wvec2 ${cpp} = wvec2((WF)U[${o}], (WF)U[${o + 1}]);`) }
        } else {
            throw new Error(`parameter '${k}' is ${Array.isArray(v) ? 'an outline' : v && v.path ? 'a path' : 'of a kind'} that scene code here does not take`)
        }
    }
    return { decls, names }
}

function prelude(unit) {
    return [
        `const WF wunit = WL(${unit.toExponential()});`,
        'auto w_mm = [&](WF x) -> WF { return x * WL(0.001) / wunit; };',
        'auto w_metres = [&](WF x) -> WF { return x / wunit; };',
        '(void)w_mm; (void)w_metres;',
    ]
}

/**
 * An object's code: the body of `ucode<i>(U, q, m)` -> (normal, distance in metres) of what it folded, q the node's
 * local point in metres. Returns { device, host } function texts.
 */
function objectFunction(i, code, params, constants, material, unit, push) {
    const P = paramDecls(params, constants, push)
    const names = Object.assign(libraryNames(), P.names, { MATERIAL: `${material}u` })
    const t = wgsl.translate(code, {
        names, hitFns: HIT_FNS, indent: '    ',
        vars: { p: { cpp: 'wp' }, C: { cpp: 'wC', hit: true } },
        assignable: new Set(['XF_S', 'XF_M', 'XF_K', 'XF_SUB']),
    })
    if (t.calls.has('children')) throw new Error("an object's code calls children(): only a group's may")
    const body = [
        '    WXf wx;',
        ...prelude(unit).map(l => '    ' + l),
        ...P.decls.map(l => '    ' + l),
        '    wvec3 wp = wvec3((WF)q[0], (WF)q[1], (WF)q[2]) / wunit;',
        '    WHit wC = w_emptyHit(0);',
        ...t.lines,
        '    *m = wC.m;',
        '    out[0] = wC.s.x; out[1] = wC.s.y; out[2] = wC.s.z; out[3] = wC.s.w * wunit;',
    ]
    const device = [`__device__ __forceinline__ V4 ucode${i}(ParamPtr U, V3 qv, unsigned int* m) {`,
        '    const float q[3] = {qv.x, qv.y, qv.z};', '    float out[4];', ...body,
        '    return v4(v3(out[0], out[1], out[2]), out[3]);', '}'].join('\n')
    const host = [`extern "C" void ucode${i}(const float* U, const double* q, double* out, unsigned int* m) {`, ...body, '}'].join('\n')
    return { device, host }
}

/**
 * A group's code around its children: `children` is the C++ of a lambda body's value lines (device) that sets
 * the children's fold. Returns the translated lines and declarations for the emitter to place.
 */
function groupCode(code, params, constants, unit, push) {
    const P = paramDecls(params, constants, push)
    const names = Object.assign(libraryNames(), P.names, { children: 'wchildren', MATERIAL: '0u' })
    const t = wgsl.translate(code, {
        names, hitFns: HIT_FNS, indent: '',
        vars: { p: { cpp: 'wp' }, C: { cpp: 'wC', hit: true } },
        assignable: new Set(['XF_S', 'XF_M', 'XF_K', 'XF_SUB']),
    })
    return { decls: [...prelude(unit), ...P.decls], lines: t.lines, calls: t.calls }
}

/**
 * A material's code (builder/codegen.js:650-700 emitMaterials: it runs once a hit, after the material's fields are in
 * m): m's baseColor, metallic, roughness, transmission, ior, specular, specularTint and tangent (read and written),
 * the hit p and the surface normal n in the world (scene units), p_local, R_local and S_local (the hit in its object's
 * frame, that frame's turn to the world and its scale: p_local * S_local is in scene units), the material's parameters
 * by name. Returns the case's lines (device and host share them) and whether the local frame is asked for.
 */
function materialCode(code, params, unit, push) {
    const P = paramDecls(params, {}, push)
    // the Hit's words are not a material's; its parameters may take their names
    const names = libraryNames()
    for (const k of ['add_new', 'fold', 'emptyHit', 'warp', 'r', 'XF_S', 'XF_M', 'XF_K', 'XF_SUB']) delete names[k]
    Object.assign(names, P.names, { p_local: 'wpl', R_local: 'wRl', S_local: 'wSl' })
    const t = wgsl.translate(code, {
        names, hitFns: new Set(), indent: '',
        vars: { m: { cpp: 'wm', hit: true }, p: { cpp: 'wp' }, n: { cpp: 'wn' } },
    })
    const local = /\b(p_local|R_local|S_local)\b/.test(code.replace(/\/\/[^\n]*/g, ''))
    return { decls: [...prelude(unit), ...P.decls], lines: t.lines, local }
}

// The scene's material hook in C++: `cases` [{mat, decls, lines, local}], `frames` the U offset of each primitive's
// object frame. Device: sceneMaterial(U, mat, p, n, m) after sceneFull (which gives the hit's primitive, once more,
// where a case asks for its local frame). Host: umaterial(U, mat, p, n, prim, m) for the oracle, which passes the
// primitive of its own hit.
function materialHook(cases, frames, unit) {
    const frameOf = frames.length ? `const int wfr[${frames.length}] = {${frames.join(', ')}};` : 'const int wfr[1] = {-1};'
    const body = (pt) => {
        const out = []
        for (const c of cases) {
            out.push(`    case ${c.mat}u: {`)
            out.push(...c.decls.map(l => '        ' + l))
            if (c.local) {
                out.push(`        ${pt}`,
                    `        const int wxo = (wprim >= 0 && wprim < ${Math.max(frames.length, 1)}) ? wfr[wprim] : -1;`,
                    '        wvec3 wpl = wp; wmat3 wRl = mk_wmat3(wvec3(1, 0, 0), wvec3(0, 1, 0), wvec3(0, 0, 1)); WF wSl = 1;',
                    '        if (wxo >= 0) {',
                    '            const wvec3 c0((WF)U[wxo + 3], (WF)U[wxo + 4], (WF)U[wxo + 5]), c1((WF)U[wxo + 6], (WF)U[wxo + 7], (WF)U[wxo + 8]), c2((WF)U[wxo + 9], (WF)U[wxo + 10], (WF)U[wxo + 11]);',
                    '            const wvec3 d = wp * wunit - wvec3((WF)U[wxo], (WF)U[wxo + 1], (WF)U[wxo + 2]);',
                    '            wSl = (WF)U[wxo + 12];',
                    '            wpl = wvec3(w_dot(d, c0), w_dot(d, c1), w_dot(d, c2)) / (wSl * wunit);',
                    '            wRl = mk_wmat3(c0, c1, c2);',
                    '        }',
                    '        (void)wpl; (void)wRl; (void)wSl;')
            }
            out.push(...c.lines.map(l => '        ' + l), '        break;', '    }')
        }
        return out
    }
    const pre = ['    WMat wm = wmatOf(m);', `    const WF wunit0 = WL(${unit.toExponential()});`,
        '    wvec3 wp = wvec3((WF)pw[0], (WF)pw[1], (WF)pw[2]) / wunit0;', '    const wvec3 wn((WF)nw[0], (WF)nw[1], (WF)nw[2]);',
        '    (void)wn;', '    ' + frameOf, '    (void)wfr;', '    switch (mat) {']
    const post = ['    default: break;', '    }', '    wmatTo(wm, m);']
    const device = ['__device__ __forceinline__ void sceneMaterial(ParamPtr U, unsigned int mat, V3 pv, V3 nv, PtMaterial& m) {',
        '    PT_SCENE_PARAMS(U);', '    const float pw[3] = {pv.x, pv.y, pv.z}, nw[3] = {nv.x, nv.y, nv.z};', ...pre,
        ...body('const int wprim = (int)sceneFull(U, pv).prim;'), ...post, '}'].join('\n')
    const host = ['extern "C" void umaterial(const float* U, unsigned int mat, const double* pw, const double* nw, int wprim0, PtMaterial& m) {',
        ...pre, ...body('const int wprim = wprim0;'), ...post, '}'].join('\n')
    return { device, host }
}

/**
 * A material's relief (builder/codegen.js:1380-1406, :1735-1760): its code sets h, the height in millimetres, and slope,
 * a bound on the height's slope, at q (the point in the frame of the object wearing it, real millimetres) and world
 * (the scene's point, millimetres); its parameters by name. Device `urelief<i>(U, q, world)` -> (h, slope); host the
 * same through pointers.
 */
function reliefFunction(i, code, params, constants, push) {
    const P = paramDecls(params, constants, push)
    const names = libraryNames()
    for (const k of ['add_new', 'fold', 'emptyHit', 'warp', 'r', 'XF_S', 'XF_M', 'XF_K', 'XF_SUB', 'mm', 'metres']) delete names[k]
    Object.assign(names, P.names)
    const t = wgsl.translate(code, {
        names, hitFns: new Set(), indent: '        ',
        vars: { q: { cpp: 'wq' }, world: { cpp: 'wworld' }, h: { cpp: 'wh' }, slope: { cpp: 'wslope' } },
    })
    const body = ['    WF wh = 0, wslope = 0;', '    {', ...P.decls.map(l => '        ' + l), ...t.lines, '    }']
    const device = [`__device__ __forceinline__ wvec2 urelief${i}(ParamPtr U, wvec3 wq, wvec3 wworld) {`, '    (void)wq; (void)wworld;',
        ...body, '    return wvec2(wh, wslope);', '}'].join('\n')
    const host = [`extern "C" void urelief${i}(const float* U, const double* q, const double* w, double* out) {`,
        '    const wvec3 wq(q[0], q[1], q[2]), wworld(w[0], w[1], w[2]);', '    (void)wq; (void)wworld;', ...body,
        '    out[0] = wh; out[1] = wslope;', '}'].join('\n')
    return { device, host }
}

// The oracle's form of a group's code: its children come from the oracle's own evaluation of them, through `kids`
// (each child's normal and distance in the group's frame, metres, and its material | primitive << 8)
function hostGroup(i, gc) {
    return [`extern "C" void ugroup${i}(const float* U, const double* q, void* self, int (*kids)(void*, const double*, double*, unsigned int*), double* out, unsigned int* m) {`,
        '    WXf wx;', ...gc.decls.map(l => '    ' + l),
        '    auto wchildren = [&](WHit wCin, wvec3 wq) -> WHit {',
        '        const double qm[3] = {wq.x * wunit, wq.y * wunit, wq.z * wunit};',
        '        double vals[4 * 64]; unsigned int ms[64];',
        '        const int n = kids(self, qm, vals, ms);',
        '        for (int j = 0; j < n; ++j)',
        '            wCin = w_fold(wCin, mk_wvec4(wx.M * wvec3(vals[4 * j], vals[4 * j + 1], vals[4 * j + 2]), vals[4 * j + 3] / wunit * wx.S), ms[j]);',
        '        return wCin;',
        '    };',
        '    wvec3 wp = wvec3(q[0], q[1], q[2]) / wunit;',
        '    WHit wC = w_emptyHit(0);',
        ...gc.lines.map(l => '    ' + l),
        '    *m = wC.m;',
        '    out[0] = wC.s.x; out[1] = wC.s.y; out[2] = wC.s.z; out[3] = wC.s.w * wunit;',
        '}'].join('\n')
}

// Is this text code (statements, not only comments)?
const isCode = src => typeof src === 'string' && wgsl.hasStatements(src)

module.exports = { objectFunction, groupCode, hostGroup, materialCode, materialHook, reliefFunction, moduleCode, skyParamNames, setShared, isCode, libraryNames }
