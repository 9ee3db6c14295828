'use strict'
// The scene document's WGSL code -- an object's, a group's or a material's statements (builder/codegen.js:286-306
// aliasPrelude, :609-648 the group template, :650 the material code) -- parsed and written out as C++ over the
// runtime of wgsl_rt.h, for hipRTC (FP32) and for the oracle's host compile (FP64).
//
// The language is WGSL's statement and expression grammar as scene code uses it: let / var / const with or without a
// type, assignments (= += -= *= /= %= &= |= ^= <<= >>=, ++, --), the phony `_ = e`, if / else, for, while, loop with
// break / continue, return, blocks; literals with their u / i / f suffixes, hexadecimal ints; vector, matrix and
// scalar constructors (vec3<f32>(..), vec3f(..), VEC3(..), mat3x3<f32>(..), f32(..), u32(..)), swizzles and
// component access, indexing, the WGSL operators at their WGSL precedence, bitcast<T>(e), and calls. Names are
// checked: a name the code does not declare must be one of the scope's (`scope`: parameters, the point, the Hit, the
// library's functions and constants), or the translation stops with the line and column.
//
// What C++ gets: a declared name is renamed u_<name> (no clash with C++ words or the runtime's own); a library name
// keeps its name, or takes the one `rename` gives it; `let x = e` is `const auto u_x = e`; `%` is wmod (WGSL's
// truncated remainder, float or integer); vector comparisons, select and the builtins are the runtime's overloads.
// WGSL's abstract numbers: a float literal is written WL(...) (an f32 on the device, a double in the oracle); an
// unsuffixed integer literal stays a C++ int, which promotes as WGSL's abstract int does in mixed arithmetic.

const TYPE_WORDS = new Set(['f32', 'u32', 'i32', 'bool', 'f16', 'vec2', 'vec3', 'vec4', 'mat2x2', 'mat3x3', 'mat4x4',
    'vec2f', 'vec3f', 'vec4f', 'vec2u', 'vec3u', 'vec4u', 'vec2i', 'vec3i', 'vec4i', 'mat2x2f', 'mat3x3f', 'mat4x4f',
    'VEC2', 'VEC3', 'VEC4', 'FLOAT', 'array'])
const KEYWORDS = new Set(['let', 'var', 'const', 'if', 'else', 'for', 'while', 'loop', 'break', 'continue', 'return',
    'true', 'false', 'fn', 'struct', 'switch', 'case', 'default', 'discard', 'continuing'])
// C++ words and the runtime's reserved names a WGSL name could collide with are never emitted bare: declared names
// are prefixed, so only these need no care
const OPS = ['<<=', '>>=', '&&', '||', '==', '!=', '<=', '>=', '<<', '>>', '+=', '-=', '*=', '/=', '%=', '&=', '|=', '^=',
    '++', '--', '->', '+', '-', '*', '/', '%', '&', '|', '^', '~', '!', '<', '>', '=', '(', ')', '{', '}', '[', ']', ',',
    ';', ':', '.', '@']

class WgslError extends Error {
    constructor(msg, tok) { super(tok ? `${msg} (line ${tok.line}, column ${tok.col})` : msg); this.tok = tok }
}

function tokenize(src) {
    const out = []
    let i = 0, line = 1, col = 1, depth = 0
    const s = String(src || '')
    const adv = n => { for (let k = 0; k < n; k++) { if (s[i] === '\n') { line++; col = 1 } else col++; i++ } }
    while (i < s.length) {
        const c = s[i], n = s[i + 1]
        if (depth === 0 && c === '/' && n === '/') { while (i < s.length && s[i] !== '\n') adv(1); continue }
        if (c === '/' && n === '*') { depth++; adv(2); continue }
        if (depth > 0) { if (c === '*' && n === '/') { depth--; adv(2) } else adv(1); continue }
        if (/\s/.test(c)) { adv(1); continue }
        const at = { line, col }
        if (/[A-Za-z_]/.test(c)) {
            let j = i; while (j < s.length && /[A-Za-z0-9_]/.test(s[j])) j++
            out.push({ t: 'id', v: s.slice(i, j), ...at }); adv(j - i); continue
        }
        if (/[0-9]/.test(c) || (c === '.' && /[0-9]/.test(n))) {
            let m = /^0[xX][0-9a-fA-F]+[iu]?/.exec(s.slice(i))
            if (m) { out.push({ t: 'num', v: m[0], ...at }); adv(m[0].length); continue }
            m = /^(?:[0-9]+\.[0-9]*|\.[0-9]+|[0-9]+)(?:[eE][+-]?[0-9]+)?[fhiu]?/.exec(s.slice(i))
            out.push({ t: 'num', v: m[0], ...at }); adv(m[0].length); continue
        }
        const op = OPS.find(o => s.startsWith(o, i))
        if (!op) throw new WgslError(`unexpected character '${c}'`, at)
        out.push({ t: 'op', v: op, ...at }); adv(op.length)
    }
    if (depth > 0) throw new WgslError('an unclosed block comment')
    out.push({ t: 'eof', v: '', line, col })
    return out
}

// ----------------------------------------------------------------------------------------------------------------
// The parser: statements and expressions into plain objects
function parse(src, module = false) {
    const T = tokenize(src)
    let i = 0
    const peek = (k = 0) => T[i + k]
    const next = () => T[i++]
    const is = (v, k = 0) => peek(k).v === v && peek(k).t !== 'num'
    const eat = v => { if (is(v)) { i++; return true } return false }
    const want = v => { if (!is(v)) throw new WgslError(`expected '${v}', found '${peek().v || 'the end'}'`, peek()); return next() }
    const ident = () => { const t = next(); if (t.t !== 'id' || KEYWORDS.has(t.v)) throw new WgslError(`expected a name, found '${t.v || 'the end'}'`, t); return t }

    function type() {
        const t = ident()
        if (!TYPE_WORDS.has(t.v)) return { name: t.v, tok: t }
        const ty = { name: t.v, tok: t, args: [] }
        if (is('<')) {
            next()
            while (!is('>')) {
                if (peek().t === 'num') ty.args.push({ num: next().v }); else ty.args.push(type())
                if (!eat(',')) break
            }
            // vec3<vec3<f32>> closes two lists with one token
            if (is('>>')) { T[i] = Object.assign({}, T[i], { v: '>' }); return ty }
            want('>')
        }
        return ty
    }
    function block() {
        want('{')
        const body = []
        while (!is('}')) { if (peek().t === 'eof') throw new WgslError("expected '}'", peek()); body.push(statement()) }
        want('}')
        return body
    }
    function decl() {
        const kw = next()
        const name = ident()
        let ty = null
        if (eat(':')) ty = type()
        let init = null
        if (eat('=')) init = expr()
        else if (kw.v !== 'var') throw new WgslError(`'${kw.v} ${name.v}' needs a value`, kw)
        if (!ty && !init) throw new WgslError(`'var ${name.v}' needs a type or a value`, kw)
        return { k: kw.v, name: name.v, tok: name, ty, init }
    }
    // a statement that can stand in a for's header: a declaration, an assignment, an increment or a call
    function simple() {
        if (is('let') || is('var') || is('const')) return decl()
        const at = peek()
        if (is('_') && is('=', 1)) { next(); next(); return { k: 'phony', e: expr(), tok: at } }
        const e = expr()
        if (is('++') || is('--')) return { k: 'incr', lhs: e, op: next().v, tok: at }
        const ops = ['=', '+=', '-=', '*=', '/=', '%=', '&=', '|=', '^=', '<<=', '>>=']
        if (peek().t === 'op' && ops.includes(peek().v)) { const op = next().v; return { k: 'assign', op, lhs: e, rhs: expr(), tok: at } }
        if (e.k !== 'call') throw new WgslError('a statement must assign, declare or call', at)
        return { k: 'expr', e, tok: at }
    }
    function statement() {
        const at = peek()
        if (is('{')) return { k: 'block', body: block(), tok: at }
        if (eat(';')) return { k: 'block', body: [], tok: at }
        if (is('if')) {
            next()
            const cond = expr()
            const then = block()
            let els = null
            if (eat('else')) els = is('if') ? [statement()] : block()
            return { k: 'if', cond, then, els, tok: at }
        }
        if (is('for')) {
            next(); want('(')
            const init = is(';') ? null : simple(); want(';')
            const cond = is(';') ? null : expr(); want(';')
            const update = is(')') ? null : simple(); want(')')
            return { k: 'for', init, cond, update, body: block(), tok: at }
        }
        if (is('while')) { next(); const cond = expr(); return { k: 'while', cond, body: block(), tok: at } }
        if (is('loop')) {
            next(); want('{')
            const body = []
            let continuing = null
            while (!is('}')) {
                if (is('continuing')) { next(); continuing = block(); continue }
                body.push(statement())
            }
            want('}')
            return { k: 'loop', body, continuing, tok: at }
        }
        if (is('break')) {
            next()
            if (eat('if')) { const cond = expr(); want(';'); return { k: 'breakif', cond, tok: at } }
            want(';'); return { k: 'break', tok: at }
        }
        if (is('continue')) { next(); want(';'); return { k: 'continue', tok: at } }
        if (is('return')) { next(); const e = is(';') ? null : expr(); want(';'); return { k: 'return', e, tok: at } }
        if (is('discard') || is('switch') || is('fn') || is('struct')) throw new WgslError(`'${peek().v}' is not supported in scene code`, at)
        const s = simple()
        want(';')
        return s
    }

    // precedence climbing over WGSL's binary operators (lowest first)
    const LEVELS = [['||'], ['&&'], ['|'], ['^'], ['&'], ['==', '!='], ['<', '>', '<=', '>='], ['<<', '>>'], ['+', '-'], ['*', '/', '%']]
    function expr(level = 0) {
        if (level === LEVELS.length) return unary()
        let a = expr(level + 1)
        while (peek().t === 'op' && LEVELS[level].includes(peek().v)) {
            const op = next()
            const b = expr(level + 1)
            a = { k: 'bin', op: op.v, a, b, tok: op }
        }
        return a
    }
    function unary() {
        const at = peek()
        if (peek().t === 'op' && ['-', '!', '~'].includes(peek().v)) { next(); return { k: 'un', op: at.v, e: unary(), tok: at } }
        if (peek().t === 'op' && ['&', '*'].includes(peek().v)) throw new WgslError('pointers are not supported in scene code', at)
        return postfix(primary())
    }
    function args() {
        want('(')
        const out = []
        while (!is(')')) { out.push(expr()); if (!eat(',')) break }
        want(')')
        return out
    }
    function primary() {
        const t = peek()
        if (t.t === 'num') { next(); return { k: 'num', v: t.v, tok: t } }
        if (eat('(')) { const e = expr(); want(')'); return { k: 'paren', e, tok: t } }
        if (t.t !== 'id') throw new WgslError(`unexpected '${t.v || 'end'}'`, t)
        if (t.v === 'true' || t.v === 'false') { next(); return { k: 'bool', v: t.v === 'true', tok: t } }
        if (t.v === 'bitcast') { next(); want('<'); const ty = type(); want('>'); return { k: 'bitcast', ty, args: args(), tok: t } }
        if (TYPE_WORDS.has(t.v)) { const ty = type(); return { k: 'ctor', ty, args: args(), tok: t } }
        if (KEYWORDS.has(t.v)) throw new WgslError(`unexpected '${t.v}'`, t)
        next()
        if (is('(')) return { k: 'call', fn: t.v, args: args(), tok: t }
        return { k: 'id', name: t.v, tok: t }
    }
    function postfix(e) {
        for (;;) {
            const t = peek()
            if (eat('.')) { const m = ident(); e = { k: 'member', e, name: m.v, tok: m }; continue }
            if (eat('[')) { const ix = expr(); want(']'); e = { k: 'index', e, i: ix, tok: t }; continue }
            return e
        }
    }

    // a module's function: fn name(a: T, ...) -> R { ... }
    function fnDecl() {
        const at = next()
        const name = ident()
        want('(')
        const params = []
        while (!is(')')) {
            const pn = ident(); want(':'); params.push({ name: pn.v, ty: type(), tok: pn })
            if (!eat(',')) break
        }
        want(')')
        const ret = eat('->') ? type() : null
        return { k: 'fn', name: name.v, params, ret, body: block(), tok: at }
    }
    const body = []
    while (peek().t !== 'eof') {
        if (module && is('fn')) body.push(fnDecl())
        else if (module && is('@')) throw new WgslError('attributes are not supported in scene code', peek())
        else body.push(statement())
    }
    return body
}

// ----------------------------------------------------------------------------------------------------------------
// C++ from the tree
const SWIZ = /^(?:[xyzw]{2,4}|[rgba]{2,4})$/
const ONE = { r: 'x', g: 'y', b: 'z', a: 'w' }
const SCALAR = { f32: 'WF', FLOAT: 'WF', f16: 'WF', u32: 'unsigned int', i32: 'int', bool: 'bool' }
function typeName(ty) {
    const n = ty.name
    if (SCALAR[n]) return SCALAR[n]
    const short = /^(vec[234])([fiu])$/.exec(n)
    if (short) return { f: 'w', i: 'wi', u: 'wu' }[short[2]] + short[1]
    if (/^VEC[234]$/.test(n)) return 'w' + n.toLowerCase()
    const vec = /^vec([234])$/.exec(n)
    if (vec) {
        const el = ty.args[0] ? ty.args[0].name : 'f32'
        const pre = { f32: 'w', f16: 'w', FLOAT: 'w', i32: 'wi', u32: 'wu', bool: 'wb' }[el]
        if (!pre) throw new WgslError(`a vector of '${el}' is not supported`, ty.tok)
        return `${pre}vec${vec[1]}`
    }
    const mat = /^mat([234])x([234])f?$/.exec(n)
    if (mat) {
        if (mat[1] !== mat[2]) throw new WgslError(`matrix '${n}' is not supported (square matrices only)`, ty.tok)
        return `wmat${mat[1]}`
    }
    if (n === 'array') {
        const el = typeName(ty.args[0]), len = ty.args[1] && ty.args[1].num
        if (!len) throw new WgslError('an array needs a constant length', ty.tok)
        return `warray<${el}, ${parseInt(len)}>`
    }
    return null
}

function literal(v, tok) {
    if (/^0[xX]/.test(v)) return v.endsWith('u') ? v : v.endsWith('i') ? `(int)${v.slice(0, -1)}` : `(int)${v}`
    const suf = v[v.length - 1]
    if (suf === 'u') return v
    if (suf === 'i') return v.slice(0, -1)
    if (suf === 'f' || suf === 'h') return `WL(${v.slice(0, -1)})`
    if (/[.eE]/.test(v)) return `WL(${v})`
    if (v.length > 10) throw new WgslError(`integer ${v} is out of range`, tok)
    return v
}

/**
 * translate(src, scope) -> C++ statements.
 * scope: { names: { wgslName: cppText } -- names the code may use undeclared (parameters, the point, the Hit, the
 *          library), hit: Set of names whose value is a Hit (their .r, .a are fields, not swizzles),
 *          hitFns: Set of functions returning a Hit, indent: '    ' }
 */
function translate(src, scope) {
    const all = parse(src, !!scope.module)
    // a module's functions and constants, in order, ahead of its statements; each callable from the code after it
    const fnDecls = all.filter(s => s.k === 'fn')
    const body = scope.module ? all.filter(s => s.k !== 'fn' && s.k !== 'const') : all
    const constDecls = scope.module ? all.filter(s => s.k === 'const') : []
    if (scope.module && body.length) throw new WgslError('a module holds functions and constants only', body[0].tok)
    const names = scope.names || {}
    const hitFns = scope.hitFns || new Set()
    const frames = [new Map()]
    for (const h of scope.hit || []) frames[0].set(h, { cpp: names[h] || h, hit: true, lib: true })
    for (const [n, v] of Object.entries(scope.vars || {})) frames[0].set(n, { cpp: v.cpp, hit: !!v.hit, lib: true })
    const look = n => { for (let k = frames.length - 1; k >= 0; k--) if (frames[k].has(n)) return frames[k].get(n); return null }
    const declare = (n, hit, tok) => {
        if (frames[frames.length - 1].has(n) && !frames[frames.length - 1].get(n).lib) throw new WgslError(`'${n}' is declared twice in one block`, tok)
        const d = { cpp: `u_${n}`, hit }
        frames[frames.length - 1].set(n, d)
        return d
    }
    const isHit = e => {
        if (e.k === 'id') { const d = look(e.name); return !!(d && d.hit) }
        if (e.k === 'call') return hitFns.has(e.fn)
        if (e.k === 'paren') return isHit(e.e)
        return false
    }
    const calls = new Set()
    function ex(e) {
        switch (e.k) {
        case 'num': return literal(e.v, e.tok)
        case 'bool': return e.v ? 'true' : 'false'
        case 'paren': return `(${ex(e.e)})`
        case 'id': {
            const d = look(e.name)
            if (d) return d.cpp
            if (Object.prototype.hasOwnProperty.call(names, e.name)) return names[e.name]
            throw new WgslError(`unknown name '${e.name}'`, e.tok)
        }
        case 'un': return `(${e.op}${ex(e.e)})`
        case 'bin': {
            if (e.op === '%') return `wmod(${ex(e.a)}, ${ex(e.b)})`
            return `(${ex(e.a)} ${e.op} ${ex(e.b)})`
        }
        case 'member': {
            const base = ex(e.e)
            if (isHit(e.e)) return `${base}.${e.name}`
            if (SWIZ.test(e.name)) return `wsw${e.name.length}(${base}, ${[...e.name].map(c => 'xyzw'.indexOf(ONE[c] || c)).join(', ')})`
            if (/^[xyzw]$/.test(e.name)) return `${base}.${e.name}`
            if (/^[rgba]$/.test(e.name)) return `${base}.${ONE[e.name]}`
            return `${base}.${e.name}`
        }
        case 'index': return `${ex(e.e)}[${ex(e.i)}]`
        case 'ctor': {
            const tn = typeName(e.ty)
            if (!tn) throw new WgslError(`unknown type '${e.ty.name}'`, e.tok)
            const a = e.args.map(ex)
            if (tn.startsWith('warray')) return `${tn}{{${a.join(', ')}}}`
            const fn = { WF: 'wf32', 'unsigned int': 'wu32', int: 'wi32', bool: 'wbool' }[tn] || `mk_${tn}`
            return `${fn}(${a.join(', ')})`
        }
        case 'bitcast': {
            const tn = typeName(e.ty)
            return `wbitcast_${tn.replace(/[^a-z0-9]/gi, '_')}(${e.args.map(ex).join(', ')})`
        }
        case 'call': {
            const d = look(e.fn)
            if (d) throw new WgslError(`'${e.fn}' is a value, not a function`, e.tok)
            if (!Object.prototype.hasOwnProperty.call(names, e.fn)) throw new WgslError(`unknown function '${e.fn}'`, e.tok)
            calls.add(e.fn)
            return `${names[e.fn]}(${e.args.map(ex).join(', ')})`
        }
        }
        throw new WgslError(`cannot translate '${e.k}'`, e.tok)
    }
    function lvalue(e) {
        if (e.k === 'id') {
            const d = look(e.name)
            if (d && d.konst) throw new WgslError(`'${e.name}' is a let or a const: it cannot be assigned`, e.tok)
            if (!d && !(scope.assignable || new Set()).has(e.name)) throw new WgslError(`'${e.name}' cannot be assigned`, e.tok)
            return ex(e)
        }
        if (e.k === 'member') {
            if (!isHit(e.e) && SWIZ.test(e.name)) throw new WgslError('WGSL cannot assign to a swizzle of several components', e.tok)
            return `${lvalue(e.e)}.${isHit(e.e) ? e.name : (ONE[e.name] || e.name)}`
        }
        if (e.k === 'index') return `${lvalue(e.e)}[${ex(e.i)}]`
        if (e.k === 'paren') return lvalue(e.e)
        throw new WgslError('this cannot be assigned', e.tok)
    }
    function simple(s) {
        switch (s.k) {
        case 'let': case 'const': case 'var': {
            const init = s.init ? ex(s.init) : null
            const hit = s.init ? isHit(s.init) : false
            const tn = s.ty ? typeName(s.ty) : null
            if (s.ty && !tn) throw new WgslError(`unknown type '${s.ty.name}'`, s.ty.tok)
            const d = declare(s.name, hit, s.tok)
            d.konst = s.k !== 'var'
            const ty = tn || (hit ? 'WHit' : 'auto')
            const q = s.k === 'var' ? '' : 'const '
            if (!init) return `${ty} ${d.cpp} = ${ty}()`
            return tn ? `${q}${ty} ${d.cpp} = wconv<${ty}>(${init})` : `${q}${ty} ${d.cpp} = ${init}`
        }
        case 'assign': {
            const l = lvalue(s.lhs), r = ex(s.rhs)
            if (s.op === '%=') return `${l} = wmod(${l}, ${r})`
            return `${l} ${s.op} ${r}`
        }
        case 'incr': return `${lvalue(s.lhs)}${s.op}`
        case 'phony': return `(void)(${ex(s.e)})`
        case 'expr': return ex(s.e)
        }
        throw new WgslError(`'${s.k}' cannot stand here`, s.tok)
    }
    function stmts(list, ind) {
        frames.push(new Map())
        const out = []
        for (const s of list) out.push(...stmt(s, ind))
        frames.pop()
        return out
    }
    function stmt(s, ind) {
        switch (s.k) {
        case 'block': return [`${ind}{`, ...stmts(s.body, ind + '    '), `${ind}}`]
        case 'if': {
            const out = [`${ind}if (${ex(s.cond)}) {`, ...stmts(s.then, ind + '    ')]
            if (s.els) out.push(`${ind}} else {`, ...stmts(s.els, ind + '    '))
            out.push(`${ind}}`)
            return out
        }
        case 'for': {
            frames.push(new Map())
            const init = s.init ? simple(s.init) : ''
            const cond = s.cond ? ex(s.cond) : ''
            const upd = s.update ? simple(s.update) : ''
            const out = [`${ind}for (${init}; ${cond}; ${upd}) {`, ...stmts(s.body, ind + '    '), `${ind}}`]
            frames.pop()
            return out
        }
        case 'while': return [`${ind}while (${ex(s.cond)}) {`, ...stmts(s.body, ind + '    '), `${ind}}`]
        case 'loop': {
            // loop { body; continuing { c } }: the continuing block runs after the body and after each continue
            if (!s.continuing) return [`${ind}for (;;) {`, ...stmts(s.body, ind + '    '), `${ind}}`]
            frames.push(new Map())
            const body = stmts(s.body, ind + '        ')
            const cont = stmts(s.continuing, ind + '    ')
            frames.pop()
            return [`${ind}for (bool wfirst_ = true;; wfirst_ = false) {`, `${ind}    if (!wfirst_) {`, ...cont.map(l => '    ' + l), `${ind}    }`,
                `${ind}    {`, ...body, `${ind}    }`, `${ind}}`]
        }
        case 'break': return [`${ind}break;`]
        case 'breakif': return [`${ind}if (${ex(s.cond)}) break;`]
        case 'continue': return [`${ind}continue;`]
        case 'return': {
            if (!scope.returns) throw new WgslError('scene code here does not return', s.tok)
            return [`${ind}${scope.returns(s.e ? ex(s.e) : null, s.tok)};`]
        }
        }
        return [`${ind}${simple(s)};`]
    }
    // the module's own names: a function u_<name>, a constant a function of none u_<name>_()
    const fns = []
    const own = {}
    for (const c of constDecls) {
        own[c.name] = `u_${c.name}_()`
        names[c.name] = own[c.name]
        const tn = c.ty ? typeName(c.ty) : null
        fns.push(`WG_FN ${tn || 'auto'} u_${c.name}_() { return ${tn ? `wconv<${tn}>(${ex(c.init)})` : ex(c.init)}; }`)
    }
    // every function declared ahead of the definitions, so each may call any other (WGSL's order is free)
    for (const f of fnDecls) {
        if (Object.prototype.hasOwnProperty.call(names, f.name) && !own[f.name]) throw new WgslError(`function '${f.name}' takes a name the library has`, f.tok)
        own[f.name] = `u_${f.name}`
        names[f.name] = own[f.name]
    }
    const protos = []
    for (const f of fnDecls) {
        const ret = f.ret ? typeName(f.ret) : 'void'
        if (!ret) throw new WgslError(`unknown type '${f.ret.name}'`, f.ret.tok)
        frames.push(new Map())
        const ps = f.params.map(pp => {
            const tn = typeName(pp.ty)
            if (!tn) throw new WgslError(`unknown type '${pp.ty.name}'`, pp.ty.tok)
            const d = declare(pp.name, false, pp.tok)
            d.konst = true
            return `${tn} ${d.cpp}`
        })
        const saved = scope.returns
        scope.returns = (e) => (e === null ? 'return' : `return ${ret === 'void' ? e : `wconv<${ret}>(${e})`}`)
        const inner = stmts(f.body, '    ')
        scope.returns = saved
        frames.pop()
        protos.push(`WG_FN ${ret} u_${f.name}(${ps.join(', ')});`)
        fns.push([`WG_FN ${ret} u_${f.name}(${ps.join(', ')}) {`, ...inner, '}'].join('\n'))
    }
    fns.splice(constDecls.length, 0, ...protos)
    const lines = stmts(body, scope.indent || '    ')
    return { lines, calls, fns, own }
}

// Does the text hold statements (not only comments and space)?
const hasStatements = src => { try { return parse(src).length > 0 } catch (e) { return true } }

module.exports = { tokenize, parse, translate, hasStatements, WgslError, typeName }
