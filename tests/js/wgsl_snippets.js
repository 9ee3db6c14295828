'use strict'
// WGSL snippets in (JSON on stdin: [{name, code}]), C++ out: one `extern "C" void snip_<name>(double* out)` each, the
// snippet run with R (a vec4, zero at the start) in scope and R's components written to out; or the translator's
// refusal for a snippet it does not take. For tests/test_usercode.py, which compiles them against oracle/wgsl_host.h.
const wgsl = require('../../shadytracer_amd/js/wgsl')
const usercode = require('../../shadytracer_amd/js/usercode')
const snippets = JSON.parse(require('fs').readFileSync(0, 'utf8'))
const out = []
for (const s of snippets) {
    try {
        const t = wgsl.translate(s.code, { names: usercode.libraryNames(), hitFns: new Set(['emptyHit', 'fold', 'add_new']), indent: '    ',
            vars: { R: { cpp: 'wR' }, p: { cpp: 'wp' }, C: { cpp: 'wC', hit: true } }, assignable: new Set(['XF_S', 'XF_M']) })
        out.push({ name: s.name, cpp: [`extern "C" void snip_${s.name}(double* out) {`, '    WXf wx;', '    wvec4 wR;', '    wvec3 wp;',
            '    WHit wC = w_emptyHit(0);', '    (void)wp; (void)wC;', ...t.lines,
            '    out[0] = wR.x; out[1] = wR.y; out[2] = wR.z; out[3] = wR.w;', '}'].join('\n') })
    } catch (e) {
        out.push({ name: s.name, error: e.message })
    }
}
process.stdout.write(JSON.stringify(out))
