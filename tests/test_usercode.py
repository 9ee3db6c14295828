"""The document's own WGSL code (objects' and groups'): the translation's semantics against WGSL's, and what it draws
against the native modifiers and closed forms, on the CPU (oracle/wgsl_host.h, FP64).

WGSL semantics pinned here by known answers written from the WGSL specification: swizzles, component-wise selects,
the truncated remainder `%` of floats and integers, integer division, round half to even, conversions clamped
toward zero, column-major matrices, loops with continue, break, continuing and `break if`, compound assignments,
abstract numbers beside f32, bit operations, any/all, mix/step/smoothstep. The scene code itself: the reference's
modifier blocks' behaviour (builder/modifiers.js) written as group code draws the same distances as the native
modifiers (pinned by their closed forms in tests/test_modifiers.py); a cone's code against its solid and brute-force
distances; sdMenger against the sponge's base-3 definition."""
import ctypes
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, 'oracle'))
import oracle  # noqa: E402
from shadytracer_amd import scenes  # noqa: E402

SNIPPETS = {
    'swizzle': ('let v = vec3f(1.0, 2.0, 3.0); R = vec4f(v.zyx, v.x + v.y);', [3, 2, 1, 3]),
    'rgba': ('let c = vec4<f32>(0.25, 0.5, 0.75, 1.0); R = vec4f(c.b, c.g, c.r, c.a);', [0.75, 0.5, 0.25, 1]),
    'select_vec': ('let a = vec2f(1.0, -2.0); let b = select(vec2f(0.0), vec2f(5.0), a > vec2f(0.0)); R = vec4f(b, select(1.0, 2.0, false), select(1.0, 2.0, true));',
                   [5, 0, 1, 2]),
    'remainder': ('R.x = -7.5 % 2.0; R.y = f32(-7 % 3); R.z = f32(7 / 2); R.w = 7.5 % -2.0;', [-1.5, -1, 3, 1.5]),
    'round_even': ('R = vec4f(round(2.5), round(3.5), round(-2.5), round(0.49999997));', [2, 4, -2, 0]),
    'conversions': ('R = vec4f(f32(u32(-3.7)), f32(u32(3.7)), f32(i32(-3.7)), f32(u32(true)));', [0, 3, -3, 1]),
    'matrix': ('let m = mat3x3<f32>(1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0, 9.0); let a = m * vec3f(1.0, 0.0, 0.0); '
               'let b = vec3f(1.0, 0.0, 0.0) * m; R = vec4f(a.z, b.y, m[1].y, determinant(m));', [3, 4, 5, 0]),
    'matrix_product': ('let a = mat3x3<f32>(0.0, 1.0, 0.0, -1.0, 0.0, 0.0, 0.0, 0.0, 1.0); let v = (a * a) * vec3f(1.0, 2.0, 3.0); '
                       'R = vec4f(v, transpose(a)[0].y);', [-1, -2, 3, -1]),
    'loops': ('var s = 0u; for (var i = 0u; i < 10u; i++) { if (i == 3u) { continue; } if (i == 7u) { break; } s += i; } '
              'var k = 1; while (k < 100) { k = k * 3; } '
              'var n = 0; loop { if (n >= 5) { break; } continuing { n += 1; } } '
              'var m = 0; loop { continuing { m += 2; break if m > 7; } } '
              'R = vec4f(f32(s), f32(k), f32(n), f32(m));', [18, 243, 5, 8]),
    'compound': ('var a = vec3f(1.0); a *= 2.0; a += vec3f(1.0, 2.0, 3.0); a.y -= 1.0; var b = 10; b %= 4; b <<= 2; '
                 'R = vec4f(a, f32(b));', [3, 3, 5, 8]),
    'abstract': ('R = vec4f(max(2, 3.5), clamp(5, 0.0, 1.0), f32(1 / 2), 1.0 / 2);', [3.5, 1, 0, 0.5]),
    'bits': ('let h = (0xF0u >> 4u) & 0x7u; let s = 1u << 5u; R = vec4f(f32(h), f32(s), f32(~0u >> 28u), f32(6u ^ 3u));', [7, 32, 15, 5]),
    'any_all': ('let v = vec3f(1.0, 2.0, 3.0); R = vec4f(select(0.0, 1.0, any(v > vec3f(2.5))), select(0.0, 1.0, all(v > vec3f(2.5))), '
                'select(0.0, 1.0, all(v >= vec3f(1.0))), 0.0);', [1, 0, 1, 0]),
    'mixes': ('R = vec4f(smoothstep(0.0, 1.0, 0.5), mix(vec3f(0.0), vec3f(2.0), 0.25).y, step(1.0, 0.5), smoothstep(0.0, 2.0, 3.0));',
              [0.5, 0.5, 0, 1]),
    'geometry': ('let n = normalize(vec3f(3.0, 0.0, 4.0)); let c = cross(vec3f(1.0, 0.0, 0.0), vec3f(0.0, 1.0, 0.0)); '
                 'R = vec4f(n.x, c.z, length(vec2f(3.0, 4.0)), distance(vec3f(1.0), vec3f(1.0, 1.0, 3.0)));', [0.6, 1, 5, 2]),
    'shadow_scopes': ('let a = 1.0; var r = 0.0; { let a = 2.0; r = a; } R = vec4f(a, r, fract(-0.25), trunc(-1.5));', [1, 2, 0.75, -1]),
    'hit_fold': ('var H = emptyHit(0); H = fold(H, vec4f(0.0, 0.0, 1.0, 2.0), 3u); H = fold(H, vec4f(1.0, 0.0, 0.0, -1.0), 4u); '
                 'H = fold(H, vec4f(0.0, 1.0, 0.0, -3.0), 5u); R = vec4f(H.s.w, f32(H.m), f32(H.a), H.insideW);', [-1, 4, 4, -1]),
    'warp_scale': ('XF_S = 2.0; var H = emptyHit(0); H = add_new(H, vec4f(1.0, 0.0, 0.0, 3.0), 1u); warp(0.5); '
                   'H = add_new(H, vec4f(1.0, 0.0, 0.0, 2.5), 2u); R = vec4f(H.s.w, H.s.x, f32(H.m), XF_S);', [2.5, 0, 2, 1]),
}
REFUSED = {
    'unknown_name': ('R.x = nothing;', "unknown name 'nothing'"),
    'unknown_fn': ('R.x = sweep(1.0);', "unknown function 'sweep'"),
    'let_assign': ('let a = 1.0; a = 2.0;', "'a' is a let or a const"),
    'swizzle_assign': ('R.xy = vec2f(1.0);', 'swizzle of several'),
    'fn_def': ('fn f() -> f32 { return 1.0; }', "'fn' is not supported"),
    'bad_token': ('R.x = 1.0 $ 2.0;', "unexpected character '$'"),
    'return': ('return;', 'does not return'),
    'twice': ('let a = 1.0; let a = 2.0;', 'declared twice'),
}


def _translate(snips):
    r = subprocess.run(['node', os.path.join(ROOT, 'tests', 'js', 'wgsl_snippets.js')], input=json.dumps(snips),
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, check=True)
    return json.loads(r.stdout)


@pytest.fixture(scope='module')
def snippet_module(tmp_path_factory):
    out = _translate([{'name': k, 'code': v[0]} for k, v in SNIPPETS.items()])
    bad = [o for o in out if 'error' in o]
    assert not bad, bad
    d = tmp_path_factory.mktemp('wgsl')
    src = d / 'snips.cpp'
    src.write_text('#include "wgsl_host.h"\n' + '\n'.join(o['cpp'] for o in out))
    so = d / 'snips.so'
    r = subprocess.run(['g++', '-O1', '-std=c++14', '-shared', '-fPIC', '-I', os.path.join(ROOT, 'oracle'),
                        '-I', os.path.join(ROOT, 'shadytracer_amd', 'csrc', 'device'), '-o', str(so), str(src)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-3000:]
    return ctypes.CDLL(str(so))


@pytest.mark.parametrize('name', sorted(SNIPPETS))
def test_translated_code_follows_wgsl(snippet_module, name):
    out = np.zeros(4)
    getattr(snippet_module, 'snip_' + name)(out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert np.allclose(out, SNIPPETS[name][1], rtol=0, atol=1e-12), (name, out)


def test_code_the_translation_does_not_take_is_refused_with_its_place():
    out = {o['name']: o for o in _translate([{'name': k, 'code': v[0]} for k, v in REFUSED.items()])}
    for k, (code, why) in REFUSED.items():
        assert 'error' in out[k], k
        assert why in out[k]['error'] and 'line 1, column' in out[k]['error'], (k, out[k]['error'])


@pytest.fixture(scope='module')
def drawn():
    return scenes.emit('usercode'), scenes.emit('usercode', native=True)


def test_the_modifiers_written_as_group_code_draw_as_the_native_ones(drawn):
    user, native = drawn
    assert len(user['scene']['user']) == 10 and not native['scene']['user'][2:]
    rng = np.random.default_rng(3)
    pts = rng.uniform([-0.2, -0.12, -0.01], [0.2, 0.16, 0.08], (40000, 3))
    # the gear (copies that overlap) is folded copy by copy by group code, as the reference's children() does, and
    # unioned by the native modifier: inside two teeth the two differ, so its neighbourhood is left out
    far = np.hypot(pts[:, 0] - 0.135, pts[:, 1] + 0.05) > 0.045
    pts = pts[far]
    du, nu, _ = oracle.scene_at(user['scene'], pts)
    dn, nn, _ = oracle.scene_at(native['scene'], pts)
    # the parameters are FP32 in both, the native ones in metres and the code's in millimetres: equal to their rounding
    # (a few parts in 1e8 of a distance of centimetres)
    assert np.abs(du - dn).max() < 2e-8
    # normals: where the native modifier's normal is analytic (round, shell, repeat, the copies), the same
    has = np.linalg.norm(nn, axis=1) > 0.5
    assert has.sum() > 1000
    assert np.abs(nu[has] - nn[has]).max() < 1e-5


def test_a_cone_of_its_own_code_is_the_cone(drawn):
    user, _ = drawn
    # the cone: base radius 10 mm at z = 0, top 4 mm at 30 mm (its node at z = 15 mm, height 30)
    rng = np.random.default_rng(5)
    pts = rng.uniform([-0.02, -0.02, -0.005], [0.02, 0.02, 0.035], (20000, 3))
    d, n, prim = oracle.scene_at(user['scene'], pts)
    rho, z = np.hypot(pts[:, 0], pts[:, 1]), pts[:, 2]
    inside = (z >= 0) & (z <= 0.03) & (rho <= 0.010 - 0.006 * z / 0.03)
    # the floor (z < 0) and nothing else nearby: where the cone is the nearest surface, its sign is the solid's
    near = (z > 0.0005) & (np.abs(d) < z - 0.0002)
    assert near.sum() > 3000
    assert np.array_equal(d[near] < 0, inside[near])
    # a bound, and exact where the side is nearest: brute-force distance to the cone's surface outside it
    t = np.linspace(0, 1, 400)
    prof = np.c_[0.010 - 0.006 * t, 0.03 * t]
    caps = [np.c_[np.linspace(0, 0.010, 200), np.zeros(200)], np.c_[np.linspace(0, 0.004, 80), np.full(80, 0.03)]]
    curve = np.vstack([prof] + caps)
    out = near & ~inside
    q = np.c_[rho[out], z[out]]
    exact = np.sqrt(((q[:, None, :] - curve[None, :, :]) ** 2).sum(-1)).min(1)
    assert np.all(d[out] <= exact + 2e-5)
    side = out & (z > 0.002) & (z < 0.028)
    # the side's segment, (0.010, 0) to (0.004, 0.030) in (rho, z): the distance to it in closed form
    a, b = np.array([0.010, 0.0]), np.array([0.004, 0.030])
    qs = np.c_[rho[side], z[side]]
    tt = np.clip(((qs - a) @ (b - a)) / ((b - a) @ (b - a)), 0, 1)
    ex = np.linalg.norm(qs - (a + tt[:, None] * (b - a)), axis=1)
    close = ex < 0.002
    assert close.sum() > 100 and np.abs(d[side][close] - ex[close]).max() < 1e-9


def _menger_inside(u, levels):
    """Membership of the sponge of side 2 about the origin by its definition: u = (p + 1) / 2 in the unit cube,
    removed where at some level two of its base-3 digits are 1."""
    if np.any((u < 0) | (u > 1)):
        return False
    for k in range(1, levels + 1):
        dig = np.floor(u * 3 ** k) % 3
        if (dig == 1).sum() >= 2:
            return False
    return True


def test_menger_sponge_code_carves_the_sponge(drawn):
    user, _ = drawn
    sc = user['scene']
    lay = sc['layout']['sponge']['xf']
    U = np.asarray(sc['params'], np.float64)
    t, R, s = U[lay:lay + 3], U[lay + 3:lay + 12].reshape(3, 3), U[lay + 12]
    rng = np.random.default_rng(7)
    loc = rng.uniform(-1.1, 1.1, (6000, 2))
    # points on the sponge's mid-height plane and off it, in its own frame (side 2 = 30 mm), to the world
    lz = rng.uniform(-1.1, 1.1, len(loc))
    L = np.c_[loc, lz] * 0.015
    W = t + (L @ R) * s
    d, _, prim = oracle.scene_at(sc, W)
    mine = [o['prim'] for o in sc['program'] if o['op'] == 'user' and sc['user'][o['index']]['node'] == 'sponge']
    # where the sponge is the surface the fold keeps (not the floor it stands on)
    keep = (np.abs(d) > 2e-6) & (prim == mine[0])
    want = np.array([_menger_inside((p + 1) / 2, 3) for p in L[keep] / 0.015])
    assert keep.sum() > 5000
    assert np.array_equal(d[keep] < 0, want)


def test_reference_scenes_with_code_translate(tmp_path):
    ref = '/root/reference/scenes'
    if not os.path.isdir(ref):
        pytest.skip('the reference scenes are not here')
    # the Menger sponge's object code (its material's code left out and listed)
    made = scenes.emit_doc(os.path.join(ref, 'infinite-detail.json'), width=32, height=18, approximate=True)
    assert [u['kind'] for u in made['scene']['user']] == ['object']
    assert all('material' in a for a in made['scene']['approximations'])
    d, _, _ = oracle.scene_at(made['scene'], np.zeros((1, 3)) + [0.0, 0.0, 1.0])
    assert np.isfinite(d).all()
    # the imported cone and rounded rectangles (g2Extrude) of imported-parts, its material code left out
    made = scenes.emit_doc(os.path.join(ref, 'imported-parts.json'), width=32, height=18, approximate=True)
    assert len(made['scene']['user']) >= 4
    d, _, _ = oracle.scene_at(made['scene'], np.random.default_rng(1).uniform(-0.1, 0.1, (500, 3)))
    assert np.isfinite(d).all()


# ---- the noise (core/noise.js) in the oracle's restatement, against definitions computed here from PCG4D ----------
def _pcg(v):
    out = (ctypes.c_uint32 * 4)()
    oracle.lib().or_pcg4d((ctypes.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in v]), out)
    return [out[i] for i in range(4)]


def _unit(cell, seed):
    return np.array([(x >> 8) / 16777216.0 for x in _pcg([int(c) for c in cell] + [seed])])


def _worley_brute(p, seed, reach=2, jitter=0.46):
    i = np.floor(p)
    best = []
    for dz in range(-reach, reach + 1):
        for dy in range(-reach, reach + 1):
            for dx in range(-reach, reach + 1):
                c = i + [dx, dy, dz]
                u = _unit(c, seed)
                q = c + 0.5 + (u[:3] - 0.5) * jitter
                best.append((np.linalg.norm(p - q), u[3]))
    best.sort(key=lambda t: t[0])
    return best[0][0], best[1][0], best[0][1], best[1][1]


@pytest.fixture(scope='module')
def noise_module(tmp_path_factory):
    rng = np.random.default_rng(11)
    snips, want = {}, {}
    # lattice points: value noise is the corner's value 2u - 1 with no slope; gradient noise is 0 with slope 1.6 g
    for k, c in enumerate([(0, 0, 0), (3, -2, 5), (-7, 11, -1)]):
        u = _unit(c, 9)
        g = u[:3] * 2 - 1
        g = g / max(np.linalg.norm(g), 1.0)
        snips['vlat%d' % k] = 'R = vnoise3(vec3f(%d.0, %d.0, %d.0), 9u);' % c
        want['vlat%d' % k] = [0, 0, 0, 2 * u[0] - 1]
        snips['glat%d' % k] = 'R = gnoise3(vec3f(%d.0, %d.0, %d.0), 9u);' % c
        want['glat%d' % k] = list(1.6 * g) + [0]
    # analytic gradients against central differences, at points off the lattice
    for k, p in enumerate(rng.uniform(-5, 5, (3, 3))):
        for fn, call in (('g', 'gnoise3(%s, 4u)'), ('v', 'vnoise3(%s, 4u)'), ('f', 'fbm3(%s, 4u, 3u, 2.0, 0.5)')):
            P = 'vec3f(%r, %r, %r)' % tuple(float(x) for x in p)
            code = ('let h = 1e-6; let a = %s; ' % (call % P) +
                    ' '.join('let d%d = (%s.w - %s.w) / (2.0 * h);' % (j, call % ('(%s + h * vec3f(%s))' % (P, e)),
                                                                      call % ('(%s - h * vec3f(%s))' % (P, e)))
                             for j, e in enumerate(('1.0, 0.0, 0.0', '0.0, 1.0, 0.0', '0.0, 0.0, 1.0'))) +
                    ' R = vec4f(d0 - a.x, d1 - a.y, d2 - a.z, 0.0);')
            snips['grad_%s%d' % (fn, k)] = code
            want['grad_%s%d' % (fn, k)] = None
    # Worley: the 27-cell search against a brute-force search of 125 cells (the 0.46 jitter keeps the nearest within 27)
    for k, p in enumerate(rng.uniform(-20, 20, (6, 3))):
        snips['worley%d' % k] = 'R = worley3(vec3f(%r, %r, %r), 21u);' % tuple(float(x) for x in p)
        want['worley%d' % k] = list(_worley_brute(p, 21))
    # crossed Worley: the nearest two of both point sets, brute force over each set's lattice
    A = np.cbrt(1 + 1 / 1.37 ** 3)
    B = A * 1.37
    T = np.array([[0.36, 0.48, -0.8], [-0.8, 0.6, 0.0], [0.48, 0.64, 0.6]]).T   # columns as given: T @ p turns p
    for k, p in enumerate(rng.uniform(-10, 10, (4, 3))):
        a = _worley_brute(p / A, 5)
        b = _worley_brute(T @ p / B, 5 + 1013)
        pts = sorted([(a[0] * A, a[2]), (a[1] * A, a[3]), (b[0] * B, b[2]), (b[1] * B, b[3])], key=lambda t: t[0])
        snips['worleyx%d' % k] = 'R = worley3x(vec3f(%r, %r, %r), 5u);' % tuple(float(x) for x in p)
        want['worleyx%d' % k] = [pts[0][0], pts[1][0], pts[0][1], pts[1][1]]
    snips['fbmlip'] = 'R = vec4f(nzFbmLip(4u, 2.0, 0.5, 4.3), nzFbmLip(3u, 2.0, 0.75, 1.0), 0.0, 0.0);'
    want['fbmlip'] = [4.3 * 4, 1 + 1.5 + 2.25, 0, 0]
    out = _translate([{'name': k, 'code': v} for k, v in snips.items()])
    assert not [o for o in out if 'error' in o], [o for o in out if 'error' in o]
    d = tmp_path_factory.mktemp('noise')
    (d / 'n.cpp').write_text('#include "wgsl_host.h"\n' + '\n'.join(o['cpp'] for o in out))
    r = subprocess.run(['g++', '-O1', '-std=c++14', '-shared', '-fPIC', '-I', os.path.join(ROOT, 'oracle'),
                        '-I', os.path.join(ROOT, 'shadytracer_amd', 'csrc', 'device'), '-o', str(d / 'n.so'), str(d / 'n.cpp')],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-3000:]
    return ctypes.CDLL(str(d / 'n.so')), want


def test_noise_follows_its_definitions(noise_module):
    M, want = noise_module
    for name, w in want.items():
        out = np.zeros(4)
        getattr(M, 'snip_' + name)(out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if w is None:
            # central differences of step 1e-6 in FP64: the analytic gradient to 1e-6
            assert np.abs(out).max() < 1e-6, (name, out)
        else:
            assert np.allclose(out, w, rtol=0, atol=1e-12), (name, out, w)


def test_material_code_runs_on_the_hit(drawn):
    # the floor's code: a checker of 20 mm squares in the world (the second colour where cos(pi x/20) cos(pi y/20) >= 0
    # at z = 0), called as the oracle calls it on a hit; the paint's grain reads the object's frame
    from shadytracer_amd import pt
    user, _ = drawn
    sc = user['scene']
    hook = oracle.user_module(sc["user"], sc["materialCode"], sc["reliefs"])[3]
    assert hook is not None
    fn = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_float), ctypes.c_uint, ctypes.POINTER(ctypes.c_double),
                          ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.POINTER(pt.PtMaterial))(hook.value)
    U = np.ascontiguousarray(np.asarray(sc['params'], np.float32))
    Up = U.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    floor = [i for i, m in enumerate(sc['materials']) if m.get('id') == 'floor'][0]

    def run(mat, p, prim=-1):
        m = pt.PtMaterial()
        src = sc['materials'][mat]
        for j in range(3):
            m.base[j] = src['base'][j]
        m.roughness, m.metallic, m.ior, m.specular = src['roughness'], src['metallic'], src['ior'], src['specular']
        P = np.asarray(p, np.float64)
        n = np.array([0.0, 0.0, 1.0])
        fn(Up, mat, P.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), prim,
           ctypes.byref(m))
        return np.array(m.base[:]), m.roughness
    rng = np.random.default_rng(2)
    for x, y in rng.uniform(-0.1, 0.1, (200, 2)):
        c, _ = run(floor, [x, y, 0.0])
        dark = np.cos(np.pi * x / 0.02) * np.cos(np.pi * y / 0.02) >= 0
        assert np.allclose(c, 0.15 if dark else 0.5, atol=1e-6), (x, y, c)
    # the paint on the cone (its own frame at 15 mm up): the same point of the object reads the same wherever the
    # object stands, and its roughness only grows
    paint = [i for i, m in enumerate(sc['materials']) if m.get('id') == 'paint'][0]
    prim = [o['prim'] for o in sc['program'] if o['op'] == 'user' and sc['user'][o['index']]['node'] == 'cone'][0]
    c1, r1 = run(paint, [0.003, 0.002, 0.010], prim)
    c2, r2 = run(paint, [0.003, 0.002, 0.010], prim)
    assert np.array_equal(c1, c2) and r1 >= sc['materials'][paint]['roughness']


def test_a_relief_moves_the_surface_by_its_height(drawn):
    # the ripple tile: a box whose top face is 6 mm up, under h = A sin(k qx) cos(k qy) (A 0.5 mm, k = 2 pi / 8 mm,
    # q the point in the tile's frame in mm) with slope bound A k. Within two amplitudes of the face the distance is
    # (d - h) / (1 + A k), beyond them d less one amplitude towards zero (builder/codegen.js:1735-1760)
    user, _ = drawn
    sc = user['scene']
    rng = np.random.default_rng(9)
    n = 4000
    x = rng.uniform(-0.075, -0.045, n)
    y = rng.uniform(-0.140, -0.110, n)
    z = rng.uniform(0.0045, 0.0075, n)
    d, nrm, _ = oracle.scene_at(sc, np.c_[x, y, z])
    A, k = 0.5, 2 * np.pi / 8
    qx, qy = (x + 0.060) * 1000, (y + 0.125) * 1000
    h = A * np.sin(k * qx) * np.cos(k * qy)
    d0 = z - 0.006
    band = np.abs(d0) <= 2 * A * 1e-3
    want = np.where(band, (d0 - 1e-3 * h) / (1 + A * k), d0 - np.sign(d0) * A * 1e-3)
    assert band.sum() > 1000 and (~band).sum() > 500
    assert np.abs(d - want).max() < 1e-9
    # under the surface where the ripple is above the point, and so on: the sign is the relief's
    assert np.array_equal(d[band] < 0, (d0 - 1e-3 * h)[band] < 0)
    # no analytic normal: the renderer takes it by differences
    assert np.abs(nrm).max() == 0


# ---- the document's functions block and its own sky (builder/sky.js: background(dir)) ------------------------------
def test_module_functions_and_constants_translate_in_any_order(tmp_path):
    import subprocess as sp
    code = ('const TWO: f32 = 2.0;\nfn twice(x: f32) -> f32 { return half(x) * TWO * TWO; }\n'
            'fn half(x: f32) -> f32 { return x / TWO; }\nfn v3of(x: f32) -> vec3f { return vec3f(x, twice(x), 0.0); }')
    js = ("const u = require(%r); const m = u.moduleCode(%s); process.stdout.write(JSON.stringify(m))"
          % (os.path.join(ROOT, 'shadytracer_amd', 'js', 'usercode.js'), json.dumps(code)))
    m = json.loads(sp.run(['node', '-e', js], stdout=sp.PIPE, check=True, text=True).stdout)
    assert m['own'] == {'TWO': 'u_TWO_()', 'twice': 'u_twice', 'half': 'u_half', 'v3of': 'u_v3of'}
    src = tmp_path / 'm.cpp'
    src.write_text('#include "wgsl_host.h"\n' + m['text'] + '\nextern "C" double go() { return u_v3of(3.0).y + u_TWO_(); }\n')
    r = sp.run(['g++', '-O1', '-std=c++14', '-shared', '-fPIC', '-I', os.path.join(ROOT, 'oracle'), '-I',
                os.path.join(ROOT, 'shadytracer_amd', 'csrc', 'device'), '-o', str(tmp_path / 'm.so'), str(src)],
               stdout=sp.PIPE, stderr=sp.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    M = ctypes.CDLL(str(tmp_path / 'm.so'))
    M.go.restype = ctypes.c_double
    assert M.go() == 3.0 * 2.0 + 2.0
    # a module holds functions and constants only; a function cannot take a library function's name
    for bad, why in (('let x = 1.0;', 'functions and constants only'), ('fn length(x: f32) -> f32 { return x; }', 'takes a name the library has')):
        js = "const u = require(%r); try { u.moduleCode(%s) } catch (e) { process.stdout.write(e.message) }" % (
            os.path.join(ROOT, 'shadytracer_amd', 'js', 'usercode.js'), json.dumps(bad))
        assert why in sp.run(['node', '-e', js], stdout=sp.PIPE, check=True, text=True).stdout


def test_the_scenes_own_sky_is_its_code():
    from shadytracer_amd import pt
    coded, native = scenes.emit('sky-code'), scenes.emit('sky-code', native=True)
    assert coded['scene']['env']['type'] == 5 and native['scene']['env']['type'] == 3
    # the code's background(dir) is the studio's L0 (1 + top d.z) in every direction (the oracle's FP64 compile)
    rng = np.random.default_rng(4)
    d = rng.normal(size=(500, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    uc, un = pt.uniforms_for(coded), pt.uniforms_for(native)
    Lc = oracle.env(uc, d, scene=coded['scene'])
    Ln = oracle.env(un, d)
    assert np.abs(Lc - Ln).max() <= 1e-4 * Ln.max()
    # ...and the pictures agree: the ball of the functions block's code and the floor under the two skies
    ac, mc, _ = oracle.render(coded['scene'], uc, 48)
    an, mn, _ = oracle.render(native['scene'], un, 48, first=5000)
    Lc, Ln = pt.luminance(ac), pt.luminance(an)
    s = np.sqrt(pt.standard_error(mc) ** 2 + pt.standard_error(mn) ** 2)
    ok = s > 1e-9
    z = (Lc - Ln)[ok] / s[ok]
    assert 0.75 < z.std() < 1.3 and np.mean(np.abs(z) < 3) > 0.97, z.std()
