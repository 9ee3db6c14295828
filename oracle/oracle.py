"""ctypes binding of the CPU oracle (oracle.c) -- test infrastructure only: tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg use it as the checker. The product never imports it.

The scene arrives as the emitter's output (shadytracer_amd/js/hipgen.js): the postfix program, the parameter
block, materials and light faces; the uniforms as the engine's PtUniforms (shadytracer_amd/pt.py)."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, '_build', 'liboracle.so')

SHAPES = ['plane', 'sphere', 'box', 'cylinder', 'capsule', 'torus', 'rectangle', 'disc', 'lens', 'prism', 'reflector', 'aspheric lens',
          'revolution', 'fresnel lens', 'extrusion', 'cone', 'round cone', 'octahedron', 'capped torus', 'link', 'tube', 'dome', 'pyramid']
JOINS = ['union', 'subtract', 'intersect']
OPS = {'prim': 0, 'join': 1, 'fold': 2, 'node': 3, 'mod': 4, 'modend': 5, 'user': 6, 'ugroup': 7, 'ufold': 8, 'ugroupend': 9, 'relief': 10}
MODS = ['twist', 'bend', 'repeat', 'mirror', 'circular', 'round', 'shell']

_lib = None


def build():
    subprocess.run(['make', '-s', '-C', HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P = ctypes.POINTER
        c_i, c_u, c_d, c_f, c_v = ctypes.c_int, ctypes.c_uint32, ctypes.c_double, ctypes.c_float, ctypes.c_void_p
        L.or_render.argtypes = [P(c_i), c_i, P(c_f), c_v, c_i, c_v, c_i, c_v, c_u, c_u, P(c_d), P(c_d), P(ctypes.c_longlong), c_i]
        L.or_pick.argtypes = [P(c_i), c_i, P(c_f), c_v, c_u, c_u, c_u, P(c_d), P(c_u)]
        L.or_scene_at.argtypes = [P(c_i), c_i, P(c_f), P(c_d), c_i, P(c_d), P(c_d), P(c_i)]
        L.or_bsdf.argtypes = [c_v, P(c_d), P(c_d), c_d, P(c_d), P(c_d), P(c_d), P(c_d), P(c_d), P(c_d), P(c_i), ctypes.c_int]
        L.or_fresnel.restype = c_d
        L.or_fresnel.argtypes = [c_d, c_d]
        L.or_cauchy.restype = c_d
        L.or_cauchy.argtypes = [c_d, c_d, c_d]
        L.or_spectral_weight.argtypes = [c_d, P(c_d)]
        L.or_mf_tables.argtypes = [P(c_d)]
        L.or_haze_tr.restype = c_d
        L.or_haze_tr.argtypes = [c_v, P(c_d), P(c_d), c_d]
        L.or_haze_draw.restype = c_d
        L.or_haze_draw.argtypes = [c_v, P(c_d), P(c_d), c_d]
        L.or_hg.restype = c_d
        L.or_hg.argtypes = [c_d, c_d]
        L.or_hg_sample.argtypes = [c_d, P(c_d), c_d, c_d, P(c_d)]
        L.or_pcg4d.argtypes = [P(c_u), P(c_u)]
        L.or_tiles.argtypes = [c_u, c_u, c_u, P(c_u), P(c_u)]
        L.or_sppm_radius.restype = c_d
        L.or_sppm_radius.argtypes = [c_d, c_d, c_u]
        L.or_cct_duv.argtypes = [c_d, c_d, P(c_d), P(c_d)]
        L.or_fresnel_unpolarised.restype = c_d
        L.or_fresnel_unpolarised.argtypes = [c_d, c_d, P(c_d)]
        L.or_mf_load.argtypes = [P(c_d)]
        L.or_set_user.argtypes = [P(c_v), c_i]
        L.or_set_material.argtypes = [c_v]
        L.or_set_reliefs.argtypes = [P(c_v), c_i]
        L.or_set_sky.argtypes = [c_v]
        L.or_mf_words.restype = c_i
        _lib = L
        _load_tables(L)
    return _lib


# The GGX albedo tables take the oracle seconds to integrate (an adaptive quadrature per entry, in FP64): computed once
# per build and kept beside the library, keyed by the library's own bytes, so a rebuilt oracle integrates them again.
TABLES = os.path.join(HERE, '_build', 'mf_tables.f64')


def _library_key():
    import hashlib
    return hashlib.sha1(open(LIB, 'rb').read()).digest()[:16]


def _load_tables(L):
    n = L.or_mf_words()
    try:
        raw = open(TABLES, 'rb').read()
        if len(raw) == 16 + 8 * n and raw[:16] == _library_key():
            t = np.frombuffer(raw[16:], np.float64).copy()
            L.or_mf_load(_dp(t))
            return
    except OSError:
        pass
    t = np.zeros(n, np.float64)
    L.or_mf_tables(_dp(t))
    try:
        with open(TABLES + '.tmp', 'wb') as f:
            f.write(_library_key() + t.tobytes())
        os.replace(TABLES + '.tmp', TABLES)
    except OSError:
        pass


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def pack_program(program):
    out = np.zeros((len(program), 8), np.int32)
    for i, o in enumerate(program):
        out[i, 0] = OPS[o['op']]
        if o['op'] == 'prim':
            out[i, 1] = SHAPES.index(o['shape']); out[i, 2] = o['xf']; out[i, 3] = o['p']; out[i, 4] = o['mat']; out[i, 5] = o['prim']
        elif o['op'] == 'join':
            out[i, 6] = JOINS.index(o['join']); out[i, 7] = o['k']
            out[i, 1] = ('smooth', 'round', 'chamfer').index(o.get('shape', 'smooth'))
        elif o['op'] == 'node':
            out[i, 2] = o['bs']; out[i, 6] = o['bit']; out[i, 7] = o['end']
        elif o['op'] == 'mod':
            out[i, 1] = MODS.index(o['kind']); out[i, 2] = o['xf']; out[i, 3] = o['p']; out[i, 7] = o['end']
        elif o['op'] == 'user':
            out[i, 1] = o['index']; out[i, 2] = o['xf']; out[i, 4] = o['mat']; out[i, 5] = o['prim']
        elif o['op'] == 'ugroup':
            out[i, 1] = o['index']; out[i, 2] = o['xf']; out[i, 7] = o['end']
        elif o['op'] == 'relief':
            out[i, 1] = o['index']; out[i, 2] = o['xf']; out[i, 3] = o['amp']
    return np.ascontiguousarray(out)


# A scene's own code (objects' and groups' WGSL, as the emitter translated it) compiled in FP64 against wgsl_host.h,
# one module per distinct source, kept under _build/user/
USER_DIR = os.path.join(HERE, '_build', 'user')
DEVICE_DIR = os.path.join(os.path.dirname(HERE), 'shadytracer_amd', 'csrc', 'device')
_user_mods = {}


def user_module(user, material=None, reliefs=(), functions=None, sky=None):
    """The compiled module of a scene's `user` list, material hook and reliefs: (module, its function pointers in index
    order, their count, the material hook or None, the reliefs' pointers, their count)."""
    import hashlib
    src = ('#include "wgsl_host.h"\n' + (functions or '') + '\n' + '\n\n'.join(u['cpp'] for u in user) + '\n' + (material or '') +
           '\n' + '\n\n'.join(r['cpp'] for r in reliefs) + '\n' + (sky or '') + '\n')
    key = hashlib.sha1(src.encode()).hexdigest()[:16]
    if key in _user_mods:
        return _user_mods[key]
    os.makedirs(USER_DIR, exist_ok=True)
    so = os.path.join(USER_DIR, 'u%s.so' % key)
    if not os.path.exists(so):
        cpp = os.path.join(USER_DIR, 'u%s.cpp' % key)
        with open(cpp, 'w') as f:
            f.write(src)
        r = subprocess.run(['g++', '-O2', '-std=c++14', '-shared', '-fPIC', '-I', HERE, '-I', DEVICE_DIR, '-o', so + '.tmp', cpp],
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError('the scene code did not compile for the oracle:\n' + r.stdout[-4000:])
        os.replace(so + '.tmp', so)
    M = ctypes.CDLL(so)
    M.wg_set_prim.argtypes = [ctypes.c_void_p]
    M.wg_set_prim(ctypes.cast(lib().or_prim_local, ctypes.c_void_p))
    fns = (ctypes.c_void_p * max(len(user), 1))()
    for i, u in enumerate(user):
        fns[i] = ctypes.cast(getattr(M, ('ucode%d' if u['kind'] == 'object' else 'ugroup%d') % i), ctypes.c_void_p)
    hook = ctypes.cast(M.umaterial, ctypes.c_void_p) if material else None
    rel = (ctypes.c_void_p * max(len(reliefs), 1))()
    for i in range(len(reliefs)):
        rel[i] = ctypes.cast(getattr(M, 'urelief%d' % i), ctypes.c_void_p)
    skyfn = ctypes.cast(M.usky, ctypes.c_void_p) if sky else None
    _user_mods[key] = (M, fns, len(user), hook, rel, len(reliefs), skyfn)
    return _user_mods[key]


class Scene:
    """A scene as the oracle reads it, with the ctypes buffers kept alive."""

    def __init__(self, scene):
        from shadytracer_amd import pt
        self.ops = pack_program(scene['program'])
        coded = scene.get('user') or scene.get('materialCode') or scene.get('reliefs') or scene.get('skyCode')
        self.user = (user_module(scene.get('user') or [], scene.get('materialCode'), scene.get('reliefs') or [], scene.get('functionsCode'),
                                 scene.get('skyCode')) if coded else None)
        self.U = np.ascontiguousarray(np.asarray(scene['params'], np.float32))
        mats = scene['materials']
        self.mats = (pt.PtMaterial * len(mats))()
        for i, m in enumerate(mats):
            for k in ('base', 'emissive'):
                for j in range(3):
                    getattr(self.mats[i], k)[j] = m[k][j]
            self.mats[i].metallic, self.mats[i].roughness, self.mats[i].transmission = m['metallic'], m['roughness'], m['transmission']
            self.mats[i].ior, self.mats[i].specular, self.mats[i].abbe, self.mats[i].face = m['ior'], m['specular'], m['abbe'], m['face']
            self.mats[i].specularTint, self.mats[i].attenuation = m.get('specularTint', 0.0), m.get('attenuation', 0.0)
        nf = len(scene['faces'])
        self.faces = (pt.PtLightFace * max(nf, 1))()
        for i, f in enumerate(scene['faces']):
            for k in ('centre', 'u', 'v', 'n', 'le'):
                for j in range(3):
                    getattr(self.faces[i], k)[j] = f[k][j]
            self.faces[i].halfx, self.faces[i].halfy, self.faces[i].shape = f['halfx'], f['halfy'], f['shape']
            self.faces[i].twoSided, self.faces[i].weight = f['twoSided'], f['weight']
        self.nmats, self.nfaces = len(mats), nf

    def args(self):
        # the scene's own code, for the call these arguments are for
        if self.user is not None:
            lib().or_set_user(self.user[1], self.user[2])
            lib().or_set_material(self.user[3])
            lib().or_set_reliefs(self.user[4], self.user[5])
            lib().or_set_sky(self.user[6])
        else:
            lib().or_set_user(None, 0)
            lib().or_set_material(None)
            lib().or_set_reliefs(None, 0)
            lib().or_set_sky(None)
        return (self.ops.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(self.ops),
                self.U.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))


def render(scene, u, samples, first=0, threads=0):
    """FP64 render: (accum [h, w, 4], moments [h, w, 2], stats dict)."""
    S = Scene(scene)
    n = u.width * u.height
    acc = np.zeros((n, 4), np.float64)
    mom = np.zeros((n, 2), np.float64)
    st = (ctypes.c_longlong * 9)()
    lib().or_render(*S.args(), ctypes.byref(S.mats), S.nmats, ctypes.byref(S.faces), S.nfaces, ctypes.byref(u),
                    first, samples, _dp(acc), _dp(mom), st, threads)
    names = ['segments', 'shadow_segments', 'steps', 'shadow_steps', 'maxiter', 'nonfinite', 'photon_segments', 'photons', 'gathers']
    return acc.reshape(u.height, u.width, 4), mom.reshape(u.height, u.width, 2), dict(zip(names, list(st)))


def sppm_radius(r1, alpha, k):
    """r_k of the progressive schedule, by its product."""
    return lib().or_sppm_radius(r1, alpha, k)


def pick(scene, u, sample=0, first=0, count=None):
    """Each camera ray's first event: (pos4 [n, 4], state [n], material [n], prim [n])."""
    S = Scene(scene)
    n = u.width * u.height if count is None else count
    pos = np.zeros((n, 4), np.float64)
    info = np.zeros((n, 2), np.uint32)
    lib().or_pick(*S.args(), ctypes.byref(u), sample, first, n, _dp(pos), info.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return pos, info[:, 0] & 0xFF, info[:, 0] >> 8, info[:, 1]


def env(u, dirs, scene=None):
    """The environment's radiance (sky and sun disc), RGB, in each direction; `scene` for a scene's own sky code."""
    if scene is not None:
        Scene(scene).args()
    dirs = np.ascontiguousarray(np.asarray(dirs, np.float64).reshape(-1, 3))
    out = np.zeros((len(dirs), 3))
    lib().or_env(ctypes.byref(u), _dp(dirs), len(dirs), _dp(out))
    return out


def scene_at(scene, pts):
    S = Scene(scene)
    pts = np.ascontiguousarray(np.asarray(pts, np.float64).reshape(-1, 3))
    d = np.zeros(len(pts)); nrm = np.zeros((len(pts), 3)); prim = np.zeros(len(pts), np.int32)
    lib().or_scene_at(*S.args(), _dp(pts), len(pts), _dp(d), _dp(nrm), prim.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return d, nrm, prim


def bsdf(material, n, view, eta, l=None, u=None, through=False):
    """eval/pdf at l and a sample from u, for one material (a PtMaterial); through: the transmission lobe below the
    surface evaluated too."""
    n = np.asarray(n, np.float64); view = np.asarray(view, np.float64)
    f = np.zeros(3); pdf = np.zeros(1); sd = np.zeros(3); sw = np.zeros(3); fl = ctypes.c_int(0)
    la = np.asarray(l, np.float64) if l is not None else None
    ua = np.asarray(u, np.float64) if u is not None else None
    lib().or_bsdf(ctypes.byref(material), _dp(n), _dp(view), eta, _dp(la) if la is not None else None, _dp(f), _dp(pdf),
                  _dp(ua) if ua is not None else None, _dp(sd), _dp(sw), ctypes.byref(fl), int(through))
    return f, pdf[0], sd, sw, fl.value


def mf_tables():
    from shadytracer_amd.pt import load_library  # noqa: F401  (sizes come from pt_types.h)
    out = np.zeros(lib().or_mf_words(), np.float64)
    lib().or_mf_tables(_dp(out))
    return out


# ---------------------------------------------------------------------------------------------------- optics bench
def optics_trace(scenes, threads=0, tallies=False):
    """The bench in FP64 on the CPU: (results as dicts, far [n, 181, 360] lm, lux [n, 256, 256] lx) -- the tallies only
    when asked."""
    from shadytracer_amd import pt
    n = len(scenes)
    arr = (pt.PtOpticsScene * n)(*scenes)
    res = (pt.PtOpticsResult * n)()
    F = np.zeros((n, pt.OPT_NG, pt.OPT_NC)) if tallies else None
    Lx = np.zeros((n, pt.OPT_LUX, pt.OPT_LUX)) if tallies else None
    lib().or_optics_trace(arr, n, res, _dp(F) if tallies else None, _dp(Lx) if tallies else None, threads)
    return [pt.result_dict(r) for r in res], F, Lx


def optics_numbers(scene, far, lux, tot):
    """The numbers of one scene from given tallies (far lm, lux lx, tot: X, Y, Z escaped, Y absorbed, photons)."""
    from shadytracer_amd import pt
    r = pt.PtOpticsResult()
    far = np.ascontiguousarray(far, np.float64); lux = np.ascontiguousarray(lux, np.float64)
    tot = np.ascontiguousarray(tot, np.float64)
    lib().or_optics_numbers(ctypes.byref(scene), _dp(far), _dp(lux), _dp(tot), ctypes.byref(r))
    return pt.result_dict(r)


def cct_duv(x, y):
    c, d = ctypes.c_double(), ctypes.c_double()
    lib().or_cct_duv(x, y, ctypes.byref(c), ctypes.byref(d))
    return c.value, d.value


def fresnel_unpolarised(cos_i, eta):
    t = ctypes.c_double()
    return lib().or_fresnel_unpolarised(cos_i, eta, ctypes.byref(t))

