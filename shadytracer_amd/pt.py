"""ctypes binding of libpt (include/pt.h): the engine as the tests and the benchmark drive it.

It mirrors the reference's engine object (lib/pathtracer/core/raytrace_main.js:121 RAYTRACER): set_scene,
set_uniforms, trace, image_stats, read_region, clear_accums. There is no CPU fallback: without the built
library or a GPU the calls raise.
"""
import ctypes
import hashlib
import json
import math
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, 'libpt.so')

F3 = ctypes.c_float * 3
F2 = ctypes.c_float * 2


class PtCamera(ctypes.Structure):
    _fields_ = [('pos', F3), ('fx', ctypes.c_float), ('right', F3), ('fy', ctypes.c_float), ('up', F3),
                ('type', ctypes.c_uint), ('fwd', F3), ('lensR', ctypes.c_float), ('focus', ctypes.c_float), ('pad', F3)]


class PtMaterial(ctypes.Structure):
    _fields_ = [('base', F3), ('metallic', ctypes.c_float), ('roughness', ctypes.c_float), ('transmission', ctypes.c_float),
                ('ior', ctypes.c_float), ('specular', ctypes.c_float), ('emissive', F3), ('abbe', ctypes.c_float),
                ('face', ctypes.c_int), ('specularTint', ctypes.c_float), ('attenuation', ctypes.c_float), ('pad', ctypes.c_float * 1)]


class PtLightFace(ctypes.Structure):
    _fields_ = [('centre', F3), ('halfx', ctypes.c_float), ('u', F3), ('halfy', ctypes.c_float), ('v', F3), ('shape', ctypes.c_uint),
                ('n', F3), ('twoSided', ctypes.c_uint), ('le', F3), ('weight', ctypes.c_float)]


class PtEnv(ctypes.Structure):
    _fields_ = [('type', ctypes.c_uint), ('L', ctypes.c_float), ('col', F3), ('sunDir', F3), ('sunCos', ctypes.c_float),
                ('sunE', ctypes.c_float), ('sunRadius', ctypes.c_float), ('sunPoint', F3), ('top', ctypes.c_float),
                ('pad', ctypes.c_float), ('sunL', F3), ('sunSin2', ctypes.c_float), ('sunOmega', ctypes.c_float),
                ('skyAerosol', ctypes.c_float), ('skyMs', ctypes.c_float), ('skyUp', ctypes.c_float)]


class PtHaze(ctypes.Structure):
    _fields_ = [('centre', F3), ('sigmaS', ctypes.c_float), ('half', F3), ('g', ctypes.c_float), ('albedo', ctypes.c_float),
                ('pad', F3)]


class PtUniforms(ctypes.Structure):
    _fields_ = [('width', ctypes.c_uint), ('height', ctypes.c_uint), ('maxbounce', ctypes.c_uint), ('maxiter', ctypes.c_uint),
                ('stepscale', ctypes.c_float), ('eps', ctypes.c_float), ('epsAuto', ctypes.c_float), ('envRadius', ctypes.c_float),
                ('nee', ctypes.c_uint), ('rr', ctypes.c_uint), ('rrStart', ctypes.c_uint), ('diffuseCap', ctypes.c_uint),
                ('fireflyClamp', ctypes.c_float), ('fireflyWhite', ctypes.c_float), ('maxaverages', ctypes.c_uint),
                ('seed', ctypes.c_uint), ('cam', PtCamera), ('env', PtEnv), ('haze', PtHaze), ('faceShare', ctypes.c_float),
                ('sppm', ctypes.c_uint), ('sppmRadius', ctypes.c_float), ('sppmPhotons', ctypes.c_uint),
                ('marchRefill', ctypes.c_uint), ('guide', ctypes.c_uint), ('guideFrac', ctypes.c_float), ('sppmAlpha', ctypes.c_float),
                ('relax', ctypes.c_float), ('measureMode', ctypes.c_uint), ('lt', ctypes.c_uint)]


class PtSceneDesc(ctypes.Structure):
    _fields_ = [('hip_src', ctypes.c_char_p), ('params', ctypes.POINTER(ctypes.c_float)), ('nparams', ctypes.c_uint32),
                ('materials', ctypes.POINTER(PtMaterial)), ('nmaterials', ctypes.c_uint32),
                ('faces', ctypes.POINTER(PtLightFace)), ('nfaces', ctypes.c_uint32), ('flags', ctypes.c_uint32)]


class PtGuideConfig(ctypes.Structure):
    _fields_ = [('boxMin', ctypes.c_float * 3), ('lr', ctypes.c_float), ('boxMax', ctypes.c_float * 3),
                ('batch', ctypes.c_uint), ('steps', ctypes.c_uint), ('stride', ctypes.c_uint), ('warmup', ctypes.c_uint),
                ('seed', ctypes.c_uint)]


class PtOpticsScene(ctypes.Structure):
    _fields_ = [('ledHalf', ctypes.c_float), ('flux', ctypes.c_float), ('xyz', ctypes.c_float * 3), ('lensZ', ctypes.c_float),
                ('lensT', ctypes.c_float), ('lensD', ctypes.c_float), ('c1', ctypes.c_float), ('k1', ctypes.c_float),
                ('c2', ctypes.c_float), ('k2', ctypes.c_float), ('ior', ctypes.c_float), ('reflF', ctypes.c_float),
                ('reflR0', ctypes.c_float), ('reflZ1', ctypes.c_float), ('rho', ctypes.c_float * 3), ('boardR', ctypes.c_float),
                ('photons', ctypes.c_uint), ('seed', ctypes.c_uint)]


class PtOpticsResult(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ('emitted', 'escaped', 'absorbed', 'efficiency', 'Imax', 'gammaMax', 'cMax')] + \
               [('beam', ctypes.c_double * 2), ('field', ctypes.c_double * 2)] + \
               [(k, ctypes.c_double) for k in ('X', 'Y', 'Z', 'x', 'y', 'cct', 'duv', 'luxMax', 'luxCentre', 'photons',
                                               'events', 'flops')]


OPT_NG, OPT_NC, OPT_LUX = 181, 360, 256


def result_dict(r):
    d = {k: getattr(r, k) for k, _ in PtOpticsResult._fields_ if k not in ('beam', 'field')}
    d['beam'], d['field'] = list(r.beam), list(r.field)
    return d


class PtOpts(ctypes.Structure):
    _fields_ = [('device', ctypes.c_int), ('blocks_per_cu', ctypes.c_uint32), ('cache_dir', ctypes.c_char_p),
                ('straggler_rounds', ctypes.c_uint32), ('flags', ctypes.c_uint32), ('reserved', ctypes.c_uint32 * 2)]


STRAGGLER_NONE = 0xFFFFFFFF
OPT_DIRECT_LAUNCH = 1        # a sample's kernels launched one by one, not replayed from the captured graph


class PtTuning(ctypes.Structure):
    _fields_ = [('blocksPerCu', ctypes.c_uint), ('refillMarch', ctypes.c_uint), ('refillOcclude', ctypes.c_uint),
                ('candidates', ctypes.c_uint), ('msPerSample', ctypes.c_double), ('msBefore', ctypes.c_double)]


class PtOpticsSource(ctypes.Structure):
    _fields_ = [('pos', ctypes.c_float * 3), ('shape', ctypes.c_uint), ('axis', ctypes.c_float * 3), ('distribution', ctypes.c_uint),
                ('u', ctypes.c_float * 3), ('halfAngle', ctypes.c_float), ('size', ctypes.c_float * 2), ('n', ctypes.c_float),
                ('flux', ctypes.c_float), ('colour', ctypes.c_float * 3), ('emitN', ctypes.c_float), ('twoSided', ctypes.c_uint),
                ('onSurface', ctypes.c_uint), ('pad', ctypes.c_uint * 2)]


class PtOpticsTrace(ctypes.Structure):
    _fields_ = [('rays', ctypes.c_uint), ('bounces', ctypes.c_uint), ('seed', ctypes.c_uint), ('detCap', ctypes.c_uint),
                ('maxDist', ctypes.c_float), ('farMode', ctypes.c_uint), ('farMaterial', ctypes.c_uint), ('nC', ctypes.c_uint),
                ('nG', ctypes.c_uint), ('farAxis', ctypes.c_float * 3), ('farC0', ctypes.c_float * 3),
                ('detectors', ctypes.c_uint * 2), ('marchSteps', ctypes.c_uint), ('pad', ctypes.c_uint)]


MAX_MATERIALS = 64
OPT_HIT_WORDS = 8


class PtOpticsTally(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ('emitted', 'escaped', 'lost', 'detected', 'volume', 'absorbed')] + \
               [(k, ctypes.c_uint64) for k in ('rays', 'escapedRays', 'lostRays', 'interactions', 'absorbedRays', 'detectedRays',
                                               'detectorHits', 'pad')] + \
               [('incident', ctypes.c_double * MAX_MATERIALS), ('leaving', ctypes.c_double * MAX_MATERIALS),
                ('hits', ctypes.c_uint64 * MAX_MATERIALS)]


class PtDevelop(ctypes.Structure):
    _fields_ = [('toneMap', ctypes.c_uint), ('autoExposure', ctypes.c_uint), ('exposureEV', ctypes.c_float), ('aeSpeed', ctypes.c_float),
                ('aeHeadroom', ctypes.c_float), ('lightOn', ctypes.c_uint), ('flipY', ctypes.c_uint), ('transfer', ctypes.c_uint),
                ('contrast', ctypes.c_float), ('saturation', ctypes.c_float), ('brightness', ctypes.c_float),
                ('measureMode', ctypes.c_uint), ('measureMin', ctypes.c_float), ('measureMax', ctypes.c_float),
                ('measureBands', ctypes.c_uint), ('wb', ctypes.c_float * 9), ('eq', ctypes.c_float * 8), ('lab', ctypes.c_float * 3),
                ('tone7', ctypes.c_float * 3), ('levels', ctypes.c_float * 5), ('cdlSlope', ctypes.c_float * 3),
                ('cdlOffset', ctypes.c_float * 3), ('cdlPower', ctypes.c_float * 3), ('vignette', ctypes.c_float * 3),
                ('grain', ctypes.c_float * 2), ('pad', ctypes.c_float * 7), ('curveLut', ctypes.c_float * 4096)]


class PtDevelopInfo(ctypes.Structure):
    _fields_ = [('ev', ctypes.c_float), ('log2Average', ctypes.c_float), ('log2High', ctypes.c_float), ('metered', ctypes.c_float)]


# the reference's toneMap numbering (core/rendertoscreen.js)
TONE_MAPS = {'linear': 0, 'neutral': 1, 'agx': 2, 'agx-punchy': 3, 'aces': 4, 'reinhard': 5, 'hable': 6, 'hable-adjustable': 7}
LIGHT_ANY, LIGHT_DITHER = 1, 512

STRUCTS = [PtUniforms, PtMaterial, PtLightFace, PtCamera, PtEnv, PtHaze, PtSceneDesc, PtOpts, PtGuideConfig, PtOpticsScene,
           PtOpticsResult, PtTuning, PtOpticsSource, PtOpticsTrace, PtOpticsTally, PtDevelop, PtDevelopInfo]

# the C ABI, name -> (restype, argtypes)
SIGNATURES = {
    'pt_abi_version': (ctypes.c_int, []),
    'pt_struct_size': (ctypes.c_size_t, [ctypes.c_int]),
    'pt_create': (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(PtOpts)]),
    'pt_destroy': (None, [ctypes.c_void_p]),
    'pt_last_error': (ctypes.c_char_p, [ctypes.c_void_p]),
    'pt_compile_scene': (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t]),
    'pt_set_scene': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PtSceneDesc)]),
    'pt_set_params': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_uint32]),
    'pt_fold_params': (ctypes.c_int, [ctypes.c_void_p]),
    'pt_set_uniforms': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PtUniforms), ctypes.c_size_t]),
    'pt_clear_accum': (ctypes.c_int, [ctypes.c_void_p]),
    'pt_trace': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    'pt_sync': (ctypes.c_int, [ctypes.c_void_p]),
    'pt_read_accum': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), ctypes.c_size_t]),
    'pt_read_stats': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.c_size_t]),
    'pt_reset_stats': (ctypes.c_int, [ctypes.c_void_p]),
    'pt_time_kernels': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double), ctypes.c_size_t]),
    'pt_pick_rays': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_uint32), ctypes.c_size_t]),
    'pt_guide_config': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PtGuideConfig), ctypes.c_size_t]),
    'pt_guide_weights': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_size_t]),
    'pt_guide_load': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_size_t]),
    'pt_guide_forward': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_size_t, ctypes.POINTER(ctypes.c_float)]),
    'pt_guide_train': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_size_t, ctypes.POINTER(ctypes.c_float),
                                      ctypes.POINTER(ctypes.c_double)]),
    'pt_guide_status': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_size_t]),
    'pt_optics_trace': (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(PtOpticsScene), ctypes.c_uint32, ctypes.POINTER(PtOpticsResult),
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    'pt_optics_error': (ctypes.c_char_p, []),
    'pt_develop_view': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PtDevelop), ctypes.c_size_t, ctypes.POINTER(ctypes.c_float),
                        ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint8), ctypes.c_size_t,
                        ctypes.POINTER(PtDevelopInfo)]),
    'pt_develop16': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PtDevelop), ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint16),
                     ctypes.c_size_t, ctypes.POINTER(PtDevelopInfo)]),
    'pt_develop_histogram': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PtDevelop), ctypes.c_size_t, ctypes.c_uint32,
                             ctypes.POINTER(ctypes.c_uint32), ctypes.c_size_t, ctypes.POINTER(PtDevelopInfo)]),
    'pt_develop': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PtDevelop), ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint8),
                                  ctypes.c_size_t, ctypes.POINTER(PtDevelopInfo)]),
    'pt_optics_run': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PtOpticsSource), ctypes.c_uint32, ctypes.POINTER(PtOpticsTrace),
                                     ctypes.c_size_t, ctypes.POINTER(PtOpticsTally), ctypes.c_size_t, ctypes.POINTER(ctypes.c_double),
                                     ctypes.c_size_t, ctypes.POINTER(ctypes.c_float), ctypes.c_uint32]),
    'pt_optics_trace_static': (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(PtOpticsScene), ctypes.c_uint32, ctypes.POINTER(PtOpticsResult),
                                              ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    'pt_get_tuning': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PtTuning), ctypes.c_size_t]),
    'pt_set_tuning': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PtTuning), ctypes.c_size_t]),
    'pt_autotune': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(PtTuning), ctypes.c_size_t]),
}
GUIDE_NP = 16169

STAT_NAMES = ['segments', 'shadow_segments', 'steps', 'shadow_steps', 'nonfinite', 'maxiter', 'paths', 'leaving',
              'gathers', 'photons', 'flops', 'shadow_flops', 'photon_segments', 'gather_cut', 'radius_bits',
              'overflow', 'guide_steps', 'guided', 'guide_records', 'flushed', 'shadowcut', 'guide_infer',
              'guide_tiles', 'guide_train']
PT_SCENE_KEEP_PICTURE = 1   # pt_scene_desc.flags (include/pt.h)
STRAGGLER_ROUNDS = 1    # PT_STRAGGLER_ROUNDS (include/pt_types.h): rounds a sample runs beyond maxbounce + 1
LUMA = (0.2126, 0.7152, 0.0722)

_lib = None


class PtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__('libpt error %d: %s' % (code, msg))
        self.code = code
        self.restored = code == 3


def load_library(path=LIB_PATH):
    """Loads libpt and binds the C ABI. Raises if the library is not built: there is no fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError('libpt.so is not built (%s): run `python -m shadytracer_amd.build`' % path)
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    for i, s in enumerate(STRUCTS):
        if lib.pt_struct_size(i) != ctypes.sizeof(s):
            raise RuntimeError('ABI mismatch: %s is %d bytes in libpt, %d here' % (s.__name__, lib.pt_struct_size(i), ctypes.sizeof(s)))
    _lib = lib
    return lib


def _set3(arr, v):
    for i in range(3):
        arr[i] = float(v[i])


def make_camera(cam):
    c = PtCamera()
    _set3(c.pos, cam['pos']); _set3(c.right, cam['right']); _set3(c.up, cam['up']); _set3(c.fwd, cam['fwd'])
    c.fx, c.fy, c.type = cam['fx'], cam['fy'], int(cam.get('type', 0))
    c.lensR, c.focus = float(cam.get('lensR', 0.0)), float(cam.get('focus', 1.0))
    return c


def make_env(e):
    v = PtEnv()
    v.type, v.L = int(e['type']), float(e['L'])
    _set3(v.col, e['col']); _set3(v.sunDir, e['sunDir']); _set3(v.sunPoint, e['sunPoint'])
    v.sunCos, v.sunE, v.sunRadius = e['sunCos'], e['sunE'], e['sunRadius']
    v.top = float(e.get('top', 0.0))
    # the sun's cone in double: sin^2 of its radius and its solid angle, from the cosine when a caller gives only that
    half = math.acos(min(max(float(e['sunCos']), -1.0), 1.0))
    v.sunSin2 = float(e.get('sunSin2', math.sin(half) ** 2))
    v.sunOmega = float(e.get('sunOmega', 4 * math.pi * math.sin(0.5 * half) ** 2))
    _set3(v.sunL, e.get('sunL', [0.0, 0.0, 0.0]))
    v.skyAerosol, v.skyMs, v.skyUp = float(e.get('skyAerosol', 0.0)), float(e.get('skyMs', 1.0)), float(e.get('skyUp', 0.0))
    return v


def make_haze(h):
    v = PtHaze()
    if h:
        _set3(v.centre, h['centre']); _set3(v.half, h['half'])
        v.sigmaS, v.g, v.albedo = h['sigmaS'], h['g'], h['albedo']
    return v


def default_uniforms():
    u = PtUniforms()
    u.width, u.height, u.maxbounce, u.maxiter = 256, 256, 10, 400
    u.stepscale, u.eps, u.epsAuto, u.envRadius = 0.9, 0.0, 1e-4, 400.0
    u.nee, u.rr, u.rrStart, u.diffuseCap = 1, 0, 3, 0
    u.fireflyClamp, u.fireflyWhite, u.maxaverages, u.seed = 0.0, 1.0, 0, 20261203
    u.faceShare, u.marchRefill, u.guideFrac = 0.5, 0, 0.9
    u.sppmRadius, u.sppmAlpha = 0.02, 2.0 / 3.0
    u.relax = 0.0
    return u


def uniforms_for(made, **over):
    """The engine uniforms of an emitted configuration ({scene, camera, settings}), with overrides."""
    s = dict(made.get('settings') or {})
    s.update(over)
    sc = made['scene']
    u = default_uniforms()
    u.width, u.height = int(s.get('width', 256)), int(s.get('height', 256))
    u.maxbounce = int(s.get('maxbounce', 10))
    u.nee = int(s.get('nee', 1))
    u.seed = int(s.get('seed', 20261203))
    u.envRadius = float(sc['envRadius'])
    # the scene's own tolerance (photometry.js autoEps: 1e-4 of its bounds' diagonal), unless the settings give one
    if sc.get('epsAuto'):
        u.epsAuto = float(sc['epsAuto'])
    u.cam = make_camera(made['camera'])
    u.env = make_env(sc['env'])
    u.haze = make_haze(sc.get('haze'))
    for k in ('maxiter', 'marchRefill', 'rr', 'rrStart', 'diffuseCap', 'maxaverages', 'sppm', 'sppmPhotons', 'measureMode', 'lt'):
        if k in s:
            setattr(u, k, int(s[k]))
    for k in ('stepscale', 'eps', 'epsAuto', 'faceShare', 'fireflyClamp', 'sppmRadius', 'sppmAlpha', 'relax'):
        if k in s:
            setattr(u, k, float(s[k]))
    return u


class Engine:
    """One libpt context on a GPU."""

    def __init__(self, device=0, blocks_per_cu=0, cache_dir=None, straggler_rounds=None, direct=False):
        """straggler_rounds: rounds beyond maxbounce + 1 for paths held up by hits from behind (None: the library's
        default, 0: none -- the flush closes them). direct: launch each sample's kernels one by one instead of
        replaying the captured graph (PT_OPT_DIRECT_LAUNCH)."""
        self.lib = load_library()
        o = PtOpts()
        o.device, o.blocks_per_cu = device, blocks_per_cu
        o.flags = OPT_DIRECT_LAUNCH if direct else 0
        o.straggler_rounds = 0 if straggler_rounds is None else (STRAGGLER_NONE if straggler_rounds == 0 else straggler_rounds)
        self.straggler_rounds = STRAGGLER_ROUNDS if straggler_rounds is None else min(straggler_rounds, 16)
        o.cache_dir = cache_dir.encode() if cache_dir else None
        h = ctypes.c_void_p()
        rc = self.lib.pt_create(ctypes.byref(h), ctypes.byref(o))
        if rc:
            raise PtError(rc, self.lib.pt_last_error(None).decode())
        self.h = h
        self.width = self.height = 0
        self._keep = []

    def _chk(self, rc):
        if rc:
            raise PtError(rc, self.lib.pt_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.lib.pt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_scene(self, scene, keep_picture=False):
        """Loads an emitted scene; keep_picture keeps the accumulation (a scene that draws the same picture)."""
        mats = (PtMaterial * len(scene['materials']))()
        for i, m in enumerate(scene['materials']):
            _set3(mats[i].base, m['base']); _set3(mats[i].emissive, m['emissive'])
            mats[i].metallic, mats[i].roughness, mats[i].transmission = m['metallic'], m['roughness'], m['transmission']
            mats[i].ior, mats[i].specular, mats[i].abbe = m['ior'], m['specular'], m['abbe']
            mats[i].face, mats[i].specularTint = int(m['face']), m.get('specularTint', 0.0)
            mats[i].attenuation = m.get('attenuation', 0.0)
        nf = len(scene['faces'])
        faces = (PtLightFace * max(nf, 1))()
        for i, f in enumerate(scene['faces']):
            for k in ('centre', 'u', 'v', 'n', 'le'):
                _set3(getattr(faces[i], k), f[k])
            faces[i].halfx, faces[i].halfy, faces[i].shape = f['halfx'], f['halfy'], int(f['shape'])
            faces[i].twoSided, faces[i].weight = int(f['twoSided']), f['weight']
        params = np.ascontiguousarray(np.asarray(scene['params'], dtype=np.float32))
        d = PtSceneDesc()
        d.hip_src = scene['hip'].encode()
        d.params = params.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        d.nparams = len(params)
        d.materials, d.nmaterials = mats, len(scene['materials'])
        d.faces, d.nfaces = faces, nf
        d.flags = PT_SCENE_KEEP_PICTURE if keep_picture else 0
        self._keep = [params, mats, faces, d]
        self._chk(self.lib.pt_set_scene(self.h, ctypes.byref(d)))
        self._scene_hash = hashlib.sha1(scene['hip'].encode()).hexdigest()[:16]

    def set_params(self, params):
        a = np.ascontiguousarray(np.asarray(params, dtype=np.float32))
        self._chk(self.lib.pt_set_params(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(a)))

    def fold_params(self):
        """The scene's kernels with its current parameter block as constants, until set_params or set_scene."""
        self._chk(self.lib.pt_fold_params(self.h))

    def set_uniforms(self, u):
        self._chk(self.lib.pt_set_uniforms(self.h, ctypes.byref(u), ctypes.sizeof(u)))
        self.width, self.height = u.width, u.height
        self._u = u

    def clear_accums(self):
        self._chk(self.lib.pt_clear_accum(self.h))

    def trace(self, samples=1):
        self._chk(self.lib.pt_trace(self.h, samples))

    def sync(self):
        self._chk(self.lib.pt_sync(self.h))

    def read_accum(self):
        """(accum [h, w, 4] mean rgb + samples, moments [h, w, 2] V, S) as float64 arrays; row 0 is the bottom."""
        n = self.width * self.height
        acc = np.zeros((n, 4), np.float64)
        mom = np.zeros((n, 2), np.float64)
        self._chk(self.lib.pt_read_accum(self.h, acc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                         mom.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n))
        return acc.reshape(self.height, self.width, 4), mom.reshape(self.height, self.width, 2)

    def stats(self):
        a = (ctypes.c_uint64 * 24)()
        self._chk(self.lib.pt_read_stats(self.h, a, 24))
        out = {k: int(a[i]) for i, k in enumerate(STAT_NAMES)}
        out['sppm_radius'] = float(np.array([out['radius_bits'] & 0xFFFFFFFF], np.uint32).view(np.float32)[0])
        return out

    def reset_stats(self):
        self._chk(self.lib.pt_reset_stats(self.h))

    # ---- the developed picture (core/rendertoscreen.js, core/autoexposure.js)
    def _develop_settings(self, tone, auto, ev, speed, headroom, dither, flip, grade):
        d = PtDevelop()
        d.toneMap = TONE_MAPS[tone] if isinstance(tone, str) else int(tone)
        d.autoExposure, d.exposureEV, d.aeSpeed, d.aeHeadroom = int(bool(auto)), ev, speed, headroom
        d.flipY = int(bool(flip))
        d.wb[:] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
        d.measureMin, d.measureMax = 1.0, 100000.0
        for k, v in (grade or {}).items():
            if v is None:
                continue
            f = getattr(d, k)
            if isinstance(f, ctypes.Array):
                f[:] = [float(x) for x in v]
            else:
                setattr(d, k, type(f)(v))
        if dither:
            d.lightOn |= LIGHT_ANY | LIGHT_DITHER
        return d

    def develop(self, tone='neutral', auto=True, ev=15.0, speed=1.0, headroom=2.0, dither=False, flip=True, grade=None):
        """The accumulation developed (pt_develop): (uint8 [h, w, 4] RGBA, rows top first with flip, and {'ev',
        'log2Average', 'log2High', 'metered'}). auto: metered from the picture, else EV100 `ev`. dither: triangular,
        one 8-bit step either way. grade: PtDevelop's darkroom words by name (lightOn, wb (rows), eq, lab, tone7,
        transfer, levels, cdlSlope, cdlOffset, cdlPower, vignette, grain, curveLut, contrast, saturation, brightness,
        measureMode, measureMin, measureMax, measureBands), as js/darkroom.js lightUniforms makes them."""
        d = self._develop_settings(tone, auto, ev, speed, headroom, dither, flip, grade)
        out = np.zeros((self.height, self.width, 4), np.uint8)
        info = PtDevelopInfo()
        self._chk(self.lib.pt_develop(self.h, ctypes.byref(d), ctypes.sizeof(d), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                      self.width * self.height, ctypes.byref(info)))
        return out, {k: getattr(info, k) for k, _ in PtDevelopInfo._fields_}

    def develop16(self, tone='neutral', auto=True, ev=15.0, speed=1.0, headroom=2.0, flip=True, grade=None):
        """develop() to 16 bits a channel (pt_develop16): (uint16 [h, w, 4] RGBA, alpha 65535, and the info); no dither."""
        d = self._develop_settings(tone, auto, ev, speed, headroom, False, flip, grade)
        out = np.zeros((self.height, self.width, 4), np.uint16)
        info = PtDevelopInfo()
        self._chk(self.lib.pt_develop16(self.h, ctypes.byref(d), ctypes.sizeof(d), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                                        self.width * self.height, ctypes.byref(info)))
        return out, {k: getattr(info, k) for k, _ in PtDevelopInfo._fields_}

    def develop_view(self, view, out_w, out_h, clip_warn=False, tone='neutral', auto=True, ev=15.0, speed=1.0, headroom=2.0,
                     dither=False, flip=True, grade=None):
        """A view of the developed picture (pt_develop_view): view = (cx, cy, sx, sy) in picture pixels, y up (the
        whole picture at the output: (w/2, h/2, w/out_w, h/out_h)); (uint8 [out_h, out_w, 4], info)."""
        d = self._develop_settings(tone, auto, ev, speed, headroom, dither, flip, grade)
        v = (ctypes.c_float * 4)(*[float(x) for x in view])
        out = np.zeros((out_h, out_w, 4), np.uint8)
        info = PtDevelopInfo()
        self._chk(self.lib.pt_develop_view(self.h, ctypes.byref(d), ctypes.sizeof(d), v, out_w, out_h, int(bool(clip_warn)),
                                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), out.shape[0] * out.shape[1], ctypes.byref(info)))
        return out, {k: getattr(info, k) for k, _ in PtDevelopInfo._fields_}

    def develop_histogram(self, tone='neutral', auto=True, ev=15.0, speed=1.0, headroom=2.0, dither=False, grade=None, step=1):
        """What develop() would write, counted over every `step`-th pixel each way (pt_develop_histogram): (uint32
        [4, 256]: R, G, B, luma, and the info develop() gives)."""
        d = self._develop_settings(tone, auto, ev, speed, headroom, dither, True, grade)
        hist = np.zeros((4, 256), np.uint32)
        info = PtDevelopInfo()
        self._chk(self.lib.pt_develop_histogram(self.h, ctypes.byref(d), ctypes.sizeof(d), int(step),
                                                hist.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), hist.size, ctypes.byref(info)))
        return hist, {k: getattr(info, k) for k, _ in PtDevelopInfo._fields_}

    # ---- the optics trace over the loaded scene (createOptics: optics/index.js)
    def optics_run(self, sources, rays=100000, bounces=64, seed=1, far=None, detectors=(), det_cap=0, max_dist=0.0,
                   march_steps=0):
        """Traces `rays` rays from `sources` (dicts as optics/sources.js describes them: type 'point' | 'disc' |
        'rectangle', pos, axis, u, size, distribution 'collimated' | 'uniform' | 'cosine', halfAngle (deg), n, flux (lm),
        colour, emitN, twoSided, onSurface; missing keys from default_source()) through the loaded scene, in metres.
        far: None, or {'mode': 'escaped' | 'detector', 'material', 'nC', 'nG', 'axis', 'c0'}; detectors: material ids
        whose surfaces stop and record rays; det_cap: hits kept. Returns the tallies in lm, per material, the far field
        ({'flux': [nG, nC] lm, ...}) and the hits ([k, 8]: position, lm, direction, material). march_steps: the
        sphere-tracing steps a ray may take over its whole path (0: 2000)."""
        srcs = [dict(default_source(), **sd) for sd in sources]
        arr = (PtOpticsSource * max(1, len(srcs)))()
        for a, sd in zip(arr, srcs):
            a.pos[:] = [float(x) for x in sd['pos']]
            a.shape = SOURCE_SHAPES[sd['type']]
            a.axis[:] = [float(x) for x in sd['axis']]
            a.distribution = DISTRIBUTIONS[sd['distribution']]
            a.u[:] = [float(x) for x in sd['u']]
            a.halfAngle = float(sd['halfAngle'])
            size = list(sd['size']) if hasattr(sd['size'], '__len__') else [sd['size'], sd['size']]
            a.size[:] = [float(size[0]), float(size[1] if len(size) > 1 else size[0])]
            a.n, a.flux, a.emitN = float(sd['n']), float(sd['flux']), float(sd['emitN'])
            a.colour[:] = [float(x) for x in sd.get('colour', (1.0, 1.0, 1.0))]
            a.twoSided, a.onSurface = int(bool(sd['twoSided'])), int(bool(sd.get('onSurface', False)))
        t = PtOpticsTrace()
        t.rays, t.bounces, t.seed, t.detCap, t.maxDist = rays, bounces, seed, det_cap, max_dist
        t.marchSteps = march_steps
        for m in detectors:
            t.detectors[m >> 5] |= 1 << (m & 31)
        F = None
        if far:
            f = dict(FAR_DEFAULT, **far)
            t.farMode = {'off': 0, 'escaped': 1, 'detector': 2}[f['mode']]
            t.farMaterial, t.nC, t.nG = f['material'], f['nC'], f['nG']
            t.farAxis[:] = [float(x) for x in f['axis']]
            t.farC0[:] = [float(x) for x in f['c0']]
            if t.farMode:
                F = np.zeros((f['nG'], f['nC']))
        H = np.zeros((max(det_cap, 1), OPT_HIT_WORDS), np.float32)
        out = PtOpticsTally()
        self._chk(self.lib.pt_optics_run(self.h, arr, len(srcs), ctypes.byref(t), ctypes.sizeof(t), ctypes.byref(out),
                                         ctypes.sizeof(out),
                                         F.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if F is not None else None,
                                         F.size if F is not None else 0,
                                         H.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), det_cap))
        res = {k: getattr(out, k) for k, _ in PtOpticsTally._fields_ if k not in ('incident', 'leaving', 'hits', 'pad')}
        res['materials'] = [{'incident': out.incident[m], 'leaving': out.leaving[m], 'hits': int(out.hits[m]),
                             'absorbed': out.incident[m] - out.leaving[m]} for m in range(MAX_MATERIALS)]
        res['far'] = None if F is None else dict(dict(FAR_DEFAULT, **far), flux=F, emitted=out.emitted)
        res['hits'] = H[:min(det_cap, out.detectorHits)].copy()
        return res

    def time_kernels(self, samples=1):
        ms = (ctypes.c_double * 12)()
        self._chk(self.lib.pt_time_kernels(self.h, samples, ms, 12))
        return dict(zip(['genrays', 'march', 'shade', 'occlude', 'resolve', 'bin', 'photon', 'surface', 'guide',
                         'guide_infer', 'guide_grad', 'guide_adam'], list(ms)))

    # ---- launch choices
    def tuning(self):
        t = PtTuning()
        self._chk(self.lib.pt_get_tuning(self.h, ctypes.byref(t), ctypes.sizeof(t)))
        return {k: getattr(t, k) for k, _ in PtTuning._fields_}

    def set_tuning(self, blocksPerCu, refillMarch, refillOcclude):
        t = PtTuning()
        t.blocksPerCu, t.refillMarch, t.refillOcclude = blocksPerCu, refillMarch, refillOcclude
        self._chk(self.lib.pt_set_tuning(self.h, ctypes.byref(t), ctypes.sizeof(t)))

    def tuning_key(self):
        """What a tuning is good for: the scene's source, the image, and the settings that shape a sample's work."""
        u = self._u
        return '%s/%dx%d/b%d/n%d/s%d:%d/l%d/g%d/x%d' % (self._scene_hash, u.width, u.height, u.maxbounce, u.nee, u.sppm,
                                                       u.sppmPhotons if (u.sppm or u.lt) else 0, u.lt, u.guide,
                                                       self.straggler_rounds)

    def autotune(self, samples=4, cache=None):
        """Picks the persistent kernels' grid and refill thresholds for the scene and uniforms set (pt_autotune) and
        returns the choice. With `cache` (a JSON file), a choice stored for the same tuning_key() is applied instead of
        timing again, and a new one is stored. Clears the accumulators either way."""
        key = self.tuning_key()
        table = {}
        if cache and os.path.exists(cache):
            try:
                table = json.load(open(cache))
            except (OSError, ValueError):
                table = {}
        if key in table:
            t = table[key]
            self.set_tuning(t['blocksPerCu'], t['refillMarch'], t['refillOcclude'])
            self.clear_accums()
            return dict(t, cached=True)
        t = PtTuning()
        self._chk(self.lib.pt_autotune(self.h, samples, ctypes.byref(t), ctypes.sizeof(t)))
        out = {k: getattr(t, k) for k, _ in PtTuning._fields_}
        if cache:
            table[key] = out
            tmp = cache + '.tmp%d' % os.getpid()
            with open(tmp, 'w') as f:
                json.dump(table, f, indent=1, sort_keys=True)
            os.replace(tmp, cache)
        return dict(out, cached=False)

    # ---- path guiding
    def guide_config(self, box, lr=1e-2, batch=1 << 16, steps=4, stride=0, warmup=16, seed=1):
        g = PtGuideConfig()
        for a in range(3):
            g.boxMin[a], g.boxMax[a] = float(box[0][a]), float(box[1][a])
        g.lr, g.batch, g.steps, g.stride, g.warmup, g.seed = lr, batch, steps, stride, warmup, seed
        self._chk(self.lib.pt_guide_config(self.h, ctypes.byref(g), ctypes.sizeof(g)))

    def guide_weights(self):
        th = np.zeros(GUIDE_NP, np.float32)
        self._chk(self.lib.pt_guide_weights(self.h, th.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), GUIDE_NP))
        return th

    def guide_load(self, theta):
        th = np.ascontiguousarray(np.asarray(theta, np.float32))
        self._chk(self.lib.pt_guide_load(self.h, th.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), th.size))

    def guide_forward(self, raw):
        raw = np.ascontiguousarray(np.asarray(raw, np.float32).reshape(-1, 10))
        out = np.zeros((len(raw), 41), np.float32)
        self._chk(self.lib.pt_guide_forward(self.h, raw.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(raw),
                                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def guide_train(self, samples):
        """One Adam step on rows of 16 floats; returns (summed gradient, loss)."""
        smp = np.ascontiguousarray(np.asarray(samples, np.float32).reshape(-1, 16))
        grad = np.zeros(GUIDE_NP, np.float32)
        loss = ctypes.c_double(0.0)
        self._chk(self.lib.pt_guide_train(self.h, smp.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(smp),
                                          grad.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(loss)))
        return grad, loss.value

    def guide_status(self, steps=4):
        out = (ctypes.c_double * (2 + steps))()
        self._chk(self.lib.pt_guide_status(self.h, out, 2 + steps))
        return {'steps': int(out[0]), 'collected': int(out[1]), 'losses': list(out)[2:]}

    def pick_rays(self):
        """Each camera ray of one sample to its first event: (pos4 [n, 4], state [n], material [n], prim [n])."""
        n = self.width * self.height
        pos = np.zeros((n, 4), np.float32)
        info = np.zeros((n, 2), np.uint32)
        self._chk(self.lib.pt_pick_rays(self.h, pos.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                        info.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), n))
        return pos, info[:, 0] & 0xFF, info[:, 0] >> 8, info[:, 1]

    def image_stats(self):
        """Luminance statistics of the accumulated image (R.imageStats, raytrace_main.js:876-956): mean, and the
        per-pixel standard error sqrt(S V / (1 - S)) (core/measure.js:49-52)."""
        acc, mom = self.read_accum()
        return image_stats(acc, mom)


def luminance(acc):
    return LUMA[0] * acc[..., 0] + LUMA[1] * acc[..., 1] + LUMA[2] * acc[..., 2]


def standard_error(mom):
    V, S = mom[..., 0], mom[..., 1]
    ok = (S > 0) & (S < 1) & (V >= 0)
    se = np.full(V.shape, np.nan)
    se[ok] = np.sqrt(S[ok] * V[ok] / (1 - S[ok]))
    return se


def image_stats(acc, mom):
    L = luminance(acc)
    se = standard_error(mom)
    have = acc[..., 3] > 0
    return {'meanLuma': float(L[have].mean()) if have.any() else 0.0, 'samples': float(acc[..., 3].max()),
            'relSE': float(np.nanmean(se[have] / np.maximum(L[have], 1e-30))) if have.any() else math.nan}


def scene_box(scene):
    """The box guiding encodes positions in: the bounding spheres of the scene's bounded top-level nodes (the emitter's
    program), a little enlarged; the environment sphere when nothing is bounded."""
    U = scene['params']
    lo, hi = [np.inf] * 3, [-np.inf] * 3
    for o in scene['program']:
        if o['op'] == 'node' and o['bs'] >= 0:
            c, r = U[o['bs']:o['bs'] + 3], U[o['bs'] + 3]
            for a in range(3):
                lo[a] = min(lo[a], c[a] - r)
                hi[a] = max(hi[a], c[a] + r)
    if not np.isfinite(lo[0]):
        R = scene['envRadius']
        return [-R] * 3, [R] * 3
    pad = [0.01 * (h - l) + 1e-3 for l, h in zip(lo, hi)]
    return [l - p for l, p in zip(lo, pad)], [h + p for h, p in zip(hi, pad)]


SOURCE_SHAPES = {'point': 0, 'disc': 1, 'rectangle': 2}
DISTRIBUTIONS = {'collimated': 0, 'uniform': 1, 'cosine': 2}
FAR_DEFAULT = {'mode': 'off', 'material': 0, 'nC': 72, 'nG': 181, 'axis': (0.0, 0.0, 1.0), 'c0': (1.0, 0.0, 0.0)}


def default_source():
    """optics/sources.js defaultSource: a 1 x 1 Lambertian square 1 up, facing down, 1000 lm, into air, one face."""
    return {'type': 'rectangle', 'pos': (0.0, 0.0, 1.0), 'axis': (0.0, 0.0, -1.0), 'u': (1.0, 0.0, 0.0), 'size': (0.5, 0.5),
            'distribution': 'cosine', 'halfAngle': 90.0, 'n': 1.0, 'flux': 1000.0, 'emitN': 1.0, 'twoSided': False,
            'onSurface': False, 'colour': (1.0, 1.0, 1.0)}


def optics_trace(scenes, far=False, lux=False, device=0):
    """The optics bench on the GPU: (results as dicts, far [n, 181, 360] lm or None, lux [n, 256, 256] lx or None,
    (walk ms, numbers ms))."""
    lib = load_library()
    n = len(scenes)
    arr = (PtOpticsScene * n)(*scenes)
    res = (PtOpticsResult * n)()
    F = np.zeros((n, OPT_NG, OPT_NC)) if far else None
    L = np.zeros((n, OPT_LUX, OPT_LUX)) if lux else None
    ms = (ctypes.c_double * 2)()
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None
    rc = lib.pt_optics_trace(device, arr, n, res, dp(F), dp(L), ms)
    if rc:
        raise PtError(rc, lib.pt_optics_error().decode())
    return [result_dict(r) for r in res], F, L, (ms[0], ms[1])

