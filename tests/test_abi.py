"""The drop-in boundary without a GPU: libpt exports every entry point include/pt.h declares, its structures agree
with the bindings (ctypes, and the N-API addon's JS packing), scenes compile for gfx950 through hipRTC, errors come
back as codes with messages, and the emitter refuses what it cannot translate."""
import ctypes
import json
import os
import re
import subprocess

import pytest

from shadytracer_amd import pt, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, 'include', 'pt.h')).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(pt_[a-z0-9_]+)\s*\(', text)))


def have_gpu():
    return os.path.exists('/dev/kfd') and any(n.startswith('renderD') for n in os.listdir('/dev/dri')) if os.path.exists('/dev/dri') else False


def test_library_exports_every_declared_entry_point():
    lib = pt.load_library()
    names = declared()
    assert len(names) >= 15, names
    for n in names:
        assert hasattr(lib, n), n
    # and the bindings bind exactly what is declared
    assert set(pt.SIGNATURES) == set(names)


def test_struct_sizes_match_the_bindings():
    lib = pt.load_library()
    for i, s in enumerate(pt.STRUCTS):
        assert lib.pt_struct_size(i) == ctypes.sizeof(s), s.__name__
    assert lib.pt_struct_size(99) == 0
    assert lib.pt_abi_version() == 4


def test_addon_loads_and_packs_the_same_structures():
    out = json.loads(subprocess.run(['node', os.path.join(ROOT, 'tests', 'js', 'shim_cpu.js')], check=True,
                                    stdout=subprocess.PIPE, text=True).stdout)
    s = out['sizes']
    assert s[0] == s[1] == ctypes.sizeof(pt.PtUniforms) and s[2] == s[3] and s[4] == s[5]
    assert out['addonKeys'] == sorted(['abiVersion', 'autotune', 'clearAccum', 'compileScene', 'create', 'destroy', 'develop', 'develop16', 'developHistogram', 'developView', 'foldParams', 'opticsRun',
                                       'opticsTrace',
                                       'readAccum', 'readStats', 'setParams', 'setScene', 'setTuning', 'setUniforms', 'structSize',
                                       'trace'])
    # the reference's lighting settings are set() names (uniforms.schema.js Lighting), none returned as unknown
    assert out['lighting'] == [] and out['elevation'] == 20
    assert out['developSize'] == [ctypes.sizeof(pt.PtDevelop)] * 2
    t = out['traceSizes']
    assert t[0] == t[1] == ctypes.sizeof(pt.PtOpticsSource) and t[2] == t[3] == ctypes.sizeof(pt.PtOpticsTrace)
    assert t[4] == t[5] == ctypes.sizeof(pt.PtOpticsTally)
    o = out['opticsSizes']
    assert o[0] == o[1] == ctypes.sizeof(pt.PtOpticsScene) and o[2] == o[3] == ctypes.sizeof(pt.PtOpticsResult)
    # fields land at their C offsets
    u = pt.PtUniforms
    assert [u.width.offset, u.height.offset, u.seed.offset, u.cam.offset, u.env.offset] == [0, 4, 60, 64, 144]
    assert [pt.PtCamera.lensR.offset, pt.PtCamera.focus.offset, ctypes.sizeof(pt.PtCamera), ctypes.sizeof(pt.PtEnv)] == [60, 64, 80, 96]
    assert out['packed'] == [640, 480, 77, 1.0, 3.0, 2, 1000.0]
    # the facade's set returns unknown names (lib/pathtracer/index.js set)
    assert out['unknown'] == ['notAUniform'] and out['maxbounce'] == 3
    assert "its code: unknown function 'myShape'" in out['refused']
    assert out['c2'] == {'prims': 35, 'materials': 35, 'faces': 1, 'flops': out['c2']['flops']}


def test_standard_scenes_compile_for_gfx950(tmp_path):
    lib = pt.load_library()
    buf = ctypes.create_string_buffer(4096)
    for name in scenes.STANDARD:
        made = scenes.emit(name)
        rc = lib.pt_compile_scene(made['scene']['hip'].encode(), str(tmp_path).encode(), buf, 4096)
        assert rc == 0, lib.pt_last_error(None)
        assert os.path.getsize(buf.value.decode()) > 10000


def test_compile_error_is_reported_with_the_log(tmp_path):
    lib = pt.load_library()
    buf = ctypes.create_string_buffer(4096)
    bad = b'__device__ float sceneDist(const float* U, V3 p) { return undefined_thing; }'
    rc = lib.pt_compile_scene(bad, str(tmp_path).encode(), buf, 4096)
    assert rc == 3
    assert b'undefined_thing' in lib.pt_last_error(None)


@pytest.mark.skipif(have_gpu(), reason='checks the no-GPU error path')
def test_create_without_a_gpu_fails_with_a_message():
    with pytest.raises(pt.PtError) as e:
        pt.Engine()
    assert e.value.code == 2 and 'device' in str(e.value)


def test_null_arguments_are_refused():
    lib = pt.load_library()
    assert lib.pt_compile_scene(None, None, None, 0) == 1
    assert lib.pt_create(None, None) == 1
    assert b'NULL' in lib.pt_last_error(None)


def test_emitter_output_is_consistent():
    made = scenes.emit('room-c2')
    sc = made['scene']
    prims = [o for o in sc['program'] if o['op'] == 'prim']
    # each primitive appears in both functions of the HIP source, and its parameters fit the block
    assert sc['hip'].count('xfIn(') == 2 * len(prims)
    assert max(o['p'] for o in prims) < len(sc['params'])
    assert len(sc['faces']) == 1 and abs(sc['faces'][0]['le'][1] - 5000) < 1e-6
    assert sc['faces'][0]['n'][2] < -0.999          # the panel faces down
    mats = sc['materials']
    assert sum(m['metallic'] == 1 for m in mats[2:34]) in range(4, 20)
    assert all(0.05 <= m['roughness'] <= 0.9 for m in mats[2:34])


def test_header_constants_match_the_binding():
    import re
    text = open(os.path.join(ROOT, 'include', 'pt_types.h')).read()
    define = lambda name: int(re.search(r'#define %s (\d+)' % name, text).group(1))
    assert define('PT_STRAGGLER_ROUNDS') == pt.STRAGGLER_ROUNDS
    assert pt.STAT_NAMES.index('flushed') == define('PT_S_FLUSHED')
    assert pt.STAT_NAMES.index('guide_records') == define('PT_S_GRECORDS')
    assert pt.STAT_NAMES.index('shadowcut') == define('PT_S_SHADOWCUT')
    assert pt.STAT_NAMES.index('guide_train') == define('PT_S_GTRAIN')
    assert len(pt.STAT_NAMES) <= define('PT_S_WORDS')
