"""Builds the native parts in-tree: libpt.so (host library + embedded device sources), the Node N-API addon,
and the code-object cache of the standard scenes (hipRTC for gfx950; no GPU needed).

    python -m shadytracer_amd.build [--no-scenes]
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, 'csrc')
DEV = os.path.join(CSRC, 'device')
INCLUDE = os.path.join(ROOT, 'include')
LIB = os.path.join(PKG, 'libpt.so')
ADDON = os.path.join(PKG, 'js', 'pt_addon.node')

# device sources in translation-unit order; the emitted scene goes between PRE and POST
DEVICE_PRE = [os.path.join(INCLUDE, 'pt_types.h'), os.path.join(DEV, 'pt_dev.h'), os.path.join(DEV, 'pt_prelude.h'),
              os.path.join(DEV, 'pt_wgsl_rt.h'), os.path.join(DEV, 'pt_wgsl_lib.h')]
DEVICE_POST = [os.path.join(DEV, f) for f in ('pt_bsdf.h', 'pt_sky.h', 'pt_kernels.h', 'pt_guide.h', 'pt_wavefront.h', 'pt_optrace.h',
                                            'pt_develop.h')]

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')


def _device_text(paths):
    out = []
    for p in paths:
        out.append('// ---- %s ----' % os.path.basename(p))
        for line in open(p).read().split('\n'):
            s = line.strip()
            if s.startswith('#include') or s == '#pragma once':
                continue
            out.append(line)
    text = '\n'.join(out)
    if ')PTSRC"' in text:
        raise RuntimeError('device source contains the raw-string delimiter')
    return text


def write_embedded():
    inc = os.path.join(CSRC, 'pt_device_src.inc')
    body = ('// generated by shadytracer_amd/build.py from csrc/device/*.h and include/pt_types.h: do not edit\n'
            'static const char* kPtDevicePre = R"PTSRC(%s)PTSRC";\n'
            'static const char* kPtDevicePost = R"PTSRC(%s)PTSRC";\n') % (_device_text(DEVICE_PRE), _device_text(DEVICE_POST))
    old = open(inc).read() if os.path.exists(inc) else None
    if old != body:
        open(inc, 'w').write(body)
    return inc


def run(cmd, **kw):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, **kw)
    if r.returncode != 0:
        raise RuntimeError('command failed (%d): %s\n%s' % (r.returncode, ' '.join(cmd), r.stdout))
    return r.stdout


def build_lib():
    write_embedded()
    tmp = LIB + '.tmp'
    run([HIPCC, '--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared', '-Wall', '-I', INCLUDE,
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-mllvm', '-amdgpu-atomic-optimizer-strategy=DPP', '-o', tmp,
         os.path.join(CSRC, 'pt_api.cpp'), os.path.join(CSRC, 'pt_optics.hip'), '-lhiprtc'])
    os.replace(tmp, LIB)
    return LIB


def build_addon():
    node_inc = '/usr/include/node'
    if not os.path.exists(os.path.join(node_inc, 'node_api.h')):
        return None
    tmp = ADDON + '.tmp'
    run(['g++', '-O2', '-std=c++17', '-fPIC', '-shared', '-Wall', '-I', node_inc, '-I', INCLUDE, '-o', tmp,
         os.path.join(PKG, 'js', 'addon.cc'), '-L', PKG, '-lpt', "-Wl,-rpath,$ORIGIN/.."])
    os.replace(tmp, ADDON)
    return ADDON


def build_scenes():
    """Emits the standard scenes with the JS emitter and compiles them into the code-object cache."""
    from . import scenes
    return scenes.precompile_all()


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    build_lib()
    print('built', LIB)
    a = build_addon()
    print('built', a if a else '(no node headers: addon skipped)')
    if '--no-scenes' not in argv:
        for path in build_scenes():
            print('scene', path)


if __name__ == '__main__':
    main()
