"""Scenes from the JavaScript emitter (js/emit.js, js/hipgen.js): a configuration name or a builder document in,
{scene, camera, settings, doc} out. Emitted scenes are memoised in-process; precompile_all() compiles the
standard ones into libpt's code-object cache so a GPU run loads them without a compile."""
import json
import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
JS = os.path.join(PKG, 'js')
STANDARD = ['plane-c1', 'room-c2', 'caustic-c3', 'furnace-plane', 'furnace-sphere', 'furnace-haze', 'sppm-slab', 'studio-floor', 'lens-prism',
            'reflector-condenser', 'turned-parts', 'gallery', 'sunlit', 'sunsky-floor', 'joins', 'optics-floor', 'optics-shell',
            'optics-slab', 'optics-ball', 'modifiers', 'usercode', 'absorb-slab', 'sky-code']
NODE = os.environ.get('NODE', 'node')

_memo = {}


def emit(name, **opts):
    """Runs the emitter for configuration `name` with options; returns the parsed JSON."""
    key = (name, json.dumps(opts, sort_keys=True))
    if key not in _memo:
        args = [NODE, os.path.join(JS, 'emit.js'), name] + ([json.dumps(opts)] if opts else [])
        r = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
        if r.returncode != 0:
            raise RuntimeError('emitter failed for %s: %s' % (name, r.stderr))
        _memo[key] = json.loads(r.stdout)
    return _memo[key]


def emit_doc(path, camera=None, **opts):
    """A builder document file through the emitter. camera: the engine's pinhole (default: the document's own view);
    opts: width, height, approximate (draw without what the engine does not, listed in scene['approximations'])."""
    if camera:
        opts['camera'] = camera
    args = [NODE, os.path.join(JS, 'emit.js'), '--doc', path] + ([json.dumps(opts)] if opts else [])
    r = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    if r.returncode != 0:
        raise RuntimeError('emitter failed for %s: %s' % (path, r.stderr))
    return json.loads(r.stdout)


def precompile_all(names=STANDARD, extra=()):
    """Compiles each scene's kernels into the code-object cache (hipRTC, gfx950, no GPU needed)."""
    import ctypes
    from .pt import load_library
    lib = load_library()
    out = []
    for name in list(names) + list(extra):
        made = emit(name) if isinstance(name, str) else emit(*name)
        buf = ctypes.create_string_buffer(4096)
        rc = lib.pt_compile_scene(made['scene']['hip'].encode(), None, buf, 4096)
        if rc:
            raise RuntimeError('compile of %s failed: %s' % (name, lib.pt_last_error(None).decode()))
        out.append(buf.value.decode())
    return out
