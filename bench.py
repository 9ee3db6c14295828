"""Benchmark: Mpath-segments/s of the wavefront SDF path tracer on one MI355X.

Workload (the plain method at scale): the 5 x 4 x 3 m room with 32 seeded primitives, 8 smooth-union pairs, GGX
materials and a 5000 cd/m^2 ceiling panel; 1920 x 1080, depth 8, NEE on, FP32 path state, FP64 accumulators. A
step is one sample: one camera path per pixel to its end (<= 9 segments), with its shadow rays, resolved into the
FP64 accumulators. `value` counts path segments (a camera ray or bounce marched to its next event and shaded) over
the timed steps; shadow rays are reported beside it.

    python bench.py [--gpus 1] [--steps 32] [--warmup 4] [--no-cpu] [--no-guide]
    python bench.py --guide [--scene caustic-c3] [--width 640 --height 360] [--train 300] [--steps 128] [--reference N]
    python bench.py --optics [--steps 8] [--no-cpu]

Prints one JSON line. After the timed steps, the default run also measures the metric's second half on the
glass-caustic room at 640 x 360 (`guided_vs_plain`: seconds to 1 % relative RMSE, guided against plain, each run
checked against an FP64 plain reference of 400 samples). With --guide it measures the learned guiding against the plain tracer instead: each runs the
same scene sample by sample; per-pixel variances come from the samples' own spread (both estimators are unbiased, the
GPU tests show their means agree), and the seconds each needs to reach 1 % relative RMSE -- RMSE over the image over
its mean, at the image's pixels and over 4 x 4 bins -- follow as (training time) + var / (0.01 mean)^2 x (seconds a
sample). The roofline is the dominant kernel's (sphere tracing: FP32 VALU), from HIP events on the
engine's stream; `cpu_baseline` times the FP64 C oracle on this host's cores on a bounded slice of the same workload.
With --optics it measures the optics bench (photons a second over the walk and numbers kernels, every number's error
against the FP64 oracle on the same photons).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3      # MI355X FP32 vector peak (published)
PEAK_BF16_TFLOPS = 2500.0     # MI355X dense BF16 matrix-core peak (published, no sparsity)
PEAK_HBM_GBS = 8000.0
# The guide network 80 -> 64 -> 64 -> 64 -> 41: a row's forward pass is 2 (80*64 + 64*64 + 64*64 + 64*41) FLOP; its
# backward pass is the weight gradients of every layer (the same count) and the input gradients of the last three
# (2 (64*64 + 64*64 + 64*41)); a training row runs both.
GUIDE_FWD_FLOPS = 2 * (80 * 64 + 64 * 64 + 64 * 64 + 64 * 41)
GUIDE_TRAIN_FLOPS = 2 * GUIDE_FWD_FLOPS + 2 * (64 * 64 + 64 * 64 + 64 * 41)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(made):
    """The oracle (FP64, OpenMP) on a bounded slice: the same room at 480 x 270 (1/16 of the pixels), 4 samples."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle
    from shadytracer_amd import pt, scenes
    small = scenes.emit('room-c2', width=480, height=270)
    u = pt.uniforms_for(small)
    threads = int(os.environ.get('OMP_NUM_THREADS', '0') or 0) or os.cpu_count() or 1
    t0 = time.time()
    _, _, st = oracle.render(small['scene'], u, 4, first=0, threads=threads)
    dt = time.time() - t0
    return {'value': st['segments'] / dt / 1e6, 'unit': 'Mpath-segments/s', 'cores': threads, 'kind': 'port',
            'reference_cpu_baseline': probe_reference_adapter(),
            'sample': 'the same room at 480x270 (1/16 of the 1920x1080 pixels), 4 samples, FP64 C oracle with OpenMP; '
                      '%d segments in %.1f s' % (st['segments'], dt)}


def probe_reference_adapter():
    """Whether this machine could run the reference's own WGSL on a software WebGPU adapter: Node >= 20 (its
    sources use ES2020+ syntax and top-level await), the Dawn `webgpu` package, and a CPU Vulkan driver (lavapipe or
    SwiftShader). Returns 'available' or 'unavailable (<the first missing item>)'."""
    import glob
    import shutil
    import subprocess
    node = shutil.which('node')
    if not node:
        return 'unavailable (no node)'
    try:
        v = subprocess.run([node, '--version'], capture_output=True, text=True, timeout=20).stdout.strip()
        major = int(v.lstrip('v').split('.')[0])
    except Exception:
        return 'unavailable (node --version failed)'
    if major < 20:
        return 'unavailable (node %s; the reference needs >= 20)' % v
    r = subprocess.run([node, '-e', "require.resolve('webgpu')"], capture_output=True, timeout=20)
    if r.returncode:
        return 'unavailable (no webgpu/Dawn package for node %s)' % v
    icds = glob.glob('/usr/share/vulkan/icd.d/*') + glob.glob('/etc/vulkan/icd.d/*')
    if not any(('lvp' in f or 'swiftshader' in f.lower()) for f in icds):
        return 'unavailable (no CPU Vulkan driver: lavapipe or SwiftShader)'
    return 'available'


def traffic_from_profiles():
    """HBM bytes per march launch, from the committed rocprofv3 --pmc summary (profiles/), if there is one."""
    p = os.path.join(ROOT, 'profiles', 'march_room-c2_1080p_traffic.json')
    if not os.path.exists(p):
        return None
    try:
        return json.load(open(p))['bytes_per_launch']
    except Exception:
        return None


def calibration_copy(mib):
    """A device-to-device copy of known size through the HIP runtime (its copy kernel): under rocprofv3 --pmc its
    FETCH_SIZE / WRITE_SIZE give the counters' scale, bytes counted per byte moved."""
    import ctypes
    hip = ctypes.CDLL('libamdhip64.so')
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)):
        raise RuntimeError('hipMalloc failed')
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)          # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def guide_roofline(e, samples=4):
    """The guide network's kernels timed one by one with HIP events on the engine's stream (pt_time_kernels) over
    `samples` guided samples, against the BF16 matrix-core peak. Inference runs every 64-row tile of the round's hits
    (rows where guiding does not apply are zero rows, computed all the same); training runs its batch's rows forward and
    backward."""
    e.reset_stats()
    km = e.time_kernels(samples)
    st = e.stats()
    out = {'peak': PEAK_BF16_TFLOPS, 'unit': 'TFLOP/s', 'samples': samples,
           'flop_per_row_forward': GUIDE_FWD_FLOPS, 'flop_per_row_training': GUIDE_TRAIN_FLOPS}
    rows = st['guide_tiles'] * 64
    if km['guide_infer'] > 0 and rows:
        f = rows * GUIDE_FWD_FLOPS
        out['pt_guide_infer'] = {'ms_per_sample': km['guide_infer'] / samples, 'rows': rows, 'useful_rows': st['guide_infer'],
                                 'achieved': f / (km['guide_infer'] * 1e-3) / 1e12,
                                 'frac': f / (km['guide_infer'] * 1e-3) / 1e12 / PEAK_BF16_TFLOPS}
    if km['guide_grad'] > 0 and st['guide_train']:
        f = st['guide_train'] * GUIDE_TRAIN_FLOPS
        out['pt_guide_grad'] = {'ms_per_sample': km['guide_grad'] / samples, 'rows': st['guide_train'],
                                'achieved': f / (km['guide_grad'] * 1e-3) / 1e12,
                                'frac': f / (km['guide_grad'] * 1e-3) / 1e12 / PEAK_BF16_TFLOPS}
    out['adam_collect_commit_ms_per_sample'] = km['guide_adam'] / samples
    out['kernel_ms_per_sample'] = {k: v / samples for k, v in km.items()}
    return out


def guide_vs_plain(scene, width, height, train=300, steps=64, reference=0, sets=(), nee=None, direct=False, frac=0.9):
    """Guided against plain on one scene: variance, time a sample, seconds to 1 % relative RMSE.

    Each estimator runs the same scene sample by sample (its own launch choices, tuned first) and its per-pixel
    variance comes from the samples' own spread. The seconds to 1 % relative RMSE -- the RMSE over the image's pixels,
    over the image mean, and again over 4 x 4 bins -- follow as (training time) + var / (0.01 mean)^2 x (seconds a
    sample): Monte Carlo error falls as 1 / sqrt(samples), so the time to an RMSE is the variance times the time a
    sample over that RMSE squared.

    reference > 0: first a long plain run of that many samples, accumulated in FP64 (the engine's accumulators), as
    the reference image. Each estimator's image at the end of its run is then compared with it: its measured RMSE
    (the reference's own error, var_ref / n_ref a pixel, taken out) beside the RMSE its variance predicts, and its
    bias over 16 x 16 bins in standard errors -- the check that the extrapolation to 1 % is of an unbiased estimator."""
    import numpy as np
    from shadytracer_amd import pt, scenes
    made = scenes.emit(scene, width=width, height=height)
    extra = {}
    if made['settings'].get('sppm'):
        # the configuration's photons a pixel, at this image size
        extra['sppmPhotons'] = max(1 << 14, int(made['settings']['sppmPhotons'] * width * height / (1280 * 720)))
    e = pt.Engine(direct=direct)
    e.set_scene(made['scene'])
    e.set_uniforms(pt.uniforms_for(made, **extra))
    e.guide_config(pt.scene_box(made['scene']), batch=1 << 16 if width * height >= 1 << 20 else 1 << 14, steps=4, warmup=32)
    H4, W4 = height // 4, width // 4

    def uniforms(guided):
        u = pt.uniforms_for(made, **extra)
        if nee is not None:
            u.nee = nee
        if guided:
            u.guide, u.guideFrac = 1, frac
        for item in sets:
            # name=value, a dotted name reaching into a structure (haze.sigmaS=0)
            name, value = item.split('=', 1)
            obj, parts = u, name.split('.')
            for part in parts[:-1]:
                obj = getattr(obj, part)
            setattr(obj, parts[-1], type(getattr(obj, parts[-1]))(float(value)))
        if not guided:
            u.guide = 0
        return u

    def run(guided, n, train_samples=0):
        e.set_uniforms(uniforms(guided))
        e.clear_accums()
        t_train = 0.0
        if train_samples:
            t0 = time.perf_counter()
            e.trace(train_samples)
            e.sync()
            t_train = time.perf_counter() - t0
        # each estimator timed with its own launch choices (the tuner's samples train the guided one further; untimed)
        tune = e.autotune(samples=2)
        e.clear_accums()
        s1 = np.zeros((height, width)); s2 = np.zeros_like(s1)
        b1 = np.zeros((H4, W4)); b2 = np.zeros_like(b1)
        prev, t, smeans = None, 0.0, []
        top = np.zeros(0)          # the largest pixel values met, for the share of the ten largest
        for k in range(1, n + 1):
            t0 = time.perf_counter()
            e.trace(1)
            e.sync()
            t += time.perf_counter() - t0
            Y = pt.luminance(e.read_accum()[0])
            x = Y if prev is None else k * Y - (k - 1) * prev
            prev = Y
            smeans.append(float(x.mean()))
            s1 += x; s2 += x * x
            top = np.sort(np.concatenate([top, np.partition(x.ravel(), -16)[-16:]]))[-16:]
            xb = x[:H4 * 4, :W4 * 4].reshape(H4, 4, W4, 4).mean(axis=(1, 3))
            b1 += xb; b2 += xb * xb
        mean = s1 / n
        var = (s2 - n * mean * mean) / (n - 1)
        # how much of the squared deviation the ten largest pixel values carry (fireflies: a share near 1)
        dev_top = (top[-10:] - float(mean.mean())) ** 2
        top10 = float(dev_top.sum() / max(var.sum() * (n - 1), 1e-300))
        bmean = b1 / n
        bvar = (b2 - n * bmean * bmean) / (n - 1)
        ts = t / n
        mu = float(mean.mean())
        to1 = lambda v: float(v) / (0.01 * mu) ** 2 * ts + t_train
        log('bench: %s %.3f ms a sample, mean %.4g, mean var %.4g, 4x4-bin var %.4g' % (
            'guided' if guided else 'plain', ts * 1e3, mu, var.mean(), bvar.mean()))
        return {'samples': n, 'ms_per_sample': ts * 1e3, 'train_s': t_train, 'mean': mu, 'var': float(var.mean()),
                'tuning': {k: tune[k] for k in ('blocksPerCu', 'refillMarch', 'refillOcclude')},
                'var_4x4': float(bvar.mean()), 'seconds_to_1pct': to1(var.mean()), 'seconds_to_1pct_4x4': to1(bvar.mean()),
                'sample_mean_sd': float(np.std(smeans, ddof=1)), 'largest_value': float(top[-1]), 'top10_share': top10,
                'stats': {k: v for k, v in e.stats().items() if k in ('segments', 'guided', 'guide_records', 'guide_steps')}}, mean, var

    ref = None
    if reference:
        t0 = time.perf_counter()
        rsum, rmean, rvar = run(False, reference)
        ref = {'samples': reference, 'seconds': time.perf_counter() - t0, 'accumulation': 'f64', 'mean': rsum['mean'],
               'rel_rmse': math.sqrt(rsum['var'] / reference) / rsum['mean']}

    def against_reference(summary, mean, var):
        # measured error of the run's image against the reference, the reference's own share taken out
        n = summary['samples']
        mse = float(np.mean((mean - rmean) ** 2)) - float(np.mean(rvar)) / reference
        pred = float(np.mean(var)) / n
        B = 16
        hb, wb = height // B, width // B
        blk = lambda img: img[:hb * B, :wb * B].reshape(hb, B, wb, B).mean(axis=(1, 3))
        # a bin's standard error from its pixels' variances (pixels of a sample are not independent: an upper bound
        # on how independent they are, so |z| here is if anything large)
        se = np.sqrt(blk(var) / (B * B) / n + blk(rvar) / (B * B) / reference)
        z = (blk(mean) - blk(rmean)) / np.maximum(se, 1e-300)
        summary['vs_reference'] = {'rel_rmse_measured': math.sqrt(max(mse, 0.0)) / ref['mean'],
                                   'rel_rmse_predicted': math.sqrt(pred) / ref['mean'],
                                   'bins_16x16_z_rms': float(np.sqrt(np.mean(z * z))),
                                   'bins_16x16_z_max': float(np.max(np.abs(z))),
                                   'image_mean_ratio': summary['mean'] / ref['mean']}

    plain, mp, vp = run(False, steps)
    guided, mg, vg = run(True, steps, train_samples=train)
    groof = guide_roofline(e)
    log('bench: guide kernels %s' % json.dumps({k: v for k, v in groof.items() if k.startswith('pt_')}))
    if ref:
        against_reference(plain, mp, vp)
        against_reference(guided, mg, vg)
    e.close()
    # the means' agreement from the spread of each run's per-sample image means: pixels of one sample are not
    # independent (they share the sample's seed-level draws), so per-pixel variances understate the mean's error
    z = abs(guided['mean'] - plain['mean']) / math.sqrt((plain['sample_mean_sd'] ** 2 + guided['sample_mean_sd'] ** 2) / steps)
    ratio = vp / np.maximum(vg, 1e-30)
    return {'metric': 'seconds to 1% rel. RMSE, guided vs plain', 'unit': 's', 'higher_is_better': False,
            'uniform_overrides': list(sets), 'value': guided['seconds_to_1pct'],
            'speedup_1pct': plain['seconds_to_1pct'] / guided['seconds_to_1pct'],
            'speedup_1pct_4x4': plain['seconds_to_1pct_4x4'] / guided['seconds_to_1pct_4x4'],
            'variance_reduction': plain['var'] / guided['var'], 'median_pixel_variance_ratio': float(np.median(ratio)),
            'mean_agreement_sigma': z, 'plain': plain, 'guided': guided, 'reference': ref, 'guide_roofline': groof,
            'n_gpus': 1, 'dtype': 'f32 (network bf16, accumulation f64)',
            'config': {'workload': '%s %dx%d%s' % (scene, width, height, '' if nee is None else ', nee %d' % nee),
                       'train_samples': train, 'samples': steps, 'guide_frac_cap': frac,
                       'photons': extra.get('sppmPhotons', 0)}}


def bench_guide(a):
    """bench.py --guide: guided against plain on one scene, the whole record as one JSON line."""
    out = guide_vs_plain(a.scene, a.width, a.height, train=a.train, steps=a.steps, reference=a.reference, sets=a.set,
                         nee=a.nee, direct=a.direct, frac=a.guide_frac)
    print(json.dumps(out), flush=True)


def optics_traffic_from_profiles():
    """HBM bytes per walk launch, from the committed rocprofv3 --pmc summary of the optics bench, if there is one."""
    p = os.path.join(ROOT, 'profiles', 'optics_walk_64scenes_traffic.json')
    if not os.path.exists(p):
        return None
    try:
        return json.load(open(p))['bytes_per_launch']
    except Exception:
        return None


def bench_optics(a):
    """The optics bench: 64 seeded luminaires (LED, plano-convex or aspheric lens, half of them in a paraboloid
    reflector; 2^18..2^24 photons each) in one persistent launch, FP64 far fields and lux maps, numbers on the GPU.
    A step is one such launch. The FP64 CPU oracle traces the same photons (same PCG4D draws) for the error of every
    number and the CPU baseline."""
    import numpy as np
    from shadytracer_amd import optics, pt
    sc = optics.bench_scenes(64)
    total = sum(s.photons for s in sc)
    log('bench: optics, 64 luminaires, %d photons, warmup %d, steps %d' % (total, a.warmup, a.steps))
    for _ in range(max(1, a.warmup)):
        pt.optics_trace(sc)
    walk, numbers, wall = [], [], []
    res = None
    for _ in range(a.steps):
        t0 = time.perf_counter()
        res, _, _, (mw, mn) = pt.optics_trace(sc)
        wall.append(time.perf_counter() - t0)
        walk.append(mw)
        numbers.append(mn)
    kernel_ms = float(np.mean(walk)) + float(np.mean(numbers))
    flops = sum(r['flops'] for r in res)
    events = sum(r['events'] for r in res)
    achieved = flops / (np.mean(walk) * 1e-3) / 1e12
    roof = {'kernel': 'pt_optics_walk', 'bound': 'valu', 'achieved': round(achieved, 3), 'peak': PEAK_FP32_TFLOPS,
            'unit': 'TFLOP/s', 'frac': round(achieved / PEAK_FP32_TFLOPS, 4), 'traffic': optics_traffic_from_profiles(),
            'flops_per_launch': flops, 'avg_launch_ms': float(np.mean(walk)), 'flops_per_surface_test': flops / events,
            'surface_tests_per_photon': events / total, 'numbers_kernel_ms': float(np.mean(numbers))}
    out = {'metric': 'Mphotons/s (optics bench, 64 luminaires, one persistent launch)', 'value': total / kernel_ms / 1e3,
           'unit': 'Mphotons/s', 'n_gpus': 1, 'steps': a.steps, 'warmup': a.warmup, 'ms_per_step': kernel_ms,
           'ms_per_step_with_uploads': float(np.mean(wall)) * 1e3, 'higher_is_better': True, 'scaling': 'weak',
           'vs_baseline': None, 'dtype': 'f32 walk, f64 tallies and numbers',
           'data': 'synthetic: 64 seeded luminaires (shadytracer_amd/optics.py bench_scenes, seed 20261203)',
           'config': {'workload': 'optics bench: 64 luminaires, Lambertian 1x1 mm LED 100 lm, plano-convex or aspheric '
                                  'lens, optional paraboloid reflector, 181x360 candela bins + 256x256 lux at 1 m, '
                                  '2^18..2^24 photons each', 'scenes': 64, 'photons': total},
           'roofline': roof, 'cpu_baseline': None}
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, 'oracle'))
        import oracle
        threads = int(os.environ.get('OMP_NUM_THREADS', '0') or 0) or os.cpu_count() or 1
        t0 = time.perf_counter()
        ores, _, _ = oracle.optics_trace(sc, threads=threads)
        dt = time.perf_counter() - t0
        out['cpu_baseline'] = {'value': total / dt / 1e6, 'unit': 'Mphotons/s', 'cores': threads, 'kind': 'port',
                               'sample': 'the whole workload (the same 64 scenes and photons), FP64 C oracle with OpenMP; '
                                         '%.1f s' % dt}
        # every number against the oracle's on the same photons (the GPU walks in FP32: a few photons change bins)
        err = {}
        for key in ('beam', 'field'):
            err[key + '_deg'] = max(abs(g[key][0] - o[key][0]) for g, o in zip(res, ores))
        err['cct_K'] = max(abs(g['cct'] - o['cct']) for g, o in zip(res, ores))
        err['duv'] = max(abs(g['duv'] - o['duv']) for g, o in zip(res, ores))
        err['Imax_rel'] = max(abs(g['Imax'] / o['Imax'] - 1) for g, o in zip(res, ores))
        err['efficiency'] = max(abs(g['efficiency'] - o['efficiency']) for g, o in zip(res, ores))
        err['luxCentre_rel'] = max(abs(g['luxCentre'] / o['luxCentre'] - 1) for g, o in zip(res, ores))
        out['max_error_vs_fp64_oracle'] = err
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--guide', action='store_true', help='guided against plain: seconds to 1% relative RMSE')
    ap.add_argument('--optics', action='store_true', help='the optics bench: 64 luminaires in one persistent launch')
    ap.add_argument('--scene', default='caustic-c3')
    ap.add_argument('--train', type=int, default=300)
    ap.add_argument('--nee', type=int, default=None, help='with --guide: light sampling on (1) or off (0)')
    ap.add_argument('--reference', type=int, default=0, help='with --guide: samples of an FP64 long-run plain reference '
                    'the runs are measured against (0: none)')
    ap.add_argument('--guide-frac', type=float, default=0.9, help='with --guide: the guided fraction\'s cap (guideFrac)')
    ap.add_argument('--no-guide', action='store_true', help='leave out the guided-vs-plain line of the default run')
    ap.add_argument('--set', action='append', default=[], help='with --guide: a uniform, name=value (haze.sigmaS=0)')
    ap.add_argument('--calibrate-mib', type=int, default=0, help='a device copy of this many MiB first (counter scale)')
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=32)
    ap.add_argument('--warmup', type=int, default=4)
    ap.add_argument('--width', type=int, default=1920)
    ap.add_argument('--height', type=int, default=1080)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-autotune', action='store_true', help='keep the library\'s default launch choices')
    ap.add_argument('--tuning', default=None, help='blocksPerCu,refillMarch,refillOcclude: set, not tuned')
    ap.add_argument('--stragglers', type=int, default=None, help='rounds beyond maxbounce + 1 (default: the library\'s)')
    ap.add_argument('--relax', type=float, default=0.0, help='over-relaxed sphere tracing factor (0: the plain march)')
    ap.add_argument('--direct', action='store_true', help='launch each sample kernel by kernel, no graph replay '
                    '(for profilers that follow kernel launches only)')
    ap.add_argument('--no-fold', action='store_true', help='keep the kernels that read the parameter block (no pt_fold_params)')
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
        print('bench.py: this build runs on one GPU; --gpus %d is not supported (samples would shard across GPUs with an '
              'all-reduce of the FP64 accumulators, not built here)' % a.gpus, file=sys.stderr)
        sys.exit(2)

    if a.guide:
        if a.width == 1920 and a.height == 1080 and a.scene == 'caustic-c3':
            a.width, a.height = 640, 360
        if a.steps == 32:
            a.steps = 128
        bench_guide(a)
        return
    if a.calibrate_mib:
        calibration_copy(a.calibrate_mib)
    if a.optics:
        if a.steps == 32:
            a.steps = 8
        bench_optics(a)
        return
    from shadytracer_amd import pt, scenes
    made = scenes.emit('room-c2', width=a.width, height=a.height)
    log('bench: room-c2 %dx%d, warmup %d, steps %d' % (a.width, a.height, a.warmup, a.steps))
    e = pt.Engine(straggler_rounds=a.stragglers, direct=a.direct)
    e.set_scene(made['scene'])
    u = pt.uniforms_for(made, relax=a.relax)
    e.set_uniforms(u)
    # the scene does not move during the run: its kernels with the parameter block as constants (untimed compile)
    if not a.no_fold:
        e.fold_params()
    # the persistent kernels' launch choices for this scene and image, by measurement (before the warm-up; untimed)
    if a.tuning:
        b, rm, ro = (int(v) for v in a.tuning.split(','))
        e.set_tuning(b, rm, ro)
        tuning = dict(e.tuning(), msPerSample=None, msBefore=None, candidates=0)
    else:
        tuning = None if a.no_autotune else e.autotune(samples=2)
    if tuning:
        log('bench: autotune %s' % tuning)
    e.clear_accums()
    e.trace(a.warmup)
    e.sync()
    log('bench: warm')
    e.reset_stats()
    e.sync()
    t0 = time.perf_counter()
    e.trace(a.steps)
    e.sync()
    dt = time.perf_counter() - t0
    st = e.stats()
    log('bench: timed %.3f s, %s' % (dt, st))

    # the dominant kernel, timed launch by launch with HIP events on the engine's stream
    e.reset_stats()
    ksamples = 4
    ms = e.time_kernels(ksamples)
    kst = e.stats()
    rounds = u.maxbounce + 1 + e.straggler_rounds
    dom = max(('march', 'occlude'), key=lambda k: ms[k])
    # FP32 operations the kernel's sphere tracing performed, counted by the kernel itself from the emitter's per-node
    # counts (nodes evaluated after bounding-sphere culling, the bound tests, each step's own arithmetic)
    flops_k = kst['flops'] if dom == 'march' else kst['shadow_flops']
    steps_k = kst['steps'] if dom == 'march' else kst['shadow_steps']
    flops_per_step = flops_k / max(steps_k, 1)
    launches = ksamples * rounds
    avg_launch_ms = ms[dom] / launches
    flops_per_launch = flops_k / launches
    achieved = flops_per_launch / (avg_launch_ms * 1e-3) / 1e12
    traffic = traffic_from_profiles()
    roof = {'kernel': 'pt_' + dom, 'bound': 'valu', 'achieved': round(achieved, 3), 'peak': PEAK_FP32_TFLOPS, 'unit': 'TFLOP/s',
            'frac': round(achieved / PEAK_FP32_TFLOPS, 4), 'traffic': traffic,
            'flops_per_launch': flops_per_launch, 'avg_launch_ms': avg_launch_ms,
            'flops_per_step': flops_per_step, 'flops_per_step_unculled': made['scene']['flops']['dist'] + 30,
            'kernel_ms_per_sample': {k: v / ksamples for k, v in ms.items()}}
    # the developed picture of the same accumulation: pt_develop's wall clock a call (the meter's two kernels, the
    # develop kernel, the sync and the 8-bit picture copied to the host), against the bytes its kernel must move
    # (32 B of FP64 accumulation read and 4 B written a pixel)
    e.develop(tone='agx', auto=True)
    t0 = time.perf_counter()
    for _ in range(10):
        e.develop(tone='agx', auto=True, speed=0.5)
    develop = {'workload': '%dx%d AgX, metered each call' % (a.width, a.height),
               'ms_per_call': (time.perf_counter() - t0) * 1e3 / 10, 'kernel_bytes': 36 * a.width * a.height,
               'includes': 'meter, develop kernel, sync, %d-byte copy to the host' % (4 * a.width * a.height)}
    log('bench: develop %.3f ms a call at %dx%d' % (develop['ms_per_call'], a.width, a.height))
    cpu = None if a.no_cpu else cpu_baseline(made)
    segs = st['segments']
    out = {
        'metric': 'Mpath-segments/s', 'value': segs / dt / 1e6, 'unit': 'Mpath-segments/s', 'n_gpus': 1, 'steps': a.steps,
        'warmup': a.warmup, 'ms_per_step': dt * 1e3 / a.steps, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32', 'data': 'synthetic: seeded procedural room (room-c2), emitted by the project\'s JS emitter',
        'config': {'workload': 'room-c2: 5x4x3 m room, 32 SDF primitives, 8 smooth unions, GGX, 0.5 m panel 5000 cd/m^2, '
                               '%dx%d, depth 8, NEE on, 1 sample per step' % (a.width, a.height),
                   'width': a.width, 'height': a.height, 'maxbounce': u.maxbounce, 'accumulation': 'f64',
                   'straggler_rounds': e.straggler_rounds, 'relax': u.relax, 'params_folded': not a.no_fold,
                   'tuning': None if not tuning else {k: tuning[k] for k in ('blocksPerCu', 'refillMarch', 'refillOcclude',
                                                                            'msPerSample', 'msBefore', 'candidates')}},
        'shadow_segments_per_s': st['shadow_segments'] / dt, 'march_steps_per_s': st['steps'] / dt,
        'mean_steps_per_segment': st['steps'] / max(segs, 1),
        'roofline': roof, 'cpu_baseline': cpu, 'develop': develop,
    }
    e.close()
    if not a.no_guide:
        # the metric's second half: guided against plain on the glass-caustic room, at a size the run affords
        log('bench: guided vs plain, caustic-c3 640x360')
        g = guide_vs_plain('caustic-c3', 640, 360, train=200, steps=48, reference=400, direct=a.direct)
        out['guided_vs_plain'] = {
            'metric': 'seconds to 1% rel. RMSE, guided vs plain', 'target': 'guided <= 1/3 of plain',
            'workload': 'caustic-c3 640x360 (glass-caustic room: haze box, dispersive glass sphere, 100 klx sun through '
                        'a 0.3 m aperture; photon map at every camera vertex past the first, light tracing at the first; '
                        '%d photons a sample)' % g['config']['photons'],
            'plain_seconds_to_1pct': g['plain']['seconds_to_1pct'], 'guided_seconds_to_1pct': g['guided']['seconds_to_1pct'],
            'guided_over_plain': g['guided']['seconds_to_1pct'] / g['plain']['seconds_to_1pct'],
            'guided_over_plain_4x4': g['guided']['seconds_to_1pct_4x4'] / g['plain']['seconds_to_1pct_4x4'],
            'plain_ms_per_sample': g['plain']['ms_per_sample'], 'guided_ms_per_sample': g['guided']['ms_per_sample'],
            'train_s': g['guided']['train_s'], 'samples': g['config']['samples'],
            'median_pixel_variance_ratio': g['median_pixel_variance_ratio'],
            'reference': g['reference'],
            'plain_vs_reference': g['plain'].get('vs_reference'), 'guided_vs_reference': g['guided'].get('vs_reference'),
            'guide_kernels': {k: v for k, v in g['guide_roofline'].items() if k.startswith('pt_')},
            'method': 'seconds to 1 %% = training + var / (0.01 mean)^2 x seconds a sample, var the per-pixel variance '
                      'of one sample from the run\'s own spread; each run checked against an FP64 plain reference of '
                      '%d samples' % 400}
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
