"""The guide network on the GPU (BF16 MFMA forward and backward, Adam, guided sampling) against the FP64 oracle
(oracle/guide.py) and against the plain tracer. Through the C ABI; every test names its case first."""
import math
import sys

import numpy as np
import pytest

import guide
import oracle
from shadytracer_amd import pt, scenes

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(11)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def setup(engine, name='room-c2', w=32, h=18, **cfg):
    made = scenes.emit(name, width=w, height=h)
    engine.set_scene(made['scene'])
    engine.set_uniforms(pt.uniforms_for(made))
    box = pt.scene_box(made['scene'])
    engine.guide_config(box, **cfg)
    return made, box


def raw_rows(n, box):
    lo, hi = np.array(box[0]), np.array(box[1])
    pos = lo + (hi - lo) * RNG.random((n, 3))
    return np.hstack([pos, unit(RNG.normal(size=(n, 3))), unit(RNG.normal(size=(n, 3))), RNG.random((n, 1))]), lo, hi


def bf16(x):
    """Round to BF16 (nearest even), as the engine stores weights and activations."""
    a = np.asarray(x, np.float32).copy()
    u = a.view(np.uint32)
    u += 0x7FFF + ((u >> 16) & 1)
    u &= 0xFFFF0000
    return u.view(np.float32).astype(np.float64)


def test_forward_matches_the_oracle(engine):
    say('case: guide forward, 1000 random vertices, random weights (scale 0.2) vs FP64 oracle')
    made, box = setup(engine)
    theta = RNG.normal(0, 0.2, guide.nparams())
    engine.guide_load(theta)
    raw, lo, hi = raw_rows(1000, box)
    got = engine.guide_forward(raw).astype(np.float64)
    X = guide.encode((raw[:, :3] - lo) / (hi - lo), raw[:, 3:6], raw[:, 6:9], raw[:, 9])
    want, _ = guide.forward(bf16(theta), X)
    # BF16 weights and activations, FP32 accumulation: a few parts in a thousand of the outputs' scale
    scale = np.abs(want).max(axis=1, keepdims=True) + 1e-3
    err = np.abs(got - want) / scale
    assert np.median(err) < 5e-3 and err.max() < 4e-2, (np.median(err), err.max())
    # the mixture's density at random directions agrees too
    w = unit(RNG.normal(size=(1000, 3)))
    qg, qo = guide.mixture_pdf(got, w), guide.mixture_pdf(want, w)
    assert np.median(np.abs(qg / qo - 1)) < 2e-2


def samples_rows(n, box, target=None):
    raw, lo, hi = raw_rows(n, box)
    pos01 = (raw[:, :3] - lo) / (hi - lo)
    if target is None:
        w = unit(RNG.normal(size=(n, 3)))
    else:
        w = guide.sample_vmf(np.tile(target, (n, 1)), np.full(n, 20.0), RNG.random(n), RNG.random(n))
    F = RNG.random(n) + 0.1 if target is None else np.ones(n)
    # the BSDF's density at w: Lambertian about the normal, or uniform over the sphere for the lobe target
    b = np.maximum(np.einsum('ij,ij->i', w, raw[:, 6:9]), 0) / np.pi if target is None else np.full(n, 1 / (4 * np.pi))
    rows = np.zeros((n, 16))
    rows[:, 0:3], rows[:, 3] = pos01, raw[:, 9]
    rows[:, 4:7], rows[:, 7] = raw[:, 3:6], F
    rows[:, 8:11], rows[:, 12:15], rows[:, 15] = raw[:, 6:9], w, b
    X = guide.encode(pos01, raw[:, 3:6], raw[:, 6:9], raw[:, 9])
    return rows, X, w, F, b


def test_gradient_and_adam_match_the_oracle(engine):
    say('case: guide backward on 512 samples vs FP64 oracle; Adam first step in closed form')
    made, box = setup(engine, lr=1e-2)
    theta = RNG.normal(0, 0.2, guide.nparams()).astype(np.float32)
    engine.guide_load(theta)
    rows, X, w, F, b = samples_rows(512, box)
    grad, loss = engine.guide_train(rows)
    want = guide.backward(bf16(theta), X, w, F, b)
    cos = grad @ want / (np.linalg.norm(grad) * np.linalg.norm(want))
    rel = np.linalg.norm(grad - want) / np.linalg.norm(want)
    assert cos > 0.995 and rel < 0.08, (cos, rel)
    o, _ = guide.forward(bf16(theta), X)
    assert abs(loss / guide.loss(o, w, F, b) - 1) < 2e-2, (loss, guide.loss(o, w, F, b))
    # Adam's first step: every weight moves by lr g / (|g| + eps), g the summed gradient the step applied
    th1 = engine.guide_weights()
    step = 1e-2 * grad.astype(np.float64) / (np.abs(grad) + 1e-8)
    assert np.allclose(theta - th1, step, rtol=1e-4, atol=1e-7), np.abs(theta - th1 - step).max()
    assert engine.stats()['guide_steps'] >= 1


def test_training_on_a_lobe_lowers_the_loss(engine):
    say('case: 60 Adam steps on directions drawn from one vMF lobe (kappa 20), the BSDF uniform')
    made, box = setup(engine, lr=1e-2)
    target = unit(np.array([0.6, 0.0, 0.8]))
    losses = []
    for k in range(60):
        rows, X, w, F, b = samples_rows(1024, box, target)
        _, loss = engine.guide_train(rows)
        losses.append(loss)
    # the combined density's floor: the lobe's own entropy, reached as c -> 1
    floor = -(guide.log_c(np.array(20.0)) - 1.0)
    say('loss %.3f -> %.3f (floor %.3f)' % (losses[0], losses[-1], floor))
    assert losses[-1] < losses[0] - 1.0 and np.mean(losses[-5:]) < floor + 0.8, (losses[0], losses[-5:], floor)


def per_sample(engine, samples):
    out, prev = [], None
    for k in range(1, samples + 1):
        engine.trace(1)
        Y = pt.luminance(engine.read_accum()[0])
        out.append(Y if prev is None else k * Y - (k - 1) * prev)
        prev = Y
    return np.array(out)


def test_guided_room_agrees_with_the_plain_tracer(engine):
    say('case: room-c2 96x54, guided (after 24 training samples) vs plain, 48 samples each')
    made = scenes.emit('room-c2', width=96, height=54)
    engine.set_scene(made['scene'])
    box = pt.scene_box(made['scene'])
    engine.set_uniforms(pt.uniforms_for(made))
    engine.guide_config(box, batch=4096, steps=2, warmup=8)
    ug = pt.uniforms_for(made)
    ug.guide, ug.guideFrac = 1, 0.5
    engine.set_uniforms(ug)
    engine.clear_accums()
    engine.trace(24)                  # training (guiding switches on after 8 steps)
    engine.clear_accums()
    engine.reset_stats()
    g = per_sample(engine, 48)
    st = engine.stats()
    assert st['guided'] > 0 and st['guide_records'] > 0 and st['nonfinite'] == 0
    up = pt.uniforms_for(made)
    engine.set_uniforms(up)
    engine.clear_accums()
    p = per_sample(engine, 48)
    regions = []
    for by in range(3):
        for bx in range(4):
            m = np.zeros((54, 96), bool)
            m[by * 18:(by + 1) * 18, bx * 24:(bx + 1) * 24] = True
            regions.append(m)
    worst = 0.0
    for m in regions + [np.ones((54, 96), bool)]:
        gm, pm = g[:, m].mean(axis=1), p[:, m].mean(axis=1)
        se = math.sqrt(gm.var(ddof=1) / len(gm) + pm.var(ddof=1) / len(pm))
        z = abs(gm.mean() - pm.mean()) / se
        worst = max(worst, z)
        assert z < 4.5, (gm.mean(), pm.mean(), se)
    say('guided vs plain: worst region at %.2f sigma; per-sample spread guided %.4g plain %.4g'
        % (worst, g.mean(axis=(1, 2)).std(), p.mean(axis=(1, 2)).std()))


def test_trained_guiding_gains_over_bsdf_sampling(engine):
    # What the trained network gains over the method without it: the room lit by its 0.5 m ceiling panel with light
    # sampling off, so that the plain tracer finds the panel only by BSDF sampling. The mixture learns where the panel
    # is; the guided image must have a fraction of the plain one's per-pixel variance, at the same mean.
    say('case: room-c2 96x54, light sampling off: guided (after 48 training samples) vs plain, 48 samples each')
    made = scenes.emit('room-c2', width=96, height=54)
    engine.set_scene(made['scene'])
    engine.guide_config(pt.scene_box(made['scene']), batch=4096, steps=4, warmup=16)
    ug = pt.uniforms_for(made)
    ug.nee, ug.guide, ug.guideFrac = 0, 1, 0.9
    engine.set_uniforms(ug)
    engine.clear_accums()
    engine.trace(48)
    engine.clear_accums()
    engine.reset_stats()
    g = per_sample(engine, 48)
    assert engine.stats()['guided'] > 0
    up = pt.uniforms_for(made)
    up.nee = 0
    engine.set_uniforms(up)
    engine.clear_accums()
    p = per_sample(engine, 48)
    vg, vp = g.var(axis=0, ddof=1), p.var(axis=0, ddof=1)
    gm, pm = g.mean(axis=(1, 2)), p.mean(axis=(1, 2))
    z = abs(gm.mean() - pm.mean()) / math.sqrt(gm.var(ddof=1) / len(gm) + pm.var(ddof=1) / len(pm))
    lit = vp > 0
    ratio = vp.mean() / vg.mean()
    per = vp[lit] / np.maximum(vg[lit], 1e-30)
    median, gains = float(np.median(per)), float(np.mean(per > 1.0))
    say('plain/guided variance: mean %.2f, median pixel %.2f, pixels gaining %.3f; means %.4g vs %.4g (%.2f sigma)'
        % (ratio, median, gains, gm.mean(), pm.mean(), z))
    assert z < 4.5, (gm.mean(), pm.mean())
    # the median pixel's ratio is the steady figure (~25), and nearly every lit pixel gains. The ratio of the mean
    # variances is not asserted: with 48 samples it rides on a handful of pixels' rare events on either side (plain's
    # BSDF hits on the panel; a guided draw weighted up to 1 / (1 - guideFrac) times a BSDF draw) and has measured
    # 0.9 to 2.4 across runs (the training draws its batches from samples appended in the order the atomics give).
    assert median > 5.0 and gains > 0.8, (ratio, median, gains)


def furnace_samples(engine, u, n):
    """Each sample's mean over the floor rows (the running FP64 mean read after every sample)."""
    engine.set_uniforms(u)
    engine.clear_accums()
    out, prev = [], None
    for k in range(1, n + 1):
        engine.trace(1)
        Y = pt.luminance(engine.read_accum()[0])[:20].mean()
        out.append(Y if prev is None else k * Y - (k - 1) * prev)
        prev = Y
    return np.array(out)


@pytest.mark.parametrize('case,opts', [('metal, rough 0.3', {'roughness': 0.3, 'metallic': 1, 'specular': 0.5}),
                                       ('plastic, rough 0.2', {'roughness': 0.2, 'specular': 0.5})])
def test_guided_furnace_keeps_its_energy(engine, case, opts):
    say('case: guided white furnace, %s, 64x64: trained 32 samples, then 64 guided against 64 plain' % case)
    made = scenes.emit('furnace-plane', width=64, height=64, **opts)
    engine.set_scene(made['scene'])
    engine.set_uniforms(pt.uniforms_for(made))
    engine.guide_config(pt.scene_box(made['scene']), batch=2048, steps=2, warmup=4)
    plain = furnace_samples(engine, pt.uniforms_for(made), 64)
    ug = pt.uniforms_for(made)
    ug.guide, ug.guideFrac = 1, 0.9
    engine.set_uniforms(ug)
    engine.clear_accums()
    engine.trace(32)
    engine.reset_stats()
    guided = furnace_samples(engine, ug, 64)
    st = engine.stats()
    se = math.sqrt(plain.var(ddof=1) / len(plain) + guided.var(ddof=1) / len(guided))
    say('plain %.3f guided %.3f, se %.3f (guided vertices %d)' % (plain.mean(), guided.mean(), se, st['guided']))
    assert st['guided'] > 0 and st['nonfinite'] == 0
    assert abs(guided.mean() - plain.mean()) < 4 * se + 1e-3 * plain.mean(), (plain.mean(), guided.mean(), se)


def test_guided_rough_glass_furnace_agrees_with_the_plain_tracer(engine):
    # rough glass (0.3) in the white furnace, guided: BSDF-drawn refractions and guide-drawn directions below the surface
    # are weighed by the mixture of the two densities with the transmission lobe evaluated (Walter et al. 2007), so the
    # guided picture keeps the plain tracer's energy -- the single-scatter 0.94 of the sphere, not more, not less
    say('case: guided white furnace, rough glass 0.3 sphere, 64x64: trained 32 samples, then 128 guided against 128 plain')
    made = scenes.emit('furnace-sphere', width=64, height=64, roughness=0.3, transmission=1)
    engine.set_scene(made['scene'])
    engine.set_uniforms(pt.uniforms_for(made))
    engine.guide_config(pt.scene_box(made['scene']), batch=2048, steps=2, warmup=4)
    u = pt.uniforms_for(made)
    engine.set_uniforms(u)
    engine.clear_accums()
    engine.trace(128)
    plain, pm = engine.read_accum()
    ug = pt.uniforms_for(made)
    ug.guide, ug.guideFrac = 1, 0.5
    engine.set_uniforms(ug)
    engine.clear_accums()
    engine.trace(32)
    engine.clear_accums()
    engine.reset_stats()
    engine.trace(128)
    guided, gm = engine.read_accum()
    st = engine.stats()
    Lp, Lg = pt.luminance(plain), pt.luminance(guided)
    se = math.sqrt(np.nansum(pt.standard_error(pm) ** 2) + np.nansum(pt.standard_error(gm) ** 2)) / Lp.size
    say('plain %.3f guided %.3f, se %.3f (guided vertices %d)' % (Lp.mean(), Lg.mean(), se, st['guided']))
    assert st['guided'] > 0 and st['nonfinite'] == 0
    assert abs(Lg.mean() - Lp.mean()) < 4 * se + 1e-3 * Lp.mean(), (Lp.mean(), Lg.mean(), se)


def fixed_theta(lobes, c):
    """All weights 0; the head's biases alone set the mixture at every vertex: lobes [(logit, mean, log kappa)], c."""
    th = np.zeros(guide.nparams(), np.float32)
    b4 = th[-41:]
    for j in range(8):
        lg, mu, lk = lobes[j % len(lobes)]
        b4[j] = lg if j < len(lobes) else -30.0
        b4[8 + 3 * j:11 + 3 * j] = mu
        b4[32 + j] = lk
    b4[40] = math.log(c / (1 - c))
    return th


@pytest.mark.parametrize('name,lobes,c', [('up, kappa 2', [(0.0, [0, 0, 1], math.log(2.0))], 0.5),
                                          ('tilted, kappa 8', [(0.0, [0.7, 0, 0.7], math.log(8.0))], 0.5),
                                          ('two lobes', [(0.0, [0, 0, 1], math.log(20.0)), (0.5, [-0.6, 0.3, 0.7], math.log(4.0))], 0.7)])
def test_fixed_mixture_furnace(engine, name, lobes, c):
    # a proposal that is known and does not learn (learning rate ~0, no warm-up): the estimator's unbiasedness alone
    say('case: white furnace, metal rough 0.3, guided by a fixed mixture (%s, c %.1f)' % (name, c))
    made = scenes.emit('furnace-plane', width=64, height=64, roughness=0.3, metallic=1, specular=0.5)
    engine.set_scene(made['scene'])
    engine.set_uniforms(pt.uniforms_for(made))
    engine.guide_config(pt.scene_box(made['scene']), batch=2048, steps=1, warmup=0, lr=1e-30)
    engine.guide_load(fixed_theta(lobes, c))
    o = engine.guide_forward(np.array([[0, 0, 0, 0, 0, 1, 0, 0, 1, 0.3]])).astype(np.float64)
    assert abs(guide.frac(o)[0] - c) < 1e-2
    plain = furnace_samples(engine, pt.uniforms_for(made), 48)
    ug = pt.uniforms_for(made)
    ug.guide, ug.guideFrac = 1, 0.9
    guided = furnace_samples(engine, ug, 48)
    se = math.sqrt(plain.var(ddof=1) / len(plain) + guided.var(ddof=1) / len(guided))
    say('plain %.3f guided %.3f, se %.3f' % (plain.mean(), guided.mean(), se))
    assert abs(guided.mean() - plain.mean()) < 4 * se + 1e-3 * plain.mean(), (plain.mean(), guided.mean(), se)
