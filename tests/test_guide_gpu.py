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
    rows = np.zeros((n, 16))
    rows[:, 0:3], rows[:, 3] = pos01, raw[:, 9]
    rows[:, 4:7], rows[:, 7] = raw[:, 3:6], F
    rows[:, 8:11], rows[:, 12:15] = raw[:, 6:9], w
    X = guide.encode(pos01, raw[:, 3:6], raw[:, 6:9], raw[:, 9])
    return rows, X, w, F


def test_gradient_and_adam_match_the_oracle(engine):
    say('case: guide backward on 512 samples vs FP64 oracle; Adam first step in closed form')
    made, box = setup(engine, lr=1e-2)
    theta = RNG.normal(0, 0.2, guide.nparams()).astype(np.float32)
    engine.guide_load(theta)
    rows, X, w, F = samples_rows(512, box)
    grad, loss = engine.guide_train(rows)
    want = guide.backward(bf16(theta), X, w, F)
    cos = grad @ want / (np.linalg.norm(grad) * np.linalg.norm(want))
    rel = np.linalg.norm(grad - want) / np.linalg.norm(want)
    assert cos > 0.995 and rel < 0.08, (cos, rel)
    o, _ = guide.forward(bf16(theta), X)
    assert abs(loss / guide.loss(o, w, F) - 1) < 2e-2, (loss, guide.loss(o, w, F))
    # Adam's first step: every weight moves by lr g / (|g| + eps), g the summed gradient the step applied
    th1 = engine.guide_weights()
    step = 1e-2 * grad.astype(np.float64) / (np.abs(grad) + 1e-8)
    assert np.allclose(theta - th1, step, rtol=1e-4, atol=1e-7), np.abs(theta - th1 - step).max()
    assert engine.stats()['guide_steps'] >= 1


def test_training_on_a_lobe_lowers_the_loss(engine):
    say('case: 60 Adam steps on directions drawn from one vMF lobe (kappa 20)')
    made, box = setup(engine, lr=1e-2)
    target = unit(np.array([0.6, 0.0, 0.8]))
    losses = []
    for k in range(60):
        rows, X, w, F = samples_rows(1024, box, target)
        _, loss = engine.guide_train(rows)
        losses.append(loss)
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
