"""The builtin modifier groups (builder/modifiers.js: twist, bend, repeat, mirror, circular array, round, shell) as the
oracle evaluates them from the emitter's program, against what fixes each independently: the untwisted box's closed
form, the twisted surface made by turning the box's own surface, the true distance to it (the value a lower bound on
it: the Lipschitz factor), the repeat's period and its last copy, the mirror's and the array's symmetries, and the
round's and shell's closed forms. Lengths in metres (the document's millimetres converted); the scene's parameters are
FP32 (the parameter block), so closed forms in FP64 agree to about 1e-7 m."""
import math

import numpy as np
import pytest

import oracle
from shadytracer_amd import scenes

MM = 1e-3


def at(name, pts, **opts):
    made = scenes.emit('modifiers', **opts) if name is None else scenes.emit('modifiers', only=name, **opts)
    d, n, prim = oracle.scene_at(made['scene'], pts)
    return d, n, prim


def fold(floor, d):
    """The scene's closest-surface fold over the floor (first) and one other value: the value of least magnitude."""
    return np.where(np.abs(d) < np.abs(floor), d, floor)


def rounded_box(q, b, r):
    """Quilez's rounded box: q relative to the centre, b the half sizes, r the rounding."""
    w = np.abs(q) - (b - r)
    return np.linalg.norm(np.maximum(w, 0.0), axis=-1) + np.minimum(w.max(axis=-1), 0.0) - r


def test_twist_of_zero_is_the_box():
    rng = np.random.default_rng(1)
    c = np.array([-135, 50, 40]) * MM
    p = c + rng.uniform(-0.05, 0.05, (500, 3))
    d, _, _ = at('twist', p, rate=0.0)
    want = fold(p[:, 2], rounded_box(p - c, np.array([14, 14, 40]) * MM, 1 * MM))      # and the floor
    # (inside the bar near its foot the floor and the bar's bottom face are equally near: the fold keeps either value)
    assert np.allclose(np.abs(d), np.abs(want), atol=2e-7)


def twisted_surface(rate, n=4000, seed=2):
    """Points of the twisted bar's surface: the box's own surface points (in the group's frame), each turned back by
    -rate z about the group's z axis, then placed at the group's origin."""
    rng = np.random.default_rng(seed)
    b, r = np.array([14, 14, 40]) * MM, 1 * MM
    # points on the rounded box's surface: random directions from the centre, found by bisection along them
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    lo, hi = np.zeros(n), np.full(n, 0.1)
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        inside = rounded_box(dirs * mid[:, None], b, r) < 0
        lo, hi = np.where(inside, mid, lo), np.where(inside, hi, mid)
    q = dirs * lo[:, None] + np.array([0, 0, 40]) * MM          # the bar sits 40 mm up in the group
    a = -rate * q[:, 2]
    g = np.stack([np.cos(a) * q[:, 0] - np.sin(a) * q[:, 1], np.sin(a) * q[:, 0] + np.cos(a) * q[:, 1], q[:, 2]], 1)
    return g + np.array([-135, 50, 0]) * MM


def test_twisted_bar_surface_and_safe_distance():
    rate = 0.039269908 / MM          # radians a metre (the document's a millimetre)
    S = twisted_surface(rate)
    d, _, _ = at('twist', S)
    assert np.abs(d).max() < 2e-7
    # the value is a lower bound on the distance to that surface: true distance from points near the bar by brute force
    rng = np.random.default_rng(3)
    c = np.array([-135, 50, 40]) * MM
    p = c + rng.uniform(-0.04, 0.04, (300, 3)) * np.array([1, 1, 1.6])
    p = p[p[:, 2] > 0.005]
    d, _, _ = at('twist', p)
    dense = twisted_surface(rate, n=60000, seed=4)
    true = np.array([np.min(np.linalg.norm(dense - x, axis=1)) for x in p])
    out = (d > 0) & (np.abs(d - p[:, 2]) > 1e-9)          # outside the bar, and the bar's value (not the floor's)
    # the surface cloud's spacing (~0.5 mm) makes the brute force a little long
    assert np.all(d[out] <= true[out] + 6e-4)
    # and not slack beyond the twist's Lipschitz factor both ways: the twist and its inverse each stretch distances by
    # at most 1 / k, so the value k d_box(twist(p)) is at least k^2 the true distance
    k = 1 / math.sqrt(1 + (rate * 22 * MM) ** 2)
    assert np.all(d[out] >= k * k * (true[out] - 6e-4) - 1e-6)


def test_repeat_is_periodic_to_its_last_copy():
    # balls 7 mm across... radius 7 mm, 20 mm apart, one copy each way along x and y at (45, 50, 7) mm
    c = np.array([45, 50, 7]) * MM
    rng = np.random.default_rng(5)
    p = c + rng.uniform(-0.009, 0.009, (200, 3))
    d0, _, _ = at('repeat', p)
    for off in ((20, 0), (-20, 0), (0, 20), (20, -20)):
        d1, _, _ = at('repeat', p + np.array([off[0], off[1], 0]) * MM)
        assert np.allclose(d0, d1, atol=2e-7), off
    # every copy's centre is inside a ball; one spacing past the last copy is outside them all
    ctrs = np.array([[45 + 20 * i, 50 + 20 * j, 7] for i in (-1, 0, 1) for j in (-1, 0, 1)]) * MM
    assert np.allclose(at('repeat', ctrs)[0], -7 * MM, atol=2e-7)
    # 40 mm along x and 13 mm up from the centre: the nearest copy is the last (20 mm along), not one at 40
    assert at('repeat', np.array([[45 + 40, 50, 20]]) * MM)[0][0] == pytest.approx(math.hypot(20, 13) * MM - 7 * MM, abs=2e-7)


def test_mirror_is_symmetric():
    rng = np.random.default_rng(6)
    p = np.array([135, 50, 25]) * MM + rng.uniform(-0.04, 0.04, (300, 3))
    p[:, 2] = np.abs(p[:, 2]) + 0.002
    q = p.copy()
    q[:, 0] = 2 * 135 * MM - p[:, 0]
    assert np.allclose(at('mirror', p)[0], at('mirror', q)[0], atol=2e-7)


def test_circular_array_turns_into_itself():
    rng = np.random.default_rng(7)
    o = np.array([45, -50, 0]) * MM
    p = o + rng.uniform(-0.035, 0.035, (300, 3))
    p[:, 2] = np.abs(p[:, 2] - o[2]) + 0.001
    t = 2 * math.pi / 8
    r = p - o
    q = o + np.stack([math.cos(t) * r[:, 0] - math.sin(t) * r[:, 1], math.sin(t) * r[:, 0] + math.cos(t) * r[:, 1], r[:, 2]], 1)
    da, db = at('circular', p)[0], at('circular', q)[0]
    assert np.allclose(da, db, atol=2e-7)
    # the first copy where the child is: its centre inside by the radius
    assert at('circular', np.array([[45 + 24, -50, 7]]) * MM)[0][0] == pytest.approx(-7 * MM, abs=2e-7)


def test_round_and_shell_closed_forms():
    rng = np.random.default_rng(8)
    c = np.array([-135, -50, 22]) * MM
    p = c + rng.uniform(-0.04, 0.04, (300, 3))
    p = p[p[:, 2] > 0.03]
    want = rounded_box(p - c, np.array([17, 17, 17]) * MM, 0.0) - 5 * MM
    assert np.allclose(at('round', p)[0], fold(p[:, 2], want), atol=2e-7)
    # the shell: |(|p - c| - 25 mm)| - 2 mm about the globe (the window's cut aside: points away from it)
    c = np.array([-45, -50, 25]) * MM
    p = c + rng.uniform(-0.03, 0.03, (300, 3))
    p = p[(p[:, 0] > c[0] + 3 * MM) & (p[:, 2] > 0.03)]       # clear of the window cut out of it
    want = np.abs(np.linalg.norm(p - c, axis=1) - 25 * MM) - 2 * MM
    assert np.allclose(at('shell', p)[0], fold(p[:, 2], want), atol=2e-7)


def test_modifier_groups_emit_with_nothing_left_out():
    made = scenes.emit('modifiers')
    assert made['scene']['approximations'] == []
    kinds = [o['kind'] for o in made['scene']['program'] if o['op'] == 'mod']
    assert sorted(set(kinds)) == ['bend', 'circular', 'mirror', 'repeat', 'round', 'shell', 'twist']
