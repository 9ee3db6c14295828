"""Absorption inside a material (materials.js attenuationDistance, tracebasics.js absorbAlong): light travelling
through it loses base^(d / attenuationDistance) per channel over the distance d, and its faces pass every channel
alike. Pinned by the closed form of a smooth slab seen head-on in a uniform sky: L (R + (1 - R)^2 (R tau^2 + tau) /
(1 - R^2 tau^2)), R = ((n - 1) / (n + 1))^2 the face's reflectance head-on and tau = base^(thickness / distance) the
inside's transmittance, summed over the passes back and forth -- on the oracle here, on the GPU in tests/test_gpu.py."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, 'oracle'))
import oracle  # noqa: E402
from shadytracer_amd import pt, scenes  # noqa: E402


def slab_closed_form(attenuation_mm, thickness_mm=20.0, base=(0.5, 0.7, 0.9), n=1.5):
    R = ((n - 1) / (n + 1)) ** 2
    tau = np.asarray(base) ** (thickness_mm / attenuation_mm)
    return R + (1 - R) ** 2 * (R * tau ** 2 + tau) / (1 - R ** 2 * tau ** 2)


def centre(acc, mom, k=8):
    h, w = acc.shape[:2]
    sl = (slice(h // 2 - k // 2, h // 2 + k // 2), slice(w // 2 - k // 2, w // 2 + k // 2))
    return acc[sl][..., :3].reshape(-1, 3), pt.standard_error(mom)[sl].reshape(-1)


def test_an_absorbing_slab_transmits_its_closed_form():
    for att in (10.0, 40.0):
        made = scenes.emit('absorb-slab', attenuation=att)
        u = pt.uniforms_for(made)
        acc, mom, st = oracle.render(made['scene'], u, 600)
        c, se = centre(acc, mom)
        got = c.mean(0) / 1000
        want = slab_closed_form(att)
        # the luminance's standard error over the centre's pixels, as a share of it, bounds each channel's here (every
        # path's weight is a power of one colour)
        rel = np.sqrt((se ** 2).sum()) / len(se) / (c @ [0.2126, 0.7152, 0.0722]).mean()
        assert st['nonfinite'] == 0
        assert np.all(np.abs(got / want - 1) < 4 * rel + 2e-4), (att, got, want, rel)
    # without absorption the faces tint at every crossing (the reference's rule before attenuationDistance): the slab is
    # not this closed form, and the absorbing one is darker for a shorter distance
    assert slab_closed_form(10.0)[0] < slab_closed_form(40.0)[0]
