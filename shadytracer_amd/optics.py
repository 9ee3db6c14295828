"""The optics bench's scenes: seeded luminaires for pt_optics_trace (include/pt.h), and an LED's colour from a CCT and
Duv. The reference builds its optics scenes in the builder and runs them through createOptics
(lib/pathtracer/optics/index.js); here a bench luminaire is the engine's PtOpticsScene, made from a seed:

  * an LED: a Lambertian square of 1 x 1 mm emitting 100 lm, its colour a CCT in 2700..6500 K and a Duv in
    -0.004..0.004 off the Planckian locus;
  * a lens on the axis: plano-convex (a spherical back face) or aspheric (a conic back face, k in -1.4..-0.3),
    diameter 12..25 mm, the back face's radius 0.55..1 of the diameter, PMMA (n 1.49) or BK7-like glass (n 1.52), the
    LED 0.55..1.1 of the focal length below it, an edge 1..2 mm thick;
  * with probability 1/2 a paraboloid reflector round it, focus at the LED, from a 2 mm hole to the lens's back
    vertex, mouth 0.5..3 mm wider than the lens, aluminium-like reflectance (0.90, 0.91, 0.88 in X, Y, Z);
  * an absorbing board at z = 0 out to the lens's rim;
  * a photon budget log-uniform in 2^18..2^24.
"""
import math

import numpy as np

from .pt import PtOpticsScene


def planck_uv(T):
    """The Planckian locus in CIE 1960 uv (Krystek 1985's rational fit), 1000..15000 K."""
    t = min(15000.0, max(1000.0, T))
    u = (0.860117757 + 1.54118254e-4 * t + 1.28641212e-7 * t * t) / (1 + 8.42420234e-4 * t + 7.08145163e-7 * t * t)
    v = (0.317398726 + 4.22806245e-5 * t + 4.20481691e-8 * t * t) / (1 - 2.89741816e-5 * t + 1.61456053e-7 * t * t)
    return u, v


def cct_duv_to_xy(T, duv=0.0):
    """xy of the colour at CCT T and Duv (the signed distance off the locus along its normal, positive above)."""
    u, v = planck_uv(T)
    h = max(1e-3, T * 1e-5)
    u0, v0 = planck_uv(T - h)
    u1, v1 = planck_uv(T + h)
    du, dv = u1 - u0, v1 - v0
    l = math.hypot(du, dv) or 1.0
    uu, vv = u + duv * dv / l, v - duv * du / l
    d = 2 * uu - 8 * vv + 4
    return 3 * uu / d, 2 * vv / d


def xyz_of_lumen(x, y):
    """X, Y, Z of one lumen of chromaticity xy."""
    return x / y, 1.0, (1 - x - y) / y


def scene(led_half=0.5e-3, flux=100.0, xy=(0.3457, 0.3585), lens_z=0.01, lens_t=0.004, lens_d=0.02, c1=0.0, k1=0.0,
          c2=-50.0, k2=0.0, ior=1.49, refl_f=0.0, refl_r0=0.002, refl_z1=0.0, rho=(0.90, 0.91, 0.88), board_r=0.0,
          photons=1 << 18, seed=1):
    s = PtOpticsScene()
    s.ledHalf, s.flux = led_half, flux
    X, Y, Z = xyz_of_lumen(*xy)
    s.xyz[0], s.xyz[1], s.xyz[2] = X, Y, Z
    s.lensZ, s.lensT, s.lensD = lens_z, lens_t, lens_d
    s.c1, s.k1, s.c2, s.k2, s.ior = c1, k1, c2, k2, ior
    s.reflF, s.reflR0, s.reflZ1 = refl_f, refl_r0, refl_z1
    for j in range(3):
        s.rho[j] = rho[j]
    s.boardR, s.photons, s.seed = board_r, photons, seed
    return s


def bare_led(photons=1 << 20, xy=(0.3457, 0.3585), seed=7):
    """An LED alone: the lens made of air (n 1: no refraction, no reflection) and wide enough never to be met."""
    return scene(lens_z=0.5, lens_t=0.01, lens_d=1.0e4, c2=0.0, ior=1.0, board_r=0.0, photons=photons, xy=xy, seed=seed)


def sag(c, k, r):
    return c * r * r / (1 + math.sqrt(max(0.0, 1 - (1 + k) * c * c * r * r)))


def bench_scenes(n=64, seed=20261203, log2_min=18, log2_max=24):
    """n seeded luminaires (the module's description)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        D = rng.uniform(0.012, 0.025)
        R2 = -D * rng.uniform(0.55, 1.0)           # the back face bulges to +z
        aspheric = rng.random() < 0.5
        k2 = rng.uniform(-1.4, -0.3) if aspheric else 0.0
        ior = 1.49 if rng.random() < 0.5 else 1.52
        c2 = 1.0 / R2
        edge = rng.uniform(0.001, 0.002)
        T = edge - sag(c2, k2, D / 2)                # centre thickness: the edge plus the back face's sag
        f = -R2 / (ior - 1)                          # thin plano-convex focal length
        z = max(0.002, rng.uniform(0.55, 1.1) * f - T / 2)
        cct, duv = rng.uniform(2700, 6500), rng.uniform(-0.004, 0.004)
        kw = dict(lens_z=z, lens_t=T, lens_d=D, c2=c2, k2=k2, ior=ior, xy=cct_duv_to_xy(cct, duv), board_r=D / 2,
                  photons=int(round(2 ** rng.uniform(log2_min, log2_max))), seed=int(rng.integers(1, 2 ** 31)))
        if rng.random() < 0.5:
            zm = z + T                                # the mouth at the lens's back vertex
            rm = D / 2 + rng.uniform(0.0005, 0.003)
            fr = (-zm + math.sqrt(zm * zm + rm * rm)) / 2   # x^2 + y^2 = 4 f (z + f) through (rm, zm)
            kw.update(refl_f=fr, refl_r0=0.002, refl_z1=zm, board_r=0.002)
        out.append(scene(**kw))
    return out
