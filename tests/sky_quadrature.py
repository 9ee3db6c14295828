"""The sky's light on a level surface facing up, lux: the oracle's environment (the disc left out) integrated by the
midpoint rule in (cos theta, phi) over the upper hemisphere, weighted by cos theta."""
import numpy as np

LUMA = np.array([0.2126, 0.7152, 0.0722])


def uniforms_with(env):
    from shadytracer_amd import pt
    u = pt.PtUniforms()
    u.env = pt.make_env(env)
    return u


def level_sky_lux(env, n=240):
    import oracle
    u = uniforms_with(dict(env, sunE=0.0))
    mu = (np.arange(n) + 0.5) / n
    ph = 2 * np.pi * (np.arange(2 * n) + 0.5) / (2 * n)
    m, p = np.meshgrid(mu, ph, indexing='ij')
    st = np.sqrt(1 - m ** 2)
    d = np.stack([st * np.cos(p), st * np.sin(p), m], -1).reshape(-1, 3)
    return 2 * np.pi * np.mean((oracle.env(u, d) @ LUMA) * d[:, 2])


def level_sun_lux(env):
    """The sun's light on the same surface: its disc's luminance times the solid angle, times the sine of its height."""
    return float(LUMA @ np.asarray(env['sunL'])) * env['sunOmega'] * max(env['sunDir'][2], 0.0)
