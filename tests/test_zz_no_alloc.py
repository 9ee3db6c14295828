"""The optics bench's entry point that allocates nothing (pt_optics_trace_static, include/pt.h) against the one that
allocates (pt_optics_trace), through the C ABI: the same scenes, the same photons, the same numbers -- twice in a row,
the second call smaller than the first, so that what the first left in the library's array must not leak into it.
The far fields and lux maps are sums of FP64 atomics whose order differs run to run: equal to rounding."""
import ctypes
import sys

import numpy as np
import pytest

from shadytracer_amd import optics, pt

pytestmark = pytest.mark.gpu
FIELDS = ('emitted', 'escaped', 'absorbed', 'efficiency', 'Imax', 'gammaMax', 'x', 'y', 'cct', 'duv', 'luxMax', 'luxCentre')


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def static_trace(scenes):
    lib = pt.load_library()
    f = lib.pt_optics_trace_static
    n = len(scenes)
    arr = (pt.PtOpticsScene * n)(*scenes)
    res = (pt.PtOpticsResult * n)()
    F = np.zeros((n, pt.OPT_NG, pt.OPT_NC))
    L = np.zeros((n, pt.OPT_LUX, pt.OPT_LUX))
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rc = f(0, arr, n, res, dp(F), dp(L))
    if rc:
        raise pt.PtError(rc, lib.pt_optics_error().decode())
    return [pt.result_dict(r) for r in res], F, L


def same(a, b):
    (ra, Fa, La), (rb, Fb, Lb) = a, b
    assert len(ra) == len(rb)
    np.testing.assert_allclose(Fa, Fb, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(La, Lb, rtol=1e-9, atol=1e-9)
    for x, y in zip(ra, rb):
        for k in FIELDS:
            assert x[k] == pytest.approx(y[k], rel=1e-7, abs=1e-12), (k, x[k], y[k])
        for k in ('beam', 'field'):
            assert np.allclose(x[k], y[k], rtol=1e-7, atol=1e-7), (k, x[k], y[k])


def test_static_optics_matches_the_allocating_entry_point_twice():
    big = optics.bench_scenes(8, seed=11, log2_min=18, log2_max=20)
    small = optics.bench_scenes(3, seed=12, log2_min=16, log2_max=17)
    for name, sc in (('8 scenes, 2^18..2^20 photons', big), ('3 scenes, 2^16..2^17 photons', small)):
        say('case: pt_optics_trace_static vs pt_optics_trace, %s' % name)
        res, F, L, _ = pt.optics_trace(sc, far=True, lux=True)
        same(static_trace(sc), (res, F, L))


def test_static_optics_refuses_what_its_array_cannot_hold():
    say('case: pt_optics_trace_static past its limits (host checks only, no launch)')
    with pytest.raises(pt.PtError, match='not in 1..8'):
        static_trace(optics.bench_scenes(9, log2_min=16, log2_max=16))
    with pytest.raises(pt.PtError, match='chunks of 4096'):
        static_trace(optics.bench_scenes(2, log2_min=27, log2_max=27))
