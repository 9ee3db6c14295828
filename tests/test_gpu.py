"""The HIP engine on the GPU, through its C ABI, against the oracle and the closed forms. Every test names its case
before it touches the GPU."""
import json
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from shadytracer_amd import pt, scenes

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


SCENE_OPTS = ('g', 'roughness', 'specular', 'metallic', 'albedo', 'mfp', 'floorRoughness', 'floorSpecular', 'sunDiameter', 'beam', 'native', 'attenuation')


def load(engine, name, **over):
    made = scenes.emit(name, **{k: v for k, v in over.items() if k in SCENE_OPTS + ('width', 'height')})
    engine.set_scene(made['scene'])
    u = pt.uniforms_for(made, **{k: v for k, v in over.items() if k not in SCENE_OPTS})
    engine.set_uniforms(u)
    engine.clear_accums()
    engine.reset_stats()
    return made, u


def test_c1_floor_reads_rho_times_L(engine):
    say('case: plane-c1 128x128 32 spp, floor = 0.8 x 1000')
    load(engine, 'plane-c1', width=128, height=128)
    engine.trace(32)
    acc, mom = engine.read_accum()
    L = pt.luminance(acc)
    assert abs(L[:32].mean() / 800 - 1) < 1e-3, L[:32].mean()
    st = engine.stats()
    assert st['nonfinite'] == 0 and st['segments'] > 128 * 128 * 32


def test_primary_ray_decisions_match_the_oracle(engine):
    say('case: room-c2 480x270, sample 0, pick vs oracle')
    made, u = load(engine, 'room-c2', width=480, height=270)
    gpos, gstate, gmat, gprim = engine.pick_rays()
    opos, ostate, omat, oprim = oracle.pick(made['scene'], u)
    same = (gstate == ostate) & ((gstate != 0) | (gprim == oprim))
    say('agreement', same.mean(), 'mismatches', int((~same).sum()), 'of', len(same))
    assert same.mean() >= 0.9999
    hit = (gstate == 0) & same
    d = np.linalg.norm(gpos[hit, :3] - opos[hit, :3], axis=1)
    assert np.quantile(d, 0.999) < 5e-3


def test_relaxed_march_meets_the_same_surfaces(engine):
    # over-relaxed sphere tracing (relax 1.6) keeps a step only where the free balls overlap, so each ray meets the
    # surface the plain march meets: primary-ray decisions against the oracle's plain march, and the C1 floor
    say('case: room-c2 480x270 relax 1.6, pick vs oracle; plane-c1 128x128 relax 1.6')
    made, u = load(engine, 'room-c2', width=480, height=270, relax=1.6)
    gpos, gstate, gmat, gprim = engine.pick_rays()
    opos, ostate, omat, oprim = oracle.pick(made['scene'], pt.uniforms_for(made))
    same = (gstate == ostate) & ((gstate != 0) | (gprim == oprim))
    say('agreement', same.mean())
    assert same.mean() >= 0.9999
    hit = (gstate == 0) & same
    assert np.quantile(np.linalg.norm(gpos[hit, :3] - opos[hit, :3], axis=1), 0.999) < 5e-3
    load(engine, 'plane-c1', width=128, height=128, relax=1.6)
    engine.trace(32)
    acc, _ = engine.read_accum()
    L = pt.luminance(acc)
    assert abs(L[:32].mean() / 800 - 1) < 1e-3, L[:32].mean()


def test_c2_per_pixel_means_agree_with_the_oracle(engine):
    say('case: room-c2 48x27, GPU 256 spp vs oracle 64 spp')
    made, u = load(engine, 'room-c2', width=48, height=27)
    engine.trace(256)
    ga, gm = engine.read_accum()
    oa, om, _ = oracle.render(made['scene'], u, 64, first=100000)
    gL, oL = pt.luminance(ga), pt.luminance(oa)
    s = np.sqrt(pt.standard_error(gm) ** 2 + pt.standard_error(om) ** 2)
    ok = s > 1e-9
    z = (gL - oL)[ok] / s[ok]
    say('z std', z.std(), 'frac |z|<3', np.mean(np.abs(z) < 3), 'means', gL.mean(), oL.mean())
    assert 0.75 < z.std() < 1.3
    assert np.mean(np.abs(z) < 3) > 0.97
    se = math.sqrt(np.nansum(s ** 2)) / s.size
    assert abs(gL.mean() - oL.mean()) < 4 * se + 1e-3 * oL.mean()


def test_studio_floor_reads_rho_E_over_pi(engine):
    # the builder's studio (envMode 1) puts its illuminance on a level surface whatever its top light: a Lambertian
    # floor of albedo 0.6 under 500 lx reads 0.6 x 500 / pi; a document in millimetres, seen from its orbit view
    say('case: studio-floor 64x48 256 spp, floor = 0.6 x 500 / pi')
    made, u = load(engine, 'studio-floor')
    _, state, _, _ = engine.pick_rays()
    engine.trace(256)
    acc, mom = engine.read_accum()
    L = pt.luminance(acc).ravel()
    # the floor's pixels, less the row beside its far edge (the 20 m disc), whose samples also meet the sky
    hit = (state == 0).reshape(u.height, u.width)
    hit[:-1] &= hit[1:]
    floor = hit.ravel()
    assert floor.sum() > 1500
    assert abs(L[floor].mean() / (0.6 * 500 / math.pi) - 1) < 2e-3, L[floor].mean()


def test_lens_and_prism_primary_rays_match_the_oracle(engine):
    say('case: lens-prism 96x64, sample 0, pick vs oracle')
    made, u = load(engine, 'lens-prism')
    gpos, gstate, gmat, gprim = engine.pick_rays()
    opos, ostate, omat, oprim = oracle.pick(made['scene'], u)
    same = (gstate == ostate) & ((gstate != 0) | (gprim == oprim))
    say('agreement', same.mean(), 'lens', int((oprim[ostate == 0] == 1).sum()), 'prism', int((oprim[ostate == 0] == 2).sum()))
    assert same.mean() >= 0.9995
    assert ((gprim == 1) & (gstate == 0)).sum() > 300 and ((gprim == 2) & (gstate == 0)).sum() > 100
    hit = (gstate == 0) & same
    d = np.linalg.norm(gpos[hit, :3] - opos[hit, :3], axis=1)
    assert np.quantile(d, 0.999) < 1e-4      # metres: the scene is 0.2 m across


def test_lens_and_prism_per_pixel_means_agree_with_the_oracle(engine):
    say('case: lens-prism 48x32, GPU 256 spp vs oracle 64 spp')
    made, u = load(engine, 'lens-prism', width=48, height=32)
    engine.trace(256)
    ga, gm = engine.read_accum()
    oa, om, _ = oracle.render(made['scene'], u, 64, first=100000)
    gL, oL = pt.luminance(ga), pt.luminance(oa)
    s = np.sqrt(pt.standard_error(gm) ** 2 + pt.standard_error(om) ** 2)
    ok = s > 1e-9
    z = (gL - oL)[ok] / s[ok]
    say('z std', z.std(), 'frac |z|<3', np.mean(np.abs(z) < 3), 'means', gL.mean(), oL.mean())
    assert 0.75 < z.std() < 1.3
    assert np.mean(np.abs(z) < 3) > 0.97
    se = math.sqrt(np.nansum(s ** 2)) / s.size
    assert abs(gL.mean() - oL.mean()) < 4 * se + 1e-3 * oL.mean()


def test_reflector_and_condenser_match_the_oracle(engine):
    # the reflector and aspheric lens builtins (Cardano's root for the parabola on the GPU, Newton's in the oracle):
    # primary decisions ray by ray, then per-pixel means of the lit scene (the lamp in the reflector's bowl)
    say('case: reflector-condenser 96x64 pick vs oracle; 48x32 GPU 256 spp vs oracle 64 spp')
    made, u = load(engine, 'reflector-condenser')
    gpos, gstate, _, gprim = engine.pick_rays()
    opos, ostate, _, oprim = oracle.pick(made['scene'], u)
    same = (gstate == ostate) & ((gstate != 0) | (gprim == oprim))
    say('agreement', same.mean(), 'reflector', int((oprim == 2).sum()), 'condenser', int((oprim == 4).sum()))
    assert same.mean() >= 0.9995
    assert ((gprim == 2) & (gstate == 0)).sum() > 300 and ((gprim == 4) & (gstate == 0)).sum() > 150
    hit = (gstate == 0) & same
    assert np.quantile(np.linalg.norm(gpos[hit, :3] - opos[hit, :3], axis=1), 0.999) < 1e-4
    made, u = load(engine, 'reflector-condenser', width=48, height=32)
    engine.trace(256)
    ga, gm = engine.read_accum()
    oa, om, _ = oracle.render(made['scene'], u, 64, first=100000)
    gL, oL = pt.luminance(ga), pt.luminance(oa)
    s = np.sqrt(pt.standard_error(gm) ** 2 + pt.standard_error(om) ** 2)
    ok = s > 1e-9
    z = (gL - oL)[ok] / s[ok]
    say('z std', z.std(), 'frac |z|<3', np.mean(np.abs(z) < 3), 'means', gL.mean(), oL.mean())
    assert 0.75 < z.std() < 1.3
    assert np.mean(np.abs(z) < 3) > 0.97
    se = math.sqrt(np.nansum(s ** 2)) / s.size
    assert abs(gL.mean() - oL.mean()) < 4 * se + 1e-3 * oL.mean()


def test_outlines_match_the_oracle(engine):
    # the revolution, Fresnel lens and extrusion builtins: outlines of segments and fillet arcs, each walked by a
    # function the emitter writes for its node; primary decisions ray by ray, then per-pixel means
    say('case: turned-parts 96x64 pick vs oracle; 48x32 GPU 256 spp vs oracle 64 spp')
    made, u = load(engine, 'turned-parts')
    gpos, gstate, _, gprim = engine.pick_rays()
    opos, ostate, _, oprim = oracle.pick(made['scene'], u)
    same = (gstate == ostate) & ((gstate != 0) | (gprim == oprim))
    say('agreement', same.mean(), 'fresnel', int((oprim == 2).sum()), 'knob', int((oprim == 3).sum()), 'plate', int((oprim == 4).sum()))
    assert same.mean() >= 0.9995
    for k, least in ((2, 120), (3, 250), (4, 180)):
        assert ((gprim == k) & (gstate == 0)).sum() > least
    hit = (gstate == 0) & same
    assert np.quantile(np.linalg.norm(gpos[hit, :3] - opos[hit, :3], axis=1), 0.999) < 1e-4
    made, u = load(engine, 'turned-parts', width=48, height=32)
    engine.trace(256)
    ga, gm = engine.read_accum()
    oa, om, _ = oracle.render(made['scene'], u, 64, first=100000)
    gL, oL = pt.luminance(ga), pt.luminance(oa)
    s = np.sqrt(pt.standard_error(gm) ** 2 + pt.standard_error(om) ** 2)
    ok = s > 1e-9
    z = (gL - oL)[ok] / s[ok]
    say('z std', z.std(), 'frac |z|<3', np.mean(np.abs(z) < 3), 'means', gL.mean(), oL.mean())
    assert 0.75 < z.std() < 1.3
    assert np.mean(np.abs(z) < 3) > 0.97
    se = math.sqrt(np.nansum(s ** 2)) / s.size
    assert abs(gL.mean() - oL.mean()) < 4 * se + 1e-3 * oL.mean()


def test_solid_builtins_match_the_oracle(engine):
    # cone, round cone, octahedron, capped torus, link, tube, dome and pyramid: primary decisions ray by ray, then
    # per-pixel means
    say('case: gallery 160x64 pick vs oracle; 80x32 GPU 256 spp vs oracle 64 spp')
    made, u = load(engine, 'gallery')
    gpos, gstate, _, gprim = engine.pick_rays()
    opos, ostate, _, oprim = oracle.pick(made['scene'], u)
    same = (gstate == ostate) & ((gstate != 0) | (gprim == oprim))
    say('agreement', same.mean(), 'per shape', [int(((oprim == k) & (ostate == 0)).sum()) for k in range(1, 9)])
    assert same.mean() >= 0.9995
    for k in range(1, 9):
        assert ((gprim == k) & (gstate == 0)).sum() > 50, k
    hit = (gstate == 0) & same
    assert np.quantile(np.linalg.norm(gpos[hit, :3] - opos[hit, :3], axis=1), 0.999) < 1e-4
    made, u = load(engine, 'gallery', width=80, height=32)
    engine.trace(256)
    ga, gm = engine.read_accum()
    oa, om, _ = oracle.render(made['scene'], u, 64, first=100000)
    gL, oL = pt.luminance(ga), pt.luminance(oa)
    s = np.sqrt(pt.standard_error(gm) ** 2 + pt.standard_error(om) ** 2)
    ok = s > 1e-9
    z = (gL - oL)[ok] / s[ok]
    say('z std', z.std(), 'frac |z|<3', np.mean(np.abs(z) < 3), 'means', gL.mean(), oL.mean())
    assert 0.75 < z.std() < 1.3
    assert np.mean(np.abs(z) < 3) > 0.97
    se = math.sqrt(np.nansum(s ** 2)) / s.size
    assert abs(gL.mean() - oL.mean()) < 4 * se + 1e-3 * oL.mean()


def test_modifier_groups_match_the_oracle(engine):
    # the builtin modifier groups (twist, bend, repeat, mirror, round, shell, two circular arrays): primary decisions ray
    # by ray against the oracle's evaluation of the same program, then per-pixel means
    say('case: modifiers 160x96 pick vs oracle; 80x48 GPU 256 spp vs oracle 64 spp')
    made, u = load(engine, 'modifiers')
    gpos, gstate, _, gprim = engine.pick_rays()
    opos, ostate, _, oprim = oracle.pick(made['scene'], u)
    same = (gstate == ostate) & ((gstate != 0) | (gprim == oprim))
    say('agreement', same.mean(), 'per primitive', [int(((oprim == k) & (ostate == 0)).sum()) for k in range(1, 10)])
    assert same.mean() >= 0.9995
    for k in range(1, 10):
        assert ((gprim == k) & (gstate == 0)).sum() > 20, k
    hit = (gstate == 0) & same
    assert np.quantile(np.linalg.norm(gpos[hit, :3] - opos[hit, :3], axis=1), 0.999) < 1e-4
    made, u = load(engine, 'modifiers', width=80, height=48)
    engine.trace(256)
    ga, gm = engine.read_accum()
    oa, om, _ = oracle.render(made['scene'], u, 64, first=100000)
    gL, oL = pt.luminance(ga), pt.luminance(oa)
    s = np.sqrt(pt.standard_error(gm) ** 2 + pt.standard_error(om) ** 2)
    ok = s > 1e-9
    z = (gL - oL)[ok] / s[ok]
    say('z std', z.std(), 'frac |z|<3', np.mean(np.abs(z) < 3), 'means', gL.mean(), oL.mean())
    assert 0.75 < z.std() < 1.3
    assert np.mean(np.abs(z) < 3) > 0.97
    se = math.sqrt(np.nansum(s ** 2)) / s.size
    assert abs(gL.mean() - oL.mean()) < 4 * se + 1e-3 * oL.mean()


def test_scene_code_matches_the_oracle(engine):
    # the document's own WGSL, translated: the modifiers as group code, a cone and a Menger sponge of object code,
    # materials whose code reads the world point and the object's frame through the noise library -- primary decisions
    # ray by ray against the oracle's FP64 compile of the same translation, then per-pixel means
    say('case: usercode 160x96 pick vs oracle; 80x48 GPU 256 spp vs oracle 64 spp')
    made, u = load(engine, 'usercode')
    assert len(made['scene']['user']) == 10 and made['scene']['materialCode']
    gpos, gstate, _, gprim = engine.pick_rays()
    opos, ostate, _, oprim = oracle.pick(made['scene'], u)
    same = (gstate == ostate) & ((gstate != 0) | (gprim == oprim))
    say('agreement', same.mean())
    assert same.mean() >= 0.999
    for k in range(0, 12):
        assert ((gprim == k) & (gstate == 0)).sum() > 20, k
    hit = (gstate == 0) & same
    assert np.quantile(np.linalg.norm(gpos[hit, :3] - opos[hit, :3], axis=1), 0.999) < 1e-4
    made, u = load(engine, 'usercode', width=80, height=48)
    engine.trace(256)
    ga, gm = engine.read_accum()
    oa, om, _ = oracle.render(made['scene'], u, 64, first=100000)
    gL, oL = pt.luminance(ga), pt.luminance(oa)
    s = np.sqrt(pt.standard_error(gm) ** 2 + pt.standard_error(om) ** 2)
    ok = s > 1e-9
    z = (gL - oL)[ok] / s[ok]
    say('z std', z.std(), 'frac |z|<3', np.mean(np.abs(z) < 3), 'means', gL.mean(), oL.mean())
    assert 0.75 < z.std() < 1.3
    assert np.mean(np.abs(z) < 3) > 0.97
    se = math.sqrt(np.nansum(s ** 2)) / s.size
    assert abs(gL.mean() - oL.mean()) < 4 * se + 1e-3 * oL.mean()


def test_scene_code_draws_as_the_native_modifiers(engine):
    # the same scene with its modifiers as group code and as the native modifier groups: one picture within the spread
    # of the two renders (the gear's overlapping teeth fold copy by copy in code, so the two differ inside them only)
    say('case: usercode vs native modifiers 80x48, 256 spp each, independent seeds')
    out = []
    for native in (False, True):
        made, u = load(engine, 'usercode', width=80, height=48, native=native, seed=7 + int(native))
        engine.trace(256)
        out.append(engine.read_accum())
    (ua, um), (na, nm) = out
    uL, nL = pt.luminance(ua), pt.luminance(na)
    s = np.sqrt(pt.standard_error(um) ** 2 + pt.standard_error(nm) ** 2)
    ok = s > 1e-9
    z = (uL - nL)[ok] / s[ok]
    say('z std', z.std(), 'means', uL.mean(), nL.mean())
    assert 0.75 < z.std() < 1.3
    assert np.mean(np.abs(z) < 3) > 0.97


def test_shaped_joins_match_the_oracle(engine):
    # rounds and chamfers (joinShape): a ball-box round, a chamfered hole, a concave round fillet and a chamfered union --
    # primary decisions ray by ray against the oracle (whose safe-step allowance differs from the engine's; the surfaces
    # are the same), then per-pixel means
    say('case: joins 160x64 pick vs oracle; 80x32 GPU 256 spp vs oracle 64 spp')
    made, u = load(engine, 'joins')
    gpos, gstate, _, gprim = engine.pick_rays()
    opos, ostate, _, oprim = oracle.pick(made['scene'], u)
    same = (gstate == ostate) & ((gstate != 0) | (gprim == oprim))
    # a hit labelled differently stopped a little elsewhere on the same surface (within eps of it, further along a
    # grazing ray) where a round or chamfer passes from one part to the other: the oracle's own label at the engine's
    # hit point must then be the engine's
    moved = (gstate == ostate) & (gstate == 0) & (gprim != oprim)
    d, _, lab = oracle.scene_at(made['scene'], gpos[moved, :3])
    say('agreement', same.mean(), 'relabelled hits', moved.sum(), 'labels at the engine hit', np.mean(lab == gprim[moved]))
    assert np.all(lab == gprim[moved]) and np.all(np.abs(d) < 4 * u.epsAuto)
    assert moved.mean() < 0.005
    assert (same | moved).mean() >= 0.9995
    hit = (gstate == 0) & same
    assert np.quantile(np.linalg.norm(gpos[hit, :3] - opos[hit, :3], axis=1), 0.999) < 1e-4
    made, u = load(engine, 'joins', width=80, height=32)
    engine.trace(256)
    ga, gm = engine.read_accum()
    oa, om, _ = oracle.render(made['scene'], u, 64, first=100000)
    gL, oL = pt.luminance(ga), pt.luminance(oa)
    s = np.sqrt(pt.standard_error(gm) ** 2 + pt.standard_error(om) ** 2)
    ok = s > 1e-9
    z = (gL - oL)[ok] / s[ok]
    say('z std', z.std(), 'frac |z|<3', np.mean(np.abs(z) < 3), 'means', gL.mean(), oL.mean())
    assert 0.75 < z.std() < 1.3
    assert np.mean(np.abs(z) < 3) > 0.97
    se = math.sqrt(np.nansum(s ** 2)) / s.size
    assert abs(gL.mean() - oL.mean()) < 4 * se + 1e-3 * oL.mean()


def test_sun_and_sky_rows_on_the_engine(engine):
    # the reference's photometry, illuminance-view and cosine-law rows (tests/photometry_rows.py) on the HIP path: each
    # row's ratio to the light its environment puts on the surface, against the fixture's ratio within both bars
    import photometry_rows as P
    for test, table in P.TABLES.items():
        for case in table:
            say('case: %s / %s, 32x32' % (test, case))
            made, u = P.emit(test, case, 32, 32)
            engine.set_scene(made['scene'])
            u.width, u.height = 32, 32
            engine.set_uniforms(u)
            engine.clear_accums()
            engine.trace(1024 if 'NEE off' in case else 128)
            acc, mom = engine.read_accum()
            want = P.answer(test, case, made)
            ours, se = P.ratio_of(acc, mom, want)
            ref, rse = P.reference(test, case)
            say('  ratio %.5f +- %.5f, reference %.5f +- %.5f' % (ours, se, ref, rse))
            if want == 0:
                assert ours == 0
                continue
            assert abs(ours - ref) < 4 * math.sqrt(se ** 2 + rse ** 2) + 1e-4, (test, case, ours, se, ref, rse)


def test_furnace_rows_on_the_engine(engine):
    # the reference's white furnace rows (tests/test_oracle.py FURNACE_ROWS) on the HIP path: a unit sphere of each
    # material in a uniform sky, the frame's coverage from the engine's own black-sphere control
    from test_oracle import FURNACE_ROWS, FURNACE_CONTROL, ref_rows
    rows = ref_rows('furnace')

    def row(opts, over, spp):
        made = scenes.emit('furnace-sphere', width=79, height=106, **opts)
        engine.set_scene(made['scene'])
        engine.set_uniforms(pt.uniforms_for(made, **over))
        engine.clear_accums()
        engine.trace(spp)
        acc, mom = engine.read_accum()
        L = pt.luminance(acc)
        return L.mean(), math.sqrt(np.nansum(pt.standard_error(mom) ** 2)) / L.size

    say('case: furnace control (black sphere), 79x106, 64 spp')
    ctl, _ = row({'albedo': 0, 'roughness': 0.5, 'specular': 0}, {}, 64)
    f_ref, f_ours = 1 - rows[FURNACE_CONTROL]['meanLuma'] / 1000, 1 - ctl / 1000
    assert abs(f_ours / f_ref - 1) < 0.005, (f_ours, f_ref)
    for case, (opts, over) in FURNACE_ROWS.items():
        say('case: furnace %s, 79x106, 256 spp' % case)
        mean, se = row(opts, over, 256)
        mean = 1000 - (1000 - mean) * f_ref / f_ours
        r = rows[case]
        rse = r['meanLuma'] * r['relSE'] / 100
        say('  %.3f +- %.3f, reference %.3f +- %.3f' % (mean, se, r['meanLuma'], rse))
        assert abs(mean - r['meanLuma']) < 4 * math.sqrt(se ** 2 + rse ** 2) + 0.02, (case, mean, se, r['meanLuma'], rse)


@pytest.mark.parametrize('case,opts', [('Lambertian', {'roughness': 0.3}), ('plastic 0.3', {'roughness': 0.3, 'specular': 0.5}),
                                       ('metal 0.3', {'roughness': 0.3, 'metallic': 1}), ('metal 0.6', {'roughness': 0.6, 'metallic': 1})])
def test_bsdf_parity_rows_on_the_engine(engine, case, opts):
    # bsdf-parity (lib/pathtracer/tests.js:1868) on the HIP path: the oblique floor under a 5-degree sun, light sampling
    # on against off -- BSDF sampling alone must find the same light
    say('case: bsdf-parity %s, 32x32, NEE on 256 spp, NEE off 4096 spp' % case)
    made = scenes.emit('sunsky-bench', width=32, height=32, oblique=True, sunDiameter=5, **opts)
    engine.set_scene(made['scene'])
    got = []
    for nee, spp in ((1, 256), (0, 4096)):
        engine.set_uniforms(pt.uniforms_for(made, nee=nee))
        engine.clear_accums()
        engine.trace(spp)
        acc, mom = engine.read_accum()
        L = pt.luminance(acc).ravel()
        got.append((L.mean(), math.sqrt(np.nansum(pt.standard_error(mom).ravel() ** 2)) / L.size))
    (a, sa), (b, sb) = got
    say('  NEE on %.2f +- %.2f, off %.2f +- %.2f' % (a, sa, b, sb))
    assert abs(a - b) < 4 * math.hypot(sa, sb), got


def test_thin_lens_rays_match_the_oracle(engine):
    # the thin lens: the same lens points and focus-plane aims as the oracle's, ray by ray, out of focus on the floor
    from test_oracle import _looking_down
    say('case: plane-c1 64x64 looking down, lens radius 0.05 m focused at 1 m, pick vs oracle')
    made, _ = load(engine, 'plane-c1', width=64, height=64)
    u = _looking_down(made, 0.05, 1.0)
    engine.set_uniforms(u)
    gpos, gstate, _, _ = engine.pick_rays()
    opos, ostate, _, _ = oracle.pick(made['scene'], u)
    assert (gstate == ostate).all() and (gstate == 0).all()
    assert np.abs(gpos[:, :3] - opos[:, :3]).max() < 1e-4
    u0 = _looking_down(made)
    p0, _, _, _ = oracle.pick(made['scene'], u0)
    assert np.linalg.norm(gpos[:, :3] - p0[:, :3], axis=1).max() > 0.04     # the lens moved them


def test_sunsky_floor_reads_rho_E_over_pi(engine):
    # the builder's default environment, the sun & sky (envMode 0): a level Lambertian floor of albedo 0.6 reads
    # 0.6 / pi (E_sun sin 40 + E_sky), the disc's light from the environment itself and the sky's by quadrature
    from sky_quadrature import level_sky_lux, level_sun_lux
    say('case: sunsky-floor 64x48 512 spp, floor = 0.6 (E_sun sin 40 + E_sky) / pi')
    made, u = load(engine, 'sunsky-floor')
    env = made['scene']['env']
    want = 0.6 / math.pi * (level_sun_lux(env) + level_sky_lux(env))
    engine.trace(512)
    acc, mom = engine.read_accum()
    L = pt.luminance(acc).ravel()
    se = math.sqrt((pt.standard_error(mom).ravel() ** 2).sum()) / L.size
    say('floor', L.mean(), 'want', want, 'se', se)
    assert abs(L.mean() - want) < 4 * se + 1e-3 * want
    assert engine.stats()['nonfinite'] == 0


def test_sunlit_gallery_matches_the_oracle(engine):
    # the gallery under the sun & sky: shadows from the disc, the sky's colour; primary decisions, then per-pixel means
    say('case: sunlit 160x64 pick vs oracle; 80x32 GPU 256 spp vs oracle 64 spp')
    made, u = load(engine, 'sunlit')
    gpos, gstate, _, gprim = engine.pick_rays()
    opos, ostate, _, oprim = oracle.pick(made['scene'], u)
    same = (gstate == ostate) & ((gstate != 0) | (gprim == oprim))
    assert same.mean() >= 0.9995
    made, u = load(engine, 'sunlit', width=80, height=32)
    engine.trace(256)
    ga, gm = engine.read_accum()
    oa, om, _ = oracle.render(made['scene'], u, 64, first=100000)
    gL, oL = pt.luminance(ga), pt.luminance(oa)
    s = np.sqrt(pt.standard_error(gm) ** 2 + pt.standard_error(om) ** 2)
    ok = s > 1e-9
    z = (gL - oL)[ok] / s[ok]
    say('z std', z.std(), 'frac |z|<3', np.mean(np.abs(z) < 3), 'means', gL.mean(), oL.mean())
    assert 0.75 < z.std() < 1.3
    assert np.mean(np.abs(z) < 3) > 0.97
    se = math.sqrt(np.nansum(s ** 2)) / s.size
    assert abs(gL.mean() - oL.mean()) < 4 * se + 1e-3 * oL.mean()
    # the colour: the sky's blue and the sun's warmth come through channel by channel
    for c in range(3):
        assert abs(ga[..., c].mean() / oa[..., c].mean() - 1) < 0.015, (c, ga[..., c].mean(), oa[..., c].mean())


# the Lambertian's 1e-3: the march's 400-step cap (the reference's maxiter) drops a path that leaves the plane within
# about 0.7 degrees of grazing before it reaches the sky sphere, ~0.03 % of the light here
def test_every_pixel_closes_every_sample(engine):
    # maxbounce + 1 rounds end the paths that bounce every round; a hit from behind moves a path on without a bounce,
    # so the straggler rounds end those and the flush deposits the rare path still going: every pixel counts every sample
    say('case: room-c2 320x180, 8 samples, every pixel counted 8 times')
    made, u = load(engine, 'room-c2', width=320, height=180)
    engine.trace(8)
    acc, _ = engine.read_accum()
    st = engine.stats()
    assert (acc[..., 3] == 8).all(), (int((acc[..., 3] != 8).sum()), float(acc[..., 3].min()))
    # one straggler round leaves ~0.2 % of this room's paths to the flush (those met two hits from behind)
    assert st['leaving'] > 0 and st['flushed'] <= 5e-3 * 8 * 320 * 180, st


@pytest.mark.parametrize('case,opts,tol', [('Lambertian', {'roughness': 1, 'specular': 0}, 1e-3),
                                           ('plastic, rough 0.5', {'roughness': 0.5, 'specular': 0.5}, 3e-3),
                                           ('metal, rough 0.5', {'roughness': 0.5, 'metallic': 1}, 5e-3)])
def test_white_furnace(engine, case, opts, tol):
    say('case: furnace plane', case)
    load(engine, 'furnace-plane', width=64, height=64, **opts)
    engine.trace(64)
    acc, mom = engine.read_accum()
    L = pt.luminance(acc)[:20]
    assert abs(L.mean() / 1000 - 1) < tol, L.mean()


@pytest.mark.parametrize('g', [0.6, 0.0, -0.5])
def test_haze_box_conserves_energy(engine, g):
    say('case: haze furnace g', g)
    load(engine, 'furnace-haze', width=64, height=64, g=g, maxbounce=250)
    engine.trace(256)
    acc, mom = engine.read_accum()
    L = pt.luminance(acc)
    se = math.sqrt(np.nansum(pt.standard_error(mom) ** 2)) / L.size
    say('haze mean', L.mean(), 'se', se)
    assert abs(L.mean() / 1000 - 1) < max(2e-4, 4 * se / 1000), (L.mean(), se)
    assert engine.stats()['nonfinite'] == 0


def test_compile_failure_keeps_the_previous_scene(engine):
    say('case: bad scene source after plane-c1')
    made, u = load(engine, 'plane-c1', width=32, height=32)
    bad = dict(made['scene'])
    bad['hip'] = made['scene']['hip'].replace('return best;', 'return bestx;')
    with pytest.raises(pt.PtError) as e:
        engine.set_scene(bad)
    assert e.value.restored and 'bestx' in str(e.value)
    engine.trace(4)
    acc, _ = engine.read_accum()
    assert abs(pt.luminance(acc)[:8].mean() / 800 - 1) < 0.02


def test_graph_replay_matches_direct_launches(engine):
    say('case: room-c2 64x36, 3 samples through the graph and launch by launch')
    made, u = load(engine, 'room-c2', width=64, height=36)
    engine.trace(3)
    a1, _ = engine.read_accum()
    engine.clear_accums()
    engine.time_kernels(3)
    a2, _ = engine.read_accum()
    L1, L2 = pt.luminance(a1), pt.luminance(a2)
    assert np.allclose(L1, L2, rtol=1e-4, atol=1e-4 * L1.max())


def test_params_move_geometry_without_a_rebuild(engine):
    say('case: sphere moved by its parameter block')
    made, u = load(engine, 'plane-c1', width=64, height=64)
    p0, s0, _, _ = engine.pick_rays()
    params = np.asarray(made['scene']['params'], np.float32).copy()
    params[2] = 0.25                      # the plane's frame lifted 0.25 m
    engine.set_params(params)
    p1, s1, _, _ = engine.pick_rays()
    hit = (s0 == 0) & (s1 == 0)
    assert np.allclose(p1[hit, 2], 0.25, atol=2e-3) and np.allclose(p0[hit, 2], 0.0, atol=2e-3)


def test_folded_params_render_the_same_scene(engine):
    # pt_fold_params compiles the scene with its parameter block as constants: the same primary decisions as the
    # oracle's, the same per-pixel means as the kernels that read the block, and a later set_params goes back to
    # those (the moved geometry shows at once)
    say('case: room-c2 480x270 folded, pick vs oracle; 48x27 folded vs unfolded means; plane-c1 fold then move')
    made, u = load(engine, 'room-c2', width=480, height=270)
    engine.fold_params()
    gpos, gstate, _, gprim = engine.pick_rays()
    opos, ostate, _, oprim = oracle.pick(made['scene'], u)
    same = (gstate == ostate) & ((gstate != 0) | (gprim == oprim))
    say('agreement', same.mean())
    assert same.mean() >= 0.9999
    made, u = load(engine, 'room-c2', width=48, height=27)
    engine.trace(128)
    ua, um = engine.read_accum()
    engine.fold_params()
    engine.clear_accums()
    engine.trace(128)
    fa, fm = engine.read_accum()
    uL, fL = pt.luminance(ua), pt.luminance(fa)
    sd = np.sqrt(pt.standard_error(um) ** 2 + pt.standard_error(fm) ** 2)
    ok = sd > 1e-9
    z = (uL - fL)[ok] / sd[ok]
    say('folded vs unfolded: identical pixels %.4f, z std %.3f' % (np.mean(uL == fL), z.std()))
    assert np.mean(np.abs(z) < 3) > 0.97 and z.std() < 1.3
    made, u = load(engine, 'plane-c1', width=64, height=64)
    engine.fold_params()
    p0, s0, _, _ = engine.pick_rays()
    params = np.asarray(made['scene']['params'], np.float32).copy()
    params[2] = 0.25
    engine.set_params(params)
    p1, s1, _, _ = engine.pick_rays()
    hit = (s0 == 0) & (s1 == 0)
    assert np.allclose(p1[hit, 2], 0.25, atol=2e-3) and np.allclose(p0[hit, 2], 0.0, atol=2e-3)


def test_node_addon_drives_the_engine(tmp_path):
    say('case: node tests/js/gpu_step.js (createPathTracer through the N-API addon), then its light files')
    r = subprocess.run(['node', os.path.join(ROOT, 'tests', 'js', 'gpu_step.js'), str(tmp_path)], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    assert abs(out['floor']['mean'] / 800 - 1) < 2e-3, out['floor']
    assert out['stats']['samples'] == 16 and out['log']['segments'] > 0
    # the render saved as light: the EXR's float32 and the PFM's are the same numbers, the floor rows 800 cd/m^2
    from test_export import _read_exr
    exr = _read_exr((tmp_path / 'c1.exr').read_bytes())['img']
    head = (tmp_path / 'c1.pfm').read_bytes().split(b'\n', 3)
    pfm = np.frombuffer(head[3], '<f4').reshape(64, 64, 3)[::-1]
    assert np.array_equal(exr.astype(np.float32), pfm)
    Y = pfm.astype(np.float64) @ np.array([0.2126, 0.7152, 0.0722])
    assert abs(Y[-16:].mean() / 800 - 1) < 2e-3
    # the compile ran on a worker thread: the event loop's 10 ms timer kept ticking through it; keepPicture kept the 16
    # samples; a step during a build was refused
    say('compile %.0f ms, timer ticks %d, samples kept %d, refused: %s' % (out['compileMs'], out['ticks'], out['kept'], out['refused']))
    assert out['compileMs'] > 50 and out['ticks'] >= 0.5 * out['compileMs'] / 10, out
    assert out['kept'] == 16
    assert out['refused'] and 'during a scene build' in out['refused']
    # set() took the reference's lighting settings and returned only the unknown name; the sun alone at 20 degrees on
    # the bench's floor reads rho E_sun sin 20 / pi, as the fixture's row does (tests.js photometry)
    import photometry_rows as P
    assert out['unknown'] == ['noSuchSetting']
    ref, rse = P.reference('photometry', 'sun only, elevation 20')
    ratio = out['sunOnly']['mean'] / out['sunOnly']['want']
    say('sun only at 20 degrees: ratio %.5f, reference %.5f +- %.5f' % (ratio, ref, rse))
    assert abs(ratio - ref) < 4 * rse + 0.01


# ---------------------------------------------------------------------------------------------------- SPPM
SLAB_FLOOR = 0.5 * 100000 * (1 - 0.04) / (1 + 0.04) / math.pi   # rho E T / pi, T = (1 - R) / (1 + R), R = 0.04


def beam_pixels(made, radius):
    u = pt.uniforms_for(made)
    pos, state, mat, prim = oracle.pick(made['scene'], u)
    return (np.hypot(pos[:, 0], pos[:, 1]) < radius).reshape(u.height, u.width)


def gpu_samples(engine, samples):
    """Each sample's own luminance image, from the running accumulator read after every sample."""
    out, prev = [], None
    for k in range(1, samples + 1):
        engine.trace(1)
        Y = pt.luminance(engine.read_accum()[0])
        out.append(Y if prev is None else k * Y - (k - 1) * prev)
        prev = Y
    return out


def oracle_samples(made, u, samples):
    """The oracle's sample k alone (its photon pass at r_k), for k = 1..samples."""
    return [pt.luminance(oracle.render(made['scene'], u, 1, first=k)[0]) for k in range(samples)]


def agree(g, o, mask, rel):
    """Per-sample means over mask: equal within 4 standard errors (sample to sample spread, map noise included)
    and rel of the mean (the kernel's blur differs between the engine's subsampled map and the oracle's exact one
    only in noise; rel covers the float rounding of long paths)."""
    gm = np.array([x[mask].mean() for x in g])
    om = np.array([x[mask].mean() for x in o])
    se = math.sqrt(gm.var(ddof=1) / len(gm) + om.var(ddof=1) / len(om))
    tol = 4 * se + rel * max(om.mean(), 1e-3)
    return abs(gm.mean() - om.mean()) / tol, gm.mean(), om.mean(), se


def test_sppm_slab_closed_form(engine):
    say('case: sppm-slab 64x48, 2^20 photons x 16 samples, floor = rho E T / pi')
    made, u = load(engine, 'sppm-slab', width=64, height=48)
    inb = beam_pixels(made, 0.2)
    engine.trace(16)
    acc, mom = engine.read_accum()
    st = engine.stats()
    Y = pt.luminance(acc)[inb]
    se = Y.std() / math.sqrt(inb.sum())
    assert abs(Y.mean() - SLAB_FLOOR) < 4 * se + 0.01 * SLAB_FLOOR, (Y.mean(), se, SLAB_FLOOR)
    frac = st['photons'] / (16 * (1 << 20))
    assert abs(frac - 0.96 / 1.04) < 2e-3, frac
    assert st['overflow'] == 0 and st['nonfinite'] == 0 and st['gathers'] >= 16 * 64 * 48 * 0.9


def test_sppm_radius_follows_the_schedule(engine):
    say('case: sppm-slab 32x24, r_k after k = 1..6 samples')
    made, u = load(engine, 'sppm-slab', width=32, height=24, sppmPhotons=1 << 14, sppmRadius=0.05, sppmAlpha=0.7)
    for k in range(1, 7):
        engine.trace(1)
        got = engine.stats()['sppm_radius']
        assert got == pytest.approx(oracle.sppm_radius(0.05, 0.7, k), rel=2e-6), k


def test_sppm_map_on_a_glossy_floor_matches_the_oracle(engine):
    say('case: sppm-slab glossy floor, 20-degree sun, 0.6 m beam, 32x24, 2^18 photons')
    opts = dict(width=32, height=24, floorRoughness=0.4, floorSpecular=0.5, sunDiameter=20, beam=0.6)
    made, u = load(engine, 'sppm-slab', sppmPhotons=1 << 18, sppmRadius=0.03, **opts)
    inb = beam_pixels(made, 0.3)
    g = gpu_samples(engine, 16)
    o = oracle_samples(made, pt.uniforms_for(made, sppmPhotons=1 << 18, sppmRadius=0.03), 16)
    r = agree(g, o, inb, 0.005)
    assert r[0] < 1, r


def test_sppm_switches_off_and_on(engine):
    say('case: sppm-slab 32x24 with the map, without, and with it again')
    made, u = load(engine, 'sppm-slab', width=32, height=24, sppmPhotons=1 << 16)
    inb = beam_pixels(made, 0.2)
    engine.trace(2)
    on = pt.luminance(engine.read_accum()[0])[inb].mean()
    engine.set_uniforms(pt.uniforms_for(made, sppm=0))
    engine.clear_accums()
    engine.trace(2)
    off = pt.luminance(engine.read_accum()[0])[inb].max()
    engine.set_uniforms(pt.uniforms_for(made, sppmPhotons=1 << 16))
    engine.clear_accums()
    engine.trace(2)
    again = pt.luminance(engine.read_accum()[0])[inb].mean()
    assert off == 0
    assert abs(on / SLAB_FLOOR - 1) < 0.05 and abs(again / SLAB_FLOOR - 1) < 0.05, (on, again)


def test_light_tracing_slab_closed_form(engine):
    # light tracing alone, and with the map on: the floor under the slab reads rho E T / pi once (the connections'),
    # and its connections are traced as the photon pass's shadow rays
    for over in (dict(lt=1, sppm=0), dict(lt=1, sppm=1)):
        say('case: sppm-slab 64x48, light tracing %s, 2^18 photons x 16 samples' % over)
        made, u = load(engine, 'sppm-slab', width=64, height=48, sppmPhotons=1 << 18, **over)
        inb = beam_pixels(made, 0.2)
        engine.trace(16)
        acc, _ = engine.read_accum()
        st = engine.stats()
        Y = pt.luminance(acc)[inb]
        se = Y.std() / math.sqrt(inb.sum())
        say('  %.1f +- %.1f against %.1f; connections %d' % (Y.mean(), se, SLAB_FLOOR, st['shadow_segments']))
        assert abs(Y.mean() - SLAB_FLOOR) < 4 * se + 0.01 * SLAB_FLOOR, (over, Y.mean(), se, SLAB_FLOOR)
        assert st['shadow_segments'] > 16 * (1 << 17) and st['nonfinite'] == 0
        assert st['gathers'] == 0      # the camera sees the floor directly: its first vertex is the connections'
        assert pt.luminance(acc)[~beam_pixels(made, 0.45)].max() == 0


def test_light_tracing_in_the_haze_matches_the_oracle(engine):
    # the beam's track through a haze box under the slab and its spot on the floor, light traced with the map on: the
    # connections from the dust and from the floor, and the camera's paths that leave them their light, sample by
    # sample against the oracle's (the same seeds), over the picture and over the beam's spot
    opts = dict(width=48, height=36, mfp=4.0, g=0.3, sunDiameter=20, beam=0.3)
    say('case: sppm-slab with haze 48x36, lt 1, sppm 1, 2^16 photons, 48 samples vs oracle')
    made, u = load(engine, 'sppm-slab', sppmPhotons=1 << 16, lt=1, **opts)
    g = gpu_samples(engine, 48)
    st = engine.stats()
    assert st['nonfinite'] == 0 and st['shadow_segments'] > 0
    o = oracle_samples(made, pt.uniforms_for(made, sppmPhotons=1 << 16, lt=1), 48)
    for name, m in (('picture', np.ones((36, 48), bool)), ('spot', beam_pixels(made, 0.25))):
        r = agree(g, o, m, 0.01)
        say('  %s: engine %.2f, oracle %.2f, se %.2f (%.2f of the tolerance)' % (name, r[1], r[2], r[3], r[0]))
        assert r[0] < 1, (name, r)


def test_light_tracing_from_a_light_face_matches_the_oracle(engine):
    # the lens and the prism under their panel, light traced: the panel's photons through the glass connect to the
    # camera, and the camera's paths leave that light to them -- sample by sample against the oracle's (the same seeds)
    say('case: lens-prism 48x36, lt 1, 2^16 photons, 48 samples vs oracle')
    made, u = load(engine, 'lens-prism', width=48, height=36, sppmPhotons=1 << 16, lt=1, sppm=0)
    g = gpu_samples(engine, 48)
    st = engine.stats()
    assert st['nonfinite'] == 0 and st['shadow_segments'] > 0
    o = oracle_samples(made, pt.uniforms_for(made, sppmPhotons=1 << 16, lt=1, sppm=0), 48)
    r = agree(g, o, np.ones((36, 48), bool), 0.01)
    say('  picture: engine %.2f, oracle %.2f, se %.2f (%.2f of the tolerance)' % (r[1], r[2], r[3], r[0]))
    assert r[0] < 1, r


@pytest.mark.parametrize('mode', [1, 2])
def test_c3_caustic_room_matches_the_oracle(engine, mode):
    # sppm 1 gathers at a camera path's first non-delta vertex after a delta event (the reference's); 2 at every
    # non-delta vertex past the first; the first is light tracing's (the room's setting, as the reference's default)
    say('case: caustic-c3 160x90, 2^18 photons, 8 samples, sppm %d, light tracing, 16 regions vs oracle' % mode)
    made, u = load(engine, 'caustic-c3', width=160, height=90, sppmPhotons=1 << 18, sppm=mode)
    regions = []
    for by in range(4):
        for bx in range(4):
            m = np.zeros((90, 160), bool)
            m[by * 90 // 4:(by + 1) * 90 // 4, bx * 40:(bx + 1) * 40] = True
            regions.append(m)
    g = gpu_samples(engine, 8)
    st = engine.stats()
    assert st['nonfinite'] == 0 and st['photons'] > 0 and st['gathers'] > 0 and st['overflow'] == 0
    o = oracle_samples(made, pt.uniforms_for(made, sppmPhotons=1 << 18, sppm=mode), 8)
    results = [agree(g, o, m, 0.03) for m in regions + [np.ones((90, 160), bool)]]
    say('worst region at %.2f of its tolerance' % max(r[0] for r in results))
    for r in results:
        assert r[0] < 1, r


def test_autotune_picks_a_setting_and_caches_it(engine, tmp_path):
    say('case: room-c2 160x90, pt_autotune with 2 samples a candidate, then the cached choice')
    made, u = load(engine, 'room-c2', width=160, height=90)
    t = engine.autotune(samples=2, cache=str(tmp_path / 'tuning.json'))
    say('tuning', t)
    assert not t['cached'] and t['candidates'] >= 10
    assert t['blocksPerCu'] in (2, 4, 6, 8, 12, 16) and t['refillMarch'] in (1, 4, 8, 16, 24, 32, 48)
    assert 0 < t['msPerSample'] <= t['msBefore']
    engine.trace(4)
    acc, _ = engine.read_accum()
    assert (acc[..., 3] == 4).all()               # the tuner's samples were cleared; the picture is the 4 traced
    again = engine.autotune(samples=2, cache=str(tmp_path / 'tuning.json'))
    assert again['cached'] and again['blocksPerCu'] == t['blocksPerCu'] and engine.tuning()['refillOcclude'] == t['refillOcclude']
    engine.set_tuning(8, 16, 16)                  # back to the library's defaults for the tests after this one


def test_an_absorbing_slab_transmits_its_closed_form(engine):
    # absorption inside (attenuationDistance): a smooth slab 20 mm thick seen head-on in a uniform sky reads
    # L (R + (1 - R)^2 (R tau^2 + tau) / (1 - R^2 tau^2)) per channel, tau = base^(20 / distance)
    from test_absorption import slab_closed_form, centre
    for att in (10.0, 40.0):
        say('case: absorb-slab 32x32 4096 spp, attenuation', att)
        made, u = load(engine, 'absorb-slab', attenuation=att)
        engine.trace(4096)
        acc, mom = engine.read_accum()
        c, se = centre(acc, mom)
        got = c.mean(0) / 1000
        want = slab_closed_form(att)
        rel = np.sqrt((se ** 2).sum()) / len(se) / (c @ [0.2126, 0.7152, 0.0722]).mean()
        say('got', got, 'want', want, 'rel se', rel)
        assert engine.stats()['nonfinite'] == 0
        assert np.all(np.abs(got / want - 1) < 4 * rel + 2e-4), (att, got, want, rel)


def test_the_scenes_own_sky_code_lights_as_the_studio(engine):
    # the sky written as the scene's code (background(dir) with its parameters) and a ball of the document's functions
    # block: the picture of the builtin studio and sphere within the spread, and the oracle's FP64 compile of it
    say('case: sky-code 64x48, code vs builtin studio 256 spp each, vs oracle 64 spp')
    out = []
    for native in (False, True):
        made, u = load(engine, 'sky-code', native=native, seed=11 + int(native))
        engine.trace(256)
        out.append(engine.read_accum())
        if not native:
            oa, om, _ = oracle.render(made['scene'], u, 64, first=100000)
    (ca, cm), (na, nm) = out
    for (a, m), (b, n) in (((ca, cm), (na, nm)), ((ca, cm), (oa, om))):
        La, Lb = pt.luminance(a), pt.luminance(b)
        s = np.sqrt(pt.standard_error(m) ** 2 + pt.standard_error(n) ** 2)
        ok = s > 1e-9
        z = (La - Lb)[ok] / s[ok]
        say('z std', z.std(), 'means', La.mean(), Lb.mean())
        assert 0.75 < z.std() < 1.3 and np.mean(np.abs(z) < 3) > 0.97
