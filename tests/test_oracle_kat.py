"""The CPU oracle and the JavaScript host pinned against the reference's recorded known answers
(tests/live/validation_summary.json, kept as tests/golden/reference_kats.json) and against closed forms.

These run on the CPU only. They pin the oracle the GPU parity tests rely on:
  * host parameter block and dt at dx = 20 / 30 / 40 (recorded f32 values, js/live/solver.js:225-238);
  * Ashburn domain facts (zbase, kt range, ktop max, building columns, nz, fluid cells; js/live/domain.js:166-190);
  * the initial state, bit-identical to the numpy closed form (tools/live/validate_live.py:132-143);
  * the p30a run: 300 steps from a cold start on the Ashburn 30 m grid, read() surface max / mean / wind mean
    (the f32 oracle against the WebGPU record, the f64-arithmetic oracle against the numba record);
  * the p30b run: campus A, backflow-clipped outlets, 2,000 + 300 steps, against the WebGPU record.
"""

import numpy as np
import pytest

from conftest import dump_domain, load_summary


@pytest.mark.parametrize('tag,dx,extra', [('p30a', 30, ['--nsw', 16, '--oclip', 0]), ('p20', 20, ['--campus', 'AB']),
                                          ('p40', 40, ['--campus', 'C', '--u10', 6, '--dir', 75])])
def test_params_and_domain_match_reference_records(tmp_path, tag, dx, extra):
    K = load_summary()['short'][tag]
    d = dump_domain(tmp_path, '--dx', dx, *extra)
    for k in ('tau0', 'cs2smag', 'gbeta', 'prt', 'kfloor', 'sigmax'):
        assert d['params'][k] == K['params'][k], k                    # exact f32 values
    for k in ('nsp', 'nsw'):
        assert d['params'][k] == K['params'][k]
    info = d['info']
    assert info['dt'] == K['info']['dt'] and info['Umax'] == K['info']['Umax']
    for k in ('nx', 'ny', 'nz', 'zbase', 'nFluid'):
        assert info[k] == K['info'][k], k
    assert info['dcRoofArea_m2'] == K['info']['dcRoofArea_m2']
    assert abs(d['heatMW'] - K['webgpu']['heatMW']) < 1e-6
    if K['campusInfo']:
        for got, want in zip(d['campusInfo'], K['campusInfo']):
            for k in ('roofArea_m2', 'inDomain_m2', 'flux_wm2', 'cells'):
                assert abs(got[k] - want[k]) < 1e-9 * max(1, abs(want[k])), k


@pytest.mark.parametrize('tag', ['p30a', 'p20', 'p40'])
def test_domain_facts(tmp_path, tag):
    R = load_summary()['domain'][tag]
    d = dump_domain(tmp_path, '--dx', R['dx'])
    kt, ktop = d['kt'].astype(int), d['ktop'].astype(int)
    assert d['zbase'] == R['zbase_js'] and d['nz'] == R['nz']
    assert [kt.min(), kt.max()] == R['kt_range']
    assert ktop.max() == R['ktop_max']
    assert int((ktop > kt).sum()) == R['building_columns']


def test_init_state_bit_identical_to_closed_form(ashburn30, O):
    o = O.Oracle(ashburn30, 'f32')
    o.init_flow()
    f = O.init_closed_form(ashburn30)
    assert np.array_equal(f.view(np.uint32), o.f.view(np.uint32))


def test_p30a_f32_oracle_matches_webgpu_record(ashburn30, O):
    K = load_summary()['short']['p30a']
    o = O.Oracle(ashburn30, 'f32')
    o.init_flow()
    o.step(K['steps'])
    th, u, v = o.surface(ashburn30['conv'])
    wind = np.hypot(u.astype(np.float64), v.astype(np.float64)).mean()
    # the same f32 expression order as the WebGPU kernel: agreement to a few f32 ulps of the field maximum
    assert abs(th.max() - K['webgpu']['thetaMax']) < 3e-6 * K['webgpu']['thetaMax']
    assert abs(th.astype(np.float64).mean() - K['webgpu']['surfMean']) < 1e-5 * K['webgpu']['surfMean']
    assert abs(wind - K['webgpu']['windMean']) < 1e-5 * K['webgpu']['windMean']
    st = o.stats()
    assert st['non_finite'] == 0 and st['n_fluid'] == K['info']['nFluid']


@pytest.mark.slow
def test_p30a_f64_oracle_matches_numba_record(ashburn30, O):
    K = load_summary()['short']['p30a']
    o = O.Oracle(ashburn30, 'f64s32')
    o.init_flow()
    o.step(K['steps'])
    th, _, _ = o.surface(ashburn30['conv'])
    assert abs(th.max() - K['numba']['theta_max']) < 1e-6 * K['numba']['theta_max']


@pytest.mark.slow
def test_p30b_f32_oracle_matches_webgpu_record(tmp_path, O):
    """p30b: campus A, backflow-clipped outlets, outflow sponge 5, 2,000 + 300 steps (js/live/README.md:284); pins
    the oracle's clipped-outlet branch and rooftop sources directly (about 3 minutes on 8 cores)"""
    K = load_summary()['short']['p30b']
    dom = dump_domain(tmp_path, '--dx', 30, '--nsw', 5, '--oclip', 1, '--campus', 'A')
    o = O.Oracle(dom, 'f32')
    o.init_flow()
    o.step(K['pre'] + K['steps'])
    th, u, v = o.surface(dom['conv'])
    wind = np.hypot(u.astype(np.float64), v.astype(np.float64)).mean()
    assert abs(th.max() - K['webgpu']['thetaMax']) < 1e-5 * K['webgpu']['thetaMax']
    assert abs(th.astype(np.float64).mean() - K['webgpu']['surfMean']) < 1e-5 * K['webgpu']['surfMean']
    assert abs(wind - K['webgpu']['windMean']) < 1e-5 * K['webgpu']['windMean']
