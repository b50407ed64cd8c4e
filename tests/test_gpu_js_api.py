"""The Node host's interface on the GPU, the reference's `api` test sequence at dx = 20 (tests/live/validate.html:281-328)
compared with its record (api20 in tests/live/validation_summary.json)."""
import json
import os
import subprocess

import pytest

from conftest import ROOT, load_summary

pytestmark = pytest.mark.gpu


def test_api_sequence_matches_reference_record():
    K = load_summary()['api20']
    p = subprocess.run(['node', os.path.join(ROOT, 'tests', 'js', 'api_test.js')], capture_output=True, text=True,
                       timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    out = json.loads(p.stdout.strip().splitlines()[-1])
    print(json.dumps({k: v for k, v in out.items() if k != 'info'}))
    assert out['badDx'] == K['badDx'] and out['badU10Max'] == K['badU10Max']
    r0, k0 = out['readNoVolNoWind'], K['readNoVolNoWind']
    assert r0['surf'] == k0['surf'] and r0['u'] is None and r0['v'] is None and r0['vol'] is None
    assert abs(r0['simTime'] - k0['simTime']) < 1e-9
    assert abs(r0['thetaMax'] - k0['thetaMax']) <= 1e-4 * k0['thetaMax']
    assert out['concurrentReads'] == K['concurrentReads']
    c, kc = out['campuses'], K['campuses']
    assert c['nFluidWith'] == kc['nFluidWith'] and c['nFluidWithout'] == kc['nFluidWithout']
    assert abs(c['heatMWwith'] - kc['heatMWwith']) < 1e-6 and abs(c['heatMWwithout'] - kc['heatMWwithout']) < 1e-6
    assert c['nonFinite'] == 0 and c['umaxDomain'] < 25
    assert abs(out['heatExistingMW800'] - K['heatExistingMW800']) < 1e-6
    assert out['afterResetHeat']['simTime'] == 0 and out['afterResetHeat']['thetaMax'] == 0
    assert abs(out['afterResetHeat']['windMean'] - K['afterResetHeat']['windMean']) <= 2e-3 * K['afterResetHeat']['windMean']
    assert out['afterResetFlow']['simTime'] == 0 and out['afterResetFlow']['thetaMax'] == 0
    assert abs(out['afterResetFlow']['windMean'] - K['afterResetFlow']['windMean']) <= 1e-5 * K['afterResetFlow']['windMean']
    assert out['flowReinits'] == K['flowReinits']
    assert out['afterDestroy'] == K['afterDestroy']
    assert out['readAfterDestroy'] == 'rejected: live solver: destroyed'
    # destroy() with a step in flight (the reference caller's pattern, js/live.js:44): that step settles, later
    # calls reject; at the addon the library solver is freed only when the worker's step has completed
    assert out['stepBeforeDestroy'] == 'resolved'
    assert out['stepAfterDestroy'] == 'rejected: live solver: destroyed'
    assert out['addonStepBeforeDestroy'] == 'resolved'
    assert out['addonStepAfterDestroy'] == 'threw: aht addon: invalid or destroyed solver handle'
    assert out['mlups'] > 1000
    # the launch-shape autotuner (no counterpart in the reference) leaves the simulation where it was
    assert out['autotune']['block'] in (64, 128, 256) and out['autotune']['levels'] in (1, 2, 4)
    assert out['autotuneKeepsState'] and out['launch'] == out['autotune']
