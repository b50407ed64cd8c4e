"""The stochastic-rounding conversions of the 16-bit stores on the device: gfx950's v_cvt_sr_f16_f32 and
v_cvt_sr_bf16_f32 (storage.h f32_to_f16_sr / f32_to_bf16_sr), lane by lane against the models storage.h states for
them (tests/native/sr_check.hip, built by `make -C ashburn-heat-twin_amd probes`)."""
import json
import os
import subprocess

import pytest

from conftest import PKG

pytestmark = pytest.mark.gpu


def test_stochastic_rounding_instructions_match_their_models():
    exe = os.path.join(PKG, 'lib', 'aht_sr_check')
    if not os.path.exists(exe):
        raise RuntimeError(f'{exe} is not built: run __graft_entry__.build() or make -C ashburn-heat-twin_amd probes')
    print('running', exe, flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    out = json.loads(r.stdout)
    assert out['bf16_checked'] > 3_000_000 and out['bf16_mismatch'] == 0
    assert out['f16_normal_checked'] > 1_500_000 and out['f16_normal_mismatch'] == 0
    assert abs(out['f16_subnormal_worst_mean_error_units']) <= 0.05
    assert r.returncode == 0
