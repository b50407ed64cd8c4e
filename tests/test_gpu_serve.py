"""The serving loop (js/serve.js) on the GPU: a burst of requests of three grid sizes is served by continuous batching
with a slot limit, and every result (FP64 time-mean surface warming field, mean warming, advected heat) is
bit-identical to the same request run alone through the plain per-solver path."""
import json
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone():
    p = subprocess.run(['node', os.path.join(ROOT, 'tests', 'js', 'serve_test.js')], capture_output=True, text=True,
                       timeout=600)
    assert p.returncode == 0, p.stderr[-3000:]
    out = json.loads(p.stdout.strip().splitlines()[-1])
    print(json.dumps(out))
    assert out['summary']['requests'] == 6
    assert all(out['identical']), out['identical']
    for a, b, m in out['nAcc']:
        assert a == b == m


def test_served_results_equal_alone_with_storage_per_request():
    """storage chosen per request (f16s for the finer grids, f32 for the coarsest, two batches stepped in turn): every
    result bit-identical to the request alone in its own storage"""
    p = subprocess.run(['node', os.path.join(ROOT, 'tests', 'js', 'serve_test.js'), '--auto'], capture_output=True,
                       text=True, timeout=600)
    assert p.returncode == 0, p.stderr[-3000:]
    out = json.loads(p.stdout.strip().splitlines()[-1])
    print(json.dumps(out))
    assert out['summary']['requests'] == 6
    assert set(out['storages']) == {0, 4}, out['storages']
    assert all(out['identical']), out['identical']
    for a, b, m in out['nAcc']:
        assert a == b == m
