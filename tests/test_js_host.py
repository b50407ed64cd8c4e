"""The Node host without a GPU: option validation (the reference's messages, js/live/solver.js:50-54), the clean
failure when no device exists, determinism of the seeded synthetic generator, and the workload geometry."""
import json
import os
import subprocess

import numpy as np

from conftest import ROOT, dump_domain, load_summary

API = os.path.join(ROOT, 'tests', 'js', 'api_test.js')


def node_eval(code):
    out = subprocess.run(['node', '-e', code], capture_output=True, text=True, check=True, cwd=ROOT).stdout
    return json.loads(out)


def test_rejection_messages_match_reference():
    K = load_summary()['api20']
    out = json.loads(subprocess.run(['node', API, '--cpu'], capture_output=True, text=True, check=True).stdout)
    assert out['badDx'] == K['badDx']
    assert out['badU10Max'] == K['badU10Max']
    assert out['badHeight'].startswith('rejected: live solver: heightM must be within 300-2000 m')
    assert out['noGpu'].startswith('rejected: aht:')          # no GPU here: the addon refuses, no CPU path


def test_synthetic_generator_is_seeded(tmp_path):
    a = dump_domain(tmp_path / 'a', '--tiny', '40,30,16,5')
    b = dump_domain(tmp_path / 'b', '--tiny', '40,30,16,5')
    c = dump_domain(tmp_path / 'c', '--tiny', '40,30,16,6')
    for k in ('kt', 'ktop', 'cd', 'q', 'prof'):
        assert np.array_equal(a[k], b[k]), k
    assert not (np.array_equal(a['ktop'], c['ktop']) and np.array_equal(a['q'], c['q']))


def test_workload_c1_geometry(tmp_path):
    d = dump_domain(tmp_path, '--workload', 'c1')
    assert (d['nx'], d['ny'], d['nz'], d['dx']) == (200, 200, 32, 30)
    kt, ktop = d['kt'].astype(int), d['ktop'].astype(int)
    assert kt.min() >= 1 and (ktop >= kt).all() and ktop.max() < d['nz'] - 8
    assert d['params']['inflow_w'] == 1 and d['params']['inflow_s'] == 1      # wind from 225 deg
    # 120 source roofs at 400 W/m2: heat equals 400 W/m2 times the rasterised source roof area
    assert abs(d['heatMW'] - 400 * d['info']['dcRoofArea_m2'] / 1e6) < 1e-6 * d['heatMW']
    assert d['info']['dt'] == 0.1 * 30 / (5 * 15 ** 0.2)


def test_campus_rasterisation_conserves_area():
    r = node_eval("""
      const D = require('./ashburn-heat-twin_amd/js/domain.js');
      const cp = [{ id: 'X', mw: 60, halls: [[-1013, 222, -713, 422]] }];
      const a = D.rasterizeCampuses(cp, 30, 200, 200);
      let area = 0, heat = 0; for (let k = 0; k < a.newArea.length; k++) { area += a.newArea[k]; heat += a.heatW[k]; }
      console.log(JSON.stringify({ area, heat, info: a.info }));""")
    assert abs(r['area'] - 300 * 200) < 1e-6
    assert abs(r['heat'] - 60e6) < 1e-3
    assert r['info'][0]['cells'] == 11 * 7          # columns 66..76, rows 107..113


def test_serving_chunk_plan():
    """the serving loop's iteration-level plan (js/serve.js planChunk): exact per-member step counts, averaging start
    relative to the chunk, and the short-first rule (members with far more work left yield most launches)"""
    code = """
const { planChunk } = require('./ashburn-heat-twin_amd/js/serve.js');
const mem = [{ cells: 320000, total: 4400, spin: 3400, done: 3390 },      // small, 1010 steps left
             { cells: 23040000, total: 26000, spin: 20000, done: 100 },   // large
             { cells: 1920000, total: 8800, spin: 6600, done: 8780 }];    // small, 20 steps left
const r = planChunk(mem, 50, 'short-first', 3e10, 8);
const f = planChunk(mem, 50, 'fifo', 3e10, 8);
const big = planChunk([mem[1]], 50, 'short-first', 3e10, 8);
console.log(JSON.stringify({ r: { n: r.n, ns: Array.from(r.ns), af: Array.from(r.af) },
  f: { n: f.n, ns: Array.from(f.ns) }, big: Array.from(big.ns) }));
"""
    out = node_eval(code)
    assert out['r'] == {'n': 50, 'ns': [50, 7, 20], 'af': [10, 19900, 0]}
    assert out['f'] == {'n': 50, 'ns': [50, 50, 20]}
    assert out['big'] == [50]                  # alone, a large member steps every launch


def test_addon_device_thread_against_cpu_standin(tmp_path):
    """the N-API addon built against a CPU stand-in of the C-ABI (tests/js/aht_standin.cpp), which aborts on any
    threading or lifetime violation: jobs issued without awaiting run one at a time in issue order, destroy() with
    a step or batch step in flight is deferred, synchronous calls on a solver or batch with a job in flight throw
    instead of running beside it, and no thread that made device calls outlives the runtime's static
    teardown at exit (libuv joins its pool threads after it; the addon's device thread is joined before it)"""
    lib, addon = tmp_path / 'libaht.so', tmp_path / 'aht_node.node'
    subprocess.run(['g++', '-O1', '-std=c++17', '-shared', '-fPIC', '-I', os.path.join(ROOT, 'include'), '-o', str(lib),
                    os.path.join(ROOT, 'tests', 'js', 'aht_standin.cpp'), '-lpthread'], check=True)
    subprocess.run(['gcc', '-O2', '-std=c11', '-shared', '-fPIC', '-Wall', '-I/usr/include/node',
                    '-I', os.path.join(ROOT, 'include'), '-o', str(addon),
                    os.path.join(ROOT, 'ashburn-heat-twin_amd', 'csrc', 'aht_napi.c'), '-L', str(tmp_path), '-laht',
                    '-lpthread', '-Wl,-rpath,$ORIGIN'], check=True)
    for mode in ('natural', 'exit'):
        r = subprocess.run(['node', os.path.join(ROOT, 'tests', 'js', 'exit_test.js'), str(addon), mode],
                           capture_output=True, text=True, timeout=60)
        assert r.returncode == 0 and 'STANDIN FAILURE' not in r.stderr, (mode, r.returncode, r.stderr)
        out = json.loads(r.stdout)
        assert out['stepBeforeDestroy'] == 'resolved' and out['stepAfterDestroy'] == 'threw'
        assert out['batchStep'] == 10 and out['bench'] == 3
        assert out['order'] == list(range(6))
        assert out['statsDuringStep'] == 'threw' and out['setParamsDuringStep'] == 'threw'
        assert out['statsAfterStep'] == 'ran'
        assert out['memberStatsDuringBatchStep'] == 'threw' and out['setBlockDuringBatchStep'] == 'threw'
        assert out['setBlockAfterBatchStep'] == 'ran'
        assert 'last step settled' in r.stderr
