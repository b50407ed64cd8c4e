"""Long runs of the synthetic workloads on the GPU against the oracle's records of the same runs.

tests/golden/long_<case>.json holds what the f32 oracle (the reference kernel's expressions and rounding) produced
for the case's protocol, written by tests/golden/make_long_reference.py: `steps` steps from rest on the seeded
synthetic domain, FP64 time means over the last `mean_steps`. The GPU rebuilds the identical domain with the Node
host (checked by the column fingerprint), runs the same protocol through the C-ABI and must give:
  * the mean near-ground warming (spatial mean of the time-mean theta at the first fluid cell) within 1e-4 relative,
    the same peak (value within 1e-4 relative, at the same column) and the subsampled field within 1e-3 rel. L2;
  * heat-budget closure (advected out / injected) within 0.5 percentage points, heat content within 1e-4 relative,
    mass imbalance within 5e-5 absolute (the bars of the reference's own validation).
Cases: c1 = the 200 x 200 x 32 config's full protocol (2,000 steps + the 10-minute mean); c2s = the 600 x 600 x 64
domain with a shortened protocol (1,000 + 300 steps); c2m = the same domain past the start-up transient (3,000 steps,
then a 1,000-step mean).
"""
import glob
import json
import os

import numpy as np
import pytest

from conftest import dump_domain, rel_l2

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
CASES = sorted(os.path.basename(p)[5:-5] for p in glob.glob(os.path.join(GOLDEN, 'long_*.json')))


@pytest.mark.parametrize('case', CASES)
def test_synthetic_long_run_matches_oracle(tmp_path, pkg, case):
    K = json.load(open(os.path.join(GOLDEN, f'long_{case}.json')))
    dom = dump_domain(tmp_path, '--workload', K['workload'])
    fp = K['domain']
    assert (dom['nx'], dom['ny'], dom['nz']) == (fp['nx'], fp['ny'], fp['nz']) and dom['dt'] == fp['dt']
    assert int(dom['kt'].astype(np.int64).sum()) == fp['kt_sum']
    assert int(dom['ktop'].astype(np.int64).sum()) == fp['ktop_sum']
    assert float(dom['q'].astype(np.float64).sum()) == fp['q_sum']
    assert float(dom['cd'].astype(np.float64).sum()) == fp['cd_sum']
    with pkg.Solver.from_dump(dom) as s:
        s.step(K['steps'] - K['mean_steps'])
        s.mean_reset()
        s.step(K['mean_steps'], acc=True)
        s3, b = s.read_mean()
        st = s.stats()
    assert st['non_finite'] == 0
    th = s3[0]
    W = K['warming']
    print(f'{case}: warming mean {th.mean():.8f} (oracle {W["mean"]:.8f}), max {th.max():.6f} at {th.argmax()} '
          f'(oracle {W["max"]:.6f} at {W["argmax"]})')
    assert abs(th.mean() - W['mean']) <= 1e-4 * abs(W['mean'])
    assert abs(th.max() - W['max']) <= 1e-4 * abs(W['max']) and int(th.argmax()) == W['argmax']
    assert abs(np.percentile(th, 99) - W['p99']) <= 1e-3 * abs(W['p99'])
    sub = K['warming_subsample']
    assert rel_l2(th[::sub['every']], np.asarray(sub['theta'])) <= 1e-3
    R = K['budget']
    print(f'{case}: budget', {k: b[k] for k in R}, 'oracle', R)
    assert b['n_acc'] == R['n_acc'] == K['mean_steps']
    assert abs(b['heat_in'] - R['heat_in']) <= 1e-9 * R['heat_in']
    assert abs(b['heat_out'] / b['heat_in'] - R['heat_out'] / R['heat_in']) <= 0.005
    assert abs(b['content'] - R['content']) <= 1e-4 * abs(R['content'])
    assert abs(b['mass_imbalance'] - R['mass_imbalance']) <= 5e-5
