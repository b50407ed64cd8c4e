"""The reference's long-horizon validation case on the GPU: 25 Aug 2024 at 30 m (U10 3.09 m/s from 210 deg,
u* 0.32, w* 1.67, blh 903 m, 400 W/m2), 2200 s spin-up then a 600 s mean accumulated every step
(tests/live/validate.html:123-159; metrics tools/live/validate_live.py:216-240, :258-290).

Compared with the recorded WebGPU and numba results (tests/golden/reference_kats.json):
  * 10-minute mean near-ground theta: relative error of the mean <= 1e-4 and the same peak (<= 1e-4 relative);
  * heat-budget closure (advected out / injected) within 0.5 percentage points; heat content within 1e-4;
  * mass imbalance within 5e-5 absolute;
  * the diagnostics of the final state (the run's last _dbg.stats() sample at step nspin + navg,
    js/live/solver.js:559-576): heat content in MJ and max theta within 1e-4 relative, mean kinetic energy per fluid
    cell (keMean = sum 0.5|j|^2/rho / n_fluid * conv^2, m2/s2) within 1e-3 relative, the fluid count exactly.
L30 runs the prototype outlet, L30c the backflow-clipped outlet; both use the default outflow-sponge width 5.
"""
import numpy as np
import pytest

from conftest import dump_domain, load_summary

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('tag', ['L30', 'L30c'])
def test_ten_minute_mean_and_budgets(tmp_path, pkg, tag):
    K = load_summary()['long'][tag]
    dom = dump_domain(tmp_path, '--dx', 30, '--nsw', K['nsw'], '--oclip', int(K['oclip']))
    assert dom['dt'] == K['dt']
    with pkg.Solver.from_dump(dom) as s:
        s.step(K['nspin'])
        s.mean_reset()
        s.step(K['navg'], acc=True)
        s3, b = s.read_mean()
        st = s.stats()
    assert st['non_finite'] == 0
    fin = K['final']
    assert fin['step'] == K['nspin'] + K['navg'] and fin['nonFinite'] == 0
    ph = pkg.stats_physical(st, dom['dx'], dom['conv'])
    print(tag, 'final stats', ph, 'record', fin)
    assert ph['n_fluid'] == dom['info']['nFluid']
    assert abs(ph['heat_content_MJ'] - fin['heatMJ']) <= 1e-4 * fin['heatMJ']
    assert abs(ph['theta_max'] - fin['thetaMax']) <= 1e-4 * fin['thetaMax']
    assert abs(ph['ke_mean'] - fin['ke']) <= 1e-3 * fin['ke']
    th = s3[0]
    want = K['surface_mean_theta']['webgpu']
    print(tag, 'surface mean theta max %.7f mean %.8f p99 %.6f' % (th.max(), th.mean(), np.percentile(th, 99)),
          'record', want)
    assert abs(th.max() - want['max']) <= 1e-4 * want['max']
    assert abs(th.mean() - want['mean']) <= 1e-4 * want['mean']
    assert abs(np.percentile(th, 99) - want['p99']) <= 1e-3 * want['p99']
    phys = pkg.budget_physical(b, dom['dx'], dom['dt'], dom['conv'])
    hb = K['heat_budget_webgpu']
    print(tag, 'budget', phys, 'record', hb)
    assert b['n_acc'] == K['navg']
    assert abs(phys['heat_injected_MW'] - hb['heat_injected_MW']) <= 1e-6 * hb['heat_injected_MW']
    closure, closure_ref = (phys['heat_advected_out_MW'] / phys['heat_injected_MW'],
                            hb['heat_advected_out_MW'] / hb['heat_injected_MW'])
    assert abs(closure - closure_ref) <= 0.005
    assert abs(phys['heat_content_MJ'] - hb['heat_content_MJ']) <= 1e-4 * hb['heat_content_MJ']
    assert abs(phys['mass_imbalance_frac'] - hb['mass_imbalance_frac']) <= 5e-5
