"""Writes tests/golden/reference_kats.json: the recorded known-answer numbers of the reference's own validation runs
(tests/live/validation_summary.json of alanhuang116/ashburn-heat-twin) that the tests compare against.

Run in a container where the reference is mounted at /root/reference; the output is a data fixture (inputs and the
reference's recorded outputs), so tests never read the reference at run time.
"""
import json
import os
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else '/root/reference/tests/live/validation_summary.json'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'reference_kats.json')


def main():
    d = json.load(open(SRC))
    out = {'source': 'tests/live/validation_summary.json (reference)', 'short': {}, 'long': {}, 'domain': {}}
    for tag in ('p30a', 'p30b', 'p20', 'p40', 'p40w', 'q30a', 'q20'):
        r, sr = d[tag]['result'], d[tag]['short_report']
        out['short'][tag] = {
            'steps': r['steps'], 'pre': r['pre'], 'inflow': r['inflow'], 'params': r['params'],
            'info': {k: r['info'][k] for k in ('nx', 'ny', 'nz', 'dx', 'dt', 'zbase', 'nFluid', 'nsp', 'nsw', 'Umax',
                                               'convU', 'heatExistingMW', 'dcRoofArea_m2')},
            'campusInfo': r.get('campusInfo', []),
            'webgpu': {'thetaMax': r['thetaMax'], 'surfMean': r['surfMean'], 'windMean': r['windMean'],
                       'heatMW': r['heatMW']},
            'numba': {'theta_max': sr['theta_max_ref'],
                      'theta_field_absmax': sr['step_outputs_all_fluid_cells']['theta']['ref_absmax']},
            'webgpu_vs_numba': {k: sr['step_outputs_all_fluid_cells'][k]['rel_l2'] for k in ('theta', 'ux', 'uy')},
        }
        if 'domain_report' in d[tag]:
            dr = d[tag]['domain_report']
            out['domain'][tag] = {k: dr[k] for k in ('dx', 'zbase_js', 'kt_range', 'ktop_max', 'building_columns', 'nz')}
    for tag, key in (('L30', 'long_report'), ('L30c', 'long_refauto_report')):
        lr = d[tag][key]
        out['long'][tag] = {
            'nspin': lr['nspin'], 'navg': lr['navg'], 'dt': lr['dt'], 'nsw': lr['nsw_webgpu'],
            'oclip': bool(lr.get('oclip_webgpu', False)),
            'surface_mean_theta': {'webgpu': lr['surface_mean_theta']['webgpu'],
                                   'numba': lr['surface_mean_theta']['numba'],
                                   'rel_l2_webgpu_vs_numba': lr['surface_mean_theta']['rel_l2']},
            'heat_budget_webgpu': lr['heat_budget_webgpu'], 'heat_budget_numba': lr['heat_budget_numba'],
            'speed_mean': {'webgpu': lr['surface_mean_wind']['speed_mean_webgpu'],
                           'numba': lr['surface_mean_wind']['speed_mean_numba']},
            # the WebGPU run's last _dbg.stats() sample (step nspin + navg): heat content MJ, max theta, keMean
            'final': d[tag]['result']['final'],
        }
    # the stability matrix (tests/live/validate.html:162-198): per dx and met/campus combination, read() + stats()
    # every 1,000 steps of a 10,000-step run
    out['stab'] = [{k: r[k] for k in ('dx', 'name', 'ok', 'checks', 'dt')} for r in d['stab']['result']['results']]
    api = d['api20']['result']
    out['api20'] = {k: api[k] for k in ('badDx', 'badU10Max', 'readNoVolNoWind', 'concurrentReads', 'campuses',
                                        'heatExistingMW800', 'afterResetHeat', 'afterResetFlow', 'flowReinits',
                                        'afterDestroy')}
    out['api20']['dt'] = api['info']['dt']
    json.dump(out, open(OUT, 'w'), indent=1)
    print('->', OUT)


if __name__ == '__main__':
    main()
