#!/usr/bin/env python3
"""Throughput of the coupled D3Q19 + D3Q7 lattice-Boltzmann step on one MI355X.

Workload (default): 600 x 600 x 64 cells at 10 m (23.04 M cells), the seeded synthetic 6 x 6 km domain with 4,000 box
buildings, 210 heat-source roofs at 400 W/m2, 12 Gaussian hills, inflow 4 m/s from 315 deg (power law 0.2), f32
two-lattice storage. The domain is built by the Node host (ashburn-heat-twin_amd/js/dump.js) and uploaded once; the
timed region is K steps with every input already in HBM, bracketed by HIP events on the solver's stream
(aht_benchmark) and by stream synchronisation for the wall clock.

One step = one launch of the fused step kernel over all fluid cells. MLUPS counts all nx*ny*nz cells, like the
reference (js/live/solver.js:482). Roofline: HBM, 208 algorithmic bytes per fluid-cell update (read + write of 19 + 7
f32 populations), 8 TB/s MI355X peak.

The CPU baseline times the oracle (a C restatement of the reference kernel, OpenMP over all host cores) on a bounded
sample of the same workload: whole steps on the same 600 x 600 x 64 domain for about --cpu-seconds.
Also measured in the same run: the 16-bit storage formats on the same domain (`reduced_precision`: f16 without theta
plane and f16 with its f32 theta plane) and the serving
half of the metric (`serving`: --serve-requests requests at --serve req/s through the Node serving loop, continuous
batching against one-at-a-time execution with the plain f32 step, and the same requests batched again with f16s
storage for the 10 m grids; about four minutes).
Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import importlib.util
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, 'ashburn-heat-twin_amd')
METRIC = 'MLUPS & % HBM roofline at 600×600×64; scenarios/hour + p50/p99 latency'
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = {0: 208, 1: 112, 2: 112, 3: 416, 4: 104}
STORAGE = ['f32', 'f16', 'bf16', 'f64', 'f16s']
# published reference number for this metric and grid: numba.cuda prototype step at 600x600x64, 10 m, RTX 4090
# (README.md:37 of the reference, data/manifest.json:9468)
REFERENCE_MLUPS_600 = 950.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_pkg():
    spec = importlib.util.spec_from_file_location('aht_pkg', os.path.join(PKG, 'aht.py'))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def build_domain(workload, out):
    subprocess.run(['node', os.path.join(PKG, 'js', 'dump.js'), '--workload', workload, '--out', out], check=True,
                   stdout=subprocess.DEVNULL)


def traffic_from_profiles(kernel_tag):
    """HBM bytes per launch measured by rocprofv3 PMC runs (FETCH_SIZE + WRITE_SIZE, copy-calibrated), if committed"""
    p = os.path.join(ROOT, 'profiles', 'traffic.json')
    if not os.path.exists(p):
        return None
    try:
        d = json.load(open(p))
        return d.get(kernel_tag, {}).get('bytes_per_launch')
    except Exception:
        return None


def run_serve(rate, requests, modes, storage, out):
    """one Node serving-loop process; returns its record (the record is written last: a failure after it shows in
    the exit status, which is reported)"""
    p = subprocess.run(['node', os.path.join(PKG, 'js', 'serve.js'), '--rate', str(rate), '--requests', str(requests),
                        '--mode', modes, '--storage', str(storage), '--out', out], stdout=subprocess.DEVNULL)
    if not os.path.exists(out):
        raise RuntimeError(f'serving loop failed (exit status {p.returncode}) before writing its record')
    return json.load(open(out)), p.returncode


def serving(rate, requests, f16s=True):
    """the serving half of the metric (scenarios/hour, p50/p99 latency): measured live by the Node serving loop
    (js/serve.js: Poisson arrivals at `rate`, the serving grid mix, 2,200 s spin-up + 10-minute mean per request,
    continuous batching against the same requests one at a time with the plain f32 step); with f16s, the same requests
    batched once more with storage per request (serve.js --storage auto: f16s with stochastic rounding for the 10 m
    grids, which carry most of the work, f32 for the coarser ones) against the same f32 sequential run, with the
    largest relative difference of any request's mean warming from its f32 result; rate 0 skips it"""
    if rate <= 0:
        return None
    d = tempfile.mkdtemp(prefix='aht_serve_')
    t0 = time.perf_counter()
    r, rc = run_serve(rate, requests, 'batch,sequential', 0, os.path.join(d, 'serve.json'))
    b, q = r['batch']['summary'], r['sequential']['summary']
    out = {'measured_in_this_run': True, 'rate_req_s': rate, 'requests': requests, 'node_exit_status': rc,
           'wall_s': None,
           'scenarios_per_hour': b['scenarios_per_hour'], 'latency_p50_s': b['latency_p50_s'],
           'latency_p99_s': b['latency_p99_s'],
           'vs_sequential': round(b['scenarios_per_hour'] / q['scenarios_per_hour'], 4),
           'batch': b, 'sequential': q}
    if f16s:
        r16, rc16 = run_serve(rate, requests, 'batch', 'auto', os.path.join(d, 'serve16.json'))
        b16 = r16['batch']['summary']
        ref = {x['id']: x['warmingMean'] for x in r['sequential']['results']}
        diff = max(abs(x['warmingMean'] - ref[x['id']]) / abs(ref[x['id']]) for x in r16['batch']['results'])
        out['batch_f16s'] = dict(b16, node_exit_status=rc16,
                                 vs_sequential_f32=round(b16['scenarios_per_hour'] / q['scenarios_per_hour'], 4),
                                 warming_mean_max_rel_diff_vs_f32=float(f'{diff:.3g}'))
    out['wall_s'] = round(time.perf_counter() - t0, 1)
    subprocess.run(['rm', '-rf', d])
    return out


def timed_step(m, dom, storage, steps, warmup, launch, tune_steps):
    """event-timed steps of one storage format on the domain: (ms per step, launch shape, stats)"""
    s = m.Solver.from_dump(dom, storage=storage)
    try:
        if launch:
            s.set_launch(*(int(x) for x in launch.split(',')))
        else:
            s.step(2)
            log(f'{STORAGE[storage]}: autotuned launch shape', s.autotune(tune_steps))
        s.step(warmup)
        s.sync()
        t0 = time.perf_counter()
        ms = s.benchmark(steps)          # HIP events on the solver's stream around exactly `steps` step launches
        s.sync()
        wall = time.perf_counter() - t0
        return ms / steps, wall * 1e3 / steps, s.launch, s.stats()
    finally:
        s.close()


def cpu_baseline(dom, seconds):
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    o = O.Oracle(dom, 'f32')
    o.init_flow()
    o.step(1)                      # first step outside the timing (page faults of the fresh lattices)
    n, t0 = 0, time.perf_counter()
    while True:
        o.step(1)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    cells = o.N * n
    return {'value': round(cells / el / 1e6, 3), 'unit': 'MLUPS', 'cores': int(O.lib().oracle_num_threads()),
            'kind': 'port',
            'sample': f'{n} whole steps of the same {o.nx}x{o.ny}x{o.nz} domain after one untimed step, '
                      f'{el:.1f} s, f32 oracle (C, OpenMP)'}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--workload', default='c2')
    ap.add_argument('--storage', type=int, default=0,
                    help='0 f32, 1 f16 deviations, 2 bf16 deviations, 3 f64 twin, 4 f16 without theta plane (SR)')
    ap.add_argument('--launch', default=None, help='fixed launch shape block,levels,window[,sweep[,waves]] (no tuning)')
    ap.add_argument('--tune-steps', type=int, default=12, help='steps per launch-shape candidate of the autotuner')
    ap.add_argument('--cpu-seconds', type=float, default=10.0)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--dump', default=None, help='existing domain dump directory')
    ap.add_argument('--serve', type=float, default=5.0,
                    help='also measure the serving loop at this arrival rate (req/s), batch and sequential; 0 = skip')
    ap.add_argument('--serve-requests', type=int, default=16)
    ap.add_argument('--f16', type=int, default=1,
                    help='also time the 16-bit storage formats (f16s, f16) on the same domain')
    a = ap.parse_args()
    if a.gpus != 1:
        print(f'bench.py: one GPU only (asked for --gpus {a.gpus}): scenarios are independent per GPU and a single '
              f'domain would be cut into y-slabs with a one-plane halo exchange per step, which this library does '
              f'not run', file=sys.stderr)
        sys.exit(2)
    m = load_pkg()
    tmp = None
    if a.dump:
        ddir = a.dump
    else:
        tmp = tempfile.mkdtemp(prefix='aht_bench_')
        ddir = tmp
        log('building domain', a.workload)
        build_domain(a.workload, ddir)
    dom = m.load_dump(ddir)
    nx, ny, nz = dom['nx'], dom['ny'], dom['nz']
    N = nx * ny * nz
    nf = int(sum(nz - int(k) for k in dom['ktop']))
    log(f'grid {nx}x{ny}x{nz} cells {N} fluid {nf}')
    ms_step, wall_step, launch, st = timed_step(m, dom, a.storage, a.steps, a.warmup, a.launch, a.tune_steps)
    ms = ms_step * a.steps
    mlups = N * a.steps / (ms * 1e3)
    bpc = BYTES_PER_CELL[a.storage]
    achieved = bpc * nf / (ms_step * 1e-3) / 1e9
    log(f'{ms_step:.4f} ms/step (events), wall {wall_step:.4f} ms/step, {mlups:.1f} MLUPS, '
        f'{achieved:.1f} GB/s, nonfinite {st["non_finite"]}')
    reduced = None
    if a.f16 and a.storage == 0:
        reduced = {}
        for st16, name in ((4, 'f16 deviations, no theta plane, stochastic rounding; f32 arithmetic'),
                           (1, 'f16 deviations with f32 theta plane; f32 arithmetic')):
            ms16, _, launch16, s16 = timed_step(m, dom, st16, a.steps, a.warmup, None, a.tune_steps)
            bpc16 = BYTES_PER_CELL[st16]
            ach16 = bpc16 * nf / (ms16 * 1e-3) / 1e9
            reduced[STORAGE[st16]] = {
                'storage': name, 'ms_per_step': round(ms16, 5), 'value': round(N / (ms16 * 1e3), 2), 'unit': 'MLUPS',
                'speedup_vs_f32': round(ms_step / ms16, 3),
                'roofline': {'bound': 'hbm', 'achieved': round(ach16, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                             'frac': round(ach16 / HBM_PEAK_GBS, 4), 'bytes_per_fluid_cell': bpc16},
                'launch': launch16, 'non_finite': s16['non_finite']}
            log(f'{STORAGE[st16]}: {ms16:.4f} ms/step, {ms_step / ms16:.3f}x f32')
    sweep = bool(launch.get('sweep')) and a.storage == 0      # two steps per launch (opt-in launch shape)
    kernel_tag = f'{"sweep2" if sweep else "step"}_{STORAGE[a.storage]}_{a.workload}'
    traffic = traffic_from_profiles(kernel_tag)
    cpu = None
    if not a.no_cpu:
        log('cpu baseline ...')
        cpu = cpu_baseline(dom, a.cpu_seconds)
        log('cpu baseline', cpu)
    if tmp:
        subprocess.run(['rm', '-rf', tmp])
    serve = None
    if a.serve > 0:
        log(f'serving sample: {a.serve_requests} requests at {a.serve} req/s, batch and sequential ...')
        serve = serving(a.serve, a.serve_requests)
        log('serving', {k: serve[k] for k in ('scenarios_per_hour', 'latency_p50_s', 'latency_p99_s', 'vs_sequential')})
        if 'batch_f16s' in serve:
            log('serving, batched with f16s for the 10 m grids', {k: serve['batch_f16s'][k] for k in (
                'scenarios_per_hour', 'latency_p50_s', 'latency_p99_s', 'vs_sequential_f32',
                'warming_mean_max_rel_diff_vs_f32')})
    out = {
        'metric': METRIC, 'value': round(mlups, 2), 'unit': 'MLUPS', 'n_gpus': 1, 'steps': a.steps,
        'warmup': a.warmup, 'ms_per_step': round(ms_step, 5), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': round(mlups / REFERENCE_MLUPS_600, 3) if (nx, ny, nz) == (600, 600, 64) else None,
        'baseline_note': 'vs the reference numba.cuda prototype, 950 MLUPS at 600x600x64 on an RTX 4090',
        'dtype': ['f32', 'f16 storage / f32 arithmetic', 'bf16 storage / f32 arithmetic', 'f64',
                  'f16 storage (no theta plane) / f32 arithmetic'][a.storage],
        'data': 'synthetic (seeded generator: ' + ('flat ground, no buildings' if a.workload == 'c2flat' and not a.dump
                                                   else 'hills, box buildings, source roofs') + ')',
        'config': {'workload': f'{a.workload}: {nx}x{ny}x{nz} at {dom["dx"]} m, fused D3Q19+D3Q7 step',
                   'cells': N, 'fluid_cells': nf, 'bytes_per_fluid_cell': bpc, 'dt_s': dom['dt'],
                   'storage': STORAGE[a.storage], 'launch': launch, 'parallelism': 'single GPU'},
        'fluid_mlups': round(nf * a.steps / (ms * 1e3), 2),
        'wall_ms_per_step': round(wall_step, 5),
        'roofline': {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': round(achieved / HBM_PEAK_GBS, 4), 'traffic': traffic,
                     'traffic_source': 'profile: rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0 (copy-calibrated) per '
                                       'launch, profiles/traffic.json; not measured in this run',
                     'steps_per_launch': 2 if sweep else 1},
        'cpu_baseline': cpu,
        'reduced_precision': reduced,
        'serving': serve,
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
