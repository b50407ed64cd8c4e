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
BYTES_PER_CELL = {0: 208, 1: 112, 2: 112, 3: 416}
STORAGE = ['f32', 'f16', 'bf16', 'f64']
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


def serving(rate):
    """the serving half of the metric (scenarios/hour, p50/p99 latency): measured live by the Node serving loop when
    rate > 0 (16 requests, batch and sequential, a few minutes), else the committed record of profiles/serving.json"""
    if rate > 0:
        out = os.path.join(tempfile.mkdtemp(prefix='aht_serve_'), 'serve.json')
        subprocess.run(['node', os.path.join(PKG, 'js', 'serve.js'), '--rate', str(rate), '--requests', '16',
                        '--mode', 'batch,sequential', '--out', out], check=True, stdout=subprocess.DEVNULL)
        d = json.load(open(out))
        subprocess.run(['rm', '-rf', os.path.dirname(out)])
        return {'measured_in_this_run': True, 'rate_req_s': rate,
                'batch': d['batch']['summary'], 'sequential': d['sequential']['summary']}
    p = os.path.join(ROOT, 'profiles', 'serving.json')
    if not os.path.exists(p):
        return None
    d = json.load(open(p))['summaries']
    return {'measured_in_this_run': False, 'source': 'profiles/serving_c5.md', 'rate_req_s': 5,
            'batch': d.get('batch_r5'), 'sequential': d.get('sequential_r5')}


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
    ap.add_argument('--storage', type=int, default=0, help='0 f32, 1 f16 deviations, 2 bf16 deviations, 3 f64 twin')
    ap.add_argument('--launch', default=None, help='fixed launch shape block,levels,window (no tuning)')
    ap.add_argument('--tune-steps', type=int, default=12, help='steps per launch-shape candidate of the autotuner')
    ap.add_argument('--cpu-seconds', type=float, default=12.0)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--dump', default=None, help='existing domain dump directory')
    ap.add_argument('--serve', type=float, default=0.0,
                    help='also measure the serving loop at this arrival rate (req/s); 0 = report the committed record')
    a = ap.parse_args()
    if a.gpus != 1:
        print(f'bench.py: one GPU only (asked for --gpus {a.gpus}): scenarios are independent per GPU and a single '
              f'domain would be cut into y-slabs with a one-plane halo exchange per step, which this build does not '
              f'run', file=sys.stderr)
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
    s = m.Solver.from_dump(dom, storage=a.storage)
    if a.launch:
        b, l, w = (int(x) for x in a.launch.split(','))
        s.set_launch(b, l, w)
    else:
        s.step(2)
        log('autotuned launch shape', s.autotune(a.tune_steps))
    s.step(a.warmup)
    s.sync()
    t0 = time.perf_counter()
    ms = s.benchmark(a.steps)          # HIP events on the solver's stream around exactly K step launches
    s.sync()
    wall = time.perf_counter() - t0
    st = s.stats()
    launch = s.launch
    s.close()
    ms_step = ms / a.steps
    mlups = N * a.steps / (ms * 1e3)
    bpc = BYTES_PER_CELL[a.storage]
    achieved = bpc * nf / (ms_step * 1e-3) / 1e9
    log(f'{ms_step:.4f} ms/step (events), wall {wall * 1e3 / a.steps:.4f} ms/step, {mlups:.1f} MLUPS, '
        f'{achieved:.1f} GB/s, nonfinite {st["non_finite"]}')
    kernel_tag = f'step_{STORAGE[a.storage]}_{a.workload}'
    traffic = traffic_from_profiles(kernel_tag)
    cpu = None
    if not a.no_cpu:
        log('cpu baseline ...')
        cpu = cpu_baseline(dom, a.cpu_seconds)
        log('cpu baseline', cpu)
    if tmp:
        subprocess.run(['rm', '-rf', tmp])
    out = {
        'metric': METRIC, 'value': round(mlups, 2), 'unit': 'MLUPS', 'n_gpus': 1, 'steps': a.steps,
        'warmup': a.warmup, 'ms_per_step': round(ms_step, 5), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': round(mlups / REFERENCE_MLUPS_600, 3) if (nx, ny, nz) == (600, 600, 64) else None,
        'baseline_note': 'vs the reference numba.cuda prototype, 950 MLUPS at 600x600x64 on an RTX 4090',
        'dtype': ['f32', 'f16 storage / f32 arithmetic', 'bf16 storage / f32 arithmetic', 'f64'][a.storage],
        'data': 'synthetic (seeded generator: hills, box buildings, source roofs)',
        'config': {'workload': f'{a.workload}: {nx}x{ny}x{nz} at {dom["dx"]} m, fused D3Q19+D3Q7 step',
                   'cells': N, 'fluid_cells': nf, 'bytes_per_fluid_cell': bpc, 'dt_s': dom['dt'],
                   'storage': STORAGE[a.storage], 'launch': launch, 'parallelism': 'single GPU'},
        'fluid_mlups': round(nf * a.steps / (ms * 1e3), 2),
        'wall_ms_per_step': round(wall * 1e3 / a.steps, 5),
        'roofline': {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': round(achieved / HBM_PEAK_GBS, 4), 'traffic': traffic},
        'cpu_baseline': cpu,
        'serving': serving(a.serve),
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
