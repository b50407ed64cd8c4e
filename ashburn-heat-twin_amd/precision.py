#!/usr/bin/env python3
"""Error of reduced-precision population storage against the FP64 twin on the same domain and steps.

Default workload: 1200 x 1200 x 96 at 5 m (138 M cells), the seeded synthetic box with 16,000 buildings and 850 source
roofs, 5,000 steps from rest; the near-ground warming is the time mean of theta at the first fluid cell of every
column over the last --mean-steps steps (FP64 accumulators on the GPU). Each storage format runs the identical
sequence; reported against the FP64 twin (f64 storage and arithmetic):
  * theta_rel_l2     : relative L2 of the final theta (the last step's pre-collision theta) over fluid cells;
  * warming_mean_rel : relative error of the spatial mean of the time-mean near-ground theta ("mean warming");
  * warming_peak_rel : relative error of its maximum; warming_rel_l2: relative L2 of the time-mean surface field;
  * ms_per_step      : event-timed step time of that storage on this grid.
Prints one JSON line per storage (progress on stderr).
Usage: python precision.py [--workload c3] [--steps 5000] [--mean-steps 2500] [--storages f32,f16,bf16]
"""
import argparse
import importlib.util
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_pkg():
    spec = importlib.util.spec_from_file_location('aht_pkg', os.path.join(HERE, 'aht.py'))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def run(m, dom, storage, steps, mean_steps, chunk=500):
    """returns (final theta over fluid cells, time-mean surface theta, ms per step)"""
    with m.Solver.from_dump(dom, storage=storage) as s:
        t_dev = 0.0
        done = 0
        spin = steps - mean_steps
        while done < spin:
            n = min(chunk, spin - done)
            t_dev += s.benchmark(n)
            done += n
            log(f'  {storage}: step {done}/{steps}')
        s.mean_reset()
        while done < steps - 1:
            n = min(chunk, steps - 1 - done)
            s.step(n, acc=True)
            done += n
            s.sync()
            log(f'  {storage}: step {done}/{steps} (averaging)')
        s.step(1, acc=True, out=True)
        out = s.read_out()
        surf3, budget = s.read_mean()
        st = s.stats()
    fl = out[0] != 0
    return out[0], fl, surf3[0], t_dev / max(spin, 1), st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--workload', default='c3')
    ap.add_argument('--steps', type=int, default=5000)
    ap.add_argument('--mean-steps', type=int, default=2500)
    ap.add_argument('--storages', default='f32,f16,bf16')
    ap.add_argument('--dump', default=None)
    a = ap.parse_args()
    m = load_pkg()
    ddir = a.dump
    if not ddir:
        ddir = tempfile.mkdtemp(prefix='aht_prec_')
        subprocess.run(['node', os.path.join(HERE, 'js', 'dump.js'), '--workload', a.workload, '--out', ddir], check=True,
                       stdout=subprocess.DEVNULL)
    dom = m.load_dump(ddir)
    log(f'grid {dom["nx"]}x{dom["ny"]}x{dom["nz"]}, {a.steps} steps, mean over the last {a.mean_steps}')
    t0 = time.time()
    th64, fl64, sm64, ms64, st64 = run(m, dom, 'f64', a.steps, a.mean_steps)
    ref = {'storage': 'f64', 'ms_per_step': round(ms64, 4), 'warming_mean': float(sm64.mean()),
           'warming_peak': float(sm64.max()), 'non_finite': st64['non_finite'], 'wall_s': round(time.time() - t0, 1)}
    print(json.dumps(ref), flush=True)
    th64f = th64[fl64].astype(np.float64)
    for stg in a.storages.split(','):
        t0 = time.time()
        th, fl, sm, ms, st = run(m, dom, stg, a.steps, a.mean_steps)
        d = th[fl64].astype(np.float64) - th64f
        res = {
            'storage': stg, 'ms_per_step': round(ms, 4), 'speedup_vs_f64': round(ms64 / ms, 3),
            'theta_rel_l2': float(np.linalg.norm(d) / np.linalg.norm(th64f)),
            'theta_max_abs_err': float(np.abs(d).max()),
            'warming_mean': float(sm.mean()),
            'warming_mean_rel': float(abs(sm.mean() - sm64.mean()) / abs(sm64.mean())),
            'warming_peak_rel': float(abs(sm.max() - sm64.max()) / abs(sm64.max())),
            'warming_rel_l2': float(np.linalg.norm(sm - sm64) / np.linalg.norm(sm64)),
            'non_finite': st['non_finite'], 'wall_s': round(time.time() - t0, 1),
        }
        print(json.dumps(res), flush=True)
    if not a.dump:
        subprocess.run(['rm', '-rf', ddir])


if __name__ == '__main__':
    main()
