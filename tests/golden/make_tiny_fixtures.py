#!/usr/bin/env python3
"""Writes the per-cell golden vectors of small synthetic domains: tests/golden/tiny_<name>.npz (+ .json).

Each fixture is one seeded domain built by the Node host (columns kt/ktop/cd/q, the parameter block, the inflow
profile) and the CPU oracle's state after 1 and after 10 steps from the initial state, in the f32 build (the
reference kernel's expressions and rounding) and, after 10 steps, the f64 build (the same expressions in double).
Populations are stored over fluid cells only, [19][Nf] and [7][Nf], in cell order. The domains are chosen so that
every boundary branch of the step is hit: inflow faces on two sides, backflow-clipped and prototype outlets, the
lid, terrain and buildings with bounce-back, log-law drag and rooftop sources.

Usage: python tests/golden/make_tiny_fixtures.py   (a few seconds)
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, 'oracle'))

CASES = {  # name: dump.js flags
    'sw_clip': ['--tiny', '14,12,9,11,5,3', '--dx', 30, '--dir', 210, '--u10', 6, '--oclip', 1, '--nsw', 3,
                '--flux', 4000],
    'n_proto': ['--tiny', '12,14,8,23,4,2', '--dx', 30, '--dir', 0, '--u10', 7, '--oclip', 0, '--nsw', 4,
                '--flux', 3000],
}


def main():
    import oracle as O
    for name, flags in CASES.items():
        d = tempfile.mkdtemp(prefix='aht_tiny_')
        subprocess.run(['node', os.path.join(ROOT, 'ashburn-heat-twin_amd', 'js', 'dump.js'), '--out', d]
                       + [str(x) for x in flags], check=True, stdout=subprocess.DEVNULL)
        dom = O.load_dump(d)
        meta = json.load(open(os.path.join(d, 'domain.json')))
        subprocess.run(['rm', '-rf', d])
        o32 = O.Oracle(dom, 'f32')
        o32.init_flow()
        fl = o32.fluid_mask()
        o32.step(1)
        f1, g1 = o32.f[:, fl].copy(), o32.g[:, fl].copy()
        o32.step(9)
        f10, g10 = o32.f[:, fl].copy(), o32.g[:, fl].copy()
        o64 = O.Oracle(dom, 'f64')
        o64.init_flow()
        o64.step(10)
        np.savez_compressed(os.path.join(HERE, f'tiny_{name}.npz'),
                            kt=dom['kt'], ktop=dom['ktop'], cd=dom['cd'], q=dom['q'], prof=dom['prof'],
                            f1=f1, g1=g1, f10=f10, g10=g10, f10_f64=o64.f[:, fl], g10_f64=o64.g[:, fl])
        keep = {k: meta[k] for k in ('nx', 'ny', 'nz', 'dx', 'dt', 'conv', 'params')}
        keep['flags'] = [str(x) for x in flags]
        keep['n_fluid'] = int(fl.sum())
        with open(os.path.join(HERE, f'tiny_{name}.json'), 'w') as fh:
            json.dump(keep, fh, indent=1)
        print(name, dom['nx'], dom['ny'], dom['nz'], 'fluid', int(fl.sum()), 'inflow',
              {k: meta['params'][k] for k in ('inflow_w', 'inflow_e', 'inflow_s', 'inflow_n')})


if __name__ == '__main__':
    main()
