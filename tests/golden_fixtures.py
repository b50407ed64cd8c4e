"""Loader of the per-cell golden fixtures tests/golden/tiny_<name>.{json,npz} (written by
tests/golden/make_tiny_fixtures.py): a domain dict in the shape of a Node host dump plus the recorded states."""
import glob
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
NAMES = sorted(os.path.basename(p)[5:-4] for p in glob.glob(os.path.join(GOLDEN, 'tiny_*.npz')))


def load(name):
    meta = json.load(open(os.path.join(GOLDEN, f'tiny_{name}.json')))
    z = np.load(os.path.join(GOLDEN, f'tiny_{name}.npz'))
    dom = dict(meta)
    for k in ('kt', 'ktop', 'cd', 'q', 'prof'):
        dom[k] = z[k]
    rec = {k: z[k] for k in ('f1', 'g1', 'f10', 'g10', 'f10_f64', 'g10_f64')}
    return dom, rec


def fluid_mask(dom):
    k = np.arange(dom['nz'])[:, None]
    return (k >= dom['ktop'][None, :].astype(np.int64)).reshape(-1)
