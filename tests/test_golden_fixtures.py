"""The committed per-cell golden vectors (tests/golden/tiny_*.npz) against the oracle and the Node host.

* The oracle, rebuilt and run from the fixture's own columns and parameters, reproduces the recorded states bit for
  bit (f32 after 1 and 10 steps, f64 after 10 steps): any change to the oracle's arithmetic fails here.
* The Node host rebuilds the fixture's domain from the recorded dump flags identically (columns, profile, params).
* The f64 and f32 records agree to f32 accuracy (relative L2 <= 1e-5 after 10 steps): the two builds share one
  expression list, so a dropped term or wrong index in one of them shows here.
"""
import json

import numpy as np
import pytest

from conftest import dump_domain, rel_l2
from golden_fixtures import NAMES, fluid_mask, load


@pytest.mark.parametrize('name', NAMES)
def test_oracle_reproduces_fixture(O, name):
    dom, rec = load(name)
    fl = fluid_mask(dom)
    o = O.Oracle(dom, 'f32')
    o.init_flow()
    o.step(1)
    assert np.array_equal(o.f[:, fl].view(np.uint32), rec['f1'].view(np.uint32))
    assert np.array_equal(o.g[:, fl].view(np.uint32), rec['g1'].view(np.uint32))
    o.step(9)
    assert np.array_equal(o.f[:, fl].view(np.uint32), rec['f10'].view(np.uint32))
    assert np.array_equal(o.g[:, fl].view(np.uint32), rec['g10'].view(np.uint32))
    o64 = O.Oracle(dom, 'f64')
    o64.init_flow()
    o64.step(10)
    assert np.array_equal(o64.f[:, fl], rec['f10_f64']) and np.array_equal(o64.g[:, fl], rec['g10_f64'])
    assert rel_l2(rec['f10'], rec['f10_f64']) <= 1e-5 and rel_l2(rec['g10'], rec['g10_f64']) <= 1e-5


@pytest.mark.parametrize('name', NAMES)
def test_node_host_rebuilds_fixture_domain(tmp_path, name):
    dom, _ = load(name)
    d = dump_domain(tmp_path, *dom['flags'])
    assert (d['nx'], d['ny'], d['nz']) == (dom['nx'], dom['ny'], dom['nz']) and d['dt'] == dom['dt']
    for k in ('kt', 'ktop', 'cd', 'q', 'prof'):
        assert np.array_equal(d[k], dom[k]), k
    assert json.dumps(d['params'], sort_keys=True) == json.dumps(dom['params'], sort_keys=True)
