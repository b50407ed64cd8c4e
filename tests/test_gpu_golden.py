"""The HIP step against the committed per-cell golden vectors (tests/golden/tiny_*.npz), through the C-ABI.

The data alone is the reference here (no oracle call): from the initial state, the f32 kernel after 1 step within
relative L2 1e-6 of the recorded f32 state, after 10 steps within 1e-5; the FP64 twin after 10 steps within 1e-12
of the recorded f64 state (reassociation only).
"""
import numpy as np
import pytest

from conftest import rel_l2
from golden_fixtures import NAMES, fluid_mask, load

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('name', NAMES)
def test_step_matches_golden_vectors(pkg, name):
    dom, rec = load(name)
    fl = fluid_mask(dom)
    with pkg.Solver.from_dump(dom) as s:
        s.step(1)
        f, g = s.read_state()
        assert rel_l2(f[:, fl], rec['f1']) <= 1e-6 and rel_l2(g[:, fl], rec['g1']) <= 1e-6
        s.step(9)
        f, g = s.read_state()
        assert rel_l2(f[:, fl], rec['f10']) <= 1e-5 and rel_l2(g[:, fl], rec['g10']) <= 1e-5
    with pkg.Solver.from_dump(dom, storage='f64') as s:
        s.step(10)
        f, g = s.read_state_f64()
        assert rel_l2(f[:, fl], rec['f10_f64']) <= 1e-12 and rel_l2(g[:, fl], rec['g10_f64']) <= 1e-12
