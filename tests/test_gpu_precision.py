"""Reduced-precision storage and the FP64 twin against the CPU oracle.

* The FP64 twin (f64 storage and arithmetic) against the oracle's f64 build (`_f64`, the generator's expressions in
  double): the initial state is bit-identical, steps agree to relative L2 <= 1e-12 (reassociation only).
* f16 / bf16 deviation storage and f16 without theta plane (f32 arithmetic) against the FP64 twin on the same seeded
  domain: the state round trip is exact to the format's resolution, and after 200 steps theta stays within the
  format's accuracy class (relative L2 <= 2e-3 for f16, <= 5e-3 for bf16, <= 2e-3 for f16 without theta plane; the
  last two round stochastically in the step). The full-size statement (1200 x 1200 x 96) is made by
  ashburn-heat-twin_amd/precision.py and test_gpu_configs.py.
"""
import numpy as np
import pytest

from conftest import dump_domain, rel_l2

pytestmark = pytest.mark.gpu


def tiny(tmp_path, wdir=210, oclip=1):
    return dump_domain(tmp_path, '--tiny', '24,20,12,7', '--dx', 30, '--dir', wdir, '--u10', 6, '--oclip', oclip,
                       '--nsw', 5, '--flux', 4000)


@pytest.mark.parametrize('wdir,oclip', [(210, 1), (30, 0)])
def test_fp64_twin_matches_f64_oracle(tmp_path, pkg, O, wdir, oclip):
    dom = tiny(tmp_path, wdir, oclip)
    o = O.Oracle(dom, 'f64')
    o.init_flow()
    with pkg.Solver.from_dump(dom, storage='f64') as s:
        f, g = s.read_state_f64()
        assert np.array_equal(f, o.f) and not g.any()
        ro = o.step(1, out=True)
        s.step(1, out=True)
        go = s.read_out()
        fl = o.fluid_mask()
        assert rel_l2(go[0, fl], ro[0, fl].astype(np.float32)) <= 1e-6      # outputs are f32
        f, g = s.read_state_f64()
        assert rel_l2(f[:, fl], o.f[:, fl]) <= 1e-12 and rel_l2(g[:, fl], o.g[:, fl]) <= 1e-12
        o.step(49)
        s.step(49)
        f, g = s.read_state_f64()
        assert rel_l2(f[:, fl], o.f[:, fl]) <= 1e-11 and rel_l2(g[:, fl], o.g[:, fl]) <= 1e-11


@pytest.mark.parametrize('storage,quant', [('f16', 2.0 ** -11), ('bf16', 2.0 ** -8), ('f16s', 2.0 ** -11)])
def test_deviation_storage_round_trip(tmp_path, pkg, O, storage, quant):
    dom = tiny(tmp_path)
    o = O.Oracle(dom, 'f32')
    o.init_flow()
    o.step(50)
    w = O.W19.astype(np.float32)[:, None]
    with pkg.Solver.from_dump(dom, storage=storage) as s:
        s.write_state(o.f, o.g)
        f, g = s.read_state()
    # deviation formats store (f - w) * 2^10, theta = sum_q g_q in f32 and (g - w7 theta) * 2^10: errors bounded by
    # half a unit of the format's precision on the deviation (plus the f32 rounding of adding the reference back,
    # and the binary16 subnormal step for tiny deviations)
    assert np.all(np.abs(f - o.f) <= quant * np.abs(o.f - w) + 1e-7 * np.abs(o.f) + 1e-9)
    w7 = np.array([0.25] + [0.125] * 6, np.float32)[:, None]
    th = o.g.astype(np.float64).sum(0).astype(np.float32)
    if storage == 'f16s':    # no theta plane: g_q * 2^10 itself in binary16
        assert np.all(np.abs(g - o.g) <= quant * np.abs(o.g) + 1e-9)
    else:
        assert np.all(np.abs(g - o.g) <= quant * np.abs(o.g - w7 * th) + 2e-7 * np.abs(o.g) + 1e-9)


@pytest.mark.parametrize('storage,tol', [('f16', 2e-3), ('bf16', 5e-3), ('f16s', 2e-3), ('f32', 1e-5)])
def test_reduced_precision_against_fp64_twin(tmp_path, pkg, storage, tol):
    dom = tiny(tmp_path)
    with pkg.Solver.from_dump(dom, storage='f64') as ref, pkg.Solver.from_dump(dom, storage=storage) as s:
        ref.step(200, out=True)
        s.step(200, out=True)
        a, b = ref.read_out(), s.read_out()
        st = s.stats()
    fl = a[0] != 0
    err = rel_l2(b[0, fl], a[0, fl])
    print(f'{storage}: theta rel L2 vs FP64 twin after 200 steps = {err:.3e}')
    assert st['non_finite'] == 0
    assert err <= tol
