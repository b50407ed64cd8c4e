"""Two steps per launch (the z sweep with LDS-staged halo planes) against one step per launch.

The bar is exact: the sweep recomputes nothing differently, it only keeps the intermediate state on chip, so states,
debug outputs and FP64 means must be bit-identical to the one-step kernel for every region shape (32 x 8 and 16 x 16
columns), level chunk (1 level up to the whole column, so chunk ends fall on every kind of plane) and tile order,
on tiny domains with inflow and outlet faces (with and without backflow clipping), buildings up to the lid, on the
reference's own Ashburn domain at 30 m, and on the 600 x 600 x 64 workload; odd step counts end with a one-step
launch, and steps that accumulate means or write debug outputs take the one-step kernel.
"""
import numpy as np
import pytest

from conftest import dump_domain

pytestmark = pytest.mark.gpu

SHAPES = [('24,20,12,7', 210, 1), ('40,28,14,3', 30, 0), ('16,36,10,5', 95, 1), ('70,9,8,11', 300, 0)]


def tiny(tmp_path, k):
    shape, wdir, oclip = SHAPES[k]
    return dump_domain(tmp_path / f'd{k}', '--tiny', shape, '--dx', 30, '--dir', wdir, '--u10', 5, '--oclip', oclip,
                       '--flux', 3000)


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def run_pair(pkg, dom, sweep, window, steps, out=False):
    """(one-step state, sweep state) after `steps` steps of the same domain"""
    res = []
    for sw in (0, sweep):
        with pkg.Solver.from_dump(dom) as s:
            s.set_launch(256, 1, window, sw)
            s.step(steps, out=out)
            f, g = s.read_state()
            res.append((f, g, s.read_out() if out else None, s.steps))
    return res


@pytest.mark.parametrize('k', range(len(SHAPES)))
@pytest.mark.parametrize('sweep,window', [(64, 1), (-64, 4), (3, 16), (-2, 0), (1, 1), (-1, 2), (5, 1)])
def test_sweep_matches_single_steps_bitwise(tmp_path, pkg, k, sweep, window):
    dom = tiny(tmp_path, k)
    (f0, g0, _, n0), (f1, g1, _, n1) = run_pair(pkg, dom, sweep, window, 9)
    assert n0 == n1 == 9
    assert same(f0, f1), (SHAPES[k], sweep, window)
    assert same(g0, g1), (SHAPES[k], sweep, window)


def test_sweep_with_outputs_and_means(tmp_path, pkg):
    """debug outputs on the last step and FP64 means: the steps that need them run one at a time"""
    dom = tiny(tmp_path, 0)
    res = []
    for sw in (0, 4, -3):
        with pkg.Solver.from_dump(dom) as s:
            s.set_launch(256, 1, 1, sw)
            s.step(10, out=True)
            o = s.read_out()
            s.mean_reset()
            s.step(5, acc=True)
            s.step(6)
            f, g = s.read_state()
            s3, b = s.read_mean()
            res.append((f, g, o, s3, b, s.steps))
    for r in res[1:]:
        assert same(r[0], res[0][0]) and same(r[1], res[0][1]) and same(r[2], res[0][2])
        assert np.array_equal(r[3], res[0][3]) and r[4] == res[0][4] and r[5] == res[0][5] == 21


def test_sweep_on_the_reference_domain(pkg, ashburn30):
    """the Ashburn static10 domain at 30 m (terrain, several thousand buildings, prototype outlet)"""
    for sweep, window in ((16, 4), (-32, 1)):
        (f0, g0, o0, _), (f1, g1, o1, _) = run_pair(pkg, ashburn30, sweep, window, 7, out=True)
        assert same(f0, f1) and same(g0, g1) and same(o0, o1), (sweep, window)


def test_sweep_at_full_size(tmp_path, pkg):
    """600 x 600 x 64 (the single-GPU benchmark workload): 6 steps, every population compared"""
    dom = dump_domain(tmp_path / 'c2', '--workload', 'c2')
    (f0, g0, _, _), (f1, g1, _, _) = run_pair(pkg, dom, 32, 16, 6)
    assert same(f0, f1) and same(g0, g1)


def test_sweep_launch_field(tmp_path, pkg):
    """the launch shape carries the sweep; out-of-range chunks are rejected; other formats ignore it"""
    dom = tiny(tmp_path, 2)
    with pkg.Solver.from_dump(dom) as s:
        s.set_launch(128, 2, 4, -16)
        assert s.launch == {'block': 128, 'levels': 2, 'window': 4, 'sweep': -16, 'waves': 0}
        with pytest.raises(pkg.AhtError):
            s.set_launch(256, 1, 1, 5000)
    with pkg.Solver.from_dump(dom, storage='f16s') as a, pkg.Solver.from_dump(dom, storage='f16s') as b:
        b.set_launch(256, 1, 1, 8)
        a.step(5)
        b.step(5)
        fa, ga = a.read_state()
        fb, gb = b.read_state()
        assert same(fa, fb) and same(ga, gb)
