"""Batched stepping of uneven domains in one launch per step (hipGraph-captured chunks).

The bar is exact: a member's state and FP64 means after a batched run must be bit-identical to the same domain run
alone through aht_step with the same steps and averaging window, whatever the mix, the chunking, the block size and
grid (online-tuned or fixed), the members' own launch shapes (which the batched launch does not use), graph capture on
or off, and changes of a member's columns or launch shape between chunks. Tiny domains of different sizes, winds and outlets.
"""
import numpy as np
import pytest

from conftest import dump_domain

pytestmark = pytest.mark.gpu

SHAPES = [('24,20,12,7', 210, 1), ('40,28,14,3', 30, 0), ('16,36,10,5', 95, 1)]


def domains(tmp_path):
    out = []
    for k, (shape, wdir, oclip) in enumerate(SHAPES):
        out.append(dump_domain(tmp_path / f'd{k}', '--tiny', shape, '--dx', 30, '--dir', wdir, '--u10', 5,
                               '--oclip', oclip, '--flux', 3000))
    return out


def alone(pkg, dom, steps, acc_from, storage=0):
    with pkg.Solver.from_dump(dom, storage=storage) as s:
        s.step(acc_from)
        s.mean_reset()
        s.step(steps - acc_from, acc=True)
        f, g = s.read_state()
        s3, b = s.read_mean()
    return f, g, s3, b


@pytest.mark.parametrize('storage', [0, 1, 2, 4])
def test_batch_matches_alone_bitwise(tmp_path, pkg, storage):
    """batched members (stream bases formed on the device, blocked layout of every format) give the bits of the
    plain step (bases from the host-built table); for the stochastic-rounding formats (bf16, f16 without theta plane)
    this also checks that a member's random bits follow its own step count"""
    doms = domains(tmp_path)
    steps = [37, 52, 45]
    acc_from = [20, 30, 0]
    ref = [alone(pkg, d, n, a, storage) for d, n, a in zip(doms, steps, acc_from)]
    # (graph, grid, block, levels per member): tuned, fixed persistent grids, one block per tile, 64/128/256 threads
    for use_graph, grid, block, levels in ((True, 0, 0, (1, 1, 1)), (False, 64, 128, (2, 1, 4)),
                                           (True, 3, 64, (1, 4, 2)), (True, 1 << 30, 256, (4, 2, 1))):
        sol = [pkg.Solver.from_dump(d, storage=storage) for d in doms]
        with pkg.Batch(storage=storage) as b:
            for s, lv in zip(sol, levels):
                s.set_launch(256, lv, 1)
                s.mean_reset()
                b.add(s)
            b.set_grid(grid, use_graph)
            b.set_block(block)
            done = [0, 0, 0]
            chunk = 10
            while any(d < n for d, n in zip(done, steps)):
                ns = [min(chunk, n - d) for d, n in zip(done, steps)]
                af = [max(0, a - d) for d, a in zip(done, acc_from)]
                b.step(chunk, ns, af)
                done = [d + k for d, k in zip(done, ns)]
            b.sync()
            for k, s in enumerate(sol):
                f, g = s.read_state()
                s3, bud = s.read_mean()
                case = (use_graph, grid, block, levels, k)
                assert np.array_equal(f.view(np.uint32), ref[k][0].view(np.uint32)), case
                assert np.array_equal(g.view(np.uint32), ref[k][1].view(np.uint32)), case
                assert np.array_equal(s3, ref[k][2]), case
                assert bud['n_acc'] == ref[k][3]['n_acc'] == steps[k] - acc_from[k]
                assert bud['heat_out'] == ref[k][3]['heat_out'] and bud['content'] == ref[k][3]['content'], case
                assert s.steps == steps[k]
            if block:
                assert b.block == block
            for s in sol:
                b.remove(s)
        for s in sol:
            s.close()


@pytest.mark.parametrize('storage', [0, 1, 2, 4])
def test_qcell_layout_matches_blocked_layout(tmp_path, pkg, storage):
    """Lattice pairs below 4 GiB are blocked on the device (f32 in 2048-cell blocks, f16 / bf16 in 256-cell blocks with
    the theta plane after the array) unless the reference's [q][cell] layout is asked for (the layout of larger
    pairs): both give the same bits, state and FP64 means alike, through the whole step, read-back and state-exchange
    path"""
    dom = domains(tmp_path)[0]
    out = []
    for layout in (0, 1):
        with pkg.Solver.from_dump(dom, storage=storage, layout=layout) as s:
            s.step(25)
            s.mean_reset()
            s.step(30, acc=True)
            f, g = s.read_state()
            s3, b = s.read_mean()
            th, u, v = s.read_surface(dom['conv'])
            st = s.stats()
            s.write_state(f, g)
            s.step(5)
            f2, g2 = s.read_state()
        out.append((f, g, s3, b, th, u, v, st, f2, g2))
    a, q = out
    for x, y in zip(a[:3] + a[4:7] + a[8:], q[:3] + q[4:7] + q[8:]):
        assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))
    assert a[3] == q[3] and a[7] == q[7]


@pytest.mark.parametrize('storage', [0, 1, 2, 4])
def test_reset_heat_zeroes_theta_populations_only(tmp_path, pkg, storage):
    """aht_reset_heat sets every D3Q7 population (and the theta plane of the 16-bit formats) to zero and leaves the
    D3Q19 populations as they were, in the blocked layout (one array per lattice pair, the D3Q7 populations of each
    block after its D3Q19 ones; this 5,760-cell domain also has a partly filled last block) as in [q][cell]."""
    dom = domains(tmp_path)[0]
    for layout in (0, 1):
        with pkg.Solver.from_dump(dom, storage=storage, layout=layout) as s:
            s.step(25)
            f0, g0 = s.read_state()
            assert np.abs(g0).max() > 0
            s.reset_heat()
            f1, g1 = s.read_state()
            assert np.array_equal(f1, f0), layout
            assert not np.any(g1), layout
            s.step(3)                       # and the step runs on from there
            assert s.stats()['non_finite'] == 0


def test_members_join_and_leave_mid_run(tmp_path, pkg):
    """continuous batching: a member added after others have stepped, another removed early"""
    doms = domains(tmp_path)
    ref0 = alone(pkg, doms[0], 40, 40)
    ref2 = alone(pkg, doms[2], 20, 20)
    s0, s1, s2 = (pkg.Solver.from_dump(d) for d in doms)
    with pkg.Batch() as b:
        b.add(s0)
        b.add(s1)
        b.step(20)
        b.remove(s1)                     # leaves after 20 steps
        b.add(s2)                        # joins now
        b.step(20)
        b.sync()
        f0, g0 = s0.read_state()
        f2, g2 = s2.read_state()
        assert np.array_equal(f0, ref0[0]) and np.array_equal(g0, ref0[1])
        assert np.array_equal(f2, ref2[0]) and np.array_equal(g2, ref2[1])
        assert s1.steps == 20 and s0.steps == 40 and s2.steps == 20
    for s in (s0, s1, s2):
        s.close()


def test_member_columns_and_levels_change_between_chunks(tmp_path, pkg):
    """A captured chunk is replayed only while every launch keeps its member and tile counts: raising a member's lowest
    fluid level (fewer tiles) between chunks re-captures, and the member, whose own launch shape also changes, still
    gives the bits of the same changes applied alone"""
    doms = domains(tmp_path)
    d0 = doms[0]
    ktop2 = np.maximum(d0['ktop'], d0['ktop'].min() + 2).astype(np.uint32)    # lowest fluid level up by two
    kt2 = np.minimum(d0['kt'], ktop2).astype(np.uint32)

    def run(batched):
        s0, s1 = pkg.Solver.from_dump(d0), pkg.Solver.from_dump(doms[1])
        b = pkg.Batch() if batched else None
        if b:
            b.set_block(128)
            b.add(s0)
            b.add(s1)
        chunk = lambda: b.step(8) if b else (s0.step(8), s1.step(8))
        chunk()
        chunk()
        if b:
            b.sync()
        s0.set_columns(kt2, ktop2, d0['cd'], d0['q'])
        chunk()
        s0.set_launch(256, 2, 1)
        chunk()
        if b:
            b.sync()
        out = [s.read_state() for s in (s0, s1)]
        if b:
            b.close()
        s0.close()
        s1.close()
        return out

    ref, got = run(False), run(True)
    for (fa, ga), (fb, gb) in zip(ref, got):
        assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32))
        assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32))


def test_autotune_restores_state(tmp_path, pkg):
    dom = domains(tmp_path)[0]
    with pkg.Solver.from_dump(dom) as s:
        s.step(5)
        f0, g0 = s.read_state()
        best = s.autotune(3)
        assert best['block'] in (64, 128, 256) and best['levels'] in (1, 2, 4) and best['window'] in (-1, 0, 1, 2, 4, 8, 16)
        assert best['waves'] in (0, 4, 5)
        assert s.launch == best
        f1, g1 = s.read_state()
        assert np.array_equal(f0, f1) and np.array_equal(g0, g1) and s.steps == 5


@pytest.mark.parametrize('block,levels,window,waves', [(256, 1, 1, 0), (128, 2, 4, 0), (64, 4, 0, 0), (192, 1, 16, 0),
                                                        (256, 4, 2, 0), (256, 1, -1, 0), (64, 2, -1, 0), (256, 1, 1, 4),
                                                        (128, 1, 8, 5), (64, 2, 1, 3), (192, 1, 1, 4)])
def test_launch_shapes_give_identical_states(tmp_path, pkg, block, levels, window, waves):
    """every launch shape (and cap of wavefronts per SIMD) updates every cell with the same arithmetic: states and FP64
    means are bit-identical to the default"""
    dom = domains(tmp_path)[2]
    with pkg.Solver.from_dump(dom) as a, pkg.Solver.from_dump(dom) as b:
        b.set_launch(block, levels, window, 0, waves)
        assert b.launch['waves'] == waves
        for s in (a, b):
            s.step(6)
            s.mean_reset()
            s.step(7, acc=True)
        fa, ga = a.read_state()
        fb, gb = b.read_state()
        assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32))
        assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32))
        (sa, ba), (sb, bb) = a.read_mean(), b.read_mean()
        assert np.array_equal(sa, sb) and ba == bb
        with pytest.raises(pkg.AhtError):
            b.set_launch(256, 1, 1, 0, 9)
