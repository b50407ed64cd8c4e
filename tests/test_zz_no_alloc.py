"""The allocation-free diagnostics entry point (aht_stats_noalloc, library-owned scratch) against aht_stats through
the C-ABI: identical results bit for bit, on two calls in a row, the second with a smaller solver (fewer partial
blocks than the first left in the shared scratch), then the larger one again."""
import pytest

from conftest import dump_domain

pytestmark = pytest.mark.gpu


def test_stats_noalloc_equals_stats(tmp_path, pkg):
    big = dump_domain(tmp_path / 'big', '--batch-member', 9)            # 150 x 150 x 32: 2,048 partial blocks
    small = dump_domain(tmp_path / 'small', '--tiny', '24,20,12,7', '--dx', 30, '--flux', 4000)
    assert big['nx'] * big['ny'] * big['nz'] > 2048 * 256 > small['nx'] * small['ny'] * small['nz']
    sb, ss = pkg.Solver.from_dump(big), pkg.Solver.from_dump(small)
    try:
        sb.step(30)
        ss.step(30)
        for s in (sb, ss, sb):
            want = s.stats()
            got = s.stats_noalloc()
            assert got == want, (got, want)
            assert got['non_finite'] == 0 and got['n_fluid'] > 0
    finally:
        sb.close()
        ss.close()
