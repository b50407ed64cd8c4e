"""The C-ABI library loads and exports every symbol include/aht.h declares (no compute calls: no GPU here), and the
ctypes structures match the header's layout."""
import ctypes
import os
import re
import subprocess

from conftest import PKG, ROOT, load_pkg


def header_functions():
    txt = open(os.path.join(ROOT, 'include', 'aht.h')).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(aht_[a-z0-9_]+)\s*\(', txt)))


def test_every_declared_symbol_is_exported():
    m = load_pkg()
    L = m.lib()
    names = header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the binding covers them all
    assert set(names) == set(m.SIGNATURES), set(names) ^ set(m.SIGNATURES)


def test_no_oracle_symbols_in_product_library():
    m = load_pkg()
    out = subprocess.run(['nm', '-D', '--defined-only', m.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert 'oracle_' not in out


def test_struct_layouts():
    m = load_pkg()
    assert ctypes.sizeof(m.Dims) == 32
    assert ctypes.sizeof(m.Params) == 64
    assert ctypes.sizeof(m.Diag) == 56
    assert ctypes.sizeof(m.Budget) == 72


def test_footprint_is_positive_and_scales():
    m = load_pkg()
    L = m.lib()
    small = L.aht_footprint_bytes(ctypes.byref(m.Dims(100, 100, 32, 0, 0, 0)))
    big = L.aht_footprint_bytes(ctypes.byref(m.Dims(600, 600, 64, 0, 0, 0)))
    assert small > 208 * 100 * 100 * 32 and big > 208 * 600 * 600 * 64
    assert big > 50 * small


def test_create_without_gpu_fails_cleanly():
    """with no device the library reports an error instead of running anything on the CPU"""
    m = load_pkg()
    try:
        import torch  # noqa: F401  (only to know whether a GPU is present)
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        return
    try:
        m.Solver(8, 8, 8)
    except m.AhtError as e:
        assert e.code in (m.AHT_EDEVICE, m.AHT_ENOMEM)
    else:
        raise AssertionError('solver created without a GPU')


def test_library_links_no_torch():
    out = subprocess.run(['ldd', os.path.join(PKG, 'lib', 'libaht.so')], capture_output=True, text=True).stdout
    assert 'torch' not in out and 'amdhip64' in out


def _create_rc(m, nx, ny, nz, storage):
    L = m.lib()
    h = ctypes.c_void_p()
    rc = L.aht_create(ctypes.byref(m.Dims(nx, ny, nz, storage, 0, 0)), ctypes.byref(h))
    if h.value:
        L.aht_destroy(h)
    return rc


def test_size_rules_are_checked_before_any_device_work():
    """argument rules of aht_create, checked without a GPU (only the rejected shapes are tried where one is present):
    nz <= 1022 (10-bit column fields); 32-bit byte offsets within a population plane and within a theta-referenced
    D3Q7 lattice (18 B per cell for f16 / bf16)"""
    m = load_pkg()
    assert _create_rc(m, 8, 8, 1023, 0) == m.AHT_EINVAL
    assert _create_rc(m, 2000, 2000, 60, m.STORE_F16) == m.AHT_EINVAL       # 18 * 240 M > 2^32
    assert _create_rc(m, 2000, 2000, 60, m.STORE_BF16) == m.AHT_EINVAL
    assert _create_rc(m, 1, 8, 8, 0) == m.AHT_EINVAL
    try:
        import torch  # noqa: F401  (only to know whether a GPU is present)
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if not has_gpu:
        # accepted shapes get as far as the device lookup: 1200 x 1200 x 96 in f64 (plane 1.1 GB) and f16 (2.5 GB)
        assert _create_rc(m, 1200, 1200, 96, m.STORE_F64) == m.AHT_EDEVICE
        assert _create_rc(m, 1200, 1200, 96, m.STORE_F16) == m.AHT_EDEVICE


def test_blocked_pull_addresses_stay_in_the_allocation(tmp_path):
    """the blocked layouts' pulled addresses (storage.h pull_shift / pull_rem / pull_voff, the header the kernels
    use), built for the host: for every cell the fast path can step and every population, the lane's address is the
    source cell's element (also for sources outside the lattice, at levels 1 and nz - 2), its offset fits 32 bits, and
    it lies inside the guarded allocation (tests/native/blocked_pull_check.hip)"""
    exe = str(tmp_path / 'blocked_pull_check')
    subprocess.run(['/opt/rocm/bin/hipcc', '-x', 'hip', '--offload-host-only', '-O2', '-std=c++17',
                    '-I', os.path.join(PKG, 'csrc'), '-I', os.path.join(ROOT, 'include'), '-o', exe,
                    os.path.join(ROOT, 'tests', 'native', 'blocked_pull_check.hip')], check=True)
    grids = [(14, 12, 15, 1), (24, 20, 12, 1), (2, 2, 3, 1), (100, 100, 32, 1), (150, 150, 32, 1), (200, 200, 48, 2),
             (300, 300, 48, 4), (2047, 3, 5, 1), (2049, 3, 5, 1), (255, 3, 4, 1), (256, 3, 4, 1), (6000, 3, 4, 1),
             (3, 2000, 5, 1), (600, 600, 64, 16)]
    args = [str(x) for g in grids for x in g]
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count('pulled addresses checked') == 3 * len(grids), r.stdout
