"""Shared test helpers: the `gpu` marker, the package module (ctypes binding of libaht.so), the CPU oracle, and
domain dumps written by the JavaScript host (the only geometry builder)."""
import importlib.util
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, 'ashburn-heat-twin_amd')
GOLDEN = os.path.join(ROOT, 'tests', 'golden')
DUMP_JS = os.path.join(PKG, 'js', 'dump.js')
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP kernels)')
    config.addinivalue_line('markers', 'slow: long CPU oracle runs')


def load_pkg():
    """the package's ctypes binding (directory name has dashes, so load it by path)"""
    if 'aht_pkg' in sys.modules:
        return sys.modules['aht_pkg']
    spec = importlib.util.spec_from_file_location('aht_pkg', os.path.join(PKG, 'aht.py'))
    m = importlib.util.module_from_spec(spec)
    sys.modules['aht_pkg'] = m
    spec.loader.exec_module(m)
    return m


def oracle_mod():
    from oracle import oracle as O
    return O


def dump_domain(out_dir, *args):
    """run the Node host's domain dumper into out_dir with CLI flags; returns the loaded dump"""
    cmd = ['node', DUMP_JS, '--out', str(out_dir)] + [str(a) for a in args]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return oracle_mod().load_dump(str(out_dir))


@pytest.fixture(scope='session')
def pkg():
    return load_pkg()


@pytest.fixture(scope='session')
def O():
    return oracle_mod()


@pytest.fixture(scope='session')
def ashburn30(tmp_path_factory):
    """the reference's p30a domain: Ashburn static10 at dx = 30, 25 Aug met with u* 0.32, nsw 16, prototype outlet"""
    d = tmp_path_factory.mktemp('p30a')
    return dump_domain(d, '--dx', 30, '--nsw', 16, '--oclip', 0)


def load_summary():
    """the reference's recorded validation numbers, kept as a fixture (tests/live/validation_summary.json subset)"""
    return json.load(open(os.path.join(GOLDEN, 'reference_kats.json')))


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))
