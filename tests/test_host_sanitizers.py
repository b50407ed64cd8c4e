"""Host code under the sanitizers (CPU only): the oracle and the N-API addon built with AddressSanitizer and
UndefinedBehaviorSanitizer (the addon also with ThreadSanitizer), and the host build of the blocked-pull address
functions with both. The HIP library's host code is not covered: it needs a GPU, and its tests there run the
uninstrumented build. Errors abort the sanitized process (-fno-sanitize-recover), so a clean exit status is the check;
leak reports are off (the addon's handles are finalised by the JS garbage collector, which need not run at exit)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ['-O1', '-g', '-fno-omit-frame-pointer', '-fsanitize=address,undefined', '-fno-sanitize-recover=undefined']


def preload(rt):
    """LD_PRELOAD with the sanitizer runtime first and whatever the environment already preloads after it"""
    cur = os.environ.get('LD_PRELOAD', '')
    return rt + (':' + cur if cur else '')


def runtime(name):
    p = subprocess.run(['gcc', f'-print-file-name={name}'], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(p) or not os.path.exists(p):
        pytest.skip(f'{name} not installed')
    return p


def test_oracle_under_asan_ubsan(tmp_path):
    """the oracle's golden-vector and closed-form tests against an ASan + UBSan build of the oracle"""
    asan = runtime('libasan.so')
    lib = tmp_path / 'liblbm_oracle.so'
    subprocess.run(['gcc'] + SAN + ['-fPIC', '-fopenmp', '-ffp-contract=off', '-fno-fast-math', '-std=c11', '-Wall',
                                    '-Wno-unused-function', '-shared', '-o', str(lib),
                                    os.path.join(ROOT, 'oracle', 'lbm_oracle.c'), '-lm'], check=True)
    env = dict(os.environ, LD_PRELOAD=preload(asan), ASAN_OPTIONS='detect_leaks=0', AHT_ORACLE_LIB=str(lib),
               OMP_NUM_THREADS='2')
    probe = ('import sys; sys.path.insert(0, %r); from oracle import oracle as O; O.lib(); '
             'm = open("/proc/self/maps").read(); print(%r in m, "libasan" in m)' % (ROOT, str(lib)))
    r = subprocess.run([sys.executable, '-c', probe], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.split() == ['True', 'True'], r.stdout + r.stderr
    r = subprocess.run([sys.executable, '-m', 'pytest', '-q', '-p', 'no:cacheprovider',
                        os.path.join(ROOT, 'tests', 'test_golden_fixtures.py'),
                        os.path.join(ROOT, 'tests', 'test_oracle_stats.py')],
                       env=env, capture_output=True, text=True, timeout=600, cwd=ROOT)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert 'AddressSanitizer' not in r.stderr and 'runtime error' not in r.stderr, r.stderr[-3000:]


def build_addon(d, flags):
    """the N-API addon over the CPU stand-in of the C-ABI (tests/js/aht_standin.cpp), both built with `flags`"""
    subprocess.run(['g++'] + flags + ['-std=c++17', '-shared', '-fPIC', '-I', os.path.join(ROOT, 'include'),
                    '-o', str(d / 'libaht.so'), os.path.join(ROOT, 'tests', 'js', 'aht_standin.cpp'), '-lpthread'],
                   check=True)
    subprocess.run(['gcc'] + flags + ['-std=c11', '-shared', '-fPIC', '-Wall', '-I/usr/include/node',
                    '-I', os.path.join(ROOT, 'include'), '-o', str(d / 'aht_node.node'),
                    os.path.join(ROOT, 'ashburn-heat-twin_amd', 'csrc', 'aht_napi.c'), '-L', str(d), '-laht',
                    '-lpthread', '-Wl,-rpath,$ORIGIN'], check=True)
    return str(d / 'aht_node.node')


@pytest.mark.parametrize('san', ['address', 'thread'])
def test_addon_under_sanitizers(tmp_path, san):
    """the addon's device thread, deferred destroy, busy-solver refusals and exit order (tests/js/exit_test.js)
    under ASan + UBSan, and under ThreadSanitizer"""
    if san == 'address':
        rt, flags, opts = runtime('libasan.so'), SAN, {'ASAN_OPTIONS': 'detect_leaks=0'}
    else:
        rt, flags, opts = runtime('libtsan.so'), ['-O1', '-g', '-fno-omit-frame-pointer', '-fsanitize=thread'], \
            {'TSAN_OPTIONS': 'halt_on_error=1'}
    addon = build_addon(tmp_path, flags)   # one directory per test: the addon loads libaht.so beside it
    env = dict(os.environ, LD_PRELOAD=preload(rt), **opts)
    for mode in ('natural', 'exit'):
        r = subprocess.run(['node', os.path.join(ROOT, 'tests', 'js', 'exit_test.js'), addon, mode], env=env,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, (mode, r.returncode, r.stderr[-3000:])
        assert 'Sanitizer' not in r.stderr and 'runtime error' not in r.stderr, (mode, r.stderr[-3000:])
        out = json.loads(r.stdout)
        assert out['order'] == list(range(6)) and out['batchStep'] == 10 and out['statsDuringStep'] == 'threw'


def test_blocked_pulls_under_asan_ubsan(tmp_path):
    """the blocked layouts' address functions (storage.h) on the host, built with ASan + UBSan (every sanitizer flag
    of the hipcc line applies to the host compilation only)"""
    runtime('libasan.so')
    exe = str(tmp_path / 'blocked_pull_check')
    host = []
    for f in SAN:
        host += ['-Xarch_host', f] if f.startswith('-fsan') or f.startswith('-fno-san') else [f]
    subprocess.run(['/opt/rocm/bin/hipcc', '-x', 'hip', '--offload-host-only', '-std=c++17'] + host +
                   ['-I', os.path.join(ROOT, 'ashburn-heat-twin_amd', 'csrc'), '-I', os.path.join(ROOT, 'include'),
                    '-o', exe, os.path.join(ROOT, 'tests', 'native', 'blocked_pull_check.hip')], check=True)
    grids = [(14, 12, 15, 1), (2, 2, 3, 1), (255, 3, 4, 1), (2049, 3, 5, 1), (600, 600, 64, 16)]
    r = subprocess.run([exe] + [str(x) for g in grids for x in g], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, ASAN_OPTIONS='detect_leaks=0'))
    assert r.returncode == 0 and r.stdout.count('pulled addresses checked') == 3 * len(grids), r.stdout + r.stderr
