"""ctypes binding of the C-ABI in include/aht.h (libaht.so, HIP for gfx950).

The Python face of the same library the Node addon wraps; used by the GPU parity tests and bench.py. There is no
CPU path: loading fails loudly when libaht.so has not been built, and every call raises AhtError on a negative code.

Host-side meaning mirrors the reference solver (js/live/solver.js): `Solver.from_dump` uploads a domain dump written
by js/dump.js (the JavaScript host is the only geometry builder), `step` enqueues like `step(n)` (:401-427),
`read_surface` is `read()`'s surface (:439-471), `stats` is `_dbg.stats()` (:559-576), `read_state`/`write_state`
are `_dbg.readState`/`writeState` (:540-547).
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, 'lib')
# AHT_LIB selects another build of the same C-ABI (e.g. one made with other compile options)
LIB_PATH = os.environ.get('AHT_LIB') or os.path.join(LIB_DIR, 'libaht.so')
ROOT = os.path.dirname(HERE)

AHT_OK, AHT_EINVAL, AHT_ENOMEM, AHT_EDEVICE, AHT_ESTATE = 0, -1, -2, -3, -4
STORE_F32, STORE_F16, STORE_BF16, STORE_F64, STORE_F16S = 0, 1, 2, 3, 4
STORAGE_NAMES = {'f32': 0, 'f16': 1, 'bf16': 2, 'f64': 3, 'f16s': 4}
STEP_ACC, STEP_OUT = 1, 2
FACE_W, FACE_E, FACE_S, FACE_N = 0, 1, 2, 3
RHO_CP = 1.16 * 1005.0


class AhtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f'aht error {code}: {msg}')
        self.code = code


class Dims(ctypes.Structure):
    _fields_ = [('nx', ctypes.c_int32), ('ny', ctypes.c_int32), ('nz', ctypes.c_int32), ('storage', ctypes.c_int32),
                ('device', ctypes.c_int32), ('block', ctypes.c_int32), ('layout', ctypes.c_int32),
                ('reserved', ctypes.c_int32)]


class Launch(ctypes.Structure):
    """launch shape of the step kernel: threads per block, cells per thread along z, tile window (1 = linear), the
    two-step sweep (0 = off; +kc: 32 x 8 column regions, -kc: 16 x 16, level chunks of kc) and the cap of wavefronts
    per SIMD (0 = none)"""
    _fields_ = [('block', ctypes.c_int32), ('levels', ctypes.c_int32), ('window', ctypes.c_int32),
                ('sweep', ctypes.c_int32), ('waves', ctypes.c_int32)]

    def as_dict(self):
        return {'block': self.block, 'levels': self.levels, 'window': self.window, 'sweep': self.sweep,
                'waves': self.waves}


PARAM_F = ('tau0', 'cs2smag', 'gbeta', 'prt', 'kfloor', 'momfac', 'srcscale', 'sigmax')
PARAM_I = ('inflow_w', 'inflow_e', 'inflow_s', 'inflow_n', 'nsp', 'nsw', 'oclip', 'pad')


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in PARAM_F] + [(n, ctypes.c_int32) for n in PARAM_I]


class Diag(ctypes.Structure):
    _fields_ = [('sum_theta', ctypes.c_double), ('sum_ke', ctypes.c_double), ('sum_drho', ctypes.c_double),
                ('n_fluid', ctypes.c_double), ('theta_max', ctypes.c_double), ('umax', ctypes.c_double),
                ('non_finite', ctypes.c_int64)]


class Budget(ctypes.Structure):
    _fields_ = [('n_acc', ctypes.c_int64), ('heat_in', ctypes.c_double), ('heat_out', ctypes.c_double),
                ('content', ctypes.c_double), ('q_w', ctypes.c_double), ('q_e', ctypes.c_double),
                ('q_s', ctypes.c_double), ('q_n', ctypes.c_double), ('mass_imbalance', ctypes.c_double)]


# every symbol include/aht.h declares, with its ctypes signature
_vp, _i32, _u32, _f32, _i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32, ctypes.c_float, ctypes.c_int64
SIGNATURES = {
    'aht_create': (ctypes.c_int, [ctypes.POINTER(Dims), ctypes.POINTER(_vp)]),
    'aht_destroy': (None, [_vp]),
    'aht_last_error': (ctypes.c_char_p, [_vp]),
    'aht_error_string': (ctypes.c_char_p, [ctypes.c_int]),
    'aht_set_columns': (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    'aht_set_params': (ctypes.c_int, [_vp, ctypes.POINTER(Params), _vp]),
    'aht_init_flow': (ctypes.c_int, [_vp]),
    'aht_reset_heat': (ctypes.c_int, [_vp]),
    'aht_reinit_cells': (ctypes.c_int, [_vp, _vp]),
    'aht_step': (ctypes.c_int, [_vp, _i32, _u32]),
    'aht_sync': (ctypes.c_int, [_vp]),
    'aht_steps_done': (_i64, [_vp]),
    'aht_benchmark': (ctypes.c_int, [_vp, _i32, ctypes.POINTER(ctypes.c_double)]),
    'aht_read_surface': (ctypes.c_int, [_vp, _f32, _vp, _vp, _vp]),
    'aht_read_volume': (ctypes.c_int, [_vp, _i32, _f32, _vp, ctypes.POINTER(_f32)]),
    'aht_stats': (ctypes.c_int, [_vp, ctypes.POINTER(Diag)]),
    'aht_stats_noalloc': (ctypes.c_int, [_vp, ctypes.POINTER(Diag)]),
    'aht_read_out': (ctypes.c_int, [_vp, _vp]),
    'aht_mean_reset': (ctypes.c_int, [_vp]),
    'aht_read_mean': (ctypes.c_int, [_vp, _vp, ctypes.POINTER(Budget)]),
    'aht_read_face_mean': (ctypes.c_int, [_vp, _i32, _vp]),
    'aht_read_state': (ctypes.c_int, [_vp, _vp, _vp]),
    'aht_write_state': (ctypes.c_int, [_vp, _vp, _vp]),
    'aht_read_state_f64': (ctypes.c_int, [_vp, _vp, _vp]),
    'aht_write_state_f64': (ctypes.c_int, [_vp, _vp, _vp]),
    'aht_footprint_bytes': (_i64, [ctypes.POINTER(Dims)]),
    'aht_batch_create': (ctypes.c_int, [_i32, _i32, ctypes.POINTER(_vp)]),
    'aht_batch_destroy': (None, [_vp]),
    'aht_batch_last_error': (ctypes.c_char_p, [_vp]),
    'aht_batch_add': (ctypes.c_int, [_vp, _vp]),
    'aht_batch_remove': (ctypes.c_int, [_vp, _vp]),
    'aht_batch_size': (ctypes.c_int, [_vp]),
    'aht_batch_step': (ctypes.c_int, [_vp, _i32, _vp, _vp, ctypes.POINTER(ctypes.c_double)]),
    'aht_batch_sync': (ctypes.c_int, [_vp]),
    'aht_batch_set_grid': (ctypes.c_int, [_vp, _i32, _i32]),
    'aht_batch_grid': (ctypes.c_int, [_vp]),
    'aht_batch_set_block': (ctypes.c_int, [_vp, _i32]),
    'aht_batch_block': (ctypes.c_int, [_vp]),
    'aht_autotune': (ctypes.c_int, [_vp, _i32, ctypes.POINTER(Launch)]),
    'aht_set_launch': (ctypes.c_int, [_vp, ctypes.POINTER(Launch)]),
    'aht_get_launch': (ctypes.c_int, [_vp, ctypes.POINTER(Launch)]),
}

_lib = None


def lib():
    """load libaht.so (built by `make -C ashburn-heat-twin_amd`); raises if it is missing"""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f'{LIB_PATH} is not built: run __graft_entry__.build() or make -C ashburn-heat-twin_amd')
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def load_dump(d):
    """read a domain dump directory written by js/dump.js"""
    meta = json.load(open(os.path.join(d, 'domain.json')))
    rd = lambda n, t: np.fromfile(os.path.join(d, n), t)
    meta['kt'] = rd('kt.u32', np.uint32)
    meta['ktop'] = rd('ktop.u32', np.uint32)
    meta['cd'] = rd('cd.f32', np.float32)
    meta['q'] = rd('q.f32', np.float32)
    meta['prof'] = rd('prof.f32', np.float32).reshape(-1, 4)
    return meta


class Solver:
    """One device-resident domain (two f32 lattices of 19 + 7 populations)."""

    def __init__(self, nx, ny, nz, storage=STORE_F32, device=0, block=0, layout=0):
        L = lib()
        storage = STORAGE_NAMES.get(storage, storage)
        self.storage = storage
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.N = self.nx * self.ny * self.nz
        self.NXY = self.nx * self.ny
        self.dims = Dims(self.nx, self.ny, self.nz, storage, device, block, layout)
        h = ctypes.c_void_p()
        rc = L.aht_create(ctypes.byref(self.dims), ctypes.byref(h))
        self.h = h
        if rc != AHT_OK:
            msg = L.aht_last_error(h).decode() if h.value else L.aht_error_string(rc).decode()
            if h.value:
                L.aht_destroy(h)
            self.h = None
            raise AhtError(rc, msg)

    # -- helpers
    def _check(self, rc):
        if rc != AHT_OK:
            raise AhtError(rc, lib().aht_last_error(self.h).decode())

    def close(self):
        if self.h:
            lib().aht_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- construction from a dump
    @classmethod
    def from_dump(cls, dom, init=True, **kw):
        s = cls(dom['nx'], dom['ny'], dom['nz'], **kw)
        s.set_columns(dom['kt'], dom['ktop'], dom['cd'], dom['q'])
        s.set_params(dom['params'], dom['prof'])
        if init:
            s.init_flow()
            s.reset_heat()
        return s

    def set_columns(self, kt, ktop, cd, q):
        self._cols = [np.ascontiguousarray(kt, np.uint32), np.ascontiguousarray(ktop, np.uint32),
                      np.ascontiguousarray(cd, np.float32), np.ascontiguousarray(q, np.float32)]
        self._check(lib().aht_set_columns(self.h, *[_ptr(a) for a in self._cols]))

    def set_params(self, params, prof):
        p = Params()
        for n in PARAM_F:
            setattr(p, n, float(params[n]))
        for n in PARAM_I[:-1]:
            setattr(p, n, int(params[n]))
        pr = np.ascontiguousarray(prof, np.float32).reshape(-1)
        if pr.size != 4 * self.nz:
            raise ValueError('prof must be [nz][4]')
        self._check(lib().aht_set_params(self.h, ctypes.byref(p), _ptr(pr)))

    def init_flow(self):
        self._check(lib().aht_init_flow(self.h))

    def reset_heat(self):
        self._check(lib().aht_reset_heat(self.h))

    def reinit_cells(self, old_ktop):
        ok = np.ascontiguousarray(old_ktop, np.uint32)
        self._check(lib().aht_reinit_cells(self.h, _ptr(ok)))

    def step(self, n=1, acc=False, out=False):
        self._check(lib().aht_step(self.h, int(n), (STEP_ACC if acc else 0) | (STEP_OUT if out else 0)))

    def sync(self):
        self._check(lib().aht_sync(self.h))

    @property
    def steps(self):
        return int(lib().aht_steps_done(self.h))

    def autotune(self, steps=20):
        """the launch shape {block, levels, window, sweep, waves} the autotuner keeps (now set); the state is left
        unchanged"""
        best = Launch()
        self._check(lib().aht_autotune(self.h, int(steps), ctypes.byref(best)))
        return best.as_dict()

    def set_launch(self, block=256, levels=1, window=1, sweep=0, waves=0):
        self._check(lib().aht_set_launch(self.h, ctypes.byref(Launch(int(block), int(levels), int(window),
                                                                     int(sweep), int(waves)))))

    @property
    def launch(self):
        l = Launch()
        self._check(lib().aht_get_launch(self.h, ctypes.byref(l)))
        return l.as_dict()

    def benchmark(self, n):
        ms = ctypes.c_double()
        self._check(lib().aht_benchmark(self.h, int(n), ctypes.byref(ms)))
        return ms.value

    def read_surface(self, conv=1.0):
        th, u, v = (np.empty(self.NXY, np.float32) for _ in range(3))
        self._check(lib().aht_read_surface(self.h, float(conv), _ptr(th), _ptr(u), _ptr(v)))
        return th, u, v

    def read_volume(self, nzv, volmax=3.0):
        vol = np.empty(nzv * self.NXY, np.uint8)
        tm = ctypes.c_float()
        self._check(lib().aht_read_volume(self.h, int(nzv), float(volmax), _ptr(vol), ctypes.byref(tm)))
        return vol, tm.value

    def stats(self):
        d = Diag()
        self._check(lib().aht_stats(self.h, ctypes.byref(d)))
        return {k: getattr(d, k) for k, _ in Diag._fields_}

    def stats_noalloc(self):
        """stats() through the allocation-free entry point (library-owned scratch, no device allocation)"""
        d = Diag()
        self._check(lib().aht_stats_noalloc(self.h, ctypes.byref(d)))
        return {k: getattr(d, k) for k, _ in Diag._fields_}

    def read_out(self):
        o = np.empty((4, self.N), np.float32)
        self._check(lib().aht_read_out(self.h, _ptr(o)))
        return o

    def mean_reset(self):
        self._check(lib().aht_mean_reset(self.h))

    def read_mean(self):
        s3 = np.empty((3, self.NXY), np.float64)
        b = Budget()
        self._check(lib().aht_read_mean(self.h, _ptr(s3), ctypes.byref(b)))
        return s3, {k: getattr(b, k) for k, _ in Budget._fields_}

    def read_face_mean(self, face):
        n = self.ny if face in (FACE_W, FACE_E) else self.nx
        o = np.empty((2, self.nz, n), np.float64)
        self._check(lib().aht_read_face_mean(self.h, int(face), _ptr(o)))
        return o

    def read_state(self):
        f = np.empty((19, self.N), np.float32)
        g = np.empty((7, self.N), np.float32)
        self._check(lib().aht_read_state(self.h, _ptr(f), _ptr(g)))
        return f, g

    def write_state(self, f=None, g=None):
        fa = None if f is None else np.ascontiguousarray(f, np.float32)
        ga = None if g is None else np.ascontiguousarray(g, np.float32)
        self._check(lib().aht_write_state(self.h, _ptr(fa), _ptr(ga)))

    def read_state_f64(self):
        f = np.empty((19, self.N), np.float64)
        g = np.empty((7, self.N), np.float64)
        self._check(lib().aht_read_state_f64(self.h, _ptr(f), _ptr(g)))
        return f, g

    def write_state_f64(self, f=None, g=None):
        fa = None if f is None else np.ascontiguousarray(f, np.float64)
        ga = None if g is None else np.ascontiguousarray(g, np.float64)
        self._check(lib().aht_write_state_f64(self.h, _ptr(fa), _ptr(ga)))


class Batch:
    """Many solvers advanced together: one persistent launch per step, chunks captured as hipGraphs."""

    def __init__(self, storage=STORE_F32, device=0):
        storage = STORAGE_NAMES.get(storage, storage)
        h = ctypes.c_void_p()
        rc = lib().aht_batch_create(device, storage, ctypes.byref(h))
        self.h = h
        if rc != AHT_OK:
            msg = lib().aht_batch_last_error(h).decode() if h.value else lib().aht_error_string(rc).decode()
            if h.value:
                lib().aht_batch_destroy(h)
            self.h = None
            raise AhtError(rc, msg)
        self.members = []

    def _check(self, rc):
        if rc != AHT_OK:
            raise AhtError(rc, lib().aht_batch_last_error(self.h).decode())

    def add(self, s):
        self._check(lib().aht_batch_add(self.h, s.h))
        self.members.append(s)

    def remove(self, s):
        self._check(lib().aht_batch_remove(self.h, s.h))
        self.members.remove(s)

    def step(self, n, nsteps=None, acc_from=None, timed=False):
        """one chunk of n launches; nsteps / acc_from per member in add order (None: all n / no means)"""
        ns = None if nsteps is None else np.ascontiguousarray(nsteps, np.int32)
        af = None if acc_from is None else np.ascontiguousarray(acc_from, np.int32)
        ms = ctypes.c_double()
        self._check(lib().aht_batch_step(self.h, int(n), _ptr(ns), _ptr(af), ctypes.byref(ms) if timed else None))
        return ms.value if timed else None

    def sync(self):
        self._check(lib().aht_batch_sync(self.h))

    def set_grid(self, blocks, use_graph=True):
        self._check(lib().aht_batch_set_grid(self.h, int(blocks), 1 if use_graph else 0))

    @property
    def grid(self):
        return int(lib().aht_batch_grid(self.h))

    def set_block(self, block):
        """threads per block of the batched step (64, 128, 256; 0 = tuned online)"""
        self._check(lib().aht_batch_set_block(self.h, int(block)))

    @property
    def block(self):
        return int(lib().aht_batch_block(self.h))

    def close(self):
        if self.h:
            lib().aht_batch_destroy(self.h)
            self.h = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def budget_physical(b, dx, dt, conv):
    """lattice-unit budget -> the reference's heat_diag units (tools/live/validate_live.py:226-240)"""
    return dict(
        heat_injected_MW=b['heat_in'] * RHO_CP * dx ** 3 / dt / 1e6,
        heat_advected_out_MW=RHO_CP * b['heat_out'] * conv * dx * dx / 1e6,
        heat_content_MJ=b['content'] * RHO_CP * dx ** 3 / 1e6,
        mass_imbalance_frac=b['mass_imbalance'],
    )


def stats_physical(d, dx, conv):
    """aht_stats diagnostics -> the units of the reference's _dbg.stats() (js/live/solver.js:559-576):
    heat content in MJ (sum theta * rho cp dx^3), mean kinetic energy per fluid cell in m2/s2 (sum 0.5|j|^2/rho over
    the fluid count, times conv^2), max theta in K, max |u| in m/s"""
    nf = max(d['n_fluid'], 1.0)
    return dict(heat_content_MJ=d['sum_theta'] * RHO_CP * dx ** 3 / 1e6, ke_mean=d['sum_ke'] / nf * conv * conv,
                theta_max=d['theta_max'], umax=d['umax'] * conv, rho_dev_mean=d['sum_drho'] / nf,
                n_fluid=d['n_fluid'], non_finite=d['non_finite'])
