"""CPU oracle bindings — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, as the reference the GPU
results are compared with or as the timed CPU baseline. The product (libaht.so and its hosts) never imports it and has no CPU fallback.

Wraps oracle/lib/liblbm_oracle.so (a C restatement of the reference kernel generator js/live/kernel.js; see
lbm_oracle.h) and the domain-dump format written by ashburn-heat-twin_amd/js/dump.js. Also restates the
reference's comparison metric rel_l2 (tools/live/validate_live.py:146-149) and the init closed form (:132-143).
"""
import ctypes
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# AHT_ORACLE_LIB selects another build of the oracle (the sanitizer build of the host tests)
LIB_PATH = os.environ.get('AHT_ORACLE_LIB') or os.path.join(HERE, 'lib', 'liblbm_oracle.so')
RHO_CP = 1.16 * 1005.0

# lattice tables (js/live/kernel.js:19-28)
C19 = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 0, 1], [-1, 0, -1], [1, 0, -1],
                [-1, 0, 1], [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1]], np.int64)
W19 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
C7 = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.int64)
W7 = np.array([1 / 4] + [1 / 8] * 6)


def build():
    """compile the oracle library (gcc)"""
    subprocess.run(['make', '-s', '-C', HERE], check=True)


class OracleDom(ctypes.Structure):
    _fields_ = [('nx', ctypes.c_int32), ('ny', ctypes.c_int32), ('nz', ctypes.c_int32), ('pad0', ctypes.c_int32),
                ('kt', ctypes.c_void_p), ('ktop', ctypes.c_void_p), ('cd', ctypes.c_void_p), ('q', ctypes.c_void_p),
                ('prof', ctypes.c_void_p)] + \
               [(n, ctypes.c_float) for n in ('tau0', 'cs2smag', 'gbeta', 'prt', 'kfloor', 'momfac', 'srcscale', 'sigmax')] + \
               [(n, ctypes.c_int32) for n in ('inflow_w', 'inflow_e', 'inflow_s', 'inflow_n', 'nsp', 'nsw', 'oclip', 'pad1')]


class OracleStats(ctypes.Structure):
    _fields_ = [('sum_theta', ctypes.c_double), ('sum_ke', ctypes.c_double), ('sum_drho', ctypes.c_double),
                ('n_fluid', ctypes.c_double), ('theta_max', ctypes.c_double), ('umax', ctypes.c_double),
                ('non_finite', ctypes.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        for sfx in ('f32', 'f64s32', 'f64'):
            getattr(L, f'oracle_init_{sfx}').argtypes = [ctypes.POINTER(OracleDom), vp]
            getattr(L, f'oracle_step_{sfx}').argtypes = [ctypes.POINTER(OracleDom), vp, vp, vp, vp, vp]
            getattr(L, f'oracle_surface_{sfx}').argtypes = [ctypes.POINTER(OracleDom), vp, vp, ctypes.c_float, vp, vp, vp]
            getattr(L, f'oracle_stats_{sfx}').argtypes = [ctypes.POINTER(OracleDom), vp, vp, ctypes.POINTER(OracleStats)]
            getattr(L, f'oracle_reinit_{sfx}').argtypes = [ctypes.POINTER(OracleDom), vp, vp, vp]
        L.oracle_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def load_dump(d):
    """read a domain dump directory (ashburn-heat-twin_amd/js/dump.js)"""
    meta = json.load(open(os.path.join(d, 'domain.json')))
    rd = lambda n, t: np.fromfile(os.path.join(d, n), t)
    meta['kt'] = rd('kt.u32', np.uint32)
    meta['ktop'] = rd('ktop.u32', np.uint32)
    meta['cd'] = rd('cd.f32', np.float32)
    meta['q'] = rd('q.f32', np.float32)
    meta['prof'] = rd('prof.f32', np.float32).reshape(-1, 4)
    return meta


PARAM_F = ('tau0', 'cs2smag', 'gbeta', 'prt', 'kfloor', 'momfac', 'srcscale', 'sigmax')
PARAM_I = ('inflow_w', 'inflow_e', 'inflow_s', 'inflow_n', 'nsp', 'nsw', 'oclip')


class Oracle:
    """Oracle state for one domain: two population lattices (A/B ping-pong, like the reference host)."""

    KINDS = {'f32': np.float32, 'f64s32': np.float32, 'f64': np.float64}

    def __init__(self, dom, kind='f32'):
        self.kind = kind
        self.dt_store = self.KINDS[kind]
        self.dom = dom
        self.nx, self.ny, self.nz = int(dom['nx']), int(dom['ny']), int(dom['nz'])
        self.N = self.nx * self.ny * self.nz
        self._keep = [np.ascontiguousarray(dom[k]) for k in ('kt', 'ktop', 'cd', 'q', 'prof')]
        self.set_columns(dom['kt'], dom['ktop'], dom['cd'], dom['q'])
        self.set_params(dom['params'], dom['prof'])
        self.f = np.zeros((19, self.N), self.dt_store)
        self.g = np.zeros((7, self.N), self.dt_store)
        self.fB = np.zeros_like(self.f)
        self.gB = np.zeros_like(self.g)
        self.steps = 0

    def _struct(self):
        s = OracleDom()
        s.nx, s.ny, s.nz = self.nx, self.ny, self.nz
        s.kt, s.ktop, s.cd, s.q, s.prof = [a.ctypes.data for a in self._cols]
        for n in PARAM_F:
            setattr(s, n, float(self._params[n]))
        for n in PARAM_I:
            setattr(s, n, int(self._params[n]))
        return s

    def set_columns(self, kt, ktop, cd, q):
        prof = self._cols[4] if hasattr(self, '_cols') else np.ascontiguousarray(self.dom['prof'], np.float32)
        self._cols = [np.ascontiguousarray(kt, np.uint32), np.ascontiguousarray(ktop, np.uint32),
                      np.ascontiguousarray(cd, np.float32), np.ascontiguousarray(q, np.float32), prof]

    def set_params(self, params, prof):
        self._params = dict(params)
        self._cols[4] = np.ascontiguousarray(prof, np.float32).reshape(-1, 4)
        self._s = None

    @property
    def s(self):
        return self._struct()

    def _fn(self, name):
        return getattr(lib(), f'oracle_{name}_{self.kind}')

    def init_flow(self):
        s = self.s
        self._fn('init')(ctypes.byref(s), self.f.ctypes.data)
        self.g[:] = 0
        self.steps = 0

    def step(self, n=1, out=False):
        """n steps; with out=True returns (theta, ux, uy, uz) of the last step, [4][N] (lbm_step do_out)"""
        s = self.s
        o = None
        for it in range(n):
            want = out and it == n - 1
            if want:
                o = np.zeros((4, self.N), self.dt_store)
            self._fn('step')(ctypes.byref(s), self.f.ctypes.data, self.fB.ctypes.data, self.g.ctypes.data,
                             self.gB.ctypes.data, o.ctypes.data if want else None)
            self.f, self.fB = self.fB, self.f
            self.g, self.gB = self.gB, self.g
            self.steps += 1
        return o

    def surface(self, conv):
        NXY = self.nx * self.ny
        th, u, v = (np.zeros(NXY, np.float32) for _ in range(3))
        self._fn('surface')(ctypes.byref(self.s), self.f.ctypes.data, self.g.ctypes.data, float(conv),
                            th.ctypes.data, u.ctypes.data, v.ctypes.data)
        return th, u, v

    def stats(self):
        st = OracleStats()
        self._fn('stats')(ctypes.byref(self.s), self.f.ctypes.data, self.g.ctypes.data, ctypes.byref(st))
        return {k: getattr(st, k) for k, _ in OracleStats._fields_}

    def reinit(self, old_ktop):
        ok = np.ascontiguousarray(old_ktop, np.uint32)
        self._fn('reinit')(ctypes.byref(self.s), ok.ctypes.data, self.f.ctypes.data, self.g.ctypes.data)

    def fluid_mask(self):
        k = np.arange(self.nz)[:, None]
        return (k >= self._cols[1][None, :].astype(np.int64)).reshape(-1)


# ---- the reference's comparison metrics (tools/live/validate_live.py) ------------------------------------------
def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))


def init_closed_form(dom):
    """the initial state as lbm3d.Sim.__init__ computes it in numpy float32 (validate_live.py:132-143)"""
    nx, ny, nz = int(dom['nx']), int(dom['ny']), int(dom['nz'])
    kt = dom['kt'].astype(np.int64).reshape(ny, nx)
    uin = dom['prof'][:, 0].astype(np.float32)
    vin = dom['prof'][:, 1].astype(np.float32)
    zagl = np.clip(np.arange(nz)[:, None, None] - kt[None], 0, nz - 1)
    U = uin[zagl]
    V = vin[zagl]
    f = np.empty((19, nz, ny, nx), np.float32)
    usq = U * U + V * V
    for q, (cx, cy, cz) in enumerate(C19):
        cu = np.float32(cx) * U + np.float32(cy) * V
        f[q] = np.float32(W19[q]) * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq)
    return f.reshape(19, -1)

