"""The oracle's diagnostics (oracle_stats: genStatsWGSL, js/live/kernel.js:580-633, reduced on the host as
_dbg.stats(), js/live/solver.js:559-576) pinned against closed forms on the CPU.

* The initial state is the equilibrium of the inflow profile at rho = 1 (genInitWGSL :524-553) with theta = 0, built
  with the weights rounded to f32 as the reference builds it (tools/live/validate_live.py:132-143). Its moments follow
  from two sums of those weights, S0 = sum_q w_q and S2 = sum_q w_q c_qx^2 (= 1 and 1/3 before rounding; the odd
  moments vanish by symmetry): rho = S0 + |U|^2 (4.5 S2 - 1.5 S0) and j = 3 S2 U, U = (u_in, v_in, 0) at the cell's
  height above terrain. So over the fluid cells: sum theta = 0, sum (rho - 1), sum 0.5 |j|^2 / rho and max |j| / rho
  follow from the profile, max theta = 0, and the fluid count is sum_columns (nz - ktop).
* A planted state: every cell at the equilibrium of its own (rho, u) drawn at random and theta_q = w_q theta with a
  random theta; one fluid cell and one solid cell set to NaN. Then sum theta, sum 0.5 rho |u|^2, sum (rho - 1),
  max(theta, 0) and max |u| are the plain sums and maxima of the drawn fields over the finite fluid cells, the NaN
  fluid cell counts once as non-finite and the solid one not at all.
The f64 build is held to 1e-10 relative (the oracle sums ~1e6 cells in sequence), the f32 build (the WebGPU kernel's
arithmetic) to f32 accuracy. A wrong lattice velocity, a dropped 1/rho, solid cells in the sums or a wrong NaN rule
fail one of these.
"""
import numpy as np
import pytest

W19 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
C19 = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 0, 1], [-1, 0, -1], [1, 0, -1],
                [-1, 0, 1], [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1]], np.float64)
W7 = np.array([1 / 4] + [1 / 8] * 6)
TOL = {'f64': 1e-10, 'f32': 2e-6}   # f64: sequential summation over ~1e6 cells


def fluid_cells(dom):
    nz, nxy = int(dom['nz']), int(dom['nx']) * int(dom['ny'])
    k = np.arange(nz)[:, None]
    return (k >= dom['ktop'].astype(np.int64)[None, :]).reshape(-1), nz, nxy


def close(a, b, tol, scale):
    return abs(a - b) <= tol * max(abs(scale), 1e-300)


@pytest.mark.parametrize('kind', ['f64', 'f32'])
def test_stats_of_initial_state_closed_form(ashburn30, O, kind):
    dom = ashburn30
    fl, nz, nxy = fluid_cells(dom)
    o = O.Oracle(dom, kind)
    o.init_flow()
    st = o.stats()
    kt = dom['kt'].astype(np.int64)
    zagl = np.clip(np.arange(nz)[:, None] - kt[None, :], 0, nz - 1).reshape(-1)
    U = dom['prof'][:, 0].astype(np.float64)[zagl]
    V = dom['prof'][:, 1].astype(np.float64)[zagl]
    u2 = (U * U + V * V)[fl]
    w32 = W19.astype(np.float32).astype(np.float64)
    S0, S2 = w32.sum(), (w32 * C19[:, 0] ** 2).sum()
    rho = S0 + u2 * (4.5 * S2 - 1.5 * S0)
    j2 = (3 * S2) ** 2 * u2
    nf = int(fl.sum())
    assert nf == int((nz - dom['ktop'].astype(np.int64)).sum()) == dom['info']['nFluid']
    assert st['n_fluid'] == nf and st['non_finite'] == 0
    assert st['sum_theta'] == 0.0 and st['theta_max'] == 0.0
    tol = TOL[kind]
    drho = (rho - 1).sum()
    # f64: the sum itself (nf * 1.5e-8 from the rounded weights) to 1e-9; f32: per-cell rounding of rho dominates
    assert abs(st['sum_drho'] - drho) <= (1e-9 * abs(drho) if kind == 'f64' else tol * nf)
    ke = (0.5 * j2 / rho).sum()
    assert close(st['sum_ke'], ke, tol, ke)
    um = np.sqrt(j2 / rho ** 2).max()
    assert close(st['umax'], um, tol, um)


@pytest.mark.parametrize('kind', ['f64', 'f32'])
def test_stats_of_planted_state(ashburn30, O, kind):
    dom = ashburn30
    fl, nz, nxy = fluid_cells(dom)
    N = nz * nxy
    rng = np.random.default_rng(7)
    rho = rng.uniform(0.9, 1.1, N)
    u = rng.uniform(-0.06, 0.06, (3, N))
    th = rng.uniform(-0.5, 3.0, N)
    cu = C19 @ u                                       # [19][N]
    usq = (u * u).sum(0)
    f = W19[:, None] * rho[None] * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq[None])
    g = W7[:, None] * th[None]
    o = O.Oracle(dom, kind)
    dt = o.f.dtype
    fluid_idx = np.flatnonzero(fl)
    solid_idx = np.flatnonzero(~fl)
    bad_fluid, bad_solid = fluid_idx[len(fluid_idx) // 3], solid_idx[len(solid_idx) // 2]
    g[:, bad_fluid] = np.nan
    f[:, bad_solid] = np.nan
    o.f[:] = f.astype(dt)
    o.g[:] = g.astype(dt)
    st = o.stats()
    ok = fl.copy()
    ok[bad_fluid] = False
    # the moments of the stored populations (exact in f64; in f32 the stored values are rounded, so the expected
    # sums come from the drawn fields at f32 accuracy)
    ke = 0.5 * rho * usq
    tol = TOL[kind] * (1 if kind == 'f64' else 10)
    assert st['non_finite'] == 1 and st['n_fluid'] == ok.sum()
    assert close(st['sum_theta'], th[ok].sum(), tol * 10, np.abs(th[ok]).sum())
    assert close(st['sum_drho'], (rho[ok] - 1).sum(), tol, np.abs(rho[ok] - 1).sum())
    assert close(st['sum_ke'], ke[ok].sum(), tol, ke[ok].sum())
    assert close(st['theta_max'], max(th[ok].max(), 0.0), tol, th[ok].max())
    assert close(st['umax'], np.sqrt(usq[ok].max()), tol * 10, np.sqrt(usq[ok].max()))
