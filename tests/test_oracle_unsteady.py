"""Pins the oracle's dual time stepping on a fixed mesh (CPU).

addTimeDifferences (DEF:1184-1241) adds V (1.5 u - 2 u^n + 0.5 u^(n-1)) / dt_p
(BDF2, grid 1) to the residual of every smoothing stage, and solveSystem
(DEF:1427-1436) scales the local time step to dt / (1 + 1.5 dt / dt_p):
- dt_p -> infinity: the unsteady smoothing step is the steady one;
- the pseudo-time iteration of a physical step converges to the BDF2
  equation R(u) + V (1.5 u - 2 u^n + 0.5 u^(n-1)) / dt_p = 0 at the
  interior nodes, with R from the oracle's own residual operation and the
  time term written here in closed form, so a wrong coefficient, sign or
  history slot fails it;
- BDF1 (timeScheme 1) converges to R(u) + V (u - u^n) / dt_p = 0.
"""
import numpy as np
import pytest

import oracle
from cases import euler, levels


def unsteady(dtp, scheme=2, **kw):
    return euler(1, flowtype=1, physicaltimestep=dtp, timescheme=scheme, **kw)


def interior(lv):
    L = lv[0]
    b = np.zeros(L.n_nodes, bool)
    b[np.unique(L.bside_index[:2])] = True
    return ~b


def test_large_physical_step_is_steady():
    lv = levels(6, 1)
    a = oracle.Oracle(unsteady(1.0e30).to_struct(), list(lv))
    b = oracle.Oracle(euler(1).to_struct(), list(lv))
    for o in (a, b):
        o.init_freestream()
        o.do_time_iterations(1, 3, True)
    ua, ub = a.get(1, "U")[:5], b.get(1, "U")[:5]
    assert np.abs(ua - ub).max() <= 1e-13 * np.abs(ub).max()
    a.close(); b.close()


def converge_step(o, tol=1e-13, maxit=400):
    prev = o.get(1, "U")[:5]
    for _ in range(maxit):
        o.do_time_iterations(1, 1, True)
        u = o.get(1, "U")[:5]
        if np.abs(u - prev).max() < tol * np.abs(u).max():
            return u
        prev = u
    raise AssertionError("inner iterations did not converge")


@pytest.mark.parametrize("scheme", [1, 2])
def test_converged_step_solves_the_time_discrete_equation(scheme):
    lv = levels(6, 1)
    dtp = 0.02
    o = oracle.Oracle(unsteady(dtp, scheme).to_struct(), list(lv))
    o.init_freestream()
    inf = np.array(euler(1).inflowfield[0])
    u0 = np.repeat(np.concatenate([inf[:4], [inf[4]]])[:, None], lv[0].n_nodes, axis=1)
    u0[1:4] *= inf[0]  # conserved momentum (rho = 1 here anyway)
    u1 = converge_step(o)
    o.advance_physical_time()
    u2 = converge_step(o)
    o.set(1, "U", np.vstack([u2, o.get(1, "U")[5:]]))
    o.apply(1, "PRESSURE")
    o.apply(1, "RHS")
    R = o.get(1, "RHS")[:5]
    V = lv[0].volume
    if scheme == 2:
        T = V * (1.5 * u2 - 2.0 * u1 + 0.5 * u0) / dtp
    else:
        T = V * (u2 - u1) / dtp
    m = interior(lv)
    assert np.abs(T[:, m]).max() > 1e-6  # the flow really moves between the steps
    err = np.abs(R[:, m] + T[:, m]).max() / np.abs(T[:, m]).max()
    assert err < 1e-9, err
    o.close()


def converge_cycles(o, tol=1e-13, maxit=600):
    prev = o.get(1, "U")[:5]
    for c in range(1, maxit + 1):
        o.cycle(c)
        u = o.get(1, "U")[:5]
        if np.abs(u - prev).max() < tol * np.abs(u).max():
            return u
        prev = u
    raise AssertionError("multigrid cycles did not converge")


@pytest.mark.parametrize("scheme", [1, 2])
def test_multigrid_step_solves_the_fine_time_discrete_equation(scheme):
    """restrictGrid adds the time term on both grids (DEF:12208, 12285), and
    the preamble of every physical step restricts u and uold to the coarse
    levels (MG:586-601).  A converged physical step of growing-V cycles over
    three levels therefore solves the fine grid's BDF equation; leaving the
    term out of either half of the restriction, or keeping a stale coarse
    uold, leaves a coarse correction at the fine solution and fails it."""
    lv = levels(6, 3)
    dtp = 0.02
    o = oracle.Oracle(unsteady(dtp, scheme, multigridscheme=3).to_struct(), list(lv))
    o.init_freestream()
    inf = np.array(euler(1).inflowfield[0])
    u0 = np.repeat(inf[:5, None], lv[0].n_nodes, axis=1)
    u1 = converge_cycles(o)
    o.advance_physical_time()
    u2 = converge_cycles(o)
    o.set(1, "U", np.vstack([u2, o.get(1, "U")[5:]]))
    o.apply(1, "PRESSURE")
    o.apply(1, "RHS")
    R = o.get(1, "RHS")[:5]
    V = lv[0].volume
    if scheme == 2:
        T = V * (1.5 * u2 - 2.0 * u1 + 0.5 * u0) / dtp
    else:
        T = V * (u2 - u1) / dtp
    m = interior(lv)
    assert np.abs(T[:, m]).max() > 1e-6
    err = np.abs(R[:, m] + T[:, m]).max() / np.abs(T[:, m]).max()
    assert err < 1e-9, err
    o.close()


def restrict(lv_c, fine):
    out = np.zeros((fine.shape[0], lv_c.n_nodes))
    for c in range(fine.shape[0]):
        np.add.at(out[c], lv_c.f2c, lv_c.rweight * fine[c])
    return out


def test_each_physical_step_restricts_u_and_uold_to_the_coarse_levels():
    """The preamble of multiGridCycles sits inside the physical-step loop
    (MG:586-601): at the start of every step the laminar viscosity is reset
    and u and uold are mapped to every coarse level with the restriction
    weights V_fine/V_coarse (mapFineToCoarse, DEF:12572-12690), restricted
    here in numpy from the level arrays."""
    lv = levels(6, 3)
    o = oracle.Oracle(unsteady(0.02, 2, multigridscheme=3).to_struct(), list(lv))
    o.init_freestream()
    for c in range(1, 4):
        o.cycle(c)
    o.advance_physical_time()
    for level in (2, 3):
        fu, fo = o.get(level - 1, "U"), o.get(level - 1, "UOLD")
        cu, co = o.get(level, "U"), o.get(level, "UOLD")
        ru, ro = restrict(lv[level - 1], fu), restrict(lv[level - 1], fo)
        assert np.abs(cu - ru).max() <= 1e-14 * np.abs(ru).max(), level
        assert np.abs(co - ro).max() <= 1e-14 * np.abs(ro).max(), level
    # after the step uold is the state the cycles reached, not the start
    assert np.abs(o.get(1, "UOLD") - o.get(1, "U")).max() == 0.0
    assert np.abs(o.get(1, "UOLD2") - o.get(1, "U")).max() > 1e-6
    o.close()
