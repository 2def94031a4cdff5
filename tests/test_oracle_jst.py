"""Pins the oracle's non-HLLC interior fluxes and its JST dissipation (CPU).

Fluxes (makeRHS, DEF:1664-1792; LLF and Roe restated with the reference's
unassigned densities/enthalpies and reversed upwind sign patched):
- consistency: F(U, U, w) = 2 F(U).w for central, LLF and Roe;
- Roe against the textbook Roe solver (Roe averages, eigen-decomposition of
  the normal Jacobian, written from the Euler equations in
  test_oracle_flux.py), LLF against the textbook Rusanov flux;
- supersonic upwinding: with every Roe wave along +n the flux is the left
  physical flux, which a reversed upwind sign fails.

JST dissipation (makeArtificialDissipation2, DEF:11484-11993), on a chain of
sides laid through a real level's nodes (unit weights, V/dt = 1) where the
scheme's terms have closed forms:
- linear U and p: the undivided Laplacian vanishes inside the chain, the
  pressure switch is zero, and only the chain ends see the fourth-difference
  term, -/+ k4 b / 3;
- quadratic p: the switch is 12 |2c / (p_{i-1} + 2 p_i + p_{i+1})| and a
  linear momentum sees only the second-difference term;
- matrix form on one side across a contact (density jump, equal velocity
  and pressure): the Roe matrix acts as |u.n| on the jump;
- on a jittered mesh: zero for a uniform state, and conservative (the node
  sums cancel) for a random one.
"""
import numpy as np
import pytest

import oracle
from cases import euler, levels
from test_oracle_flux import G, phys_flux, random_state, roe_textbook


def cases(ft, **kw):
    base = dict(fluxtype=ft, secondorderdissipationfactor=0.4, fourthorderdissipationfactor=0.2)
    base.update(kw)
    return euler(1, **base)


@pytest.mark.parametrize("ft", [1, 2, 3])
def test_simple_flux_consistency(ft):
    P = cases(ft).to_struct()
    rng = np.random.default_rng(ft)
    for _ in range(100):
        u, p = random_state(rng, mach=rng.uniform(0.05, 2.5))
        w = rng.normal(size=3) * rng.uniform(0.01, 3.0)
        f = oracle.simple_flux(P, ft, u, p, u, p, w)
        n = w / np.linalg.norm(w)
        ref, _ = phys_flux(u, n)
        ref *= 2 * np.linalg.norm(w)
        assert np.allclose(f, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max()), (f, ref)


def test_roe_flux_is_textbook_roe():
    P = cases(3).to_struct()
    rng = np.random.default_rng(7)
    for _ in range(200):
        ul, pl = random_state(rng, mach=rng.uniform(0.0, 1.8))
        ur, pr = random_state(rng, mach=rng.uniform(0.0, 1.8))
        w = rng.normal(size=3) * rng.uniform(0.1, 2.0)
        n = w / np.linalg.norm(w)
        f = oracle.simple_flux(P, 3, ul, pl, ur, pr, w)
        ref = 2 * np.linalg.norm(w) * roe_textbook(ul, ur, n)
        assert np.allclose(f, ref, rtol=1e-10, atol=1e-11 * np.abs(ref).max()), (f, ref)


def test_llf_flux_is_rusanov():
    P = cases(2).to_struct()
    rng = np.random.default_rng(8)
    for _ in range(200):
        ul, pl = random_state(rng, mach=rng.uniform(0.0, 1.8))
        ur, pr = random_state(rng, mach=rng.uniform(0.0, 1.8))
        w = rng.normal(size=3) * rng.uniform(0.1, 2.0)
        n = w / np.linalg.norm(w)
        fl, _ = phys_flux(ul, n)
        fr, _ = phys_flux(ur, n)
        lam = max(abs(ul[1:4] @ n / ul[0]) + np.sqrt(G * pl / ul[0]),
                  abs(ur[1:4] @ n / ur[0]) + np.sqrt(G * pr / ur[0]))
        ref = 2 * np.linalg.norm(w) * (0.5 * (fl + fr) - 0.5 * lam * (ur - ul))
        f = oracle.simple_flux(P, 2, ul, pl, ur, pr, w)
        assert np.allclose(f, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max()), (f, ref)


def test_roe_supersonic_upwind():
    P = cases(3).to_struct()
    rng = np.random.default_rng(9)
    n = np.array([0.6, 0.0, 0.8])
    for _ in range(50):
        rho = rng.uniform(0.8, 1.2)
        p = rng.uniform(0.8, 1.2)
        c = np.sqrt(G * p / rho)
        vel = n * c * rng.uniform(2.0, 3.0)
        ul = np.concatenate([[rho], rho * vel, [p / (G - 1) + 0.5 * rho * vel @ vel]])
        rho2, p2 = rho * rng.uniform(0.9, 1.1), p * rng.uniform(0.9, 1.1)
        ur = np.concatenate([[rho2], rho2 * vel, [p2 / (G - 1) + 0.5 * rho2 * vel @ vel]])
        f = oracle.simple_flux(P, 3, ul, p, ur, p2, 0.5 * n)
        ref, _ = phys_flux(ul, n)
        assert np.allclose(f, ref, rtol=1e-12, atol=1e-12), (f, ref)


# ---- JST on a chain of sides -------------------------------------------

def chain(m):
    """A real level (so the boundary records are valid) whose sides are
    replaced by the chain 0-1-...-(m-1) with unit weights along x."""
    L = levels(6, 1)[0]
    side = np.array([np.arange(m - 1), np.arange(1, m)], dtype=L.side_index.dtype)
    w = np.zeros((3, m - 1))
    w[0] = 1.0
    C = type(L)(**{**L.__dict__})
    C.side_index = side
    C.side_w = w
    return C


def jst(iv, L, u, p, dt):
    o = oracle.Oracle(iv.to_struct(), [L])
    o.set(1, "U", u)
    o.set(1, "PRESSURE", p[None])
    o.set(1, "DT", dt[None])
    o.apply(1, "DISSIPATION")
    d = o.get(1, "DISSIPATION")
    o.close()
    return d


def base_state(n):
    u = np.zeros((6, n))
    u[0] = 1.0
    u[1] = 0.5
    u[4] = 2.5
    return u


def test_jst_linear_field_only_chain_ends():
    m = 12
    L = chain(m)
    n = L.n_nodes
    k4 = 0.2
    # no second difference: at the chain ends the one-sided pressure sums
    # would switch it on
    iv = cases(1, secondorderdissipationfactor=0.0, fourthorderdissipationfactor=k4)
    u = base_state(n)
    b = np.array([0.01, -0.02, 0.03, 0.015, 0.05])
    i = np.arange(m)
    for c in range(5):
        u[c, :m] = u[c, 0] + b[c] * i
    p = np.full(n, 1.0)
    p[:m] = 1.0 + 0.004 * i
    dt = L.volume.copy()  # V/dt = 1
    d = jst(iv, L, u, p, dt)[:, :m]
    # L_0 = U_1 - U_0 = b (E + p for the energy) and L_{m-1} = -b, zero
    # between; conn 1 at the ends and 2 inside, so d0 = 1/3 on the end sides
    bb = b.copy()
    bb[4] += 0.004
    ref = np.zeros((5, m))
    ref[:, 0] = -k4 * bb / 3
    ref[:, 1] = k4 * bb / 3
    ref[:, m - 2] = -k4 * bb / 3
    ref[:, m - 1] = k4 * bb / 3
    assert np.allclose(d, ref, rtol=0, atol=1e-14), np.abs(d - ref).max()


def test_jst_pressure_switch_second_difference():
    m = 12
    L = chain(m)
    n = L.n_nodes
    k2 = 0.4
    iv = cases(1, secondorderdissipationfactor=k2, fourthorderdissipationfactor=0.0)
    u = base_state(n)
    i = np.arange(m)
    bmom = 0.02
    u[1, :m] = 0.5 + bmom * i
    cq = 0.003
    p = np.full(n, 1.0)
    p[:m] = 1.0 + cq * i * i
    dt = L.volume.copy()
    d = jst(iv, L, u, p, dt)
    nu = np.zeros(m)
    for k in range(1, m - 1):
        nu[k] = 12 * abs(2 * cq / (p[k - 1] + 2 * p[k] + p[k + 1]))
    nu[0] = 12 * abs((p[1] - p[0]) / (p[0] + p[1]))
    nu[m - 1] = 12 * abs((p[m - 2] - p[m - 1]) / (p[m - 1] + p[m - 2]))
    conn = np.full(m, 2.0)
    conn[0] = conn[-1] = 1.0
    for k in range(2, m - 2):
        d2l = k2 * max(nu[k - 1], nu[k]) / (conn[k - 1] + conn[k])
        d2r = k2 * max(nu[k], nu[k + 1]) / (conn[k] + conn[k + 1])
        ref = bmom * (d2l - d2r)
        assert abs(d[1, k] - ref) <= 1e-14, (k, d[1, k], ref)


def test_jst_matrix_contact_is_scaled_by_normal_velocity():
    L = chain(2)
    n = L.n_nodes
    k4 = 0.3
    iv = cases(3, usematrixdissipation=True, fourthorderdissipationfactor=k4)
    u = base_state(n)
    vel = np.array([0.4, -0.3, 0.2])
    p = np.full(n, 1.0)
    for k, rho in ((0, 1.0), (1, 1.3)):
        u[0, k] = rho
        u[1:4, k] = rho * vel
        u[4, k] = p[k] / (G - 1) + 0.5 * rho * vel @ vel
    dt = L.volume.copy()
    d = jst(iv, L, u, p, dt)
    jump = u[:5, 1] - u[:5, 0]
    # L_0 = jump, L_1 = -jump; |w| = 1 along x; |A| jump = |u_x| jump
    ref0 = -k4 * abs(vel[0]) * 2 * jump
    assert np.allclose(d[:, 0], ref0, rtol=1e-12, atol=1e-14), (d[:, 0], ref0)
    assert np.allclose(d[:, 1], -ref0, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("ft,matrix", [(1, False), (3, False), (3, True)])
def test_jst_uniform_and_conservative(ft, matrix):
    lv = levels(6, 1, wall=3)  # all farfield: the freestream start is uniform
    iv = cases(ft, usematrixdissipation=matrix)
    o = oracle.Oracle(iv.to_struct(), list(lv))
    o.init_freestream()
    o.apply(1, "PRESSURE_ABS")
    o.apply(1, "TIMESTEPS")
    o.apply(1, "DISSIPATION")
    d = o.get(1, "DISSIPATION")
    assert np.abs(d).max() < 1e-15
    rng = np.random.default_rng(3)
    u = o.get(1, "U")
    u[:5] *= 1 + 0.05 * rng.uniform(-1, 1, size=u[:5].shape)
    o.set(1, "U", u)
    o.apply(1, "PRESSURE_ABS")
    o.apply(1, "TIMESTEPS")
    o.apply(1, "DISSIPATION")
    d = o.get(1, "DISSIPATION")
    assert np.abs(d).max() > 1e-6
    assert np.abs(d.sum(axis=1)).max() < 1e-13 * np.abs(d).max() * d.shape[1]
    o.close()
