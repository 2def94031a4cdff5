"""Pins the oracle's flux functions with closed forms (CPU).

- HLLC (makeRHS interior, DEF:1794-2088) consistency: F(U, U, w) equals the
  physical flux |2w| F(U).n with n = w/|w|; rotational symmetry
  F(U_R, U_L, -w) = -F(U_L, U_R, w); supersonic upwinding: with every wave
  moving along +n the flux is the left physical flux;
- farfield boundary flux (DEF:2258-2446) against a textbook Roe solver
  (Roe 1981: Roe averages, eigen-decomposition of the normal Jacobian,
  no entropy fix), written here from the Euler equations, not from the
  reference.
"""
import numpy as np

import oracle
from cases import euler

G = 1.4


def phys_flux(u, n):
    rho, m, E = u[0], np.asarray(u[1:4]), u[4]
    vel = m / rho
    p = (G - 1.0) * (E - 0.5 * rho * vel @ vel)
    vn = vel @ n
    return np.concatenate([[rho * vn], m * vn + p * n, [(E + p) * vn]]), p


def random_state(rng, mach=0.6):
    rho = rng.uniform(0.5, 2.0)
    vel = rng.normal(size=3)
    vel *= mach / np.linalg.norm(vel)
    p = rng.uniform(0.8, 2.0)
    c = np.sqrt(G * p / rho)
    vel *= c
    E = p / (G - 1.0) + 0.5 * rho * vel @ vel
    return np.concatenate([[rho], rho * vel, [E]]), p


def test_hllc_consistency():
    P = euler(1).to_struct()
    rng = np.random.default_rng(0)
    for _ in range(200):
        u, p = random_state(rng, mach=rng.uniform(0.05, 2.5))
        w = rng.normal(size=3) * rng.uniform(0.01, 3.0)
        f = oracle.hllc_flux(P, u, p, u, p, w)
        n = 2 * w / np.linalg.norm(2 * w)
        ref, _ = phys_flux(u, n)
        ref *= np.linalg.norm(2 * w)
        assert np.allclose(f, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max()), (f, ref)


def test_hllc_rotational_symmetry():
    P = euler(1).to_struct()
    rng = np.random.default_rng(1)
    for _ in range(200):
        ul, pl = random_state(rng, mach=rng.uniform(0.0, 1.5))
        ur, pr = random_state(rng, mach=rng.uniform(0.0, 1.5))
        w = rng.normal(size=3)
        f1 = oracle.hllc_flux(P, ul, pl, ur, pr, w)
        f2 = oracle.hllc_flux(P, ur, pr, ul, pl, -w)
        assert np.allclose(f1, -f2, rtol=1e-11, atol=1e-11 * np.abs(f1).max())


def test_hllc_supersonic_upwind():
    P = euler(1).to_struct()
    rng = np.random.default_rng(2)
    for _ in range(100):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        rho, p = rng.uniform(0.8, 1.2), rng.uniform(0.8, 1.2)
        c = np.sqrt(G * p / rho)
        vel = 3.0 * c * n
        E = p / (G - 1.0) + 0.5 * rho * vel @ vel
        ul = np.concatenate([[rho], rho * vel, [E]])
        ur = ul * np.concatenate([[1.05], [1.05] * 3, [1.1]])
        prr = (G - 1) * (ur[4] - 0.5 * ur[1:4] @ ur[1:4] / ur[0])
        w = 0.5 * n * rng.uniform(0.1, 2.0)
        f = oracle.hllc_flux(P, ul, p, ur, prr, w)
        ref, _ = phys_flux(ul, n)
        assert np.allclose(f, ref * np.linalg.norm(2 * w), rtol=1e-12)


def roe_textbook(ul, ur, n):
    """Roe's approximate Riemann flux through unit normal n (no entropy fix)."""
    def prim(u):
        rho = u[0]
        vel = u[1:4] / rho
        p = (G - 1.0) * (u[4] - 0.5 * rho * vel @ vel)
        H = (u[4] + p) / rho
        return rho, vel, p, H
    rl, vl, pl, Hl = prim(ul)
    rr, vr, pr, Hr = prim(ur)
    sl, sr = np.sqrt(rl), np.sqrt(rr)
    vel = (sl * vl + sr * vr) / (sl + sr)
    H = (sl * Hl + sr * Hr) / (sl + sr)
    q2 = vel @ vel
    c2 = max((G - 1.0) * (H - 0.5 * q2), 1.0e-5)
    c = np.sqrt(c2)
    vn = vel @ n
    # right eigenvectors and wave strengths of the normal Jacobian
    du = ur - ul
    drho, dp = rr - rl, pr - pl
    dvn = (vr - vl) @ n
    dv = (vr - vl)
    rho_hat = sl * sr
    a1 = (dp - rho_hat * c * dvn) / (2 * c2)
    a5 = (dp + rho_hat * c * dvn) / (2 * c2)
    a2 = drho - dp / c2
    r1 = np.concatenate([[1.0], vel - c * n, [H - c * vn]])
    r5 = np.concatenate([[1.0], vel + c * n, [H + c * vn]])
    r2 = np.concatenate([[1.0], vel, [0.5 * q2]])
    dvt = dv - dvn * n
    r3 = np.concatenate([[0.0], rho_hat * dvt, [rho_hat * (vel @ dvt)]])
    diss = abs(vn - c) * a1 * r1 + abs(vn + c) * a5 * r5 + abs(vn) * (a2 * r2 + r3)
    fl, _ = phys_flux(ul, n)
    fr, _ = phys_flux(ur, n)
    return 0.5 * (fl + fr) - 0.5 * diss


def test_farfield_roe_textbook():
    iv = euler(1, machnumber=0.6, alpha=3.0)
    P = iv.to_struct()
    inf = np.array(iv.inflowfield[0])  # rho, u, v, w, e, p
    uinf = np.concatenate([[inf[0]], inf[0] * inf[1:4], [inf[0] * inf[4]]])
    rng = np.random.default_rng(3)
    for _ in range(200):
        u, p = random_state(rng, mach=rng.uniform(0.1, 0.9))
        u = 0.5 * u + 0.5 * uinf
        p = (G - 1) * (u[4] - 0.5 * u[1:4] @ u[1:4] / u[0])
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        f = oracle.farfield_flux(P, 3, u, p, n)
        ref = roe_textbook(u, uinf, n)
        assert np.allclose(f, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max()), (f, ref)
