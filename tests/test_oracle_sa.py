"""Pins the oracle's Spalart-Allmaras terms with closed forms (CPU).

makeSASourceTerm (DEF:9285-9370), with the velocity zero so that the SA
convection vanishes and rhs(:,6) holds the source alone:
- log layer: nu~ = kappa u_tau d Re, Omega = u_tau / (kappa d), chi >> 1
  gives r = 1, fw = 1 and ft2 = 0, and the source reduces to
  V Re u_tau^2 (cw1 kappa^2 - cb1) = V Re u_tau^2 kappa^2 (1 + cb2) / sigma,
  the balance the model's cw1 is defined by (Spalart & Allmaras 1994); a
  wrong cw1, fw, r, S~ or Reynolds-number scaling fails it;
- far from walls (d -> infinity) the destruction vanishes and the source
  is the production -V cb1 (1 - ft2) Omega nu~ with the trip function
  ft2 = ct3 exp(-ct4 chi^2), plus the conservative-convection correction
  -V nu~ div(u) (the convection is discretised as div(u nu~));
- nu~ = 0 gives no source.
makeSADiffusion (DEF:7436-7808): -(V / (sigma Re)) [div((nu + nu~) grad nu~)
  + cb2 |grad nu~|^2]; for nu~ linear in x and nu constant the Green-Gauss
  operators are exact at every node, so the term is
  -(1 + cb2) V |b|^2 / (sigma Re).
makeTurbulentTimeSteps (DEF:12878-13099): against a brute-force numpy
evaluation over the side list, as the mean-flow time steps.
"""
import dataclasses

import numpy as np

import oracle
from cases import levels, rans

KAPPA, CB1, CB2, SIGMA, CT3, CT4 = 0.41, 0.1355, 0.622, 2.0 / 3.0, 1.1, 2.0
G = 1.4


def still_state(n, nut):
    """Zero velocity, rho = 1, p = 1, the given nu~."""
    u = np.zeros((6, n))
    u[0] = 1.0
    u[4] = 1.0 / (G - 1.0)
    u[5] = nut
    return u


def source(o, u, mu, vort, divg):
    o.set(1, "U", u)
    o.set(1, "LAMINAR_VISCOSITY", mu[None])
    o.set(1, "VORTICITY", vort[None])
    o.set(1, "DIVERGENCE", divg[None])
    o.apply(1, "PRESSURE")
    o.apply(1, "RHS")
    return o.get(1, "RHS")[5]


def test_sa_source_log_layer():
    Re = 1.0e6
    lv = levels(6, 1, first_cell=0.01)
    L = lv[0]
    o = oracle.Oracle(rans(reynoldsnumber=Re).to_struct(), [L])
    o.init_freestream()
    d = L.wall_dist
    m = d >= 0.05
    ut = 0.5
    nut = np.where(m, KAPPA * ut * d * Re, 0.0)
    vort = np.where(m, ut / (KAPPA * np.maximum(d, 1e-30)), 0.0)
    mu = np.full(L.n_nodes, 0.1 / Re * 1.0001)  # just above the laminarViscosity guard
    r = source(o, still_state(L.n_nodes, nut), mu, vort, np.zeros(L.n_nodes))
    ref = L.volume * Re * ut ** 2 * KAPPA ** 2 * (1 + CB2) / SIGMA
    assert np.abs(r[m] / ref[m] - 1).max() < 5e-4, np.abs(r[m] / ref[m] - 1).max()
    o.close()


def test_sa_source_away_from_walls():
    Re = 1.0e6
    lv = levels(6, 1, first_cell=0.01)
    L = dataclasses.replace(lv[0], wall_dist=np.full(lv[0].n_nodes, 1.0e8))
    o = oracle.Oracle(rans(reynoldsnumber=Re).to_struct(), [L])
    o.init_freestream()
    rng = np.random.default_rng(1)
    n = L.n_nodes
    nut = rng.uniform(0.2, 2.0, n)
    vort = rng.uniform(0.1, 3.0, n)
    divg = rng.uniform(-0.5, 0.5, n)
    mu = np.full(n, 1.0 / Re)  # chi = nu~
    r = source(o, still_state(n, nut), mu, vort, divg)
    ft2 = CT3 * np.exp(-CT4 * nut ** 2)
    ref = -L.volume * CB1 * (1 - ft2) * vort * nut - L.volume * nut * divg
    assert np.abs(r - ref).max() <= 1e-12 * np.abs(ref).max()
    r0 = source(o, still_state(n, np.zeros(n)), mu, vort, divg)
    assert np.abs(r0).max() == 0.0
    o.close()


def test_sa_diffusion_linear_field():
    Re = 5.0e5
    lv = levels(6, 1, jitter=0.15)
    L = lv[0]
    o = oracle.Oracle(rans(reynoldsnumber=Re).to_struct(), [L])
    o.init_freestream()
    b = np.array([0.7, -0.4, 1.3])
    x = L.coords
    u = still_state(L.n_nodes, 3.0 + b @ x)
    o.set(1, "U", u)
    o.set(1, "LAMINAR_VISCOSITY", np.full((1, L.n_nodes), 2.0 / Re))
    o.apply(1, "SA_DIFFUSION")
    t = o.get(1, "TURB_DIFFUSION")[0]
    ref = -(1 + CB2) * L.volume * (b @ b) / (SIGMA * Re)
    assert np.abs(t / ref - 1).max() < 1e-10, np.abs(t / ref - 1).max()
    o.close()


def test_sa_timesteps_brute_force():
    Re = 1.0e6
    lv = levels(5, 1, first_cell=0.02)
    L = lv[0]
    iv = rans(reynoldsnumber=Re)
    o = oracle.Oracle(iv.to_struct(), [L])
    o.init_freestream()
    rng = np.random.default_rng(4)
    n = L.n_nodes
    u = o.get(1, "U")
    u[:5] *= 1 + 0.05 * rng.uniform(-1, 1, size=u[:5].shape)
    u[5] = rng.uniform(0.0, 5.0, n)
    o.set(1, "U", u)
    mu = (1 + rng.uniform(0, 1, n)) / Re
    vort = rng.uniform(0, 2, n)
    divg = rng.uniform(-0.3, 0.3, n)
    o.set(1, "LAMINAR_VISCOSITY", mu[None])
    o.set(1, "VORTICITY", vort[None])
    o.set(1, "DIVERGENCE", divg[None])
    o.apply(1, "SA_TIMESTEPS")
    dtt = o.get(1, "DT_TURB")[0]
    tk = iv.turbk
    vis = (mu + 3.0 * CB2 * u[5] / Re) / SIGMA
    lam = np.zeros(n)
    lv_ = np.zeros(n)
    i1, i2 = L.side_index
    w = L.side_w
    al = (w * w).sum(0)
    for a, bb in ((i1, i2), (i2, i1)):
        vn = tk * np.sqrt(al) + np.abs((w * u[1:4, a]).sum(0) / u[0, a])
        np.add.at(lam, a, vn)
        np.add.at(lv_, a, 2 * vis[bb] * al)
    b1, b2 = L.bside_index[0], L.bside_index[1]
    wb = L.bside_w
    alb = (wb * wb).sum(0)
    vn1 = tk * np.sqrt(alb) + np.abs((wb * u[1:4, b1]).sum(0) / u[0, b1])
    vn2 = tk * np.sqrt(alb) + np.abs((wb * u[1:4, b2]).sum(0) / u[0, b2])
    np.add.at(lam, b1, 2 * vn1 + vn2)
    np.add.at(lam, b2, 2 * vn2 + vn1)
    d1, d2 = 2 * vis[b1] * alb, 2 * vis[b2] * alb
    np.add.at(lv_, b1, 2 * d1 + d2)
    np.add.at(lv_, b2, 2 * d2 + d1)
    d = 0.5 / (lam / L.volume + lv_ / L.volume ** 2)
    di = np.where(L.wall_dist > 1e-10, 1.0 / np.maximum(L.wall_dist, 1e-300), 1e10) ** 2
    cw1 = CB1 / KAPPA ** 2 + (1 + CB2) / SIGMA
    S = np.abs(divg) + CB1 * (vort + (1 + CT3) * u[5] / KAPPA ** 2 * di / Re) + cw1 * u[5] * di / Re
    ref = d / (1 + 0.5 * S * d)
    assert np.abs(dtt / ref - 1).max() < 1e-12, np.abs(dtt / ref - 1).max()
    o.close()
