"""Pins the oracle's edge loops and multigrid transfer with invariants (CPU).

Each check is chosen so that a plausible transcription error fails it:
- freestream preservation of the full residual (interior + wall/farfield
  boundary loops, HLLC-1 and HLLC-2, all levels): needs the side-coefficient
  closure, the +/- scatter signs and the 3:1 boundary weighting right;
- conservation: with boundary weights zeroed the interior scatter sums to 0;
- Green-Gauss gradient (computeGradient, DEF:2923-3245) exact for linear
  fields at every node;
- LocalMinMax (DEF:3248-3459) and makeTimeSteps (DEF:13335-13610) against
  brute-force numpy evaluations over the side list;
- viscous term (DEF:5198-5893): constant stress gives zero momentum residual
  at interior nodes; vorticity/divergence exact for linear velocity;
- restriction (DEF:12572-12690) conserves sum V u; growing-V cycles converge
  to a fixed point where the coarse-grid correction vanishes.
"""
import numpy as np
import pytest

import oracle
from cases import euler, laminar, levels, rans


def make(iv, lv):
    return oracle.Oracle(iv.to_struct(), list(lv))


@pytest.mark.parametrize("order", [1, 2])
def test_freestream_preservation(order):
    lv = levels(6, 3, wall=3)
    o = make(euler(order), lv)
    o.init_freestream()
    for l in (1, 2, 3):
        o.apply(l, "PRESSURE")
        o.apply(l, "RHS")
        r = o.get(l, "RHS")[:5]
        scale = np.abs(lv[l - 1].side_w).max()
        assert np.abs(r).max() < 1e-13 * max(1.0, 1.0 / scale), (l, np.abs(r).max())


def test_interior_conservation():
    lv = list(levels(6, 1))
    L = lv[0]
    L0 = type(L)(**{**L.__dict__})
    L0.bside_w = np.zeros_like(L.bside_w)
    o = make(euler(2), [L0])
    o.init_freestream()
    rng = np.random.default_rng(0)
    u = o.get(1, "U")
    u[:5] *= 1 + 0.05 * rng.uniform(-1, 1, size=u[:5].shape)
    o.set(1, "U", u)
    o.apply(1, "PRESSURE")
    o.apply(1, "RHS")
    r = o.get(1, "RHS")[:5]
    assert np.abs(r.sum(axis=1)).max() < 1e-13 * np.abs(r).max() * r.shape[1]


def linear_state(x, y, z):
    rho = 1.0 + 0.1 * x - 0.2 * y + 0.05 * z
    u = 0.5 + 0.3 * x + 0.1 * y - 0.2 * z
    v = -0.1 + 0.2 * x - 0.3 * y + 0.1 * z
    w = 0.2 - 0.1 * x + 0.05 * y + 0.3 * z
    E = 2.5 + 0.4 * x + 0.1 * y - 0.3 * z
    return np.array([rho, rho * u, rho * v, rho * w, E, np.zeros_like(x)])


GRAD = np.array([[0.1, -0.2, 0.05], [0.3, 0.1, -0.2], [0.2, -0.3, 0.1],
                 [-0.1, 0.05, 0.3], [0.4, 0.1, -0.3]])  # rows rho,u,v,w,E


def test_green_gauss_exact_for_linear_fields():
    lv = levels(6, 1)
    L = lv[0]
    o = make(euler(2), lv)
    x, y, z = L.coords
    o.set(1, "U", linear_state(x, y, z))
    o.apply(1, "PRESSURE")
    o.apply(1, "RHS")
    g = o.get(1, "GRADIENT")
    for q in range(5):
        for d in range(3):
            err = np.abs(g[5 * d + q] - GRAD[q, d]).max()
            assert err < 1e-11, (q, d, err)


def test_local_minmax_brute_force():
    lv = levels(5, 1)
    L = lv[0]
    o = make(euler(2), lv)
    o.init_freestream()
    rng = np.random.default_rng(1)
    u = o.get(1, "U")
    u[:5] *= 1 + 0.1 * rng.uniform(-1, 1, size=u[:5].shape)
    o.set(1, "U", u)
    o.apply(1, "PRESSURE")
    o.apply(1, "RHS")
    g = o.get(1, "GRADIENT")
    q = np.array([u[0], u[1] / u[0], u[2] / u[0], u[3] / u[0], u[4]])
    i1, i2 = L.side_index
    for c in range(5):
        mn = np.full(L.n_nodes, np.inf)
        mx = np.full(L.n_nodes, -np.inf)
        for a, b in ((i1, i2), (i2, i1)):
            np.minimum.at(mn, a, np.minimum(q[c][a], q[c][b]))
            np.maximum.at(mx, a, np.maximum(q[c][a], q[c][b]))
        assert np.abs(g[15 + c] - (mn - q[c])).max() < 1e-14
        assert np.abs(g[20 + c] - (mx - q[c])).max() < 1e-14


def test_timesteps_brute_force():
    lv = levels(5, 1, first_cell=0.02)
    L = lv[0]
    iv = laminar(reynoldsnumber=100.0)
    o = make(iv, lv)
    o.init_freestream()
    rng = np.random.default_rng(2)
    u = o.get(1, "U")
    u[:5] *= 1 + 0.05 * rng.uniform(-1, 1, size=u[:5].shape)
    o.set(1, "U", u)
    mu = 0.01 * (1 + rng.uniform(0, 1, size=L.n_nodes))
    o.set(1, "LAMINAR_VISCOSITY", mu)
    o.apply(1, "PRESSURE")
    o.apply(1, "TIMESTEPS")
    dt = o.get(1, "DT")[0]
    p = o.get(1, "PRESSURE")[0]
    c2 = np.maximum(1.4 * p / u[0], 1e-3)
    vis = np.maximum(1.0 / iv.reynoldsnumber, mu)
    lam = np.zeros(L.n_nodes)
    lv_ = np.zeros(L.n_nodes)
    i1, i2 = L.side_index
    w = L.side_w
    al = (w * w).sum(0)
    for a, b in ((i1, i2), (i2, i1)):
        vn = np.abs((w * u[1:4, a]).sum(0) / u[0, a])
        np.add.at(lam, a, vn + np.sqrt(c2[a] * al))
        np.add.at(lv_, a, 2 * vis[b] * al / u[0, b])
    b1, b2 = L.bside_index[0], L.bside_index[1]
    wb = L.bside_w
    alb = (wb * wb).sum(0)
    vn1 = np.abs((wb * u[1:4, b1]).sum(0) / u[0, b1])
    vn2 = np.abs((wb * u[1:4, b2]).sum(0) / u[0, b2])
    cc1, cc2 = np.sqrt(c2[b1] * alb), np.sqrt(c2[b2] * alb)
    d1, d2 = 2 * vis[b1] * alb / u[0, b1], 2 * vis[b2] * alb / u[0, b2]
    np.add.at(lam, b1, 2 * vn1 + vn2 + 2 * cc1 + cc2)
    np.add.at(lam, b2, 2 * vn2 + vn1 + 2 * cc2 + cc1)
    np.add.at(lv_, b1, 2 * d1 + d2)
    np.add.at(lv_, b2, 2 * d2 + d1)
    ref = 0.5 / (lam / L.volume + lv_ / L.volume ** 2)
    assert np.abs(dt / ref - 1).max() < 1e-12


def test_viscous_constant_stress():
    lv = levels(6, 1)
    L = lv[0]
    iv = laminar(reynoldsnumber=1000.0, wallsareisentropic=True)
    o = make(iv, lv)
    x, y, z = L.coords
    rho = np.ones_like(x)
    vel = np.array([0.5 + 0.3 * x + 0.1 * y - 0.2 * z, -0.1 + 0.2 * x - 0.3 * y + 0.1 * z,
                    0.2 - 0.1 * x + 0.05 * y + 0.3 * z])
    p = np.full_like(x, 1.0 / (1.4 * 0.25))  # constant T
    E = p / 0.4 + 0.5 * rho * (vel ** 2).sum(0)
    o.set(1, "U", np.array([rho, *vel, E, 0 * x]))
    o.apply(1, "PRESSURE")
    o.apply(1, "DISSIPATION")
    o.apply(1, "VISCOUS")
    d = o.get(1, "DISSIPATION")
    interior = np.arange(L.n_nodes) >= L.n_bnodes
    mom = np.abs(d[1:4, interior]).max()
    assert mom < 1e-12 * np.abs(d[1:4]).max() + 1e-15, mom
    vort = np.sqrt((0.1 - 0.2) ** 2 + (0.1 - 0.05) ** 2 + (-0.1 + 0.2) ** 2)
    assert np.abs(o.get(1, "VORTICITY")[0][interior] - vort).max() < 1e-12
    assert np.abs(o.get(1, "DIVERGENCE")[0][interior] - (0.3 - 0.3 + 0.3)).max() < 1e-12


def test_restriction_conserves_volume_integral():
    lv = levels(6, 3)
    o = make(euler(1), lv)
    o.init_freestream()
    rng = np.random.default_rng(5)
    u = o.get(1, "U")
    u[:5] *= 1 + 0.1 * rng.uniform(-1, 1, size=u[:5].shape)
    o.set(1, "U", u)
    o.map_u_fine_to_coarse(1)
    o.map_u_fine_to_coarse(2)
    for l in (1, 2, 3):
        s = (o.get(l, "U")[:5] * lv[l - 1].volume).sum(1)
        assert np.allclose(s, (u[:5] * lv[0].volume).sum(1), rtol=1e-13)


def test_growing_v_fixed_point():
    lv = levels(6, 3)
    o = make(euler(1, multigridscheme=3), lv)
    o.init_freestream()
    for c in range(1, 401):
        o.cycle(c)
    m = o.monitor()
    assert m["res"][0] < 1e-12, m["res"]
    u0 = o.get(1, "U")
    o.cycle(401)
    assert np.abs(o.get(1, "U") - u0)[:5].max() < 1e-11


def test_boundary_weighting_exact_for_linear_flux():
    """The boundary loop's 3:1 weighting (boundaryTerm 1, DEF:2233-2243):
    with the central interior flux (fluxType 1, no dissipation) and every
    boundary side of wall type (physical flux), a state whose fluxes are
    linear in position (density linear, velocity and pressure constant)
    has the residual V div F exactly at every node, the boundary nodes
    included; a 4:0 or 2:2 boundary weighting would not."""
    lv = levels(6, 1, jitter=0.15)
    L0 = lv[0]
    bi = L0.bside_index.copy()
    bi[2] = 1
    L = type(L0)(**{**L0.__dict__, "bside_index": bi})
    iv = euler(1, fluxtype=1, secondorderdissipationfactor=0.0, fourthorderdissipationfactor=0.0)
    o = make(iv, [L])
    o.init_freestream()
    b = np.array([0.3, -0.2, 0.4])
    vel = np.array([0.5, 0.1, -0.2])
    p0 = 1.0
    x = L.coords
    rho = 1.0 + b @ x
    u = np.zeros((6, L.n_nodes))
    u[0] = rho
    u[1:4] = rho * vel[:, None]
    u[4] = p0 / 0.4 + 0.5 * rho * (vel @ vel)
    o.set(1, "U", u)
    o.apply(1, "PRESSURE")
    o.apply(1, "RHS")
    r = o.get(1, "RHS")[:5]
    # div F: d(rho u_j)/dx_j = vel.b, momentum vel_k (vel.b), energy (|vel|^2 / 2)(vel.b)
    vb = vel @ b
    ref = np.vstack([np.full(L.n_nodes, vb), vel[:, None] * vb * np.ones(L.n_nodes),
                     np.full(L.n_nodes, 0.5 * (vel @ vel) * vb)]) * L.volume
    err = np.abs(r - ref).max() / np.abs(ref).max()
    assert err < 1e-12, err
    o.close()


def test_threads_give_the_same_bits():
    """The oracle's OpenMP side-flux loops (the timed CPU comparator) leave
    every result independent of the thread count."""
    lv = levels(6, 2, first_cell=0.01)
    out = []
    for threads in (1, 3):
        o = make(rans(multigridscheme=3), lv)
        o.set_threads(threads)
        o.init_freestream()
        o.cycle(1)
        o.cycle(2)
        out.append(o.get(1, "U"))
        o.close()
    assert np.array_equal(out[0], out[1])
