"""Mesh derivatives of the viscous and Spalart-Allmaras terms of the level-1
residual and of the drag with its friction term (flite_mesh_adjoint and
flite_drag_mesh_gradient with ReynoldsNumber > 0, flite_drag_state_gradient
with friction), on the GPU.

The reference has no differentiated code, so these are pinned by central
differences of the oracle's residual and drag (the CPU restatement of the
reference's):
  * each mesh quantity on its own: volumes, wall distances and coordinates
    along a random direction; side and boundary-side weights one entry at a
    time, where the entry must equal the central difference or a one-sided
    one (the nu~ convection carries |w.a|, whose kink can lie within the
    step, and moving one weight alone opens the dual cells);
  * the drag's state gradient along random directions of the state;
  * the whole coordinate chain L(X) = drag - lam^T R at a fixed state, with
    the metrics and wall distances rebuilt by the preprocessor from
    X +- h dX, against fp_mesh_vjp of the library's shares plus the wall
    distance's transpose.
Laminar (Re 200) and RANS-SA (Re 1e4, nu~ of a few laminar viscosities)
about smooth states on a jittered bump box, M 0.5-0.7 with 2 degrees of
sideslip (every farfield face then has |u.n| bounded away from the Roe
flux's kink).
"""
import dataclasses

import numpy as np
import pytest

from flite3d_amd import mesh
from flite3d_amd.device import Solver
from oracle import Oracle

from cases import laminar, rans
from test_adjoint import smooth_state

pytestmark = pytest.mark.gpu

H = 1e-6


def inputs(kind, order):
    if kind == "laminar":
        return laminar(hllcfluxorder=order, reynoldsnumber=200.0, wallsareisentropic=True, beta=2.0)
    return rans(hllcfluxorder=order, reynoldsnumber=1.0e4, beta=2.0)


def make(kind, order, n=5):
    """Raw mesh, levels, a smooth state and random multipliers."""
    raw = mesh.bump_box(n, jitter=0.15, seed=0)
    lv, _ = mesh.build_levels(*raw, 2)
    iv = inputs(kind, order)
    s = Solver(iv.to_struct(), list(lv))
    s.init_freestream()
    inf = np.array(iv.inflowfield[0])
    base = np.zeros((6, lv[0].n_nodes))
    base[:5] = inf[:5, None]
    base[1:4] *= inf[0]
    base[5] = s.get(1, "U")[5]
    s.close()
    u = smooth_state(lv, base)
    if kind == "rans":
        # nu~ = 0 on the walls, as the wall condition sets it: elsewhere the
        # source's (nu~ / d)^2 at the clamped d = 1e-10 of a wall node would
        # dominate lam^T R by 1e16 and hide every other term
        x, y, z = lv[0].coords
        u[5] = np.where(lv[0].wall_dist > 0, 3.0 * (1.0 + 0.3 * np.sin(1.7 * x + 0.9 * y - 1.3 * z + 0.2)), 0.0)
    nc = 6 if kind == "rans" else 5
    lam = np.random.default_rng(61).standard_normal((nc, lv[0].n_nodes))
    return raw, lv, u, lam


def evaluate(kind, order, lv, u, lam):
    """(lam^T R, drag) at state u on the fine level of lv, by the oracle: R
    as the tangent and adjoint linearise it (makeRHS, the viscous
    dissipation and, with SA, the nu~ convection, diffusion and source), the
    drag of its monitor (pressure and friction)."""
    o = Oracle(inputs(kind, order).to_struct(), [lv[0]])
    o.init_freestream()
    o.set(1, "U", u)
    o.apply(1, "PRESSURE")
    o.set(1, "DISSIPATION", np.zeros((5, lv[0].n_nodes)))  # makeViscosityTerm subtracts into it
    o.apply(1, "VISCOUS")
    if kind == "rans":
        o.apply(1, "SA_DIFFUSION")
    o.apply(1, "RHS")
    nc = lam.shape[0]
    r = o.get(1, "RHS")[:nc].copy()
    r[:5] += o.get(1, "DISSIPATION")[:5]
    if kind == "rans":
        r[5] += o.get(1, "TURB_DIFFUSION")[0]
    drag = float(o.monitor()["forces"][1])
    o.close()
    return float(np.sum(lam * r)), drag


def shares(kind, order, lv, u, lam):
    s = Solver(inputs(kind, order).to_struct(), list(lv))
    s.set(1, "U", u)
    out = s.mesh_adjoint(lam)
    jout = s.drag_mesh_gradient()
    s.close()
    return out, jout


def replaced(lv, name, x):
    return [dataclasses.replace(lv[0], **{name: np.ascontiguousarray(x)})] + list(lv[1:])


def check_entries(f, x0, g, name, rng, k=4, h=1e-8, cols=None):
    """g's entries against central or one-sided differences of f(x) at
    k random columns (of cols if given); at most 3 one-sided."""
    f0 = f(x0)
    one_sided = 0
    for e in rng.choice(x0.shape[1] if cols is None else cols, k, replace=False):
        for q in range(3):
            d = np.zeros_like(x0)
            d[q, e] = 1.0
            fp, fm = f(x0 + h * d), f(x0 - h * d)
            an = g[q, e]
            tol = 1e-4 * max(abs(an), 1e-3)
            c_ok = abs((fp - fm) / (2 * h) - an) <= tol
            s_ok = abs((fp - f0) / h - an) <= tol or abs((f0 - fm) / h - an) <= tol
            assert c_ok or s_ok, (name, e, q, (fp - fm) / (2 * h), (fp - f0) / h, (f0 - fm) / h, an)
            one_sided += int(not c_ok)
    assert one_sided <= 3, (name, one_sided)


def check_direction(f, x0, g, name, rng, mask=None):
    d = rng.standard_normal(x0.shape) * np.abs(x0).max()
    if mask is not None:
        d = d * mask
    fd = (f(x0 + H * d) - f(x0 - H * d)) / (2 * H)
    an = float(np.sum(g * d))
    print(name, fd, an, flush=True)
    assert abs(an) > 0, name
    assert abs(fd - an) <= 1e-6 * abs(an), (name, fd, an)


@pytest.mark.parametrize("kind,order", [("laminar", 1), ("rans", 1), ("rans", 2)])
def test_viscous_mesh_adjoint_matches_central_differences(kind, order):
    """d(lam^T R) with respect to each mesh quantity of the viscous and SA
    terms (plus the convective ones already pinned inviscid)."""
    print(f"[case] viscous mesh adjoint {kind} order={order}", flush=True)
    _, lv, u, lam = make(kind, order)
    (gw, gwb, gv, gx, gd), _ = shares(kind, order, lv, u, lam)
    rng = np.random.default_rng(62)

    def lamR(name):
        return lambda x: evaluate(kind, order, replaced(lv, name, x), u, lam)[0]

    L0 = lv[0]
    check_direction(lamR("volume"), L0.volume, gv, "volume", rng)
    if kind == "rans":
        # wall nodes sit at distance 0, where the source's clamp at 1e-10
        # makes the one-sided derivatives differ: move the others only
        check_direction(lamR("wall_dist"), L0.wall_dist, gd, "wall_dist", rng, mask=L0.wall_dist > 0)
    else:
        assert np.all(gd == 0.0)
    if order == 2:
        check_direction(lamR("coords"), L0.coords, gx, "coords", rng)
    check_entries(lamR("side_w"), L0.side_w, gw, "side_w", rng)
    check_entries(lamR("bside_w"), L0.bside_w, gwb, "bside_w", rng)


@pytest.mark.parametrize("kind", ["laminar", "rans"])
def test_drag_with_friction_gradients_match_central_differences(kind):
    """The drag with its friction term: d/dU along random directions, and
    d/d(volumes, weights) one quantity at a time."""
    print(f"[case] drag with friction {kind}", flush=True)
    _, lv, u, lam = make(kind, 1)
    rng = np.random.default_rng(63)
    s = Solver(inputs(kind, 1).to_struct(), list(lv))
    s.set(1, "U", u)
    drag, dJdU = s.drag_gradient()
    s.close()
    nc = 6 if kind == "rans" else 5
    assert dJdU.shape == (nc, lv[0].n_nodes)
    for _ in range(3):
        du = rng.standard_normal((nc, lv[0].n_nodes)) * np.abs(u[:nc]).max(axis=1, keepdims=True)
        up, um = u.copy(), u.copy()
        up[:nc] += H * du
        um[:nc] -= H * du
        fd = (evaluate(kind, 1, lv, up, lam)[1] - evaluate(kind, 1, lv, um, lam)[1]) / (2 * H)
        an = float(np.sum(dJdU * du))
        print("dJ/dU", fd, an, flush=True)
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
    _, (drag2, jw, jwb, jv) = shares(kind, 1, lv, u, lam)
    assert abs(drag2 - drag) <= 1e-14 * abs(drag)

    def J(name):
        return lambda x: evaluate(kind, 1, replaced(lv, name, x), u, lam)[1]

    # the drag reads the weights of the sides at the counted wall nodes
    # (z > 0.01 above the ground plane): draw the entries from those
    L0 = lv[0]
    counted = (L0.wall_dist == 0.0) & (L0.coords[2] > 0.01)
    sides = np.nonzero(counted[L0.side_index[0]] | counted[L0.side_index[1]])[0]
    bsides = np.nonzero(counted[L0.bside_index[0]] | counted[L0.bside_index[1]])[0]
    assert len(sides) >= 4 and len(bsides) >= 4
    check_direction(J("volume"), L0.volume, jv, "drag/volume", rng)
    check_entries(J("side_w"), L0.side_w, jw, "drag/side_w", rng, cols=sides)
    check_entries(J("bside_w"), L0.bside_w, jwb, "drag/bside_w", rng, cols=bsides)


@pytest.mark.parametrize("kind,order", [("laminar", 1), ("rans", 2)])
def test_viscous_coordinate_gradient_matches_central_differences(kind, order):
    """d/dX of drag - lam^T R through every mesh quantity the preprocessor
    rebuilds from X: weights, volumes, coordinates and wall distances."""
    print(f"[case] viscous coordinate gradient {kind} order={order}", flush=True)
    raw, lv, u, lam = make(kind, order)
    xyz, tets, btri, bind = raw
    (gw, gwb, gv, gxd, gd), (_, jw, jwb, jv) = shares(kind, order, lv, u, lam)
    gxn = -gxd + mesh.wall_distance_vjp(lv[0], -gd)
    gx = mesh.mesh_vjp(xyz, tets, btri, bind, jw - gw, jwb - gwb, gvol=jv - gv, gxn=gxn)
    rng = np.random.default_rng(64)
    for _ in range(2):
        d = rng.standard_normal(xyz.shape)
        vals = []
        for sgn in (1.0, -1.0):
            lvp, _ = mesh.build_levels(xyz + sgn * H * d, tets, btri, bind, 2)
            r, dr = evaluate(kind, order, lvp, u, lam)
            vals.append(dr - r)
        fd = (vals[0] - vals[1]) / (2 * H)
        an = float(np.sum(gx * d))
        print(fd, an, flush=True)
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-12), (fd, an)
