"""Tangent and adjoint of the level-1 residual on the GPU.

The reference has no differentiated code, so these are pinned by the
mathematics: the tangent against central finite differences of the
residual the library itself computes (FLITE_OP_RHS, parity-checked against
the oracle elsewhere), the adjoint against the tangent by the dot-product
identity <lam, J du> = <J^T lam, du>, and the drag gradient against finite
differences of the drag.  Euler, first-order HLLC, M=0.8, alpha=1.25 deg
(the discrete-adjoint case's conditions) on a jittered 8^3 bump box and a
12^3 one after a few cycles, so that every flux branch and both boundary
kinds (wall and farfield) are exercised.
"""
import numpy as np
import pytest

from flite3d_amd.device import Solver

from cases import euler, levels

pytestmark = pytest.mark.gpu

H = 1e-7


def make(n, cycles):
    lv = levels(n, 2, jitter=0.15)
    iv = euler(1, machnumber=0.8, alpha=1.25, cflnumber=2.0)
    s = Solver(iv.to_struct(), list(lv))
    s.init_freestream()
    for c in range(1, cycles + 1):
        s.cycle(c)
    return s, lv


def residual(s, u):
    s.set(1, "U", u)
    s.apply(1, "PRESSURE")
    s.apply(1, "RHS")
    return s.get(1, "RHS")[:5]


def directions(rng, u, k):
    """Unit-norm random perturbations, scaled per component to the state."""
    out = []
    for _ in range(k):
        d = rng.standard_normal((5, u.shape[1])) * np.abs(u[:5]).max(axis=1, keepdims=True)
        out.append(d / np.linalg.norm(d))
    return out


@pytest.mark.parametrize("n,cycles", [(8, 0), (12, 5)])
def test_tangent_matches_central_differences(n, cycles):
    print(f"[case] tangent FD n={n} cycles={cycles}", flush=True)
    s, lv = make(n, cycles)
    u = s.get(1, "U")
    rng = np.random.default_rng(1)
    for du in directions(rng, u, 4):
        jt = s.tangent(du)
        up, um = u.copy(), u.copy()
        up[:5] += H * du
        um[:5] -= H * du
        fd = (residual(s, up) - residual(s, um)) / (2 * H)
        s.set(1, "U", u)
        err = np.abs(fd - jt).max() / np.abs(jt).max()
        assert err < 1e-5, err
    s.close()


@pytest.mark.parametrize("n,cycles", [(8, 0), (12, 5)])
def test_adjoint_dot_product(n, cycles):
    print(f"[case] adjoint dot product n={n} cycles={cycles}", flush=True)
    s, lv = make(n, cycles)
    u = s.get(1, "U")
    rng = np.random.default_rng(2)
    for du, lam in zip(directions(rng, u, 4), directions(rng, u, 4)):
        a = np.sum(lam * s.tangent(du))
        b = np.sum(s.adjoint(lam) * du)
        assert abs(a - b) <= 1e-12 * max(abs(a), abs(b)) + 1e-300, (a, b)
    s.close()


def test_adjoint_columns_are_jacobian_rows():
    """J^T e_k for unit vectors e_k equals row k of J built column by
    column from the tangent (small mesh, a few nodes)."""
    print("[case] adjoint vs explicit Jacobian rows", flush=True)
    s, lv = make(8, 2)
    u = s.get(1, "U")
    n = u.shape[1]
    rng = np.random.default_rng(3)
    for node in rng.choice(n, 3, replace=False):
        for comp in (0, 4):
            e = np.zeros((5, n))
            e[comp, node] = 1.0
            row = s.adjoint(e)  # J^T e = row (node, comp) of J
            for j in rng.choice(n, 4, replace=False):
                for c in range(5):
                    d = np.zeros((5, n))
                    d[c, j] = 1.0
                    assert abs(s.tangent(d)[comp, node] - row[c, j]) <= 1e-12 * (1 + abs(row[c, j]))
    s.close()


def test_drag_gradient_matches_central_differences():
    print("[case] drag gradient FD", flush=True)
    s, lv = make(12, 5)
    u = s.get(1, "U")
    drag0, g = s.drag_gradient()
    assert np.isfinite(drag0) and np.abs(g).max() > 0
    rng = np.random.default_rng(4)
    for du in directions(rng, u, 4):
        up, um = u.copy(), u.copy()
        up[:5] += H * du
        um[:5] -= H * du
        s.set(1, "U", up)
        dp, _ = s.drag_gradient()
        s.set(1, "U", um)
        dm, _ = s.drag_gradient()
        fd = (dp - dm) / (2 * H)
        an = np.sum(g * du)
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-12), (fd, an)
    s.set(1, "U", u)
    s.close()


def test_linearisation_rejects_unsupported_settings():
    from flite3d_amd.device import DeviceError
    from cases import rans
    lv = levels(8, 2, first_cell=0.005)
    s = Solver(rans().to_struct(), list(lv))
    s.init_freestream()
    with pytest.raises(DeviceError):
        s.tangent(np.zeros((5, lv[0].n_nodes)))
    s.close()
