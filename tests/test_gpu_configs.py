"""The benchmark configurations at their own workloads: GPU against the oracle.

Each test builds the configuration's mesh and runs its physics through the
C-ABI (libflite3d_amd.so) and through the oracle on the same inputs:

1. Euler, 20^3 jittered bump box, 3 levels, 200 V-cycles: HLLC-1 with the
   reference's 3 stages, and the Roe-1 + JST / 5-stage extension; states to
   1e-9, drag to 1e-10, monitor residuals every cycle to 1e-8.
2. RANS-SA, 128^3 graded from 1e-6 (2.1M nodes): one fine residual (HLLC-2 +
   SA convection and source), viscous term, time steps, SA time steps and SA
   diffusion, per node and component, normalised by max(|oracle|, sum over
   the node's terms of |term|), to 1e-11 or to ten times the oracle's own
   change when its sides are summed in another order, whichever is larger:
   on this grading the Green-Gauss gradients cancel to 1e-6 of their terms,
   and reordering the sums alone moves the oracle's viscous term by 3e-7 of
   the node scale (its residual by 5e-11); freestream preserved with the wall
   made farfield.
3. Discrete adjoint, Euler 64^3: 16 unit directions (seed 1), the
   tangent/adjoint dot product to 1e-12 and the tangent against central
   differences (h = 1e-7) of the oracle's residual.
4. 64 conditions (seed 2) on the jittered 64^3 box in one context, FP32
   coarse levels: four conditions equal their single runs bit for bit; one
   condition with FP64 coarse levels equals the oracle to 1e-9.
5. RANS-SA, 200^3 graded from 1e-6 (8.1M nodes), 5 levels: one fine
   residual per node as in 2, and freestream preservation.  Its
   sensitivities (HLLC-2 + SA): the tangent/adjoint dot product of the
   6-component residual (to 1e-12 of |lam| |J du|), and the coordinate gradient of drag - lam^T R at a
   fixed state against central differences with the fine level rebuilt
   from the deformed coordinates (2 degrees of sideslip, so that no
   farfield face sits on the Roe flux's |u.n| kink).
"""
import dataclasses
import os

import numpy as np
import pytest

from flite3d_amd import device, mesh
from flite3d_amd.device import Solver
from flite3d_amd.params import InputVariables
from oracle import Oracle as _Oracle

from cases import euler, rans

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(1200)]


class Oracle(_Oracle):
    """The oracle with its side-flux and node loops on up to 16 host threads:
    its results do not depend on the thread count (test_oracle_fields.py),
    and the 128^3 and 200^3 cases then take seconds instead of minutes."""

    def __init__(self, params, levels):
        super().__init__(params, levels)
        self.set_threads(min(16, os.cpu_count() or 1))

RK5 = [0.25, 1.0 / 6.0, 0.375, 0.5, 1.0]


def rel_err(a, b):
    return np.array([np.abs(a[c] - b[c]).max() / max(np.abs(b[c]).max(), 1e-300) for c in range(b.shape[0])])


def node_err(a, b, mag):
    """max over nodes of |a - b| / max(|b|, sum of |terms|, 1e-300), per component."""
    den = np.maximum(np.maximum(np.abs(b), mag), 1e-300)
    return (np.abs(a - b) / den).max(axis=1)


def say(msg):
    print(f"[case] {msg}", flush=True)


# ---- configuration 1 ------------------------------------------------------------
@pytest.fixture(scope="module")
def box20():
    return mesh.bump_levels(20, 3, jitter=0.15, seed=0)


@pytest.mark.parametrize("variant", ["hllc1_rk3", "roe1_jst_rk5"])
def test_config1_euler_20_200_cycles(box20, variant):
    say(f"config 1 {variant}: 20^3 jittered, 3 levels, 200 cycles")
    lv = box20
    if variant == "roe1_jst_rk5":
        P = euler(1, fluxtype=3, machnumber=0.5, alpha=2.0, cflnumber=2.0).to_struct(rk=RK5)
    else:
        P = euler(1, machnumber=0.5, alpha=2.0, cflnumber=2.0).to_struct()
    s, o = Solver(P, list(lv)), Oracle(P, list(lv))
    s.init_freestream()
    o.init_freestream()
    for c in range(1, 201):
        s.cycle(c)
        o.cycle(c)
        ms, mo = s.monitor(), o.monitor()
        assert np.allclose(ms["res"][:5], mo["res"][:5], rtol=1e-8, atol=1e-300), (c, ms["res"], mo["res"])
        if c == 1:
            r1 = mo["res"][0]
    e = rel_err(s.get(1, "U")[:5], o.get(1, "U")[:5])
    assert e.max() <= 1e-9, e
    dg, do = ms["forces"][1], mo["forces"][1]
    assert abs(dg - do) <= 1e-10 * abs(do), (dg, do)
    assert mo["res"][0] < r1  # the residual falls (the V-cycle of scheme 1 stalls later, as in the reference)
    s.close()
    o.close()


# ---- configuration 2 ------------------------------------------------------------
def rans_state(L, rng_seed):
    """Freestream times a smooth perturbation, nu~ smooth and positive, on
    every node (the residual operations see no boundary conditions)."""
    x, y, z = L.coords
    P = rans().to_struct()
    inf = np.array(P.inflow_field[0])
    u = np.zeros((6, L.n_nodes))
    u[0] = inf[0] * (1 + 0.02 * np.sin(3 * x + 2 * y) * np.cos(2 * z))
    vel = inf[1:4, None] * (1 + 0.05 * np.cos(2 * x - y + z))[None, :]
    vel[2] += 0.02 * np.sin(4 * z)
    u[1:4] = u[0] * vel
    p = (1.0 / (1.4 * P.mach_number ** 2)) * (1 + 0.02 * np.cos(x + 3 * y - z))
    u[4] = p / 0.4 + 0.5 * u[0] * (vel ** 2).sum(axis=0)
    u[5] = 2.0 + 1.5 * np.sin(5 * x * y + z)
    return u


def fine_operations(lv_fine, iv, what, gpu=Solver):
    """GPU and oracle on one fine level at the same state, in a smoothing
    step's order: PRESSURE, TIMESTEPS, VISCOUS, SA_TIMESTEPS, SA_DIFFUSION,
    RHS; returns per quantity (gpu, oracle, magnitudes)."""
    P = iv.to_struct()
    s, o = gpu(P, [lv_fine]), Oracle(P, [lv_fine])
    u = rans_state(lv_fine, 0)
    z = np.zeros((1, lv_fine.n_nodes))
    out = {}
    say(f"{what}: time steps")
    for x in (s, o):
        x.init_freestream()  # the laminar viscosity the time steps read: 1/Re
        x.set(1, "U", u)
        x.apply(1, "PRESSURE")
        x.apply(1, "TIMESTEPS")
        x.apply(1, "DISSIPATION")
    out["dt"] = (s.get(1, "DT"), o.get(1, "DT"), z)
    say(f"{what}: viscous term")
    s.apply(1, "VISCOUS")
    mag = o.magnitudes(1, "VISCOUS")
    out["viscous"] = (s.get(1, "DISSIPATION")[1:5], o.get(1, "DISSIPATION")[1:5], mag[1:5])
    for f in ("LAMINAR_VISCOSITY", "VORTICITY", "DIVERGENCE"):
        out[f] = (s.get(1, f), o.get(1, f), z)
    say(f"{what}: SA time steps, SA diffusion")
    for x in (s, o):
        x.apply(1, "SA_TIMESTEPS")
    out["dt_turb"] = (s.get(1, "DT_TURB"), o.get(1, "DT_TURB"), z)
    s.apply(1, "SA_DIFFUSION")
    mag = o.magnitudes(1, "SA_DIFFUSION")
    out["sa_diffusion"] = (s.get(1, "TURB_DIFFUSION"), o.get(1, "TURB_DIFFUSION"), mag[5:6])
    say(f"{what}: residual")
    s.apply(1, "RHS")
    mag = o.magnitudes(1, "RHS")
    out["rhs"] = (s.get(1, "RHS")[:6], o.get(1, "RHS")[:6], mag)
    s.close()
    o.close()
    return out


def reordered(L, seed=0):
    """The level with its sides and boundary sides in another order: the
    same mesh, every sum over sides taken in a different order."""
    rng = np.random.default_rng(seed)
    ps, pb = rng.permutation(L.n_sides), rng.permutation(L.n_bsides)
    return dataclasses.replace(L, side_index=L.side_index[:, ps].copy(), side_w=L.side_w[:, ps].copy(),
                               bside_index=L.bside_index[:, pb].copy(), bside_w=L.bside_w[:, pb].copy())


def check_fine_operations(out, sens):
    """GPU against the oracle, per node and component: within 1e-11 of the
    node scale, or within ten times what reordering the oracle's own sums
    changes (sens: the oracle on the reordered level)."""
    for k, (a, b, mag) in out.items():
        e = node_err(a, b, mag)
        s = node_err(sens[k][1], b, mag)
        tol = np.maximum(1e-11, 10.0 * s)
        print(f"config check {k}: gpu {e.max():.2e}, oracle reordered {s.max():.2e}", flush=True)
        assert np.all(e <= tol), (k, e, s)
        # the magnitudes bound the sums they normalise (triangle inequality)
        if mag.any():
            assert np.all(np.abs(b) <= mag * (1 + 1e-12) + 1e-300), k


def freestream_residual(lv_fine, iv, what):
    """Residual of the freestream with every boundary side farfield: zero up
    to rounding, relative to the node's sum of |terms| (from the oracle at
    the same state).  The oracle's own value on this grading grows with the
    mesh: 3e-14 at 32^3, 7e-14 at 64^3."""
    say(f"{what}: freestream preservation, wall sides made farfield")
    bi = lv_fine.bside_index.copy()
    bi[2][bi[2] == 1] = 3
    L = dataclasses.replace(lv_fine, bside_index=bi)
    P = iv.to_struct()
    s, o = Solver(P, [L]), Oracle(P, [L])
    inf = np.array(P.inflow_field[0])
    u = np.zeros((6, L.n_nodes))
    u[:5] = inf[:5, None]
    for x in (s, o):
        x.set(1, "U", u)
        x.apply(1, "PRESSURE")
    s.apply(1, "RHS")
    mag = o.magnitudes(1, "RHS")
    r = s.get(1, "RHS")[:5]
    s.close()
    o.close()
    return (np.abs(r) / np.maximum(mag[:5], 1e-300)).max()


@pytest.fixture(scope="module")
def box128():
    say("config 2 mesh: 128^3 graded from 1e-6, 4 levels")
    return mesh.bump_levels(128, 4, first_cell=1e-6, ratio=1.15, progress=say)


def test_config2_rans_128_fine_operations(box128):
    lv = box128
    assert lv[0].n_nodes == 2146689 and lv[0].n_sides == 14827904
    iv = rans(reynoldsnumber=6.0e6, machnumber=0.7, alpha=1.0, multigridscheme=3)
    out = fine_operations(lv[0], iv, "config 2")
    say("config 2: the oracle on the reordered level")
    sens = fine_operations(reordered(lv[0]), iv, "config 2 reordered", gpu=Oracle)
    check_fine_operations(out, sens)
    assert freestream_residual(lv[0], iv, "config 2") <= 1e-12


def test_config2_rans_128_multigrid_cycle(box128):
    """One growing-V cycle over the four levels from the freestream: the
    restriction, coarse smoothing and prolongation at this size, states to
    1e-9 of the oracle."""
    say("config 2: one growing-V cycle, 4 levels")
    lv = box128
    P = rans(reynoldsnumber=6.0e6, machnumber=0.7, alpha=1.0, multigridscheme=3).to_struct()
    s, o = Solver(P, list(lv)), Oracle(P, list(lv))
    s.init_freestream()
    o.init_freestream()
    s.cycle(1)
    o.cycle(1)
    e = rel_err(s.get(1, "U")[:6], o.get(1, "U")[:6])
    assert e.max() <= 1e-9, e
    ms, mo = s.monitor(), o.monitor()
    assert np.allclose(ms["res"][:6], mo["res"][:6], rtol=1e-8), (ms["res"], mo["res"])
    s.close()
    o.close()


# ---- configuration 3 ------------------------------------------------------------
def test_config3_adjoint_64():
    say("config 3: 64^3 Euler M=0.8 alpha=1.25, 300 cycles, 16 directions")
    raw = mesh.bump_box(64)
    lv, _ = mesh.build_levels(*raw, 3)
    P = euler(1, machnumber=0.8, alpha=1.25, multigridscheme=3).to_struct()
    s = Solver(P, lv)
    s.init_freestream()
    for c in range(1, 301):
        s.cycle(c)
    u = s.get(1, "U")
    n = u.shape[1]
    o = Oracle(P, [lv[0]])

    def resid(x):
        o.set(1, "U", x)
        o.apply(1, "PRESSURE")
        o.apply(1, "RHS")
        return o.get(1, "RHS")[:5]

    rng = np.random.default_rng(1)
    h = 1e-7
    fd_err, dot_err = [], []
    for k in range(16):
        du = rng.standard_normal((5, n))
        du /= np.linalg.norm(du)
        lam = rng.standard_normal((5, n))
        lam /= np.linalg.norm(lam)
        s.set(1, "U", u)
        jt = s.tangent(du)
        ja = s.adjoint(lam)
        a, b = float(np.sum(lam * jt)), float(np.sum(ja * du))
        dot_err.append(abs(a - b) / max(abs(a), abs(b)))
        up, um = u.copy(), u.copy()
        up[:5] += h * du
        um[:5] -= h * du
        fd = (resid(up) - resid(um)) / (2 * h)
        fd_err.append(float(np.linalg.norm(fd - jt) / np.linalg.norm(jt)))
    print(f"config 3: dot {max(dot_err):.2e}, tangent vs oracle FD {max(fd_err):.2e}", flush=True)
    assert max(dot_err) <= 1e-12, dot_err
    assert max(fd_err) <= 1e-5, fd_err
    s.close()
    o.close()


# ---- configuration 4 ------------------------------------------------------------
def test_config4_batch_64_conditions():
    say("config 4: 64 conditions on the jittered 64^3 box, FP32 coarse")
    lv = mesh.bump_levels(64, 4, jitter=0.15, seed=0)
    rng = np.random.default_rng(2)
    mach = rng.uniform(0.3, 0.8, 64)
    alpha = rng.uniform(-4.0, 8.0, 64)

    def iv(m, a):
        return InputVariables(reynoldsnumber=0.0, machnumber=float(m), alpha=float(a), fluxtype=4,
                              hllcfluxorder=2, automaticlipping=1, cflnumber=2.0, multigridscheme=3,
                              restartnumber=0)

    ps = [iv(m, a).to_struct(coarse_fp32=True) for m, a in zip(mach, alpha)]
    cycles = 3
    s = Solver(ps, lv)
    s.init_freestream()
    for c in range(1, cycles + 1):
        s.cycle(c)
    pick = (0, 21, 42, 63)
    batch = {}
    for b in pick:
        s.select(b)
        batch[b] = s.get(1, "U")
    s.close()
    for b in pick:
        say(f"config 4: condition {b} alone")
        t = Solver([ps[b]], lv)
        t.init_freestream()
        for c in range(1, cycles + 1):
            t.cycle(c)
        assert np.array_equal(t.get(1, "U"), batch[b]), b
        t.close()
    say("config 4: condition 0 with FP64 coarse levels against the oracle")
    P = iv(mach[0], alpha[0]).to_struct()
    t, o = Solver(P, lv), Oracle(P, lv)
    t.init_freestream()
    o.init_freestream()
    for c in range(1, cycles + 1):
        t.cycle(c)
        o.cycle(c)
    e = rel_err(t.get(1, "U")[:5], o.get(1, "U")[:5])
    assert e.max() <= 1e-9, e
    t.close()
    o.close()


# ---- configuration 5 ------------------------------------------------------------
@pytest.fixture(scope="module")
def box200():
    say("config 5 mesh: 200^3 graded from 1e-6, 5 levels")
    raw = mesh.bump_box(200, first_cell=1e-6, ratio=1.15)
    lv, _ = mesh.build_levels(*raw, 5, progress=say)
    return raw, lv


def test_config5_rans_200_fine_residual(box200):
    _, lv = box200
    assert lv[0].n_nodes == 8120601 and lv[0].n_sides == 56360600
    iv = rans(reynoldsnumber=1.0e7, machnumber=0.75, alpha=2.0, multigridscheme=3)
    P = iv.to_struct()
    say("config 5: fine residual with all five levels uploaded (nested ordering)")
    s, o = Solver(P, list(lv)), Oracle(P, [lv[0]])
    u = rans_state(lv[0], 0)
    for x in (s, o):
        x.init_freestream()
        x.set(1, "U", u)
        x.apply(1, "PRESSURE")
        x.apply(1, "DISSIPATION")
        x.apply(1, "VISCOUS")
    s.apply(1, "RHS")
    mag = o.magnitudes(1, "RHS")
    out = {"rhs": (s.get(1, "RHS")[:6], o.get(1, "RHS")[:6], mag)}
    s.close()
    o.close()
    say("config 5: the oracle on the reordered level")
    r = Oracle(P, [reordered(lv[0])])
    r.init_freestream()
    r.set(1, "U", u)
    r.apply(1, "PRESSURE")
    r.apply(1, "DISSIPATION")
    r.apply(1, "VISCOUS")
    r.apply(1, "RHS")
    sens = {"rhs": (None, r.get(1, "RHS")[:6], None)}
    r.close()
    check_fine_operations(out, sens)
    assert freestream_residual(lv[0], iv, "config 5") <= 1e-12


def test_config5_rans_200_sensitivities(box200):
    """d(drag)/dX's two halves at 200^3: the state part through the adjoint
    (its dot product with the tangent), and the mesh part, the coordinate
    gradient of L(X) = drag - lam^T R(U, X) at a fixed state U and vector
    lam, through every mesh quantity the preprocessor rebuilds (weights,
    volumes, coordinates, wall normals and areas, wall distances), against a
    central difference of the oracle's residual and drag along a smooth
    deformation of the bump: dz = sin(pi x) sin(pi y) (1 - z), which moves
    the wall and its graded cells together."""
    raw, lv = box200
    xyz, tets, btri, bind = raw
    L = lv[0]
    n = L.n_nodes
    iv = rans(reynoldsnumber=1.0e7, machnumber=0.75, alpha=2.0, beta=2.0, multigridscheme=3)
    P = iv.to_struct()
    u = rans_state(L, 0)
    u[5] = np.where(L.wall_dist > 0, u[5], 0.0)  # the wall condition's nu~ = 0
    rng = np.random.default_rng(5)
    lam = rng.standard_normal((6, n))
    lam /= np.linalg.norm(lam)
    say("config 5 sensitivities: tangent/adjoint dot product (6 components, HLLC-2 + SA)")
    s = Solver(P, [L])
    s.set(1, "U", u)
    du = rng.standard_normal((6, n)) * np.abs(u).max(axis=1, keepdims=True)
    du /= np.linalg.norm(du)
    jt, ja = s.tangent(du), s.adjoint(lam)
    a, b = float(np.sum(lam * jt)), float(np.sum(ja * du))
    # random vectors of 48M entries: their dot product cancels to ~1e-4 of
    # the Cauchy-Schwarz bound |lam| |J du|, which is the scale of the sums'
    # rounding
    scale = max(np.linalg.norm(lam) * np.linalg.norm(jt), np.linalg.norm(ja) * np.linalg.norm(du))
    print(f"config 5 dot product: {a:.15e} {b:.15e}; |a - b| / |a| {abs(a - b) / abs(a):.2e}, "
          f"/ (|lam| |J du|) {abs(a - b) / scale:.2e}", flush=True)
    assert abs(a - b) <= 1e-12 * scale, (a, b, scale)
    say("config 5 sensitivities: mesh shares and their chain to the coordinates")
    gw, gwb, gv, gxd, gd = s.mesh_adjoint(lam)
    drag, jw, jwb, jv = s.drag_mesh_gradient()
    s.close()
    gx = mesh.mesh_vjp(xyz, tets, btri, bind, jw - gw, jwb - gwb, gvol=jv - gv,
                       gxn=-gxd + mesh.wall_distance_vjp(L, -gd))
    x, y, z = xyz.T
    d = np.zeros_like(xyz)
    d[:, 2] = np.sin(np.pi * x) * np.sin(np.pi * y) * (1.0 - z)
    an = float(np.sum(gx * d))
    h = 1e-5
    vals = []
    for sgn in (1.0, -1.0):
        say(f"config 5 sensitivities: fine level rebuilt at X {'+' if sgn > 0 else '-'} h dX")
        lvp, _ = mesh.build_levels(xyz + sgn * h * d, tets, btri, bind, 1)
        o = Oracle(P, [lvp[0]])  # the difference is of the oracle's residual and drag
        o.init_freestream()
        o.set(1, "U", u)
        o.apply(1, "PRESSURE")
        o.set(1, "DISSIPATION", np.zeros((5, n)))  # makeViscosityTerm subtracts into it
        o.apply(1, "VISCOUS")
        o.apply(1, "SA_DIFFUSION")
        o.apply(1, "RHS")
        r = o.get(1, "RHS")[:6].copy()
        r[:5] += o.get(1, "DISSIPATION")[:5]
        r[5] += o.get(1, "TURB_DIFFUSION")[0]
        dr = float(o.monitor()["forces"][1])
        o.close()
        vals.append(dr - float(np.sum(lam * r)))
        del lvp
    fd = (vals[0] - vals[1]) / (2 * h)
    print(f"config 5 coordinate gradient: fd {fd:.12e} analytic {an:.12e} rel {abs(fd - an) / abs(an):.2e} "
          f"(drag {drag:.9e})", flush=True)
    assert abs(fd - an) <= 1e-5 * abs(an), (fd, an)
