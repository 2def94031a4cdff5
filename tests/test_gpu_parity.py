"""GPU path (libflite3d_amd.so through its C-ABI) against the CPU oracle.

Bar: per component, max_i |gpu - oracle| <= tol * max(max_i |oracle|, floor)
with tol = 1e-11 for single operations on the FP64 path (the reference's
per-node residual tolerance), and tol = 1e-9 for states after whole V-cycles,
where rounding differences of the summation order are amplified by the
nonlinear limiters over several hundred kernel launches.
"""
import numpy as np
import pytest

from oracle import Oracle
from flite3d_amd.device import Solver

from cases import euler, laminar, levels, rans

pytestmark = pytest.mark.gpu


def rel_err(a, b, floor=1e-30):
    a = np.atleast_2d(a)
    b = np.atleast_2d(b)
    out = []
    for c in range(b.shape[0]):
        sc = max(np.abs(b[c]).max(), floor)
        out.append(np.abs(a[c] - b[c]).max() / sc)
    return np.array(out)


def node_check(a, b, mag, what, tol=1e-11):
    """Per node and component: |a - b| <= tol max(|b|, the node's sum of
    |terms| from the oracle), so that no small-residual node hides behind a
    component's global maximum."""
    den = np.maximum(np.maximum(np.abs(b), mag), 1e-300)
    e = (np.abs(a - b) / den).max(axis=1)
    assert e.max() < tol, (what, e)


def pair(iv, lv, fp32=False):
    P = iv.to_struct(coarse_fp32=fp32)
    return Solver(P, list(lv)), Oracle(iv.to_struct(), list(lv))


def perturb(s, o, level, rng, amp=0.02):
    """Same smooth-plus-noise state on both sides."""
    u = o.get(level, "U")
    n = u.shape[1]
    u = u * (1.0 + amp * rng.uniform(-1, 1, size=u.shape))
    s.set(level, "U", u)
    o.set(level, "U", u)


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("level", [1, 2])
def test_rhs_euler(order, level):
    print(f"[case] rhs euler order={order} level={level}", flush=True)
    lv = levels(8, 3)
    s, o = pair(euler(order), lv)
    s.init_freestream(); o.init_freestream()
    perturb(s, o, level, np.random.default_rng(10 + level))
    for x in (s, o):
        x.apply(level, "PRESSURE")
        x.apply(level, "RHS")
    assert np.abs(s.get(level, "PRESSURE") - o.get(level, "PRESSURE")).max() < 1e-13
    e = rel_err(s.get(level, "RHS")[:5], o.get(level, "RHS")[:5])
    assert e.max() < 1e-11, e
    node_check(s.get(level, "RHS")[:5], o.get(level, "RHS")[:5], o.magnitudes(level, "RHS")[:5], "rhs")
    if order == 2 and level == 1:
        eg = rel_err(s.get(1, "GRADIENT"), o.get(1, "GRADIENT"))
        assert eg.max() < 1e-11, eg


def test_timesteps_and_viscous():
    print("[case] timesteps/viscous laminar", flush=True)
    lv = levels(8, 3, first_cell=0.01)
    s, o = pair(laminar(), lv)
    s.init_freestream(); o.init_freestream()
    rng = np.random.default_rng(3)
    perturb(s, o, 1, rng)
    for x in (s, o):
        x.apply(1, "PRESSURE")
        x.apply(1, "TIMESTEPS")
        x.apply(1, "DISSIPATION")
        x.apply(1, "VISCOUS")
    assert rel_err(s.get(1, "DT"), o.get(1, "DT")).max() < 1e-11
    for f in ("DISSIPATION", "LAMINAR_VISCOSITY", "VORTICITY", "DIVERGENCE"):
        e = rel_err(s.get(1, f), o.get(1, f))
        assert e.max() < 1e-11, (f, e)
    e = rel_err(s.get(1, "WALL_STRESS"), o.get(1, "WALL_STRESS"))
    assert e.max() < 1e-11, e


def test_sa_terms():
    print("[case] SA terms", flush=True)
    lv = levels(8, 3, first_cell=0.005)
    s, o = pair(rans(), lv)
    s.init_freestream(); o.init_freestream()
    rng = np.random.default_rng(4)
    perturb(s, o, 1, rng)
    u = o.get(1, "U")
    u[5] = 0.1 + 5.0 * rng.uniform(0, 1, size=u.shape[1])
    s.set(1, "U", u); o.set(1, "U", u)
    for x in (s, o):
        x.apply(1, "PRESSURE")
        x.apply(1, "TIMESTEPS")
        x.apply(1, "VISCOUS")
        x.apply(1, "SA_TIMESTEPS")
        x.apply(1, "SA_DIFFUSION")
        x.apply(1, "RHS")
    for f in ("DT", "DISSIPATION", "DT_TURB", "TURB_DIFFUSION", "RHS"):
        e = rel_err(s.get(1, f), o.get(1, f))
        assert e.max() < 1e-11, (f, e)
    # per node: the residual, the viscous term (components 2-5 of the
    # dissipation) and the SA diffusion (its magnitudes sit in row 6)
    node_check(s.get(1, "RHS")[:6], o.get(1, "RHS")[:6], o.magnitudes(1, "RHS"), "rhs")
    node_check(s.get(1, "DISSIPATION")[1:5], o.get(1, "DISSIPATION")[1:5], o.magnitudes(1, "VISCOUS")[1:5], "viscous")
    node_check(s.get(1, "TURB_DIFFUSION"), o.get(1, "TURB_DIFFUSION"), o.magnitudes(1, "SA_DIFFUSION")[5:6],
               "sa diffusion")


@pytest.mark.parametrize("case", ["euler1", "euler2", "rans"])
def test_time_iterations(case):
    print(f"[case] time iterations {case}", flush=True)
    if case == "rans":
        lv, iv = levels(8, 3, first_cell=0.005), rans()
    else:
        lv, iv = levels(8, 3), euler(int(case[-1]))
    s, o = pair(iv, lv)
    s.init_freestream(); o.init_freestream()
    for x in (s, o):
        x.do_time_iterations(1, 2, True)
    e = rel_err(s.get(1, "U")[: (6 if case == "rans" else 5)], o.get(1, "U")[: (6 if case == "rans" else 5)])
    assert e.max() < 1e-11, e


@pytest.mark.parametrize("case", ["euler1", "rans"])
def test_restrict_prolong(case):
    print(f"[case] restrict/prolong {case}", flush=True)
    if case == "rans":
        lv, iv = levels(8, 3, first_cell=0.005), rans()
    else:
        lv, iv = levels(8, 3), euler(1)
    s, o = pair(iv, lv)
    s.init_freestream(); o.init_freestream()
    nc = 6 if case == "rans" else 5
    for x in (s, o):
        x.do_time_iterations(1, 1, True)
        x.restrict(1)
    for f in ("U", "SOURCE", "RHS"):
        e = rel_err(s.get(2, f)[:nc], o.get(2, f)[:nc], floor=1e-12)
        assert e.max() < 1e-10, (f, e)
    for x in (s, o):
        x.do_time_iterations(2, 1, False)
        x.prolongate(2)
    e = rel_err(s.get(1, "U")[:nc], o.get(1, "U")[:nc])
    assert e.max() < 1e-10, e


@pytest.mark.parametrize("case,scheme", [("euler1", 1), ("euler2", 3), ("rans", 1), ("rans", 3)])
def test_cycles_and_monitor(case, scheme):
    print(f"[case] cycles {case} scheme {scheme}", flush=True)
    if case == "rans":
        lv, iv = levels(8, 3, first_cell=0.005), rans(multigridscheme=scheme)
    else:
        lv, iv = levels(8, 3), euler(int(case[-1]), multigridscheme=scheme)
    s, o = pair(iv, lv)
    s.init_freestream(); o.init_freestream()
    for c in range(1, 6):
        s.cycle(c); o.cycle(c)
        ms, mo = s.monitor(), o.monitor()
        assert np.allclose(ms["res"], mo["res"], rtol=1e-8, atol=1e-16), (c, ms["res"], mo["res"])
        assert np.allclose(ms["forces"], mo["forces"], rtol=1e-8, atol=1e-14), (c, ms["forces"], mo["forces"])
    nc = 6 if case == "rans" else 5
    e = rel_err(s.get(1, "U")[:nc], o.get(1, "U")[:nc])
    assert e.max() < 1e-9, e


@pytest.mark.parametrize("lanes", ["1", "4", "16"])
def test_gather_lanes(lanes, monkeypatch):
    """Every lanes-per-node variant of the gather kernels (forced on all
    levels) against the oracle: the single operations to the per-operation
    bar, then three growing-V cycles with restriction on every level."""
    print(f"[case] gather lanes {lanes}", flush=True)
    monkeypatch.setenv("FLITE_GATHER_LANES", lanes)
    lv = levels(8, 3, first_cell=0.005)
    s, o = pair(rans(multigridscheme=3), lv)
    s.init_freestream(); o.init_freestream()
    rng = np.random.default_rng(7)
    perturb(s, o, 1, rng)
    u = o.get(1, "U")
    u[5] = 0.1 + 5.0 * rng.uniform(0, 1, size=u.shape[1])
    s.set(1, "U", u); o.set(1, "U", u)
    for x in (s, o):
        x.apply(1, "PRESSURE")
        x.apply(1, "TIMESTEPS")
        x.apply(1, "VISCOUS")
        x.apply(1, "SA_TIMESTEPS")
        x.apply(1, "SA_DIFFUSION")
        x.apply(1, "RHS")
    for f in ("DT", "DISSIPATION", "DT_TURB", "TURB_DIFFUSION", "RHS", "GRADIENT"):
        e = rel_err(s.get(1, f), o.get(1, f))
        assert e.max() < 1e-11, (f, e)
    s.init_freestream(); o.init_freestream()
    for c in (1, 2, 3):
        s.cycle(c); o.cycle(c)
    ms, mo = s.monitor(), o.monitor()
    assert np.allclose(ms["res"], mo["res"], rtol=1e-8, atol=1e-16), (ms["res"], mo["res"])
    e = rel_err(s.get(1, "U")[:6], o.get(1, "U")[:6])
    assert e.max() < 1e-9, e


@pytest.mark.parametrize("env", [
    {"FLITE_GATHER_LANES": "1"},                                   # LDS-staged node blocks on every level
    {"FLITE_GATHER_LANES": "1", "FLITE_GATHER_KERNEL": "node"},    # per-node gathers only
    {"FLITE_FLUX_KERNEL": "side"},                                 # per-side flux records and gather
    {"FLITE_FLUX_KERNEL": "side", "FLITE_GATHER_LANES": "4"},
    {"FLITE_SIDE_BLOCK": "512"},                                   # 512-side blocks of at most 224 nodes
])
def test_kernel_variants(env, monkeypatch):
    """Every kernel path the library selects (side blocks with per-node
    partial sums or per-side flux records; LDS-staged or per-node gathers)
    against the oracle over three growing-V RANS cycles."""
    print(f"[case] kernel variant {env}", flush=True)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    lv = levels(8, 3, first_cell=0.005)
    s, o = pair(rans(multigridscheme=3), lv)
    s.init_freestream(); o.init_freestream()
    for c in (1, 2, 3):
        s.cycle(c); o.cycle(c)
    ms, mo = s.monitor(), o.monitor()
    assert np.allclose(ms["res"], mo["res"], rtol=1e-8, atol=1e-16), (ms["res"], mo["res"])
    e = rel_err(s.get(1, "U")[:6], o.get(1, "U")[:6])
    assert e.max() < 1e-9, e


def test_autotune():
    """flite_autotune picks 1, 4 or 16 lanes and a side-block shape per level
    by timing; the cycles that follow, on the chosen launch shapes, match the
    oracle."""
    print("[case] autotune lanes, rans 8^3", flush=True)
    lv = levels(8, 3, first_cell=0.005)
    s, o = pair(rans(multigridscheme=3), lv)
    s.init_freestream(); o.init_freestream()
    lanes = s.autotune(2)
    assert len(lanes) == 3 and all(x in (1, 4, 16) for x in lanes), lanes
    assert [s.lanes(k + 1) for k in range(3)] == lanes
    assert all(s.side_block(k + 1) in ((256, 128), (512, 224)) for k in range(3))
    for c in (1, 2, 3):
        s.cycle(c); o.cycle(c)
    ms, mo = s.monitor(), o.monitor()
    assert np.allclose(ms["res"], mo["res"], rtol=1e-8, atol=1e-16), (ms["res"], mo["res"])
    e = rel_err(s.get(1, "U")[:6], o.get(1, "U")[:6])
    assert e.max() < 1e-9, e


# ---- fluxType 1-3 with the JST dissipation (makeArtificialDissipation2) ----
JST = {
    "central": dict(fluxtype=1),
    "llf": dict(fluxtype=2),
    "roe": dict(fluxtype=3),
    "roe_matrix": dict(fluxtype=3, usematrixdissipation=True, hartenscorrectionfactor=0.1),
}


def jst_case(name, **kw):
    base = dict(secondorderdissipationfactor=0.5, fourthorderdissipationfactor=1 / 64.0, cflnumber=1.5)
    base.update(JST[name])
    base.update(kw)
    return euler(1, **base)


@pytest.mark.parametrize("name", list(JST))
@pytest.mark.parametrize("level", [1, 2])
def test_jst_dissipation_and_rhs(name, level):
    print(f"[case] jst dissipation/rhs {name} level={level}", flush=True)
    lv = levels(8, 3)
    s, o = pair(jst_case(name), lv)
    s.init_freestream(); o.init_freestream()
    perturb(s, o, level, np.random.default_rng(20 + level), amp=0.05)
    for x in (s, o):
        x.apply(level, "PRESSURE_ABS")
        x.apply(level, "TIMESTEPS")
        x.apply(level, "DISSIPATION")
        x.apply(level, "RHS")
    assert rel_err(s.get(level, "DT"), o.get(level, "DT")).max() < 1e-11
    d = o.get(level, "DISSIPATION")
    assert np.abs(d).max() > 1e-8  # the JST terms are really there
    e = rel_err(s.get(level, "DISSIPATION"), d)
    assert e.max() < 1e-11, e
    e = rel_err(s.get(level, "RHS")[:5], o.get(level, "RHS")[:5])
    assert e.max() < 1e-11, e


@pytest.mark.parametrize("name", list(JST))
def test_jst_time_iterations(name):
    print(f"[case] jst time iterations {name}", flush=True)
    lv = levels(8, 3)
    s, o = pair(jst_case(name), lv)
    s.init_freestream(); o.init_freestream()
    for x in (s, o):
        x.do_time_iterations(1, 2, True)
        x.do_time_iterations(1, 1, False)
    e = rel_err(s.get(1, "U")[:5], o.get(1, "U")[:5])
    assert e.max() < 1e-11, e


@pytest.mark.parametrize("name,scheme,visc", [("central", 1, False), ("llf", 3, False), ("roe", 1, False),
                                               ("roe_matrix", 3, False), ("roe", 1, True)])
def test_jst_cycles(name, scheme, visc):
    """Whole V-cycles: the JST runs on every level (coarse factor below
    grid 1), in the restriction's coarse and fine parts too."""
    print(f"[case] jst cycles {name} scheme {scheme} viscous={visc}", flush=True)
    if visc:
        lv = levels(8, 3, first_cell=0.005)
        iv = rans(multigridscheme=scheme, secondorderdissipationfactor=0.5,
                  fourthorderdissipationfactor=1 / 64.0, **JST[name])
    else:
        lv = levels(8, 3)
        iv = jst_case(name, multigridscheme=scheme)
    s, o = pair(iv, lv)
    s.init_freestream(); o.init_freestream()
    for c in range(1, 6):
        s.cycle(c); o.cycle(c)
        ms, mo = s.monitor(), o.monitor()
        assert np.allclose(ms["res"], mo["res"], rtol=1e-8, atol=1e-16), (c, ms["res"], mo["res"])
        assert np.allclose(ms["forces"], mo["forces"], rtol=1e-8, atol=1e-14), (c, ms["forces"], mo["forces"])
    nc = 6 if visc else 5
    e = rel_err(s.get(1, "U")[:nc], o.get(1, "U")[:nc])
    assert e.max() < 1e-9, e


def test_five_stage_runge_kutta_cycles():
    """Five stages (Jameson's factors) through the same stage kernel, with
    the Roe flux and JST: GPU against the oracle over V-cycles."""
    print("[case] roe + jst, 5-stage RK cycles", flush=True)
    lv = levels(8, 3)
    iv = jst_case("roe", multigridscheme=1, cflnumber=2.0)
    rk = [0.25, 1 / 6.0, 0.375, 0.5, 1.0]
    s = Solver(iv.to_struct(rk=rk), list(lv))
    o = Oracle(iv.to_struct(rk=rk), list(lv))
    s.init_freestream(); o.init_freestream()
    for c in range(1, 6):
        s.cycle(c); o.cycle(c)
        ms, mo = s.monitor(), o.monitor()
        assert np.allclose(ms["res"], mo["res"], rtol=1e-8, atol=1e-16), (c, ms["res"], mo["res"])
    e = rel_err(s.get(1, "U")[:5], o.get(1, "U")[:5])
    assert e.max() < 1e-9, e
    # and the stage count matters: 3 stages give another state
    s3 = Solver(iv.to_struct(), list(lv))
    s3.init_freestream()
    for c in range(1, 6):
        s3.cycle(c)
    assert rel_err(s3.get(1, "U")[:5], o.get(1, "U")[:5]).max() > 1e-6
    s.close(); o.close(); s3.close()


# ---- dual time stepping on a fixed mesh (flowType 1) ----
@pytest.mark.parametrize("case", ["euler_bdf2", "euler_bdf1", "rans_bdf2"])
def test_unsteady_time_iterations(case):
    print(f"[case] unsteady time iterations {case}", flush=True)
    scheme = 1 if case.endswith("bdf1") else 2
    if case.startswith("rans"):
        lv = levels(8, 3, first_cell=0.005)
        iv = rans(flowtype=1, physicaltimestep=0.05, timescheme=scheme)
        nc = 6
    else:
        lv = levels(8, 3)
        iv = euler(1, flowtype=1, physicaltimestep=0.05, timescheme=scheme)
        nc = 5
    s, o = pair(iv, lv)
    s.init_freestream(); o.init_freestream()
    for step in range(2):
        for x in (s, o):
            x.do_time_iterations(1, 2, True)
            x.do_time_iterations(2, 1, True)
        for l in (1, 2):
            e = rel_err(s.get(l, "U")[:nc], o.get(l, "U")[:nc])
            assert e.max() < 1e-11, (step, l, e)
        s.advance_physical_time(); o.advance_physical_time()


@pytest.mark.parametrize("case,scheme", [("euler", 3), ("rans", 1)])
def test_unsteady_cycles(case, scheme):
    """Two physical steps of three V-cycles each: BDF2 on level 1, BDF1 on
    the coarse levels with the uold each step's preamble restricts, the time
    term and the modified local time step in the smoother and in both halves
    of the restriction."""
    print(f"[case] unsteady cycles {case} scheme {scheme}", flush=True)
    if case == "rans":
        lv = levels(8, 3, first_cell=0.005)
        iv = rans(flowtype=1, physicaltimestep=0.05, multigridscheme=scheme)
        nc = 6
    else:
        lv = levels(8, 3)
        iv = euler(2, flowtype=1, physicaltimestep=0.05, multigridscheme=scheme)
        nc = 5
    s, o = pair(iv, lv)
    s.init_freestream(); o.init_freestream()
    c = 0
    for step in range(2):
        for _ in range(3):
            c += 1
            s.cycle(c); o.cycle(c)
            ms, mo = s.monitor(), o.monitor()
            assert np.allclose(ms["res"], mo["res"], rtol=1e-8, atol=1e-16), (c, ms["res"], mo["res"])
            assert np.allclose(ms["forces"], mo["forces"], rtol=1e-8, atol=1e-14), (c, ms["forces"], mo["forces"])
        s.advance_physical_time(); o.advance_physical_time()
        for l in (2, 3):  # the next step's preamble: coarse u and uold restricted
            for f in ("U", "UOLD"):
                e = rel_err(s.get(l, f)[:nc], o.get(l, f)[:nc])
                assert e.max() < 1e-9, (step, l, f, e)
    e = rel_err(s.get(1, "U")[:nc], o.get(1, "U")[:nc])
    assert e.max() < 1e-9, e
    # the time term acts: the same cycles of a steady run end elsewhere
    st = Solver(euler(2, multigridscheme=scheme).to_struct() if case == "euler" else
                rans(multigridscheme=scheme).to_struct(), list(lv))
    st.init_freestream()
    for k in range(1, 7):
        st.cycle(k)
    assert rel_err(st.get(1, "U")[:nc], o.get(1, "U")[:nc]).max() > 1e-6
    s.close(); o.close(); st.close()


def test_unsteady_argument_errors():
    import dataclasses
    from flite3d_amd.device import DeviceError
    lv = levels(8, 2)
    P = euler(1).to_struct()
    P.flow_type = 3
    with pytest.raises(DeviceError):
        Solver(P, list(lv))
    P = euler(1, flowtype=2, physicaltimestep=0.05).to_struct()
    P.physical_time_step = 0.0
    with pytest.raises(DeviceError):
        Solver(P, list(lv))
    s = Solver(euler(1).to_struct(), list(lv))
    with pytest.raises(DeviceError):
        s.advance_physical_time()  # steady run: no physical time
    with pytest.raises(DeviceError):
        s.replace_mesh(list(lv))  # a fixed mesh cannot be replaced
    s.close()
    s = Solver(euler(1, flowtype=2, physicaltimestep=0.05).to_struct(), list(lv))
    s.init_freestream()
    with pytest.raises(DeviceError):  # the moved mesh keeps the level's counts
        s.replace_mesh([lv[1], lv[1]])
    with pytest.raises(DeviceError):  # nor may the tangent run on a moving mesh
        s.tangent(np.zeros((5, lv[0].n_nodes)))
    s.close()
