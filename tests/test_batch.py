"""Batched flow conditions on one mesh (one context, n_batch conditions,
every launch covering all of them).

Each condition of a batch must come out exactly as the same condition run
alone: the batch shares the mesh data and the launches, not the arithmetic
(same kernels, same summation order), so the comparison is bit for bit.
"""
import numpy as np
import pytest

from flite3d_amd.device import DeviceError, Solver

from cases import euler, levels, rans

pytestmark = pytest.mark.gpu


def run(params, lv, cycles):
    s = Solver(params, list(lv))
    s.init_freestream()
    mons = []
    nb = len(params) if isinstance(params, list) else 1
    for c in range(1, cycles + 1):
        s.cycle(c)
        row = []
        for b in range(nb):
            s.select(b)
            row.append(s.monitor())
        mons.append(row)
    us = []
    for b in range(nb):
        s.select(b)
        us.append(s.get(1, "U"))
    s.close()
    return us, mons


@pytest.mark.parametrize("order,fp32", [(1, False), (2, True)])
def test_euler_batch_equals_single_runs(order, fp32):
    print(f"[case] euler batch order={order} fp32={fp32}", flush=True)
    lv = levels(8, 3)
    conds = [(0.5, 2.0), (0.75, -3.0), (0.35, 6.5)]
    ps = [euler(order, machnumber=m, alpha=a, multigridscheme=3).to_struct(coarse_fp32=fp32) for m, a in conds]
    ub, mb = run(ps, lv, 3)
    for b, p in enumerate(ps):
        us, ms = run(p, lv, 3)
        assert np.array_equal(ub[b], us[0]), (b, np.abs(ub[b] - us[0]).max())
        for c in range(3):
            assert np.array_equal(mb[c][b]["res"], ms[c][0]["res"])
            assert np.array_equal(mb[c][b]["forces"], ms[c][0]["forces"])
    # the conditions really differ
    assert not np.array_equal(ub[0], ub[1])


def test_rans_batch_equals_single_runs():
    print("[case] rans batch", flush=True)
    lv = levels(8, 3, first_cell=0.005)
    ps = [rans(machnumber=0.7, alpha=1.0, multigridscheme=3).to_struct(),
          rans(machnumber=0.6, alpha=3.0, multigridscheme=3).to_struct()]
    ub, mb = run(ps, lv, 2)
    for b, p in enumerate(ps):
        us, ms = run(p, lv, 2)
        assert np.array_equal(ub[b], us[0]), (b, np.abs(ub[b] - us[0]).max())
        assert np.array_equal(mb[-1][b]["forces"], ms[-1][0]["forces"])


def test_batch_rejects_differing_settings():
    lv = levels(8, 2)
    ps = [euler(1).to_struct(), euler(2).to_struct()]
    with pytest.raises(DeviceError):
        Solver(ps, list(lv))


@pytest.mark.parametrize("case", ["roe_jst", "unsteady", "moving"])
def test_jst_and_unsteady_batch_equals_single_runs(case):
    """The JST node records and the physical-time states (uold, uold2) are
    per condition too: each condition of the batch equals its single run."""
    print(f"[case] batch {case}", flush=True)
    lv = levels(8, 3)
    conds = [(0.5, 2.0), (0.7, -2.0)]
    if case == "roe_jst":
        kw = dict(fluxtype=3, usematrixdissipation=True, secondorderdissipationfactor=0.5,
                  fourthorderdissipationfactor=1 / 64.0, cflnumber=1.5)
    elif case == "moving":
        kw = dict(flowtype=2, physicaltimestep=0.05)
    else:
        kw = dict(flowtype=1, physicaltimestep=0.05)
    ps = [euler(1, machnumber=m, alpha=a, multigridscheme=3, **kw).to_struct() for m, a in conds]
    meshes = [list(lv)]
    if case == "moving":
        from test_moving_mesh import moving
        rng = np.random.default_rng(31)
        dx = np.array([0.002, 0.0, -0.001])
        meshes = [moving(lv, rng, 2 + k, dx, consistent=True) for k in range(2)]

    def go(params):
        s = Solver(params, meshes[0])
        s.init_freestream()
        nb = len(params) if isinstance(params, list) else 1
        for c in range(1, 5):
            s.cycle(c)
            if case != "roe_jst" and c % 2 == 0:
                s.advance_physical_time()
                if case == "moving" and c == 2:
                    s.replace_mesh(meshes[1])  # carries every condition's u, uold, uold2
        us = []
        for b in range(nb):
            s.select(b)
            us.append(s.get(1, "U"))
        s.close()
        return us

    ub = go(ps)
    for b, p in enumerate(ps):
        assert np.array_equal(ub[b], go(p)[0]), b
    assert not np.array_equal(ub[0], ub[1])
