"""Boundary and scheme branches of the reference on the GPU against the
oracle: boundaryTerm 2 (finite-volume boundary closure), a supersonic
freestream (the M > 1 farfield branch and supersonic upwinding), symmetry
(indicator 2) and inviscid-wall (indicator 9) boundaries, isothermal
walls, a single grid (multigridScheme 0), and the FP32-coarse mode, whose
coarse corrections are FP32 by design and so are held to a looser bar.

Bars as in test_gpu_parity: states after whole V-cycles within 1e-9 of the
component scale, monitor sums and forces within 1e-8 relative.
"""
import numpy as np
import pytest

from oracle import Oracle
from flite3d_amd.device import Solver

from cases import euler, laminar, levels, rans

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return max(float(np.abs(a[c] - b[c]).max() / max(np.abs(b[c]).max(), 1e-30)) for c in range(b.shape[0]))


def cycles_match(iv, lv, n=3, tol=1e-9, mon_tol=1e-8, fp32=False):
    s = Solver(iv.to_struct(coarse_fp32=fp32), list(lv))
    o = Oracle(iv.to_struct(), list(lv))
    s.init_freestream()
    o.init_freestream()
    nc = 6 if iv.turbulencemodel == 1 else 5
    for c in range(1, n + 1):
        s.cycle(c)
        o.cycle(c)
        ms, mo = s.monitor(), o.monitor()
        assert np.allclose(ms["res"], mo["res"], rtol=mon_tol, atol=1e-16), (c, ms["res"], mo["res"])
        assert np.allclose(ms["forces"], mo["forces"], rtol=mon_tol, atol=1e-14), (c, ms["forces"], mo["forces"])
    e = rel_err(s.get(1, "U")[:nc], o.get(1, "U")[:nc])
    s.close()
    o.close()
    assert e < tol, e
    return e


def test_boundary_term_fv():
    print("[case] boundaryTerm 2, euler HLLC-2, growing V", flush=True)
    cycles_match(euler(2, boundaryterm=2, multigridscheme=3), levels(8, 3))


def test_supersonic_freestream():
    print("[case] supersonic freestream M 1.5", flush=True)
    cycles_match(euler(1, machnumber=1.5, alpha=3.0, cflnumber=1.0), levels(8, 3))


@pytest.mark.parametrize("wall", [2, 9])
def test_symmetry_and_inviscid_wall(wall):
    print(f"[case] wall indicator {wall}, euler HLLC-2", flush=True)
    cycles_match(euler(2, multigridscheme=3), levels(8, 3, wall=wall))


def test_inviscid_wall_isothermal_laminar():
    print("[case] wall indicator 9, laminar, isothermal", flush=True)
    cycles_match(laminar(wallsareisentropic=False, multigridscheme=3), levels(8, 3, first_cell=0.01, wall=9))


def test_isothermal_laminar():
    print("[case] laminar, isothermal walls, growing V", flush=True)
    cycles_match(laminar(wallsareisentropic=False, multigridscheme=3), levels(8, 3, first_cell=0.01))


def test_rans_symmetry_wall():
    print("[case] rans, wall indicator 2", flush=True)
    cycles_match(rans(multigridscheme=3), levels(8, 3, first_cell=0.005, wall=2))


def test_single_grid():
    print("[case] single grid, multigridScheme 0", flush=True)
    cycles_match(euler(2, multigridscheme=0, numberofgridstouse=1), levels(8, 1), n=4)


def test_fp32_coarse_mode():
    """FP32 coarse levels: the fine state follows the FP64 oracle to the
    precision of the coarse corrections (FP32 rounding of increments)."""
    print("[case] rans, FP32 coarse levels", flush=True)
    e = cycles_match(rans(multigridscheme=3), levels(8, 3, first_cell=0.005), tol=1e-5, mon_tol=1e-4,
                     fp32=True)
    assert e > 0.0  # the FP32 path really differs from the FP64 oracle
