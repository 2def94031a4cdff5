"""flite_monitor_noalloc: the monitor without any allocation, through the
C-ABI, against flite_monitor on the same contexts.

Both entry points run the same kernels in the same order, so their results
agree bit for bit; the no-allocation one keeps its block records in the
library's file-scope device array.  Two calls in a row on two contexts, the
second on a smaller mesh, check that the array's contents from the first
call do not leak into the second; the device's free memory is the same
before and after each call.
"""
import ctypes

import numpy as np
import pytest

from flite3d_amd import device
from flite3d_amd.device import DeviceError, Solver

from cases import euler, levels, rans

pytestmark = pytest.mark.gpu


def noalloc(s):
    res = (ctypes.c_double * 7)()
    mx = (ctypes.c_double * 7)()
    idx = (ctypes.c_int * 7)()
    forces = (ctypes.c_double * 9)()
    rc = device.lib().flite_monitor_noalloc(s.ctx, res, mx, idx, forces)
    return rc, dict(res=np.array(res[:]), maxres=np.array(mx[:]), maxidx=np.array(idx[:]),
                    forces=np.array(forces[:]))


def free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def check(s):
    ref = s.monitor()
    before = free_bytes()
    rc, got = noalloc(s)
    after = free_bytes()
    assert rc == 0, device.lib().flite_last_error().decode()
    assert before == after, (before, after)
    for k in ("res", "maxres", "maxidx", "forces"):
        assert np.array_equal(got[k], ref[k]), (k, got[k], ref[k])
    return got


def test_monitor_noalloc_equals_monitor_on_two_calls():
    print("[case] monitor_noalloc: RANS 8^3 three levels, then Euler 5^3 two levels", flush=True)
    big = Solver(rans().to_struct(), list(levels(8, 3, first_cell=0.005)))
    big.init_freestream()
    for c in range(1, 3):
        big.cycle(c)
    g1 = check(big)
    assert g1["res"][0] > 0 and g1["forces"][1] != 0
    small = Solver(euler(1).to_struct(), list(levels(5, 2)))
    small.init_freestream()
    small.cycle(1)
    g2 = check(small)  # the array still holds the first call's records
    assert not np.array_equal(g1["res"], g2["res"])
    big.close()
    small.close()


def test_monitor_noalloc_needs_a_built_context():
    print("[case] monitor_noalloc: unbuilt context is refused", flush=True)
    s = Solver(euler(1).to_struct(), list(levels(5, 2)))
    rc, _ = noalloc(s)  # the levels are uploaded but not built yet
    assert rc != 0
    assert "built context" in device.lib().flite_last_error().decode()
    s.init_freestream()
    check(s)
    s.close()
