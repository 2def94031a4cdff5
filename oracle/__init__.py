"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this package, and only to check results against / as the CPU comparator.  The product package
flite3d_amd never imports it.
"""
import ctypes
import os

import numpy as np

from flite3d_amd.params import FIELDS, OPS, FliteLevelView, FliteParams, level_view

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from . import build
            build.build()
        L = ctypes.CDLL(LIB)
        P = ctypes.POINTER
        d, i, vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        L.or_create.restype = vp
        L.or_create.argtypes = [P(FliteParams), i]
        L.or_level_upload.argtypes = [vp, i, P(FliteLevelView)]
        L.or_level_geometry.argtypes = [vp, i, P(FliteLevelView)]
        L.or_destroy.argtypes = [vp]
        L.or_last_error.restype = ctypes.c_char_p
        L.or_field_get.argtypes = [vp, i, i, P(d), i]
        L.or_field_set.argtypes = [vp, i, i, P(d), i]
        L.or_apply.argtypes = [vp, i, i]
        L.or_set_threads.argtypes = [vp, i]
        L.or_magnitudes.argtypes = [vp, i, i, P(d), i]
        L.or_init_freestream.argtypes = [vp]
        L.or_init_from_state.argtypes = [vp]
        L.or_advance_physical_time.argtypes = [vp]
        L.or_do_time_iterations.argtypes = [vp, i, i, i, i, i]
        L.or_restrict.argtypes = [vp, i, i]
        L.or_prolongate.argtypes = [vp, i]
        L.or_map_u_fine_to_coarse.argtypes = [vp, i]
        L.or_cycle.argtypes = [vp, i]
        L.or_monitor.argtypes = [vp, P(d), P(d), P(i), P(d)]
        L.or_hllc_flux.argtypes = [P(FliteParams), P(d), d, P(d), d, P(d), P(d)]
        L.or_farfield_flux.argtypes = [P(FliteParams), i, P(d), d, P(d), P(d)]
        L.or_hllc_side.argtypes = [P(FliteParams), i, i, i, P(d), d, P(d), d, P(d), P(d), P(d), P(d), P(d)]
        L.or_led.restype = d
        L.or_led.argtypes = [d, d, d, i]
        L.or_simple_flux.argtypes = [P(FliteParams), i, P(d), d, P(d), d, P(d), P(d)]
        _lib = L
    return _lib


FIELD_NCOMP = {0: 6, 1: 6, 2: 6, 3: 5, 4: 1, 5: 1, 6: 1, 7: 1, 8: 1, 9: 1,
               10: 6, 11: 1, 12: 25, 13: 3, 14: 6, 15: 6}


class OracleError(RuntimeError):
    pass


class Oracle:
    """CPU restatement with the same calls as flite3d_amd.device.Solver."""

    def __init__(self, params, levels):
        self.L = lib()
        self.params = params
        self.levels = levels
        self.ctx = self.L.or_create(ctypes.byref(params), len(levels))
        if not self.ctx:
            raise OracleError(self.L.or_last_error().decode())
        for k, lv in enumerate(levels):
            v = level_view(lv, k == 0)
            self._check(self.L.or_level_upload(self.ctx, k + 1, ctypes.byref(v)))

    def _check(self, rc):
        if rc != 0:
            raise OracleError(f"oracle error {rc}: " + self.L.or_last_error().decode())

    def replace_mesh(self, levels):
        """A moving mesh's next geometry (MG:525-548), the solution kept."""
        for k, lv in enumerate(levels):
            v = level_view(lv, k == 0)
            self._check(self.L.or_level_geometry(self.ctx, k + 1, ctypes.byref(v)))
        self.levels = levels

    def close(self):
        if self.ctx:
            self.L.or_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def n(self, level):
        return self.levels[level - 1].n_nodes

    def get(self, level, field):
        f = FIELDS[field] if isinstance(field, str) else field
        n = self.n(level)
        out = np.zeros((FIELD_NCOMP[f], n))
        self._check(self.L.or_field_get(self.ctx, level, f,
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n))
        return out

    def set(self, level, field, arr):
        f = FIELDS[field] if isinstance(field, str) else field
        n = self.n(level)
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(FIELD_NCOMP[f], n)
        self._check(self.L.or_field_set(self.ctx, level, f,
                                        a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n))

    def apply(self, level, op):
        self._check(self.L.or_apply(self.ctx, level, OPS[op] if isinstance(op, str) else op))

    def set_threads(self, threads):
        """OpenMP threads of the side-flux loops (results do not depend on it)."""
        self._check(self.L.or_set_threads(self.ctx, int(threads)))

    def magnitudes(self, level, op):
        """Per node and component, sum of |term| over what the operation adds
        up (the per-node normaliser of the parity tests); runs the op."""
        out = np.zeros((6, self.n(level)))
        self._check(self.L.or_magnitudes(self.ctx, level, OPS[op] if isinstance(op, str) else op,
                                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), self.n(level)))
        return out

    def init_freestream(self):
        self._check(self.L.or_init_freestream(self.ctx))

    def init_from_state(self):
        self._check(self.L.or_init_from_state(self.ctx))

    def do_time_iterations(self, level, steps, calc_diss, calc_turb=True, iteration=1):
        self._check(self.L.or_do_time_iterations(self.ctx, level, steps, int(calc_diss),
                                                 int(calc_turb), iteration))

    def restrict(self, fine_level, calc_turb=True):
        self._check(self.L.or_restrict(self.ctx, fine_level, int(calc_turb)))

    def prolongate(self, coarse_level):
        self._check(self.L.or_prolongate(self.ctx, coarse_level))

    def map_u_fine_to_coarse(self, fine_level):
        self._check(self.L.or_map_u_fine_to_coarse(self.ctx, fine_level))

    def cycle(self, number):
        self._check(self.L.or_cycle(self.ctx, number))

    def advance_physical_time(self):
        self._check(self.L.or_advance_physical_time(self.ctx))

    def monitor(self):
        res = (ctypes.c_double * 7)()
        mx = (ctypes.c_double * 7)()
        idx = (ctypes.c_int * 7)()
        forces = (ctypes.c_double * 9)()
        rc = self.L.or_monitor(self.ctx, res, mx, idx, forces)
        if rc not in (0, -3):
            self._check(rc)
        return dict(res=np.array(res[:]), maxres=np.array(mx[:]),
                    maxidx=np.array(idx[:]), forces=np.array(forces[:]), rc=rc)


def hllc_flux(params, ul, pl, ur, pr, w):
    f = (ctypes.c_double * 5)()
    a = (ctypes.c_double * 5)(*ul)
    b = (ctypes.c_double * 5)(*ur)
    ww = (ctypes.c_double * 3)(*w)
    lib().or_hllc_flux(ctypes.byref(params), a, pl, b, pr, ww, f)
    return np.array(f[:])


def hllc_side(params, order, autoclip, nsig, ul, pl, ur, pr, w, dx, g1, g2):
    """One interior HLLC side with explicit end gradients and bounds."""
    f = (ctypes.c_double * 5)()
    arr = lambda v, k: (ctypes.c_double * k)(*[float(x) for x in v])
    rc = lib().or_hllc_side(ctypes.byref(params), order, autoclip, nsig, arr(ul, 5), pl, arr(ur, 5), pr,
                            arr(w, 3), arr(dx, 3), arr(g1, 25), arr(g2, 25), f)
    if rc != 0:
        raise OracleError(lib().or_last_error().decode())
    return np.array(f[:])


def led(a, bmin, bmax, nsig):
    return lib().or_led(a, bmin, bmax, nsig)


def simple_flux(params, flux_type, ul, pl, ur, pr, w):
    """Interior flux of fluxType 1 (central), 2 (LLF) or 3 (Roe)."""
    f = (ctypes.c_double * 5)()
    a = (ctypes.c_double * 5)(*ul)
    b = (ctypes.c_double * 5)(*ur)
    ww = (ctypes.c_double * 3)(*w)
    rc = lib().or_simple_flux(ctypes.byref(params), flux_type, a, pl, b, pr, ww, f)
    if rc != 0:
        raise OracleError(lib().or_last_error().decode())
    return np.array(f[:])


def farfield_flux(params, indicator, state, p, n):
    f = (ctypes.c_double * 5)()
    a = (ctypes.c_double * 5)(*state)
    nn = (ctypes.c_double * 3)(*n)
    rc = lib().or_farfield_flux(ctypes.byref(params), indicator, a, p, nn, f)
    if rc != 0:
        raise OracleError(lib().or_last_error().decode())
    return np.array(f[:])
