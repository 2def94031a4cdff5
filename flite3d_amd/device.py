"""Python host binding of the HIP library (libflite3d_amd.so) via its C-ABI.

Mirrors the calls the Fortran host makes through flite3d_amd/fortran/
flite_gpu.f90; each method names the reference routine it stands for.  The
library must be built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .params import FIELD_NCOMP, FIELDS, OPS, FliteLevelView, FliteParams, level_view

_bound = None


class DeviceError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"flite3d_amd error {code}: {msg}")
        self.code = code


def lib():
    global _bound
    if _bound is None:
        L = _native.device()
        P = ctypes.POINTER
        d, i, vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        L.flite_ctx_create.argtypes = [P(FliteParams), i, i, P(vp)]
        L.flite_batch_select.argtypes = [vp, i]
        L.flite_level_upload.argtypes = [vp, i, P(FliteLevelView)]
        L.flite_level_geometry.argtypes = [vp, i, P(FliteLevelView)]
        L.flite_ctx_destroy.argtypes = [vp]
        L.flite_ctx_destroy.restype = None
        L.flite_last_error.restype = ctypes.c_char_p
        L.flite_field_get.argtypes = [vp, i, i, P(d), i]
        L.flite_field_set.argtypes = [vp, i, i, P(d), i]
        L.flite_state_set.argtypes = [vp, i, P(d), i, i, i]
        L.flite_state_get.argtypes = [vp, i, P(d), i, i, i]
        L.flite_apply.argtypes = [vp, i, i]
        L.flite_init_freestream.argtypes = [vp]
        L.flite_init_from_state.argtypes = [vp]
        L.flite_advance_physical_time.argtypes = [vp]
        L.flite_do_time_iterations.argtypes = [vp, i, i, i, i, i]
        L.flite_restrict.argtypes = [vp, i, i]
        L.flite_prolongate.argtypes = [vp, i]
        L.flite_map_u_fine_to_coarse.argtypes = [vp, i]
        L.flite_cycle.argtypes = [vp, i]
        L.flite_monitor.argtypes = [vp, P(d), P(d), P(i), P(d)]
        L.flite_monitor_noalloc.argtypes = [vp, P(d), P(d), P(i), P(d)]
        L.flite_sync.argtypes = [vp]
        L.flite_tangent_apply.argtypes = [vp, P(d), i, P(d), i]
        L.flite_adjoint_apply.argtypes = [vp, P(d), i, P(d), i]
        L.flite_tangent_apply_device.argtypes = [vp, vp, vp]
        L.flite_adjoint_apply_device.argtypes = [vp, vp, vp]
        L.flite_drag_state_gradient.argtypes = [vp, P(d), P(d), i]
        L.flite_metric_adjoint.argtypes = [vp, P(d), i, P(d), i, P(d), i]
        L.flite_mesh_adjoint.argtypes = [vp, P(d), i, P(d), i, P(d), i, P(d), P(d), i, P(d)]
        L.flite_drag_mesh_gradient.argtypes = [vp, P(d), P(d), i, P(d), i, P(d)]
        L.flite_set_coarse_precision.argtypes = [vp, i]
        L.flite_autotune.argtypes = [vp, i, P(i)]
        L.flite_level_lanes.argtypes = [vp, i, P(i)]
        L.flite_level_side_block.argtypes = [vp, i, P(i), P(i)]
        L.flite_drag_metric_gradient.argtypes = [vp, P(d), P(d), i]
        _bound = L
    return _bound


class Solver:
    """One flite_ctx: the multigrid levels of one case on one GPU."""

    def __init__(self, params, levels):
        """params: one FliteParams, or a list of them for a batch of flow
        conditions on the same mesh (they may differ only in the freestream)."""
        self.L = lib()
        plist = list(params) if isinstance(params, (list, tuple)) else [params]
        self.n_batch = len(plist)
        arr = (FliteParams * self.n_batch)(*plist)
        self.params = arr
        self.levels = levels
        ctx = ctypes.c_void_p()
        self._check(self.L.flite_ctx_create(arr, len(levels), self.n_batch, ctypes.byref(ctx)))
        self.ctx = ctx
        for k, lv in enumerate(levels):
            v = level_view(lv, k == 0)
            self._check(self.L.flite_level_upload(self.ctx, k + 1, ctypes.byref(v)))

    def _check(self, rc):
        if rc != 0:
            raise DeviceError(rc, self.L.flite_last_error().decode())

    def replace_mesh(self, levels):
        """The next physical step's mesh of a moving-mesh run (flowType 2,
        mgnsg3d_mg_par.F90:525-548): new geometry for every level, the
        solution arrays kept."""
        for k, lv in enumerate(levels):
            v = level_view(lv, k == 0)
            self._check(self.L.flite_level_geometry(self.ctx, k + 1, ctypes.byref(v)))
        self.levels = levels

    def close(self):
        if getattr(self, "ctx", None):
            self.L.flite_ctx_destroy(self.ctx)
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
        self._check(self.L.flite_field_get(
            self.ctx, level, f, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n))
        return out

    def set(self, level, field, arr):
        f = FIELDS[field] if isinstance(field, str) else field
        n = self.n(level)
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(FIELD_NCOMP[f], n)
        self._check(self.L.flite_field_set(
            self.ctx, level, f, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n))

    def apply(self, level, op):
        self._check(self.L.flite_apply(self.ctx, level, OPS[op] if isinstance(op, str) else op))

    def init_freestream(self):
        """setUpInitialField2 (DEF:3713-3996) + multiGridCycles preamble (MG:586-602)."""
        self._check(self.L.flite_init_freestream(self.ctx))

    def init_from_state(self):
        self._check(self.L.flite_init_from_state(self.ctx))

    def advance_physical_time(self):
        """End of a physical time step (MG:889-890): uold2 <- uold <- u."""
        self._check(self.L.flite_advance_physical_time(self.ctx))

    def do_time_iterations(self, level, steps, calc_diss, calc_turb=True, iteration=1):
        """doTimeIterations (DEF:966-1000)."""
        self._check(self.L.flite_do_time_iterations(self.ctx, level, steps, int(calc_diss),
                                                    int(calc_turb), iteration))

    def restrict(self, fine_level, calc_turb=True):
        """restrictGrid (DEF:12141-12430)."""
        self._check(self.L.flite_restrict(self.ctx, fine_level, int(calc_turb)))

    def prolongate(self, coarse_level):
        """prolongateGrid (DEF:12432-12570)."""
        self._check(self.L.flite_prolongate(self.ctx, coarse_level))

    def map_u_fine_to_coarse(self, fine_level):
        """mapFineToCoarse of u (DEF:12572-12690)."""
        self._check(self.L.flite_map_u_fine_to_coarse(self.ctx, fine_level))

    def cycle(self, number):
        """oneCycle (MG:907-1196)."""
        self._check(self.L.flite_cycle(self.ctx, number))

    def set_coarse_precision(self, fp32):
        """Run the coarse levels in FP32 (True) or FP64 (False) from now on."""
        self._check(self.L.flite_set_coarse_precision(self.ctx, 1 if fp32 else 0))

    def autotune(self, repeats=3):
        """Pick each level's lanes per node by timing; returns them."""
        out = (ctypes.c_int * len(self.levels))()
        self._check(self.L.flite_autotune(self.ctx, repeats, out))
        return list(out)

    def lanes(self, level):
        v = ctypes.c_int()
        self._check(self.L.flite_level_lanes(self.ctx, level, ctypes.byref(v)))
        return v.value

    def side_block(self, level):
        """(sides per block, node cap) of the level's flux side blocks."""
        a, b = ctypes.c_int(), ctypes.c_int()
        self._check(self.L.flite_level_side_block(self.ctx, level, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def select(self, batch):
        """Condition that get/set/monitor act on (batched contexts)."""
        self._check(self.L.flite_batch_select(self.ctx, batch))

    def sync(self):
        self._check(self.L.flite_sync(self.ctx))

    def monitor(self):
        """writeResiduals reductions (MG:1271-1445) + calculateLiftAndDrag (DEF:14710-14870)."""
        res = (ctypes.c_double * 7)()
        mx = (ctypes.c_double * 7)()
        idx = (ctypes.c_int * 7)()
        forces = (ctypes.c_double * 9)()
        rc = self.L.flite_monitor(self.ctx, res, mx, idx, forces)
        if rc not in (0, -3):
            self._check(rc)
        return dict(res=np.array(res[:]), maxres=np.array(mx[:]), maxidx=np.array(idx[:]),
                    forces=np.array(forces[:]), rc=rc)


    # ---- linearisation of the level-1 residual (HLLC-1/2, inviscid)
    def _vec5(self, x):
        n = self.n(1)
        return np.ascontiguousarray(np.asarray(x, dtype=np.float64)[:5].reshape(5, n))

    def tangent(self, du):
        """(dR/dU) du of the level-1 residual at the current state: 5
        components, or 6 (nu~ last) with Spalart-Allmaras."""
        n = self.n(1)
        nc = 6 if self.params[0].turbulence_model == 1 else 5
        x = np.ascontiguousarray(np.asarray(du, dtype=np.float64)[:nc].reshape(nc, n))
        out = np.zeros((nc, n))
        D = ctypes.POINTER(ctypes.c_double)
        self._check(self.L.flite_tangent_apply(self.ctx, x.ctypes.data_as(D), n, out.ctypes.data_as(D), n))
        return out

    def adjoint(self, lam):
        """(dR/dU)^T lam of the level-1 residual at the current state: 5
        components, or 6 (nu~ last) with Spalart-Allmaras."""
        n = self.n(1)
        nc = 6 if self.params[0].turbulence_model == 1 else 5
        x = np.ascontiguousarray(np.asarray(lam, dtype=np.float64)[:nc].reshape(nc, n))
        out = np.zeros((nc, n))
        D = ctypes.POINTER(ctypes.c_double)
        self._check(self.L.flite_adjoint_apply(self.ctx, x.ctypes.data_as(D), n, out.ctypes.data_as(D), n))
        return out

    def tangent_device(self, du_ptr, out_ptr):
        """flite_tangent_apply_device: du and out are device addresses of
        (nc, N) float64 arrays in host numbering (e.g. tensor.data_ptr())."""
        self._check(self.L.flite_tangent_apply_device(self.ctx, ctypes.c_void_p(du_ptr), ctypes.c_void_p(out_ptr)))

    def adjoint_device(self, lam_ptr, out_ptr):
        """flite_adjoint_apply_device, as tangent_device."""
        self._check(self.L.flite_adjoint_apply_device(self.ctx, ctypes.c_void_p(lam_ptr), ctypes.c_void_p(out_ptr)))

    def drag_gradient(self):
        """(drag, dDrag/dU) at the current state: the pressure drag, plus the
        friction with ReynoldsNumber > 0; 6 components with Spalart-Allmaras."""
        n = self.n(1)
        out = np.zeros((6 if self.params[0].turbulence_model == 1 else 5, n))
        drag = ctypes.c_double()
        D = ctypes.POINTER(ctypes.c_double)
        self._check(self.L.flite_drag_state_gradient(self.ctx, ctypes.byref(drag), out.ctypes.data_as(D), n))
        return drag.value, out


    # ---- derivatives with respect to the level-1 metrics
    def _sizes1(self):
        lv = self.levels[0]
        return lv.side_index.shape[1], lv.bside_index.shape[1]

    def metric_adjoint(self, lam):
        """(d(lam^T R)/d side_w, d(lam^T R)/d bside_w) at the current state,
        shaped like the uploaded level's side_w (3, E) and bside_w (3, Eb)."""
        n = self.n(1)
        ne, nbs = self._sizes1()
        x = self._vec5(lam)
        gw = np.zeros((3, ne))
        gwb = np.zeros((3, nbs))
        D = ctypes.POINTER(ctypes.c_double)
        self._check(self.L.flite_metric_adjoint(self.ctx, x.ctypes.data_as(D), n, gw.ctypes.data_as(D), ne,
                                                gwb.ctypes.data_as(D), nbs))
        return gw, gwb

    def mesh_adjoint(self, lam):
        """d(lam^T R) with respect to side_w (3, E), bside_w (3, Eb), the
        dual volumes (N,), the coordinates directly (3, N) and, with
        Spalart-Allmaras, the wall distances (N,; zeros otherwise), at the
        current state (first- or second-order HLLC, inviscid, laminar or
        RANS-SA)."""
        n = self.n(1)
        ne, nbs = self._sizes1()
        nc = 6 if self.params[0].turbulence_model == 1 else 5
        x = np.ascontiguousarray(np.asarray(lam, dtype=np.float64)[:nc].reshape(nc, n))
        gw = np.zeros((3, ne))
        gwb = np.zeros((3, nbs))
        gv = np.zeros(n)
        gx = np.zeros((3, n))
        gd = np.zeros(n)
        D = ctypes.POINTER(ctypes.c_double)
        self._check(self.L.flite_mesh_adjoint(self.ctx, x.ctypes.data_as(D), n, gw.ctypes.data_as(D), ne,
                                              gwb.ctypes.data_as(D), nbs, gv.ctypes.data_as(D),
                                              gx.ctypes.data_as(D), n, gd.ctypes.data_as(D)))
        return gw, gwb, gv, gx, gd

    def drag_mesh_gradient(self):
        """(drag, dDrag/d side_w (3, E), dDrag/d bside_w (3, Eb), dDrag/d
        volume (N,)) at the current state (friction included with
        ReynoldsNumber > 0)."""
        n = self.n(1)
        ne, nbs = self._sizes1()
        gw = np.zeros((3, ne))
        gwb = np.zeros((3, nbs))
        gv = np.zeros(n)
        drag = ctypes.c_double()
        D = ctypes.POINTER(ctypes.c_double)
        self._check(self.L.flite_drag_mesh_gradient(self.ctx, ctypes.byref(drag), gw.ctypes.data_as(D), ne,
                                                    gwb.ctypes.data_as(D), nbs, gv.ctypes.data_as(D)))
        return drag.value, gw, gwb, gv

    def drag_metric_gradient(self):
        """(pressure drag, dDrag/d bside_w (3, Eb)) at the current state."""
        ne, nbs = self._sizes1()
        gwb = np.zeros((3, nbs))
        drag = ctypes.c_double()
        D = ctypes.POINTER(ctypes.c_double)
        self._check(self.L.flite_drag_metric_gradient(self.ctx, ctypes.byref(drag), gwb.ctypes.data_as(D), nbs))
        return drag.value, gwb


def _bind_extra():
    L = lib()
    if getattr(L, "_extra_bound", False):
        return L
    P = ctypes.POINTER
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.flite_timing_enable.argtypes = [vp, ctypes.c_char_p]
    L.flite_timing_get.argtypes = [vp, P(d), P(i)]
    L.flite_cycle_evaluations.argtypes = [vp, P(i)]
    L.flite_level_sizes.argtypes = [vp, i, P(i), P(i), P(i)]
    L._extra_bound = True
    return L


def timing_enable(solver, name):
    """Record HIP events around every launch of kernel class `name` (e.g. 'flux_l1')."""
    L = _bind_extra()
    solver._check(L.flite_timing_enable(solver.ctx, None if name is None else name.encode()))


def timing_get(solver):
    L = _bind_extra()
    ms = ctypes.c_double()
    cnt = ctypes.c_int()
    solver._check(L.flite_timing_get(solver.ctx, ctypes.byref(ms), ctypes.byref(cnt)))
    return ms.value, cnt.value


def cycle_evaluations(solver):
    L = _bind_extra()
    out = (ctypes.c_int * len(solver.levels))()
    solver._check(L.flite_cycle_evaluations(solver.ctx, out))
    return list(out)


def level_sizes(solver, level):
    L = _bind_extra()
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    solver._check(L.flite_level_sizes(solver.ctx, level, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
    return a.value, b.value, c.value
