"""torch.autograd.Function wrappers over the C-ABI sensitivities, checked
with torch.autograd.gradcheck (central differences of the forward against
the backward), on the GPU through the library:

  Residual: U -> R(U), the level-1 inviscid residual (FLITE_OP_RHS),
            first order after a few cycles and second order (MUSCL + LED)
            about a smooth non-uniform state; backward = flite_adjoint_apply.
  ShapeObjective: X -> drag(U, X) - lam^T R(U, X) at a fixed state U and
            vector lam, with the metrics rebuilt by the preprocessor from X;
            backward = flite_metric_adjoint + flite_drag_metric_gradient,
            chained through fp_metric_vjp.

The tensors live on the host (float64); the work runs on the GPU.  Tiny
jittered bump boxes keep gradcheck's per-element differences cheap.
"""
import numpy as np
import pytest
import torch

from flite3d_amd import mesh
from flite3d_amd.device import Solver

from cases import euler

pytestmark = pytest.mark.gpu


def params(order=1):
    return euler(order, machnumber=0.8, alpha=1.25, cflnumber=2.0, multigridscheme=3).to_struct()


class Residual(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, solver):
        full = np.zeros((6, u.shape[1]))
        full[:5] = u.detach().numpy()
        solver.set(1, "U", full)
        solver.apply(1, "PRESSURE")
        solver.apply(1, "RHS")
        ctx.solver = solver
        ctx.u = full
        return torch.from_numpy(solver.get(1, "RHS")[:5].copy())

    @staticmethod
    def backward(ctx, g):
        s = ctx.solver
        s.set(1, "U", ctx.u)
        return torch.from_numpy(s.adjoint(g.detach().numpy())), None


class ShapeObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, raw, u, lam):
        _, tets, btri, bind = raw
        xyz = x.detach().numpy()
        lv, _ = mesh.build_levels(xyz, tets, btri, bind, 2)
        s = Solver(params(), lv)
        s.set(1, "U", u)
        s.apply(1, "PRESSURE")
        s.apply(1, "RHS")
        r = s.get(1, "RHS")[:5]
        drag, _ = s.drag_gradient()
        gw, gwb = s.metric_adjoint(lam)
        _, gjwb = s.drag_metric_gradient()
        s.close()
        ctx.gx = mesh.metric_vjp(xyz, tets, btri, bind, -gw, gjwb - gwb)
        return torch.tensor(drag - float(np.sum(lam * r)), dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        return g * torch.from_numpy(ctx.gx), None, None, None


@pytest.fixture(scope="module")
def small():
    raw = mesh.bump_box(3, jitter=0.15, seed=0)
    lv, _ = mesh.build_levels(*raw, 2)
    s = Solver(params(), lv)
    s.init_freestream()
    for c in range(1, 4):
        s.cycle(c)
    u = s.get(1, "U")
    s.close()
    return raw, lv, u


def test_residual_function_gradcheck(small):
    print("[case] autograd residual gradcheck n=3", flush=True)
    _, lv, u = small
    s = Solver(params(), lv)
    x = torch.from_numpy(u[:5].copy()).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda t: Residual.apply(t, s), (x,), eps=1e-6, atol=1e-7, rtol=1e-5)
    s.close()


def test_second_order_residual_function_gradcheck(small):
    """HLLC-2 with LED clipping, about a smooth state with no equal
    neighbour values (a limiter switch within eps would make the central
    differences straddle a kink)."""
    print("[case] autograd second-order residual gradcheck n=3", flush=True)
    from test_adjoint import smooth_state
    _, lv, u = small
    s = Solver(params(2), lv)
    s.init_freestream()
    u2 = smooth_state(lv, s.get(1, "U"))
    x = torch.from_numpy(u2[:5].copy()).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda t: Residual.apply(t, s), (x,), eps=1e-6, atol=1e-7, rtol=1e-5)
    s.close()


def test_shape_function_gradcheck(small):
    print("[case] autograd shape gradcheck n=3", flush=True)
    raw, lv, u = small
    lam = np.random.default_rng(7).standard_normal((5, u.shape[1]))
    x = torch.from_numpy(raw[0].copy()).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda t: ShapeObjective.apply(t, raw, u, lam), (x,), eps=1e-6, atol=1e-7,
                                    rtol=1e-5, fast_mode=True)
