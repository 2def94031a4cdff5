"""Transpose of the preprocessor's metrics with respect to the node
coordinates (fp_metric_vjp), on the CPU.

The median-dual side weights are sums of cross products of differences of
vertex, edge-midpoint, face-centroid and tet-centroid positions
(PreProSource/gen3d_grid.f90:786-855, 2092-2332; boundary sides
PreProSource/gen3d_boundary.f90:289-372): quadratic in the coordinates.
Central differences of a quadratic are exact up to rounding, so the
transpose is held to 1e-10 against them; translation invariance (the
transposed gradient sums to zero over the nodes) and homogeneity of degree
two (sum_i x_i . g_i = 2 <g_w, w>) pin it independently of the
differences.
"""
import numpy as np

from flite3d_amd import mesh


def case(n=5, jitter=0.15):
    xyz, tets, btri, bind = mesh.bump_box(n, jitter=jitter, seed=0)
    lv, old = mesh.build_levels(xyz, tets, btri, bind, 1)
    return (xyz, tets, btri, bind), lv[0]


def metrics(xyz, tets, btri, bind):
    lv, _ = mesh.build_levels(xyz, tets, btri, bind, 1)
    return lv[0].side_w, lv[0].bside_w


def test_vjp_matches_central_differences():
    (xyz, tets, btri, bind), L0 = case()
    rng = np.random.default_rng(3)
    for part in ("both", "interior", "boundary"):
        gw = rng.standard_normal(L0.side_w.shape) * (part != "boundary")
        gwb = rng.standard_normal(L0.bside_w.shape) * (part != "interior")
        gx = mesh.metric_vjp(xyz, tets, btri, bind, gw, gwb)
        d = rng.standard_normal(xyz.shape)
        h = 1e-3
        wp, bp = metrics(xyz + h * d, tets, btri, bind)
        wm, bm = metrics(xyz - h * d, tets, btri, bind)
        fd = (np.sum(gw * (wp - wm)) + np.sum(gwb * (bp - bm))) / (2 * h)
        an = float(np.sum(gx * d))
        assert abs(fd - an) <= 1e-10 * max(abs(an), 1.0), (part, fd, an)


def test_vjp_translation_invariance_and_homogeneity():
    (xyz, tets, btri, bind), L0 = case(4, 0.1)
    rng = np.random.default_rng(5)
    gw = rng.standard_normal(L0.side_w.shape)
    gwb = rng.standard_normal(L0.bside_w.shape)
    gx = mesh.metric_vjp(xyz, tets, btri, bind, gw, gwb)
    scale = np.abs(gx).sum()
    assert np.abs(gx.sum(axis=0)).max() <= 1e-12 * scale
    # w(s x) = s^2 w(x): sum_i x_i . dF/dx_i = 2 F for F = <gw, w> + <gwb, wb>
    F = np.sum(gw * L0.side_w) + np.sum(gwb * L0.bside_w)
    assert abs(np.sum(gx * xyz) - 2 * F) <= 1e-12 * max(abs(F), 1.0) * 10


def test_vjp_rejects_bad_input():
    (xyz, tets, btri, bind), L0 = case(3, 0.0)
    bad = tets.copy()
    bad[0, 0] = xyz.shape[0] + 5
    try:
        mesh.metric_vjp(xyz, bad, btri, bind, np.zeros(L0.side_w.shape), np.zeros(L0.bside_w.shape))
    except RuntimeError as e:
        assert "out of range" in str(e)
    else:
        raise AssertionError("accepted a tet with an out-of-range node")


def volumes(xyz, tets, btri, bind):
    lv, _ = mesh.build_levels(xyz, tets, btri, bind, 1)
    return lv[0].volume


def test_volume_vjp():
    """fp_mesh_vjp's dual-volume part: the volumes are cubic in the
    coordinates, so central differences with a small step, translation
    invariance and homogeneity of degree three (sum_i x_i . g_i =
    3 <g_V, V>) pin it; the direct coordinate term passes through in the
    input node order."""
    (xyz, tets, btri, bind), L0 = case()
    rng = np.random.default_rng(4)
    gv = rng.standard_normal(L0.n_nodes)
    zw, zb = np.zeros_like(L0.side_w), np.zeros_like(L0.bside_w)
    gx = mesh.mesh_vjp(xyz, tets, btri, bind, zw, zb, gvol=gv)
    d = rng.standard_normal(xyz.shape)
    h = 1e-5
    fd = (gv @ volumes(xyz + h * d, tets, btri, bind) - gv @ volumes(xyz - h * d, tets, btri, bind)) / (2 * h)
    an = float(np.sum(gx * d))
    assert abs(fd - an) <= 1e-7 * abs(an), (fd, an)
    assert np.abs(gx.sum(axis=0)).max() <= 1e-11 * np.abs(gx).max()
    assert abs(float(np.sum(xyz * gx)) - 3 * float(gv @ L0.volume)) <= 1e-10 * abs(float(gv @ L0.volume))
    # the direct term: level numbering in, input numbering out
    _, old = mesh.build_levels(xyz, tets, btri, bind, 1)
    gn = rng.standard_normal((3, L0.n_nodes))
    gx2 = mesh.mesh_vjp(xyz, tets, btri, bind, zw, zb, gxn=gn)
    ref = np.zeros_like(xyz)
    ref[np.asarray(old)] = gn.T
    assert np.array_equal(gx2, ref)


def test_wall_distance_vjp():
    """mesh.wall_distance_vjp against central differences of the
    preprocessor's wall distance: the distance to the nearest wall node,
    smooth while that node stays the same."""
    xyz, tets, btri, bind = mesh.bump_box(5, jitter=0.15, seed=0)
    lv, old = mesh.build_levels(xyz, tets, btri, bind, 1)
    L0 = lv[0]
    rng = np.random.default_rng(5)
    gwd = rng.standard_normal(L0.n_nodes)
    gx = mesh.wall_distance_vjp(L0, gwd)
    d = rng.standard_normal(xyz.shape)
    h = 1e-7
    lp, _ = mesh.build_levels(xyz + h * d, tets, btri, bind, 1)
    lm, _ = mesh.build_levels(xyz - h * d, tets, btri, bind, 1)
    assert np.array_equal(lp[0].wall_node, L0.wall_node) and np.array_equal(lm[0].wall_node, L0.wall_node)
    fd = float(np.sum(gwd * (lp[0].wall_dist - lm[0].wall_dist))) / (2 * h)
    an = float(np.sum(gx * d[np.asarray(old)].T))
    assert abs(an) > 0
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
