"""Bump-box meshes and the multigrid level data built from them.

The synthetic case: a unit box split into n^3 hexahedra, each split into the
6 Kuhn tetrahedra sharing its main diagonal; the wall (indicator 1) is the
z = 0 face carrying a Gaussian bump, the other five faces are farfield
(indicator 3).  Coordinates are generated here with numpy; the median-dual
geometry and the agglomerated levels come from the C++ preprocessor
(``libflite3d_prepro.so``), which restates Gen3d's producers
(PreProSource/gen3d_grid.f90:2092-2332, gen3d_boundary.f90:289-620,
gen3d_grid.f90:2358-3611).

Mesh recipe (fixed here so every run builds the same grid):
* structured parameters (xi, eta, zeta) in [0,1]^3 on an (n+1)^3 lattice;
  zeta is uniform, or graded from ``first_cell`` by ``ratio`` for as long as
  the next graded cell is smaller than the uniform spacing that would fill
  the rest of [0,1] with the remaining cells (n=128, 1e-6, 1.15 gives 69
  graded cells and then 0.015205 spacing);
* jitter: ``rng = numpy.random.default_rng(seed)``,
  ``d = rng.uniform(-jitter, jitter, size=(N, 3))`` drawn for every lattice
  node in lexicographic order (i fastest, then j, then k) and applied to the
  strictly interior nodes only, scaled by the local lattice spacing of each
  parameter direction;
* bump: z_wall(x, y) = h exp(-((x-1/2)^2 + (y-1/2)^2) / (2 sigma^2)), and
  z = z_wall + zeta (1 - z_wall), x = xi, y = eta;
* boundary triangles: each boundary quad split along the diagonal from its
  lowest to its highest lattice corner (the Kuhn faces), ordered wall face
  first, then x=0, x=1, y=0, y=1, z=1, oriented with the right-hand normal
  pointing into the domain.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os

import numpy as np

from . import _native


def graded_spacing(n, first_cell, ratio):
    """Parameter values zeta_0..zeta_n for the graded distribution."""
    if first_cell is None:
        return np.linspace(0.0, 1.0, n + 1)
    cells = []
    s = 0.0
    h = first_cell
    while len(cells) < n:
        rest = (1.0 - s) / (n - len(cells))
        if h >= rest:
            break
        cells.append(h)
        s += h
        h *= ratio
    m = n - len(cells)
    if m > 0:
        rest = (1.0 - s) / m
        cells.extend([rest] * m)
    z = np.concatenate([[0.0], np.cumsum(cells)])
    z[-1] = 1.0
    return z


def bump_box(n, jitter=0.0, seed=0, first_cell=None, ratio=1.15, h=0.05,
             sigma=0.1, wall_indicator=1):
    """Return (xyz[N,3], tets[T,4], btri[B,3], bind[B]) with 0-based nodes."""
    n1 = n + 1
    xi = np.linspace(0.0, 1.0, n1)
    zeta = graded_spacing(n, first_cell, ratio)
    I, J, K = np.meshgrid(np.arange(n1), np.arange(n1), np.arange(n1),
                          indexing="ij")
    # lexicographic order: i fastest
    I = I.transpose(2, 1, 0).ravel()
    J = J.transpose(2, 1, 0).ravel()
    K = K.transpose(2, 1, 0).ravel()
    px = xi[I].copy()
    py = xi[J].copy()
    pz = zeta[K].copy()
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        d = rng.uniform(-jitter, jitter, size=(n1 ** 3, 3))
        interior = (I > 0) & (I < n) & (J > 0) & (J < n) & (K > 0) & (K < n)
        dz = np.diff(zeta)
        hx = 1.0 / n
        hz = np.minimum(dz[np.clip(K - 1, 0, n - 1)], dz[np.clip(K, 0, n - 1)])
        px[interior] += d[interior, 0] * hx
        py[interior] += d[interior, 1] * hx
        pz[interior] += d[interior, 2] * hz[interior]
    zw = h * np.exp(-((px - 0.5) ** 2 + (py - 0.5) ** 2) / (2.0 * sigma ** 2))
    z = zw + pz * (1.0 - zw)
    xyz = np.stack([px, py, z], axis=1)

    def nid(i, j, k):
        return i + n1 * (j + n1 * k)

    hi, hj, hk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n),
                             indexing="ij")
    hi = hi.transpose(2, 1, 0).ravel()
    hj = hj.transpose(2, 1, 0).ravel()
    hk = hk.transpose(2, 1, 0).ravel()
    tets = []
    for p in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
        a = np.zeros((3, hi.size), dtype=np.int64)
        verts = [nid(hi, hj, hk)]
        for ax in p:
            a[ax] += 1
            verts.append(nid(hi + a[0], hj + a[1], hk + a[2]))
        tets.append(np.stack(verts, axis=1))
    tets = np.stack(tets, axis=1).reshape(-1, 4).astype(np.int32)

    tris, inds = [], []
    u, v = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    u = u.T.ravel()
    v = v.T.ravel()

    def face(fix_axis, fix_val, indicator, inward):
        # quad corners (a,b) in the two free axes, diagonal (0,0)-(1,1)
        free = [ax for ax in range(3) if ax != fix_axis]

        def node(da, db):
            c = [None, None, None]
            c[fix_axis] = np.full(u.shape, fix_val)
            c[free[0]] = u + da
            c[free[1]] = v + db
            return nid(c[0], c[1], c[2])
        t1 = np.stack([node(0, 0), node(1, 0), node(1, 1)], axis=1)
        t2 = np.stack([node(0, 0), node(1, 1), node(0, 1)], axis=1)
        t = np.stack([t1, t2], axis=1).reshape(-1, 3)
        p0, p1, p2 = xyz[t[:, 0]], xyz[t[:, 1]], xyz[t[:, 2]]
        nrm = np.cross(p1 - p0, p2 - p0)
        flip = (nrm @ np.asarray(inward, dtype=float)) < 0
        t[flip] = t[flip][:, [0, 2, 1]]
        tris.append(t)
        inds.append(np.full(t.shape[0], indicator, dtype=np.int32))

    face(2, 0, wall_indicator, (0, 0, 1))
    face(0, 0, 3, (1, 0, 0))
    face(0, n, 3, (-1, 0, 0))
    face(1, 0, 3, (0, 1, 0))
    face(1, n, 3, (0, -1, 0))
    face(2, n, 3, (0, 0, -1))
    btri = np.concatenate(tris).astype(np.int32)
    bind = np.concatenate(inds)
    return xyz, tets, btri, bind


@dataclasses.dataclass
class Level:
    """One multigrid level in the solver's record layout, 0-based indices.

    Arrays follow the .sol records (mgnsg3d_def_par.F90:791-865,
    13843-13862, 13867-13906, 13991-14007, 444-545): ``side_index[2, E]``,
    ``side_w[3, E]``, ``bside_index[3, Eb]`` (i1, i2, indicator),
    ``bside_w[3, Eb]``, node arrays of length N, and for coarse levels the
    fine->coarse ``f2c`` map with the ``rweight`` = V_fine/V_coarse weights.
    """
    n_nodes: int
    n_bnodes: int
    side_index: np.ndarray
    side_w: np.ndarray
    bside_index: np.ndarray
    bside_w: np.ndarray
    volume: np.ndarray
    coords: np.ndarray
    wall_dist: np.ndarray
    wall_node: np.ndarray
    reg: np.ndarray
    node_ind_array: np.ndarray
    node_ind: np.ndarray
    seed: np.ndarray
    f2c: np.ndarray | None = None
    rweight: np.ndarray | None = None
    # moving meshes (flowType 2), as the solver holds them after reading a
    # step's .sol: GCL weight per side [E], per boundary side [2, Eb],
    # previous volumes [N] and coordinate movement [3, N] (level 1)
    gcl: np.ndarray | None = None
    bgcl: np.ndarray | None = None
    volume_p: np.ndarray | None = None
    volume_p2: np.ndarray | None = None
    coord_move: np.ndarray | None = None

    @property
    def n_sides(self):
        return self.side_index.shape[1]

    @property
    def n_bsides(self):
        return self.bside_index.shape[1]


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


_PROGRESS = ctypes.CFUNCTYPE(None, ctypes.c_char_p)


def build_levels(xyz, tets, btri, bind, n_levels, progress=None):
    """Run the preprocessor; returns (levels, old_index).  progress(msg),
    if given, is called after each phase of the build."""
    lib = _native.prepro()
    cb = _PROGRESS(lambda m: progress(m.decode())) if progress else None
    lib.fp_set_progress(ctypes.cast(cb, ctypes.c_void_p) if cb else None)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    tets = np.ascontiguousarray(tets, dtype=np.int32)
    btri = np.ascontiguousarray(btri, dtype=np.int32)
    bind = np.ascontiguousarray(bind, dtype=np.int32)
    m = lib.fp_mesh_build(xyz.shape[0], _ptr(xyz, ctypes.c_double),
                          tets.shape[0], _ptr(tets, ctypes.c_int),
                          btri.shape[0], _ptr(btri, ctypes.c_int),
                          _ptr(bind, ctypes.c_int), n_levels)
    lib.fp_set_progress(None)
    if not m:
        raise RuntimeError("preprocessor failed: " +
                           lib.fp_last_error().decode())
    try:
        levels = []
        for lv in range(lib.fp_mesh_nlevels(m)):
            nn, ns, nbs, nbn = (ctypes.c_int(), ctypes.c_int(), ctypes.c_int(),
                                ctypes.c_int())
            lib.fp_level_sizes(m, lv, ctypes.byref(nn), ctypes.byref(ns),
                               ctypes.byref(nbs), ctypes.byref(nbn))
            nn, ns, nbs, nbn = nn.value, ns.value, nbs.value, nbn.value
            side_index = np.empty((2, ns), np.int32)
            side_w = np.empty((3, ns), np.float64)
            bside_index = np.empty((3, nbs), np.int32)
            bside_w = np.empty((3, nbs), np.float64)
            volume = np.empty(nn)
            coords = np.empty((3, nn))
            wall_dist = np.empty(nn)
            wall_node = np.empty(nn, np.int32)
            reg = np.empty(21, np.int32)
            node_ind_array = np.empty(nbn, np.int32)
            node_ind = np.empty(nbn, np.int32)
            seed = np.empty(nn, np.int32)
            f2c = rweight = None
            if lv > 0:
                nf = levels[-1].n_nodes
                f2c = np.empty(nf, np.int32)
                rweight = np.empty(nf)
            rc = lib.fp_level_copy(
                m, lv, _ptr(side_index, ctypes.c_int),
                _ptr(side_w, ctypes.c_double),
                _ptr(bside_index, ctypes.c_int),
                _ptr(bside_w, ctypes.c_double), _ptr(volume, ctypes.c_double),
                _ptr(coords, ctypes.c_double), _ptr(wall_dist, ctypes.c_double),
                _ptr(wall_node, ctypes.c_int), _ptr(reg, ctypes.c_int),
                _ptr(node_ind_array, ctypes.c_int),
                _ptr(node_ind, ctypes.c_int), _ptr(seed, ctypes.c_int),
                None if f2c is None else _ptr(f2c, ctypes.c_int),
                None if rweight is None else _ptr(rweight, ctypes.c_double))
            if rc != 0:
                raise RuntimeError(lib.fp_last_error().decode())
            levels.append(Level(nn, nbn, side_index, side_w, bside_index,
                                bside_w, volume, coords, wall_dist, wall_node,
                                reg, node_ind_array, node_ind, seed, f2c,
                                rweight))
        old = np.empty(levels[0].n_nodes, np.int32)
        lib.fp_fine_old_index(m, _ptr(old, ctypes.c_int))
        return levels, old
    finally:
        lib.fp_mesh_free(m)


def metric_vjp(xyz, tets, btri, bind, gw, gwb):
    """Transpose of the fine level's metrics with respect to the coordinates:
    (d side_w/d xyz)^T gw + (d bside_w/d xyz)^T gwb for the fine level that
    build_levels makes from the same mesh.  gw is (3, n_sides) and gwb
    (3, n_bsides) in that level's side order (the layout of side_w,
    bside_w); returns an (n_nodes, 3) array in the input node order."""
    lib = _native.prepro()
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    tets = np.ascontiguousarray(tets, dtype=np.int32)
    btri = np.ascontiguousarray(btri, dtype=np.int32)
    bind = np.ascontiguousarray(bind, dtype=np.int32)
    gw = np.ascontiguousarray(gw, dtype=np.float64)
    gwb = np.ascontiguousarray(gwb, dtype=np.float64)
    gx = np.zeros((xyz.shape[0], 3))
    rc = lib.fp_metric_vjp(xyz.shape[0], _ptr(xyz, ctypes.c_double),
                           tets.shape[0], _ptr(tets, ctypes.c_int),
                           btri.shape[0], _ptr(btri, ctypes.c_int),
                           _ptr(bind, ctypes.c_int),
                           _ptr(gw, ctypes.c_double), _ptr(gwb, ctypes.c_double),
                           _ptr(gx, ctypes.c_double))
    if rc != 0:
        raise RuntimeError("metric_vjp failed: " + lib.fp_last_error().decode())
    return gx


def mesh_vjp(xyz, tets, btri, bind, gw, gwb, gvol=None, gxn=None):
    """fp_mesh_vjp: as metric_vjp plus the dual volumes' and a direct
    coordinate term's transposes.  gvol is (n_nodes,) and gxn (3, n_nodes)
    in the level's node numbering (as the library reports them); returns an
    (n_nodes, 3) array in the input node order."""
    lib = _native.prepro()
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    tets = np.ascontiguousarray(tets, dtype=np.int32)
    btri = np.ascontiguousarray(btri, dtype=np.int32)
    bind = np.ascontiguousarray(bind, dtype=np.int32)
    gw = np.ascontiguousarray(gw, dtype=np.float64)
    gwb = np.ascontiguousarray(gwb, dtype=np.float64)
    D = ctypes.POINTER(ctypes.c_double)
    gv = np.ascontiguousarray(gvol, dtype=np.float64) if gvol is not None else None
    gn = np.ascontiguousarray(np.asarray(gxn, dtype=np.float64).T) if gxn is not None else None
    gx = np.zeros((xyz.shape[0], 3))
    rc = lib.fp_mesh_vjp(xyz.shape[0], _ptr(xyz, ctypes.c_double),
                         tets.shape[0], _ptr(tets, ctypes.c_int),
                         btri.shape[0], _ptr(btri, ctypes.c_int),
                         _ptr(bind, ctypes.c_int),
                         _ptr(gw, ctypes.c_double), _ptr(gwb, ctypes.c_double),
                         _ptr(gv, ctypes.c_double) if gv is not None else ctypes.cast(None, D),
                         _ptr(gn, ctypes.c_double) if gn is not None else ctypes.cast(None, D),
                         _ptr(gx, ctypes.c_double))
    if rc != 0:
        raise RuntimeError("mesh_vjp failed: " + lib.fp_last_error().decode())
    return gx


def wall_distance_vjp(level, gwd):
    """Transpose of the wall distance d_i = |x_i - x_w(i)| (w(i) the nearest
    wall node the build found): (3, N) coordinate shares in the level's
    numbering, to pass as (part of) mesh_vjp's gxn.  Nodes at the wall
    (d = 0) carry no share: the SA source clamps d at 1e-10."""
    x = level.coords
    w = level.wall_node
    d = level.wall_dist
    gx = np.zeros_like(x, dtype=np.float64)
    on = d > 1e-10
    i = np.nonzero(on)[0]
    v = (x[:, i] - x[:, w[i]]) * (np.asarray(gwd)[i] / d[i])
    np.add.at(gx, (slice(None), i), v)
    np.subtract.at(gx, (slice(None), w[i]), v)
    return gx


def bump_levels(n, n_levels, jitter=0.0, seed=0, first_cell=None,
                ratio=1.15, h=0.05, wall_indicator=1, progress=None):
    xyz, tets, btri, bind = bump_box(n, jitter=jitter, seed=seed,
                                     first_cell=first_cell, ratio=ratio,
                                     h=h, wall_indicator=wall_indicator)
    if progress:
        progress(f"bump box {n}^3: {xyz.shape[0]} nodes, {tets.shape[0]} tets")
    levels, _ = build_levels(xyz, tets, btri, bind, n_levels, progress)
    return levels

