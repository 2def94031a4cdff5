"""<case>.sol_1 grid files and <case>.res_1 / .rst_1 solution files.

The solver reads its preprocessed grid from a Fortran unformatted sequential
file (every READ is one record framed by 4-byte little-endian length markers;
reals are 8 bytes, integers 4).  The record order is the one the reference's
reader consumes:

  readComputationData (SolverSource/mgnsg3d_mg_par.F90:1198-1268)
    numberOfGrids
    (domain, numberOfGridDomains)
    per level, readGridComputationData (SolverSource/mgnsg3d_def_par.F90:13812-14255):
      readBoundarySolverData (DEF:791-865)
        numberOfBoundarySides
        Eb x (i1, i2, indicator, w_b(3), GCL(2), GCLprev(2))
        (numberOfBoundaryNodes, numberOfNormals)
        21 x nodeIndicatorRegister(k), k = -20..0
        numberOfNormals x (nodeIndicatorArray(i), normtan(i,1:6))
        numberOfEngineInletSides (+ records)
      numberOfSides; E x (i1, i2, w(3), GCL, GCLprev, length)
      numberOfNodes
        fine:   N x (idx, V, Vprev, Vprev2, wallDist, wallDistNode, x(3), xprev(3), xprev2(3))
        coarse: N x (idx, V, wallDist, wallDistNode, seedNode)
      fine only: numberOfTripNodes (+ records), numberOfTripFieldNodes (+ records)
      readGridParallelizationData for one domain (DEF:594-690):
        (numberOfComSides, numberOfBoundaryComSides), nSend(1), nRecv(1)
      coarse only:
        readProlongationOperatorData (DEF:452-478): nIntFine, nIntCoarse,
          nIntFine x (fine, coarse), numberOfComNodes (+ records)
        readRestrictionOperatorData (DEF:516-542): nIntFine, nIntCoarse,
          nIntFine x (fine, V_f / V_c), numberOfComNodes (+ records)
        readComParallelizationData (DEF:700-745) for one domain

The solution file written by writeGridResults (DEF:15331-15345) and read
back as a restart by setUpInitialField (DEF:3744-3767, 3853-3858) is N, then
one record rho(N), (u(:,j)/rho, j = 2..nvar).

This writer fills what a steady single-domain run needs: zero GCL weights,
previous volumes and coordinates equal to the current ones, no trip lines,
no engines, zero normal/tangent records (read only for moving walls,
DEF:823-827), and for coarse levels the side "length" is the distance
between the two seed nodes (the hot path does not read it).
"""
from __future__ import annotations

import numpy as np

from .mesh import Level

_I4 = np.dtype("<i4")
_F8 = np.dtype("<f8")


_TRANS = np.dtype([("local", "<i4"), ("global", "<i4")])  # one plot-register record


class _Writer:
    def __init__(self, f):
        self.f = f

    def rec(self, *fields):
        """One record of scalars: ints are int32, floats float64."""
        parts = []
        for x in fields:
            if isinstance(x, (int, np.integer)):
                parts.append(np.array([x], _I4).tobytes())
            else:
                parts.append(np.array([x], _F8).tobytes())
        self.raw(b"".join(parts))

    def raw(self, payload):
        n = np.array([len(payload)], _I4).tobytes()
        self.f.write(n + payload + n)

    def table(self, dtype, arr):
        """Many records of one structured layout, framed in one write."""
        size = dtype.itemsize
        framed = np.empty(len(arr), np.dtype([("h", _I4), ("p", dtype), ("t", _I4)]))
        framed["h"] = size
        framed["t"] = size
        framed["p"] = arr
        self.f.write(framed.tobytes())


class _Reader:
    def __init__(self, buf):
        self.b = memoryview(buf)
        self.o = 0

    def _payload(self):
        n = int(np.frombuffer(self.b[self.o:self.o + 4], _I4)[0])
        p = self.b[self.o + 4:self.o + 4 + n]
        t = int(np.frombuffer(self.b[self.o + 4 + n:self.o + 8 + n], _I4)[0])
        if t != n:
            raise ValueError(f"corrupt record at byte {self.o}: markers {n} != {t}")
        self.o += 8 + n
        return p

    def ints(self, k=1):
        v = np.frombuffer(self._payload(), _I4, count=k)
        return v if k > 1 else int(v[0])

    def table(self, dtype, count):
        """`count` records of at least dtype.itemsize bytes each (all one size)."""
        if count == 0:
            return np.empty(0, dtype)
        n = int(np.frombuffer(self.b[self.o:self.o + 4], _I4)[0])
        if n < dtype.itemsize:
            raise ValueError(f"record of {n} bytes shorter than {dtype.itemsize}")
        fields = [("h", _I4), ("p", dtype)]
        if n > dtype.itemsize:
            fields.append(("pad", "V%d" % (n - dtype.itemsize)))
        full = np.dtype(fields + [("t", _I4)])
        a = np.frombuffer(self.b[self.o:self.o + full.itemsize * count], full, count=count)
        if not (np.all(a["h"] == n) and np.all(a["t"] == n)):
            raise ValueError(f"corrupt record table at byte {self.o}")
        self.o += full.itemsize * count
        return a["p"].copy()


_BSIDE = np.dtype([("i", _I4, 3), ("w", _F8, 3), ("gcl", _F8, 2), ("gclp", _F8, 2)])
_NORMAL = np.dtype([("node", _I4), ("nt", _F8, 6)])
_SIDE = np.dtype([("i", _I4, 2), ("w", _F8, 3), ("gcl", _F8), ("gclp", _F8), ("len", _F8)])
_FNODE = np.dtype([("idx", _I4), ("v", _F8, 3), ("wd", _F8), ("wn", _I4), ("x", _F8, 9)])
_CNODE = np.dtype([("idx", _I4), ("v", _F8), ("wd", _F8), ("wn", _I4), ("seed", _I4)])
_PROL = np.dtype([("f", _I4), ("c", _I4)])
_REST = np.dtype([("f", _I4), ("w", _F8)])


def write_sol(path, levels: list[Level]):
    """Write `levels` (0-based mesh.Level list, fine first) as <case>.sol_1."""
    fine = levels[0]
    with open(path, "wb") as f:
        W = _Writer(f)
        W.rec(len(levels))
        W.rec(1, 1)
        for li, L in enumerate(levels):
            nbs = L.n_bsides
            b = np.zeros(nbs, _BSIDE)
            b["i"][:, :2] = L.bside_index[:2].T + 1
            b["i"][:, 2] = L.bside_index[2]
            b["w"] = L.bside_w.T
            if getattr(L, "bgcl", None) is not None:
                # a moving mesh's weights as the solver holds them: GCLprev
                # = GCL, so 1.5 GCL - 0.5 GCLprev (timeScheme 2) is GCL too
                b["gcl"] = b["gclp"] = np.asarray(L.bgcl).T
            W.rec(nbs)
            W.table(_BSIDE, b)
            nia = np.asarray(L.node_ind_array, np.int64)
            W.rec(int(L.n_bnodes), int(len(nia)))
            for k in range(21):
                W.rec(int(L.reg[k]))
            nr = np.zeros(len(nia), _NORMAL)
            nr["node"] = nia + 1
            W.table(_NORMAL, nr)
            W.rec(0)  # engine inlet sides
            ne = L.n_sides
            s = np.zeros(ne, _SIDE)
            s["i"] = L.side_index.T + 1
            s["w"] = L.side_w.T
            X = fine.coords[:, np.asarray(L.seed, np.int64)]  # seeds are level-1 nodes
            d = X[:, L.side_index[1]] - X[:, L.side_index[0]]
            s["len"] = np.sqrt((d * d).sum(0))
            if getattr(L, "gcl", None) is not None:
                s["gcl"] = s["gclp"] = L.gcl
            W.rec(ne)
            W.table(_SIDE, s)
            n = L.n_nodes
            W.rec(n)
            if li == 0:
                t = np.zeros(n, _FNODE)
                t["idx"] = np.arange(1, n + 1)
                t["v"] = L.volume[:, None]
                t["wd"] = L.wall_dist
                t["wn"] = L.wall_node + 1
                x = L.coords.T
                t["x"] = np.concatenate([x, x, x], axis=1)
                if getattr(L, "volume_p", None) is not None:
                    t["v"][:, 1] = L.volume_p
                    t["v"][:, 2] = L.volume_p2
                if getattr(L, "coord_move", None) is not None:
                    # previous coordinates x - dx and x - 2 dx: the movement
                    # is dx for timeScheme 1 and 2 alike
                    m = np.asarray(L.coord_move).T
                    t["x"] = np.concatenate([x, x - m, x - 2 * m], axis=1)
                W.table(_FNODE, t)
                W.rec(0)  # trip nodes
                W.rec(0)  # trip field nodes
                W.rec(0, 0)
                W.rec(0)
                W.rec(0)
            else:
                t = np.zeros(n, _CNODE)
                t["idx"] = np.arange(1, n + 1)
                t["v"] = L.volume
                t["wd"] = L.wall_dist
                t["wn"] = L.wall_node + 1
                t["seed"] = L.seed + 1
                W.table(_CNODE, t)
                W.rec(0, 0)
                W.rec(0)
                W.rec(0)
                nf = len(L.f2c)
                W.rec(nf)
                W.rec(n)
                p = np.zeros(nf, _PROL)
                p["f"] = np.arange(1, nf + 1)
                p["c"] = L.f2c + 1
                W.table(_PROL, p)
                W.rec(0)
                W.rec(nf)
                W.rec(n)
                r = np.zeros(nf, _REST)
                r["f"] = np.arange(1, nf + 1)
                r["w"] = L.rweight
                W.table(_REST, r)
                W.rec(0)
                W.rec(0)  # com send nodes (total), then per domain
                W.rec(0)
                W.rec(0)  # com receive nodes
                W.rec(0)


def read_sol(path) -> list[Level]:
    """Read a single-domain <case>.sol_1 into 0-based mesh.Level records.

    Coarse-level coordinates are not stored in the file (the reference never
    sets them, DEF:14002-14011); they are returned as zeros."""
    with open(path, "rb") as f:
        R = _Reader(f.read())
    ngrid = R.ints()
    dom, ndom = R.ints(2)
    if ndom != 1:
        raise NotImplementedError(f"{ndom}-domain grid files: only single-domain files are read")
    levels = []
    for li in range(ngrid):
        nbs = R.ints()
        b = R.table(_BSIDE, nbs)
        nb, nnorm = R.ints(2)
        reg = np.array([R.ints() for _ in range(21)], np.int32)
        nr = R.table(np.dtype([("node", _I4)]), nnorm)
        neng = R.ints()
        if neng:
            raise NotImplementedError("engine inlet sides")
        ne = R.ints()
        s = R.table(_SIDE, ne)
        n = R.ints()
        if li == 0:
            t = R.table(_FNODE, n)
            order = t["idx"] - 1
            vol = np.empty(n); vol[order] = t["v"][:, 0]
            wd = np.empty(n); wd[order] = t["wd"]
            wn = np.empty(n, np.int64); wn[order] = t["wn"] - 1
            xyz = np.empty((3, n)); xyz[:, order] = t["x"][:, :3].T
            seed = np.arange(n, dtype=np.int64)
            ntrip = R.ints()
            if ntrip:
                raise NotImplementedError("trip lines")
            R.ints()
            R.ints(2); R.ints(); R.ints()
            f2c = rw = None
        else:
            t = R.table(_CNODE, n)
            order = t["idx"] - 1
            vol = np.empty(n); vol[order] = t["v"]
            wd = np.empty(n); wd[order] = t["wd"]
            wn = np.empty(n, np.int64); wn[order] = t["wn"] - 1
            seed = (t["seed"] - 1).astype(np.int64)
            xyz = np.zeros((3, n))
            R.ints(2); R.ints(); R.ints()
            nf = R.ints(); R.ints()
            p = R.table(_PROL, nf)
            ncom = R.ints()
            if ncom:
                raise NotImplementedError("multi-domain prolongation")
            f2c = np.empty(nf, np.int64); f2c[p["f"] - 1] = p["c"] - 1
            nf2 = R.ints(); R.ints()
            r = R.table(_REST, nf2)
            R.ints()
            rw = np.empty(nf2); rw[r["f"] - 1] = r["w"]
            R.ints(); R.ints(); R.ints(); R.ints()
        nia = (nr["node"] - 1).astype(np.int64)
        node_ind = np.zeros(n, np.int32)
        levels.append(Level(
            n_nodes=n, n_bnodes=nb,
            side_index=(s["i"].T - 1).astype(np.int64).copy(),
            side_w=s["w"].T.copy(),
            bside_index=np.vstack([b["i"][:, :2].T - 1, b["i"][:, 2:].T]).astype(np.int64),
            bside_w=b["w"].T.copy(),
            volume=vol, coords=xyz, wall_dist=wd, wall_node=wn, reg=reg,
            node_ind_array=nia, node_ind=node_ind, seed=seed, f2c=f2c, rweight=rw))
    if R.o != len(R.b):
        raise ValueError(f"{len(R.b) - R.o} trailing bytes after the last level")
    return levels


def write_plotreg(path, n_nodes, domains):
    """The preprocessor's plot register (gen3d_grid.f90:4293-4294,
    4807-4815): (N, numberOfDomains), then per domain its node count and one
    (local, global) record per node, 1-based.  `domains` lists each domain's
    global node numbers in local order."""
    with open(path, "wb") as f:
        W = _Writer(f)
        W.rec(n_nodes, len(domains))
        for g in domains:
            g = np.asarray(g, np.int64)
            W.rec(len(g))
            t = np.empty(len(g), _TRANS)
            t["local"] = np.arange(1, len(g) + 1)
            t["global"] = g
            W.table(_TRANS, t)


def makeplot(plotreg, result_paths, out_path, nvar):
    """makeplot (mgnsg3d_mg_par.F90:1599-1707): the per-domain result files
    scattered into one N x nvar field through the plot register, written as
    <case>.unk: a record N, then one record u(N, nvar) column-major."""
    with open(plotreg, "rb") as f:
        R = _Reader(f.read())
    n, nd = (int(x) for x in R.ints(2))
    if nd != len(result_paths):
        raise ValueError(f"plot register has {nd} domains, {len(result_paths)} result files given")
    u = np.zeros((nvar, n))
    for d in range(nd):
        nl = R.ints()
        t = R.table(_TRANS, nl)
        trans = np.stack([t["local"], t["global"]], axis=1).astype(np.int64)
        with open(result_paths[d], "rb") as f:
            Q = _Reader(f.read())
        nt = Q.ints()
        unk = np.frombuffer(Q._payload(), _F8).reshape(nvar, nt)
        u[:, trans[:, 1] - 1] = unk[:, trans[:, 0] - 1]
    with open(out_path, "wb") as f:
        W = _Writer(f)
        W.rec(n)
        W.raw(u.astype(_F8).tobytes())
    return u


def write_results(path, u, turbulence_model):
    """writeGridResults (DEF:15331-15345): N, then rho, (u_j / rho, j=2..nvar)."""
    u = np.asarray(u, np.float64)
    nvar = 6 if turbulence_model == 1 else 5
    n = u.shape[1]
    rho = u[0]
    body = np.concatenate([rho] + [u[j] / rho for j in range(1, nvar)])
    with open(path, "wb") as f:
        W = _Writer(f)
        W.rec(n)
        W.raw(body.astype(_F8).tobytes())


def read_results(path, turbulence_model):
    """A .res_1 / .rst_1 file back into conserved u[nvar, N] (DEF:3744-3767,
    3853-3858: components 2..nvar are multiplied by rho)."""
    with open(path, "rb") as f:
        R = _Reader(f.read())
    n = R.ints()
    nvar = 6 if turbulence_model == 1 else 5
    body = np.frombuffer(R._payload(), _F8).reshape(nvar, n).copy()
    body[1:] *= body[0]
    return body
