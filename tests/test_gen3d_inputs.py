"""Gen3d's input files: <case>.plt (unformatted) and <case>.bco (text).

- A .plt written by a Fortran program with Gen3d's own record statements
  (the READs of gen3d_generator.f90:382, gen3d_elements.f90:340,
  gen3d_struct.f90:75, gen3d_boundary.f90:1396-1398, turned into WRITEs) is
  read back exactly: a unit cube of six Kuhn tetrahedra, its twelve boundary
  triangles on six surface segments;
- a .bco with the reference's three-column surface lines, line segments and
  a rotating-object block is read as the reference reads it
  (gen3d_boundary.f90:1344-1377);
- the bump box written as .plt/.bco and read back builds the same levels as
  the arrays it came from, and the command-line driver writes the .sol_1
  the solver reads.
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from flite3d_amd import gen3d, mesh, solfile

FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")

FORTRAN = """
program writeplt
  implicit none
  integer :: ne, nn, nb, e, k, p, b
  integer, allocatable :: tet(:,:), f(:,:), el(:), seg(:)
  real(8), allocatable :: x(:,:)
  open(11, file='cube.txt', status='old')
  read(11,*) ne, nn, nb
  allocate(tet(ne,4), x(nn,3), f(nb,3), el(nb), seg(nb))
  do e = 1, ne
    read(11,*) (tet(e,k), k=1,4)
  end do
  do p = 1, nn
    read(11,*) (x(p,k), k=1,3)
  end do
  do b = 1, nb
    read(11,*) (f(b,k), k=1,3), el(b), seg(b)
  end do
  close(11)
  open(10, file='cube.plt', form='unformatted', status='replace')
  write(10) ne, nn, nb
  write(10) ((tet(e,k), e=1,ne), k=1,4)
  write(10) ((x(p,k), p=1,nn), k=1,3)
  write(10) ((f(b,k), b=1,nb), k=1,3), (el(b), b=1,nb), (seg(b), b=1,nb)
  close(10)
end program writeplt
"""


def cube():
    """Unit cube, nodes i + 2j + 4k, six Kuhn tets, twelve boundary
    triangles (right-hand normal into the cube), segments 1..6 by face."""
    xyz = np.array([[i, j, k] for k in (0, 1) for j in (0, 1) for i in (0, 1)], float)
    nid = lambda v: int(v[0] + 2 * v[1] + 4 * v[2])
    tets = []
    for p in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
        v = np.zeros(3, int)
        t = [nid(v)]
        for ax in p:
            v[ax] += 1
            t.append(nid(v))
        tets.append(t)
    tets = np.array(tets)
    tris, segs = [], []
    faces = [(2, 0, (0, 0, 1)), (0, 0, (1, 0, 0)), (0, 1, (-1, 0, 0)), (1, 0, (0, 1, 0)),
             (1, 1, (0, -1, 0)), (2, 1, (0, 0, -1))]
    for s, (ax, val, inward) in enumerate(faces, start=1):
        free = [a for a in range(3) if a != ax]

        def node(a, b):
            v = [0, 0, 0]
            v[ax], v[free[0]], v[free[1]] = val, a, b
            return nid(v)
        for t in ([node(0, 0), node(1, 0), node(1, 1)], [node(0, 0), node(1, 1), node(0, 1)]):
            p0, p1, p2 = xyz[t]
            if np.cross(p1 - p0, p2 - p0) @ np.array(inward) < 0:
                t = [t[0], t[2], t[1]]
            tris.append(t)
            segs.append(s)
    tris = np.array(tris)
    elem = np.array([next(e for e, tt in enumerate(tets) if set(f) <= set(tt)) for f in tris])
    return xyz, tets, tris, elem, np.array(segs)


@pytest.mark.skipif(not os.path.exists(FLANG), reason="no flang")
def test_plt_written_by_fortran_is_read_exactly(tmp_path):
    xyz, tets, tris, elem, segs = cube()
    with open(tmp_path / "cube.txt", "w") as f:
        f.write(f"{len(tets)} {len(xyz)} {len(tris)}\n")
        for t in tets:
            f.write(" ".join(str(v + 1) for v in t) + "\n")
        for p in xyz:
            f.write(" ".join(repr(float(v)) for v in p) + "\n")
        for t, e, s in zip(tris, elem, segs):
            f.write(" ".join(str(v + 1) for v in t) + f" {e + 1} {s}\n")
    (tmp_path / "w.f90").write_text(FORTRAN)
    subprocess.run([FLANG, "-O0", "w.f90", "-o", "w"], cwd=tmp_path, check=True, capture_output=True)
    subprocess.run(["./w"], cwd=tmp_path, check=True, capture_output=True)
    m = gen3d.read_plt(tmp_path / "cube.plt")
    assert np.array_equal(m.tets, tets) and np.array_equal(m.xyz, xyz)
    assert np.array_equal(m.btri, tris) and np.array_equal(m.elem, elem) and np.array_equal(m.segment, segs)
    # and our writer produces the Fortran program's bytes
    gen3d.write_plt(tmp_path / "ours.plt", m)
    assert (tmp_path / "ours.plt").read_bytes() == (tmp_path / "cube.plt").read_bytes()


def test_bco_as_the_reference_reads_it(tmp_path):
    (tmp_path / "c.bco").write_text(
        "Boundary conditions for case c\n"
        "   3   2\n"
        "Surface   Indicator  WallBound\n"
        "   1   1   0\n"
        "   2   3   1\n"
        "   3   2   0\n"
        "Curve   Flag\n"
        "   1   0\n"
        "   2   1\n"
        "Rotating objects\n"
        "   1   0.5 0.5 0.0   0.0 0.0 1.0\n")
    b = gen3d.read_bco(tmp_path / "c.bco")
    assert b.surface == {1: (1, 0), 2: (3, 1), 3: (2, 0)}
    assert b.line == {1: 0, 2: 1}
    assert b.rotating == [[0.5, 0.5, 0.0, 0.0, 0.0, 1.0]]
    m = gen3d.PltMesh(np.zeros((3, 3)), np.zeros((0, 4), np.int32), np.array([[0, 1, 2]] * 3, np.int32),
                      np.zeros(3, np.int32), np.array([3, 1, 2], np.int32))
    assert list(gen3d.face_indicators(m, b)) == [2, 1, 3]
    m.segment[0] = 7
    with pytest.raises(ValueError):
        gen3d.face_indicators(m, b)


def test_bump_box_through_plt_and_bco_builds_the_same_levels(tmp_path):
    m, surface = gen3d.bump_case(5, jitter=0.1, seed=3)
    gen3d.write_plt(tmp_path / "bump.plt", m)
    gen3d.write_bco(tmp_path / "bump.bco", surface)
    lv, old = gen3d.levels_from_case(tmp_path / "bump.plt", tmp_path / "bump.bco", 2)
    xyz, tets, btri, bind = mesh.bump_box(5, jitter=0.1, seed=3)
    ref, old_ref = mesh.build_levels(xyz, tets, btri, bind, 2)
    assert np.array_equal(old, old_ref)
    for a, b in zip(lv, ref):
        for f in ("side_index", "side_w", "bside_index", "bside_w", "volume", "coords", "reg"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), f
    # every tet holding a face has the face's three nodes
    for f, e in zip(m.btri, m.elem):
        assert set(f) <= set(m.tets[e])
    # the command-line driver writes the solver's .sol_1 and the plot register
    r = subprocess.run([sys.executable, "-m", "flite3d_amd.gen3d", str(tmp_path / "bump"), "--levels", "2"],
                       capture_output=True, text=True, cwd=os.path.dirname(os.path.dirname(__file__)))
    assert r.returncode == 0, r.stderr
    back = solfile.read_sol(tmp_path / "bump.sol_1")
    assert len(back) == 2 and back[0].n_nodes == lv[0].n_nodes
    assert np.array_equal(back[0].side_index, lv[0].side_index)
    assert np.allclose(back[1].volume, lv[1].volume, rtol=0, atol=0)
    assert os.path.exists(tmp_path / "plotreg.reg")
