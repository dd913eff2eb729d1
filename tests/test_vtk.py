"""VTK output (writemesh, vtk.jl:1-82): sub-cell layout, point order and
field values read back from the written .vtu file (no GPU)."""
import numpy as np
import pytest

import canary_amd.problems as P
from canary_amd.vtk import VTK_LINE, VTK_PIXEL, VTK_VOXEL, read_vtu, write_state, writemesh


@pytest.mark.parametrize("dim,N", [(2, 3), (3, 2)])
def test_state_roundtrip_and_cells(tmp_path, dim, N):
    s = P.bubble3d(Ne=2, N=N) if dim == 3 else P.bubble2d(Ne=3, N=N)
    fn = write_state(str(tmp_path / "step"), s, s.Q)
    pts, conn, types, fields = read_vtu(fn)
    Nq = N + 1
    assert pts.shape == (s.nelem * Nq ** dim, 3)
    assert conn.shape == (s.nelem * N ** dim, 2 ** dim)
    assert set(types) == {VTK_VOXEL if dim == 3 else VTK_PIXEL}
    ix = 11 if dim == 3 else 6
    for c in range(dim):
        assert np.array_equal(pts[:, c], s.vgeo[:, ix + c, :].reshape(-1, order="F"))
    rho = s.Q[:, dim, :]
    assert np.array_equal(fields["ρ"], rho.reshape(-1, order="F"))
    assert np.allclose(fields["E"], (s.Q[:, dim + 1, :] / rho - 300.0).reshape(-1, order="F"),
                       rtol=0, atol=0)
    # every sub-cell is an axis-aligned box with its corners in VTK pixel /
    # voxel order: x fastest, then y, then z
    P0 = pts[conn]
    for c in range(dim):
        bit = 1 << c
        hi = [v for v in range(2 ** dim) if v & bit]
        lo = [v for v in range(2 ** dim) if not v & bit]
        assert np.all(P0[:, hi, c] > P0[:, lo, c])
    # the first sub-cell of element 1 uses nodes (0,0[,0]), (1,0), (0,1), (1,1) ...
    want = [0, 1, Nq, Nq + 1] if dim == 2 else \
        [0, 1, Nq, Nq + 1, Nq * Nq, Nq * Nq + 1, Nq * Nq + Nq, Nq * Nq + Nq + 1]
    assert list(conn[0]) == want


def test_line_cells_and_realelems(tmp_path):
    x = np.array([[0.0, 0.5, 1.0], [1.0, 1.5, 2.0], [2.0, 2.5, 3.0]]).T  # (Nq=3, nelem=3)
    fn = writemesh(str(tmp_path / "line"), x, fields=(("f", x ** 2),), realelems=[1, 3])
    pts, conn, types, fields = read_vtu(fn)
    assert set(types) == {VTK_LINE}
    assert conn.tolist() == [[0, 1], [1, 2], [6, 7], [7, 8]]
    assert np.array_equal(fields["f"], (x ** 2).reshape(-1, order="F"))


def test_rejects_mismatched_field(tmp_path):
    x = np.zeros((3, 3, 2))
    with pytest.raises(ValueError):
        writemesh(str(tmp_path / "bad"), x, x, fields=(("f", np.zeros(5)),))


def _reference_cells(Nq, dim, realelems):
    """Sub-cell connectivity by the loops of vtk.jl:29-79 as written there:
    1-based element, node and cell numbers, the cell of (i, j[, k]) at
    i + (j-1) Nr [+ (k-1) Nr Ns] + (e-1) Nsubcells, its corners
    offset .+ ind[i:i+1, j:j+1(, k:k+1)][:] (column-major), then shifted to
    the file's 0-based point numbers."""
    Nr = Nq - 1
    nsub = Nr ** dim
    cells = {}

    def ind(*c):  # LinearIndices((1:Nq, ...)) of a 1-based node (i, j, k)
        return sum((c[a] - 1) * Nq ** a for a in range(dim)) + 1

    for e in realelems:
        offset = (e - 1) * Nq ** dim
        rng = range(1, Nq)
        if dim == 2:
            for j in rng:
                for i in rng:
                    corners = [ind(ii, jj) for jj in (j, j + 1) for ii in (i, i + 1)]
                    cells[i + (j - 1) * Nr + (e - 1) * nsub] = [offset + c for c in corners]
        else:
            for k in rng:
                for j in rng:
                    for i in rng:
                        corners = [ind(ii, jj, kk) for kk in (k, k + 1) for jj in (j, j + 1)
                                   for ii in (i, i + 1)]
                        cells[i + (j - 1) * Nr + (k - 1) * Nr * Nr + (e - 1) * nsub] = \
                            [offset + c for c in corners]
    # the reference allocates Nsubcells * length(realelems) cells and fills
    # them by that index, so for a contiguous 1:nreal the order is the index
    return [[c - 1 for c in cells[key]] for key in sorted(cells)]


@pytest.mark.parametrize("dim,N", [(2, 1), (2, 4), (3, 1), (3, 3)])
def test_cells_follow_the_reference_loops(tmp_path, dim, N):
    s = P.bubble3d(Ne=2, N=N) if dim == 3 else P.bubble2d(Ne=2, N=N)
    nreal = s.nelem - 1  # the last element's points are written, its cells not
    fn = write_state(str(tmp_path / "cells"), s, s.Q, realelems=range(1, nreal + 1))
    pts, conn, types, _ = read_vtu(fn)
    assert pts.shape[0] == s.nelem * (N + 1) ** dim
    assert conn.tolist() == _reference_cells(N + 1, dim, range(1, nreal + 1))


def test_shallow_water_fields(tmp_path):
    """shallow_water.jl:730-741: h, b, U/(h+b), V/(h+b)."""
    from canary_amd.vtk import plot_name
    s = P.swe2d(Ne=3, N=3)
    fn = write_state(str(tmp_path / "swe"), s, s.Q)
    pts, conn, types, fields = read_vtu(fn)
    assert sorted(fields) == ["U", "V", "b", "h"]
    H = s.Q[:, 2, :] + s.Q[:, 3, :]
    assert np.array_equal(fields["h"], s.Q[:, 2, :].reshape(-1, order="F"))
    assert np.array_equal(fields["U"], (s.Q[:, 0, :] / H).reshape(-1, order="F"))
    assert set(types) == {VTK_PIXEL}
    assert plot_name("swe", 2, "Array", 0, 7) == "viz/swe2D_Array_rank_0000_step_00007"
    assert plot_name("euler", 3, "CuArray", 3, 120) == \
        "viz/euler3D_set2c_CuArray_rank_0003_step_00120"
