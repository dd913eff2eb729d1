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
