"""Host mesh preparation against the reference's own fixtures
(test/test_mesh.jl), transcribed value for value.  Julia A[i,j] is A[i-1,j-1]
here (column-major arrays, 1-based values)."""
import numpy as np
import pytest

import canary_amd.mesh as M
from canary_amd._lib import CanaryError


def test_linearpartition():  # test_mesh.jl:1-6
    assert M.linearpartition(1, 1, 1) == range(1, 2)
    assert M.linearpartition(20, 1, 1) == range(1, 21)
    assert M.linearpartition(10, 1, 2) == range(1, 6)
    assert M.linearpartition(10, 2, 2) == range(6, 11)


@pytest.mark.parametrize("x,bits,h", [
    ([0, 0], 1, [0, 0]), ([0, 1], 1, [0, 1]), ([1, 1], 1, [1, 0]), ([1, 0], 1, [1, 1]),
    ([0, 0], 2, [0, 0]), ([1, 0], 2, [0, 1]), ([1, 1], 2, [0, 2]), ([0, 1], 2, [0, 3]),
    ([0, 2], 2, [1, 0]), ([0, 3], 2, [1, 1]), ([1, 3], 2, [1, 2]), ([1, 2], 2, [1, 3]),
    ([2, 2], 2, [2, 0]), ([2, 3], 2, [2, 1]), ([3, 3], 2, [2, 2]), ([3, 2], 2, [2, 3]),
    ([3, 1], 2, [3, 0]), ([2, 1], 2, [3, 1]), ([2, 0], 2, [3, 2]), ([3, 0], 2, [3, 3]),
    ([14, 3, 4], 64, [0x0, 0x0, 0xe25]),
])
def test_hilbertcode(x, bits, h):  # test_mesh.jl:8-31
    assert list(M.hilbertcode(x, bits=bits)) == h


def test_hilbert_is_a_curve():
    """Consecutive codes of a 2^b x 2^b grid are unit steps apart (the defining
    property of a Hilbert curve), and the map is a bijection."""
    b = 4
    pts = {}
    for i in range(2 ** b):
        for j in range(2 ** b):
            H = M.hilbertcode([i, j], bits=b)
            pts[int(H[0]) * 2 ** b + int(H[1])] = (i, j)
    assert sorted(pts) == list(range(4 ** b))
    for k in range(4 ** b - 1):
        (a, c), (d, e) = pts[k], pts[k + 1]
        assert abs(a - d) + abs(c - e) == 1


def test_centroidtocode():  # test_mesh.jl:33-60
    etc = np.zeros((2, 4, 6), order="F")
    etc[:, :, 0] = [[2, 3, 2, 3], [4, 4, 5, 5]]
    etc[:, :, 1] = [[3, 4, 3, 4], [4, 4, 5, 5]]
    etc[:, :, 2] = [[4, 5, 4, 5], [4, 4, 5, 5]]
    etc[:, :, 3] = [[2, 3, 2, 3], [5, 5, 6, 6]]
    etc[:, :, 4] = [[3, 4, 3, 4], [5, 5, 6, 6]]
    etc[:, :, 5] = [[4, 5, 4, 5], [5, 5, 6, 6]]
    expect = np.array([[0x0000000000000000, 0x1555555555555555, 0xffffffffffffffff,
                        0x5555555555555555, 0x6aaaaaaaaaaaaaaa, 0xaaaaaaaaaaaaaaaa],
                       [0x0000000000000000, 0x5555555555555555, 0xffffffffffffffff,
                        0x5555555555555555, 0xaaaaaaaaaaaaaaaa, 0xaaaaaaaaaaaaaaaa]],
                      dtype=np.uint64)
    assert np.array_equal(M.centroidtocode(etc), expect)
    one = np.zeros((2, 4, 1), order="F")
    one[:, :, 0] = [[2, 3, 2, 3], [4, 4, 5, 5]]
    assert np.array_equal(M.centroidtocode(one), np.zeros((2, 1), np.uint64))


def test_vertsortandorder():  # test_mesh.jl:62-92
    assert M.vertsortandorder(1) == ((1,), 1)
    assert M.vertsortandorder(1, 2) == ((1, 2), 1)
    assert M.vertsortandorder(2, 1) == ((1, 2), 2)
    for v, o in [((1, 2, 3), 1), ((3, 1, 2), 2), ((2, 3, 1), 3), ((2, 1, 3), 4), ((3, 2, 1), 5),
                 ((1, 3, 2), 6)]:
        assert M.vertsortandorder(*v) == ((1, 2, 3), o)
    with pytest.raises(CanaryError):
        M.vertsortandorder(2, 1, 1)
    for v, o in [((1, 2, 3, 4), 1), ((1, 3, 2, 4), 2), ((2, 1, 3, 4), 3), ((2, 4, 1, 3), 4),
                 ((3, 1, 4, 2), 5), ((3, 4, 1, 2), 6), ((4, 2, 3, 1), 7), ((4, 3, 2, 1), 8)]:
        assert M.vertsortandorder(*v) == ((1, 2, 3, 4), o)
    with pytest.raises(CanaryError):
        M.vertsortandorder(1, 3, 3, 1)


def r(a, b, step=1):
    return np.arange(a, b + step / 2, step, dtype=np.float64)


def test_brickmesh_1d():  # test_mesh.jl:94-127
    etv, etc, etb, fc = M.brickmesh((r(4, 7),), (False,))
    assert np.array_equal(etv, [[1, 2, 3], [2, 3, 4]])
    assert np.array_equal(etb, [[1, 0, 0], [0, 0, 1]])
    assert fc == []
    assert np.array_equal(etc[:, :, 0], [[4, 5]]) and np.array_equal(etc[:, :, 2], [[6, 7]])
    etv, etc, etb, fc = M.brickmesh((r(4, 7),), (True,))
    assert np.array_equal(etb, np.zeros((2, 3)))
    assert fc == [[3, 2, 1]]


def test_brickmesh_2d():  # test_mesh.jl:129-152
    etv, etc, etb, fc = M.brickmesh((r(2, 5), r(4, 6)), (False, True))
    expect = np.array([[1, 2, 5, 6], [2, 3, 6, 7], [3, 4, 7, 8], [5, 6, 9, 10], [6, 7, 10, 11],
                       [7, 8, 11, 12]]).T
    assert np.array_equal(etv, expect)
    assert np.array_equal(etb, [[1, 0, 0, 1, 0, 0], [0, 0, 1, 0, 0, 1], [0] * 6, [0] * 6])
    assert fc == [[4, 4, 1, 2], [5, 4, 2, 3], [6, 4, 3, 4]]
    assert np.array_equal(etc[:, :, 0], [[2, 3, 2, 3], [4, 4, 5, 5]])
    assert np.array_equal(etc[:, :, 4], [[3, 4, 3, 4], [5, 5, 6, 6]])


def test_brickmesh_3d():  # test_mesh.jl:154-190
    etv, etc, etb, fc = M.brickmesh((r(-1, 1, 2), r(-1, 1, 2), r(-1, 1, 1)), (True, True, True))
    assert np.array_equal(etv, np.array([[1, 5], [2, 6], [3, 7], [4, 8], [5, 9], [6, 10],
                                         [7, 11], [8, 12]]))
    assert np.array_equal(etb, np.zeros((6, 2)))
    assert fc == [[1, 2, 1, 3, 5, 7], [1, 4, 1, 2, 5, 6], [2, 2, 5, 7, 9, 11],
                  [2, 4, 5, 6, 9, 10], [2, 6, 1, 2, 3, 4]]
    assert np.array_equal(etc[:, :, 0], [[-1, 1, -1, 1, -1, 1, -1, 1], [-1, -1, 1, 1, -1, -1, 1, 1],
                                         [-1, -1, -1, -1, 0, 0, 0, 0]])
    assert np.array_equal(etc[:, :, 1], [[-1, 1, -1, 1, -1, 1, -1, 1], [-1, -1, 1, 1, -1, -1, 1, 1],
                                         [0, 0, 0, 0, 1, 1, 1, 1]])


def test_brickmesh_boundary_numbers():  # test_mesh.jl:192-204
    _, _, etb, _ = M.brickmesh((r(-1, 1), r(-1, 1), r(-1, 1)), (False, False, False),
                               boundary=[[11, 13, 15], [12, 14, 16]])
    assert np.array_equal(etb, [[11, 0, 11, 0, 11, 0, 11, 0], [0, 12, 0, 12, 0, 12, 0, 12],
                                [13, 13, 0, 0, 13, 13, 0, 0], [0, 0, 14, 14, 0, 0, 14, 14],
                                [15, 15, 15, 15, 0, 0, 0, 0], [0, 0, 0, 0, 16, 16, 16, 16]])


@pytest.mark.parametrize("x,p,b", [((r(1, 1000),), (False,), [1, 2]),
                                   ((r(-1, 10, 2), r(-1, 1, 1), r(-4, 1, 1)), (True, False, True),
                                    [[1, 3, 5], [2, 4, 6]])])
def test_brickmesh_parts(x, p, b):  # test_mesh.jl:206-245
    etv, etc, etb, _ = M.brickmesh(x, p, boundary=np.array(b).reshape(2, -1))
    parts = [M.brickmesh(x, p, part=j, numparts=50, boundary=np.array(b).reshape(2, -1))
             for j in range(1, 51)]
    assert np.array_equal(etv, np.concatenate([q[0] for q in parts], axis=1))
    assert np.array_equal(etc, np.concatenate([q[1] for q in parts], axis=2))
    assert np.array_equal(etb, np.concatenate([q[2] for q in parts], axis=1))


def test_connect_1d_periodic():  # test_mesh.jl:248-285
    m = M.connectmesh(*M.partition(*M.brickmesh((r(0, 10),), (True,))))
    for e in range(10):
        assert np.array_equal(m["elemtocoord"][:, :, e], [[e, e + 1]])
    assert np.array_equal(m["elemtoelem"], [[10, 1, 2, 3, 4, 5, 6, 7, 8, 9],
                                            [2, 3, 4, 5, 6, 7, 8, 9, 10, 1]])
    assert np.array_equal(m["elemtoface"], [[2] * 10, [1] * 10])
    assert np.all(m["elemtoordr"] == 1) and np.all(m["elemtobndy"] == 0)
    assert m["realelems"] == range(1, 11) and len(m["ghostelems"]) == 0
    assert len(m["sendelems"]) == 0 and m["nabrtorank"] == []


def test_connect_2d_hilbert_order():  # test_mesh.jl:287-334
    m = M.connectmesh(*M.partition(*M.brickmesh((r(0, 4), r(5, 9)), (False, True))))
    etc = [[[0, 1, 0, 1], [5, 5, 6, 6]], [[1, 2, 1, 2], [5, 5, 6, 6]], [[1, 2, 1, 2], [6, 6, 7, 7]],
           [[0, 1, 0, 1], [6, 6, 7, 7]], [[0, 1, 0, 1], [7, 7, 8, 8]], [[0, 1, 0, 1], [8, 8, 9, 9]],
           [[1, 2, 1, 2], [8, 8, 9, 9]], [[1, 2, 1, 2], [7, 7, 8, 8]], [[2, 3, 2, 3], [7, 7, 8, 8]],
           [[2, 3, 2, 3], [8, 8, 9, 9]], [[3, 4, 3, 4], [8, 8, 9, 9]], [[3, 4, 3, 4], [7, 7, 8, 8]],
           [[3, 4, 3, 4], [6, 6, 7, 7]], [[2, 3, 2, 3], [6, 6, 7, 7]], [[2, 3, 2, 3], [5, 5, 6, 6]],
           [[3, 4, 3, 4], [5, 5, 6, 6]]]
    for e in range(16):
        assert np.array_equal(m["elemtocoord"][:, :, e], etc[e]), e
    assert np.array_equal(m["elemtoelem"], [
        [1, 1, 4, 4, 5, 6, 6, 5, 8, 7, 10, 9, 14, 3, 2, 15],
        [2, 15, 14, 3, 8, 7, 10, 9, 12, 11, 11, 12, 13, 13, 16, 16],
        [6, 7, 2, 1, 4, 5, 8, 3, 14, 9, 12, 13, 16, 15, 10, 11],
        [4, 3, 8, 5, 6, 1, 2, 7, 10, 15, 16, 11, 12, 9, 14, 13]])
    assert np.array_equal(m["elemtoface"], [
        [1, 2, 2, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2],
        [1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 1, 1, 2],
        [4] * 16, [3] * 16])
    assert np.all(m["elemtoordr"] == 1)


def test_mappings_1d():  # test_mesh.jl:337-356
    m = M.connectmesh(*M.partition(*M.brickmesh((r(0, 4),), (True,))))
    vM, vP = M.mappings(3, m["elemtoelem"], m["elemtoface"], m["elemtoordr"])
    assert np.array_equal(vM.reshape(-1, order="F"), [1, 4, 5, 8, 9, 12, 13, 16])
    assert np.array_equal(vP.reshape(-1, order="F"), [16, 5, 4, 9, 8, 13, 12, 1])


def test_mappings_2d():  # test_mesh.jl:358-396
    m = M.connectmesh(*M.partition(*M.brickmesh((r(-1, 1), r(0, 1)), (False, True))))
    vM, vP = M.mappings(2, m["elemtoelem"], m["elemtoface"], m["elemtoordr"])
    assert np.array_equal(vM.reshape(-1, order="F"), [1, 4, 7, 3, 6, 9, 1, 2, 3, 7, 8, 9, 10, 13,
                                                       16, 12, 15, 18, 10, 11, 12, 16, 17, 18])
    assert np.array_equal(vP.reshape(-1, order="F"), [1, 4, 7, 10, 13, 16, 7, 8, 9, 1, 2, 3, 3, 6,
                                                       9, 12, 15, 18, 16, 17, 18, 10, 11, 12])


def test_mappings_3d():  # test_mesh.jl:398-436
    m = M.connectmesh(*M.partition(*M.brickmesh((r(0, 1), r(0, 1), r(-1, 1)),
                                                (False, True, False))))
    vM, vP = M.mappings(2, m["elemtoelem"], m["elemtoface"], m["elemtoordr"])
    fmask = np.array([[1, 3, 1, 7, 1, 19], [4, 6, 2, 8, 2, 20], [7, 9, 3, 9, 3, 21],
                      [10, 12, 10, 16, 4, 22], [13, 15, 11, 17, 5, 23], [16, 18, 12, 18, 6, 24],
                      [19, 21, 19, 25, 7, 25], [22, 24, 20, 26, 8, 26], [25, 27, 21, 27, 9, 27]])
    Np = 27
    fm = fmask.reshape(-1, order="F")
    assert np.array_equal(vM.reshape(-1, order="F"), np.concatenate([fm, fm + Np]))
    c = [fmask[:, 0], fmask[:, 1], fmask[:, 3], fmask[:, 2], fmask[:, 4], fmask[:, 4] + Np,
         fmask[:, 0] + Np, fmask[:, 1] + Np, fmask[:, 3] + Np, fmask[:, 2] + Np, fmask[:, 5],
         fmask[:, 5] + Np]
    assert np.array_equal(vP.reshape(-1, order="F"), np.concatenate(c))


def test_partition_identity():  # test_mesh.jl:457-464
    mesh = M.brickmesh((r(-1, 1, 2), r(-1, 1, 2), r(-2, 2, 1)), (True, True, True))
    out = M.partition(*mesh)
    for a, b in zip(mesh[:3], out[:3]):
        assert np.array_equal(a, b)
    assert mesh[3] == out[3]


@pytest.mark.parametrize("Ne,per", [((6, 5, 4), (False, False, False)),
                                    ((5, 1, 3), (True, True, False)),
                                    ((8, 8), (True, False))])
def test_connectivity_invariants(Ne, per):
    """Properties every conforming brick connectivity has: vmapP is an
    involution on connected faces, matched nodes coincide (up to the
    period), boundary faces map to themselves, and the Hilbert order is a
    permutation of the Cartesian elements."""
    L = [1.0 * n for n in Ne]
    x = tuple(np.linspace(0, L[k], Ne[k] + 1) for k in range(len(Ne)))
    d = len(Ne)
    m = M.connectmesh(*M.partition(*M.brickmesh(x, per)))
    N = 3
    vM, vP = M.mappings(N, m["elemtoelem"], m["elemtoface"], m["elemtoordr"])
    Xg = M.creategrid(m["elemtocoord"], M.lglpoints(N)[0])
    flat = [g.reshape(-1, order="F") for g in Xg]
    vMf, vPf = vM.reshape(-1, order="F") - 1, vP.reshape(-1, order="F") - 1
    Nfp, nface = (N + 1) ** (d - 1), 2 * d
    nelem = m["elemtoelem"].shape[1]
    n_i, f_i, e_i = np.meshgrid(np.arange(Nfp), np.arange(nface), np.arange(nelem), indexing="ij")
    e2 = m["elemtoelem"][f_i, e_i] - 1
    f2 = m["elemtoface"][f_i, e_i] - 1
    partner = (n_i + Nfp * (f2 + nface * e2)).reshape(-1, order="F")
    bnd = np.repeat(m["elemtobndy"].reshape(-1, order="F"), (N + 1) ** (d - 1))
    conn = bnd == 0
    for c in range(d):
        dx = flat[c][vMf] - flat[c][vPf]
        if per[c]:
            dx = np.where(np.isclose(np.abs(dx), L[c]), 0.0, dx)
        assert np.allclose(dx[conn], 0.0, atol=1e-12)
    # involution: the partner face node sees this node as its P side
    assert np.array_equal(vMf[partner], vPf)
    assert np.array_equal(vPf[partner], vMf)
    assert np.array_equal(vPf[~conn], vMf[~conn])
    cen = m["elemtocoord"].mean(axis=1)
    assert len({tuple(np.round(c, 9)) for c in cen.T}) == cen.shape[1]


def _fmask(dim, N):
    """Face mask the way src/mesh.jl:978-982 states it: the 1-based node
    numbers reshaped to (Nq,)*dim, dimension f//2 fixed at 0 or N, the rest
    flattened column-major."""
    Nq = N + 1
    p = np.arange(1, Nq ** dim + 1).reshape((Nq,) * dim, order="F")
    cols = []
    for f in range(2 * dim):
        sl = [slice(None)] * dim
        sl[f // 2] = (f % 2) * N
        cols.append(p[tuple(sl)].ravel(order="F"))
    return np.stack(cols, axis=1)  # (Nfp, nface)


@pytest.mark.parametrize("dim,N", [(2, 1), (2, 4), (2, 7), (3, 1), (3, 3), (3, 4), (3, 7)])
def test_vmapM_is_always_the_canonical_face_mask(dim, N):
    """The contract behind a null vmapM: every vmapM the mesh code builds
    (walls, periodic, warped, a rank share with ghosts) is
    Np*(e-1) + fmask[:, f]."""
    from canary_amd import problems as P
    from states import small_setup, warped_setup
    fm = _fmask(dim, N)
    per = (True,) + (False,) * (dim - 1)
    shared = P.rank_share(small_setup(0, dim, N, Ne=4, periodic=per), 3, 1)
    for vM, Np, ne in [(s.vmapM, s.Np, s.nelem) for s in
                       (small_setup(0, dim, N, Ne=3), small_setup(0, dim, N, Ne=3, periodic=per),
                        warped_setup(0, dim, N, Ne=3))] + [(shared.vmapM, (N + 1) ** dim,
                                                            shared.nreal)]:
        want = fm[:, :, None] + Np * np.arange(ne)[None, None, :]
        assert np.array_equal(np.asarray(vM).reshape(want.shape, order="F"), want)
