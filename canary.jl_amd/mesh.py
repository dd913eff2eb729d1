"""Host preparation with the reference's `Canary` module API (src/Canary.jl:10-14).

Arrays are NumPy arrays in Fortran (column-major) order with the reference's
shapes, so `A[i, j]` here is `A[i+1, j+1]` in Julia, and index-valued arrays
hold the reference's 1-based Int64 values.  All work is done by the C++
library (libcanary_dg.so).
"""
import ctypes as C
from collections import namedtuple

import numpy as np

from . import _lib
from ._lib import check, lib

neither, left, right, both = (_lib.ENDPT_NEITHER, _lib.ENDPT_LEFT, _lib.ENDPT_RIGHT,
                              _lib.ENDPT_BOTH)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _f(shape, dtype=np.float64):
    return np.zeros(shape, dtype=dtype, order="F")


def linearpartition(n, p, np_):
    """src/mesh.jl:13 -- 1-based inclusive range of part p of np_."""
    return range((p - 1) * n // np_ + 1, p * n // np_ + 1)


# ----------------------------------------------------------------- operators
def legendre(n, endpt=neither, maxits=30):
    """GaussQuadrature.legendre(Float64, n, endpt): quadrature.jl:126-214."""
    x, w = np.zeros(n), np.zeros(n)
    check(lib().canary_gauss_rule(int(n), int(endpt), int(maxits), _p(x), _p(w)), "legendre")
    return x, w


def lglpoints(N):
    """src/operators.jl:10-13"""
    r, w = np.zeros(N + 1), np.zeros(N + 1)
    check(lib().canary_lglpoints(int(N), _p(r), _p(w)), "lglpoints")
    return r, w


def lgpoints(N):
    """src/operators.jl:21-24"""
    r, w = np.zeros(N + 1), np.zeros(N + 1)
    check(lib().canary_lgpoints(int(N), _p(r), _p(w)), "lgpoints")
    return r, w


def baryweights(r):
    r = np.ascontiguousarray(r, dtype=np.float64)
    wb = np.zeros(len(r))
    check(lib().canary_baryweights(len(r), _p(r), _p(wb)), "baryweights")
    return wb


def spectralderivative(r):
    """src/operators.jl:84-104 -- D[j, k] = l_k'(r_j)."""
    r = np.ascontiguousarray(r, dtype=np.float64)
    D = _f((len(r), len(r)))
    check(lib().canary_spectralderivative(len(r), _p(r), _p(D)), "spectralderivative")
    return D


def interpolationmatrix(rsrc, rdst):
    rsrc = np.ascontiguousarray(rsrc, dtype=np.float64)
    rdst = np.ascontiguousarray(rdst, dtype=np.float64)
    I = _f((len(rdst), len(rsrc)))
    check(lib().canary_interpolationmatrix(len(rsrc), _p(rsrc), len(rdst), _p(rdst), _p(I)),
          "interpolationmatrix")
    return I


# --------------------------------------------------------------------- mesh
def hilbertcode(Y, bits=64):
    """src/mesh.jl:35-88"""
    Y = np.ascontiguousarray(np.asarray(Y, dtype=np.uint64))
    H = np.zeros(len(Y), dtype=np.uint64)
    check(lib().canary_hilbertcode(len(Y), _p(Y), int(bits), _p(H)), "hilbertcode")
    return H


def centroidtocode(elemtocoord):
    """src/mesh.jl:107-141 on one rank; elemtocoord (d, nvert, nelem)."""
    etc = np.asfortranarray(elemtocoord, dtype=np.float64)
    d, nvert, nelem = etc.shape
    code = _f((d, nelem), np.uint64)
    check(lib().canary_centroidtocode(d, nvert, nelem, _p(etc), _p(code)), "centroidtocode")
    return code


def vertsortandorder(*v):
    """src/mesh.jl:595-719 -> ((sorted...), o)"""
    a = np.array(v, dtype=np.int64)
    out = np.zeros(len(v), dtype=np.int64)
    o = C.c_int(0)
    rc = lib().canary_vertsortandorder(len(v), _p(a), _p(out), C.byref(o))
    check(rc, "vertsortandorder")
    return tuple(int(x) for x in out), o.value


class _Mesh:
    """RAII wrapper of the opaque C mesh."""

    def __init__(self, ptr):
        if not ptr:
            raise _lib.CanaryError(_lib.ERR_ARG, "mesh")
        self.ptr = ptr

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().canary_mesh_free(self.ptr)
            self.ptr = None

    def size(self, what):
        return lib().canary_mesh_size(self.ptr, what)

    def get(self, name, shape, dtype=np.int64):
        a = _f(shape, dtype)
        check(lib().canary_mesh_get(self.ptr, name.encode(), _p(a)), "mesh_get " + name)
        return a

    def arrays(self):
        d, ne = self.size(1), self.size(0)
        nvert, nface, nfv = 2 ** d, 2 * d, 2 ** (d - 1)
        etv = self.get("elemtovert", (nvert, ne))
        etc = self.get("elemtocoord", (d, nvert, ne), np.float64)
        etb = self.get("elemtobndy", (nface, ne))
        nfc = self.size(2)
        fca = self.get("faceconnections", (2 + nfv, nfc)) if nfc else np.zeros((2 + nfv, 0), np.int64)
        fcs = [list(int(v) for v in fca[:, k]) for k in range(nfc)]
        return etv, etc, etb, fcs

    @staticmethod
    def from_arrays(etv, etc, etb, fcs):
        etc = np.asfortranarray(etc, dtype=np.float64)
        etv = np.asfortranarray(etv, dtype=np.int64)
        etb = np.asfortranarray(etb, dtype=np.int64)
        d = etc.shape[0]
        nelem = etc.shape[2]
        w = 2 + 2 ** (d - 1)
        fca = np.asfortranarray(np.array(fcs, dtype=np.int64).reshape(len(fcs), w).T) \
            if len(fcs) else np.zeros((w, 0), np.int64, order="F")
        st = C.c_int(0)
        ptr = lib().canary_mesh_from_arrays(d, nelem, _p(etv), _p(etc), _p(etb), len(fcs),
                                            _p(fca), C.byref(st))
        check(st.value, "mesh_from_arrays")
        return _Mesh(ptr)


def brickmesh(x, periodic, part=1, numparts=1, boundary=None):
    """src/mesh.jl:259-317 -> (elemtovert, elemtocoord, elemtobndy, faceconnections)."""
    d = len(x)
    coords = np.concatenate([np.asarray(xi, dtype=np.float64) for xi in x])
    nco = np.array([len(xi) for xi in x], dtype=np.int64)
    per = np.array([1 if p else 0 for p in periodic], dtype=np.int32)
    bnd = None if boundary is None else np.asfortranarray(np.asarray(boundary, dtype=np.int64).reshape(2, d))
    st = C.c_int(0)
    ptr = lib().canary_brickmesh(d, _p(nco), _p(coords), _p(per), int(part), int(numparts),
                                 _p(bnd) if bnd is not None else None, C.byref(st))
    check(st.value, "brickmesh")
    return _Mesh(ptr).arrays()


def partition(elemtovert, elemtocoord, elemtobndy, faceconnections):
    """src/mesh.jl:499-580 on one rank: elements in Hilbert-curve order."""
    m = _Mesh.from_arrays(elemtovert, elemtocoord, elemtobndy, faceconnections)
    check(lib().canary_mesh_partition(m.ptr), "partition")
    return m.arrays()


def connectmesh(elemtovert, elemtocoord, elemtobndy, faceconnections):
    """src/mesh.jl:751-956 on one rank; returns the reference's named tuple as a dict."""
    m = _Mesh.from_arrays(elemtovert, elemtocoord, elemtobndy, faceconnections)
    check(lib().canary_mesh_connect(m.ptr), "connectmesh")
    d, ne = m.size(1), m.size(0)
    nface = 2 * d
    etv, etc, etb, _ = m.arrays()
    return dict(elems=range(1, ne + 1), realelems=range(1, ne + 1), ghostelems=range(ne + 1, ne + 1),
                sendelems=np.zeros(0, np.int64), elemtocoord=etc,
                elemtoelem=m.get("elemtoelem", (nface, ne)),
                elemtoface=m.get("elemtoface", (nface, ne)),
                elemtoordr=m.get("elemtoordr", (nface, ne)), elemtobndy=etb,
                nabrtorank=[], nabrtorecv=[], nabrtosend=[])


def mappings(N, elemtoelem, elemtoface, elemtoordr):
    """src/mesh.jl:973-1001 -> (vmapM, vmapP), each (Nfp, nface, nelem), 1-based."""
    e2e = np.asfortranarray(elemtoelem, dtype=np.int64)
    e2f = np.asfortranarray(elemtoface, dtype=np.int64)
    e2o = np.asfortranarray(elemtoordr, dtype=np.int64)
    nface, nelem = e2e.shape
    d = nface // 2
    Nfp = (N + 1) ** (d - 1)
    vM, vP = _f((Nfp, nface, nelem), np.int64), _f((Nfp, nface, nelem), np.int64)
    check(lib().canary_mappings(int(N), d, nelem, _p(e2e), _p(e2f), _p(e2o), _p(vM), _p(vP)),
          "mappings")
    return vM, vP


# ------------------------------------------------------------------- metric
def creategrid(elemtocoord, r):
    """creategrid!(x[, y[, z]], elemtocoord, r): src/metric.jl:11-107."""
    etc = np.asfortranarray(elemtocoord, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    d, _, nelem = etc.shape
    Nq = len(r)
    X = np.zeros(d * Nq ** d * nelem)
    check(lib().canary_creategrid(d, Nq, nelem, _p(etc), _p(r), _p(X)), "creategrid")
    shape = (Nq,) * d + (nelem,)
    return tuple(np.reshape(X[c * Nq ** d * nelem:(c + 1) * Nq ** d * nelem], shape, order="F")
                 for c in range(d))


Metric1 = namedtuple("Metric1", "J xix sJ nx")
Metric2 = namedtuple("Metric2", "J xix etax xiy etay sJ nx ny")
Metric3 = namedtuple("Metric3", "J xix etax zetax xiy etay zetay xiz etaz zetaz sJ nx ny nz")


def computemetric(X, D):
    """computemetric!(...): src/metric.jl:120-400.  X = (x[, y[, z]]) grids
    of shape (Nq,)*d + (nelem,).  Face arrays come back as (Nfp, nface, nelem)
    (1D: (1, 2, nelem))."""
    d = len(X)
    D = np.asfortranarray(D, dtype=np.float64)
    Nq = D.shape[0]
    nelem = X[0].shape[-1]
    Xc = np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1, order="F") for x in X])
    nn = Nq ** d * nelem
    nfp = Nq ** (d - 1)
    ns = nfp * 2 * d * nelem
    vol = np.zeros((1 + d * d) * nn)
    surf = np.zeros((1 + d) * ns)
    check(lib().canary_computemetric(d, Nq, nelem, _p(Xc), _p(D), _p(vol), _p(surf)),
          "computemetric")
    vshape = (Nq,) * d + (nelem,)
    sshape = (nfp, 2 * d, nelem)
    v = [np.reshape(vol[i * nn:(i + 1) * nn], vshape, order="F") for i in range(1 + d * d)]
    s = [np.reshape(surf[i * ns:(i + 1) * ns], sshape, order="F") for i in range(1 + d)]
    cls = {1: Metric1, 2: Metric2, 3: Metric3}[d]
    return cls(*(v + s))


def computegeometry(dim, N, elemtocoord, vmapM):
    """computegeometry (euler3d_set2c.jl:101-139 / euler2d_set2c.jl:97-134)
    -> (vgeo (Np, nvgeo, nelem), sgeo (nsgeo, Nfp, nface, nelem))."""
    etc = np.asfortranarray(elemtocoord, dtype=np.float64)
    vM = np.asfortranarray(vmapM, dtype=np.int64)
    nelem = etc.shape[2]
    Nq = N + 1
    Np, Nfp, nface = Nq ** dim, Nq ** (dim - 1), 2 * dim
    nvgeo, nsgeo = (14, 5) if dim == 3 else (8, 4)
    vgeo = _f((Np, nvgeo, nelem))
    sgeo = _f((nsgeo, Nfp, nface, nelem))
    check(lib().canary_computegeometry(dim, N, nelem, _p(etc), _p(vM), _p(vgeo), _p(sgeo)),
          "computegeometry")
    return vgeo, sgeo
