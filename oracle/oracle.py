"""ctypes front end of the CPU oracle (liboracle.so built from dg_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker / the timed CPU baseline.  It
works on a `Setup` (canary.jl_amd/problems.py) and NumPy arrays in the
reference's column-major layout.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


class OPhys(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("gamma", "p0", "R_gas", "c_p", "c_v", "gravity", "swe_gravity", "delta_nl")]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        V, I64, D, INT = C.c_void_p, C.c_int64, C.c_double, C.c_int
        PP = C.POINTER(OPhys)
        for name in ("oracle_euler3d_volumerhs", "oracle_euler2d_volumerhs", "oracle_swe_volumerhs",
                     "oracle_euler2dte_volumerhs"):
            getattr(L, name).argtypes = [INT, V, V, V, V, I64, PP, INT]
        L.oracle_euler2dte_fluxrhs.argtypes = [INT, V, V, V, V, I64, V, V, V, PP, INT]
        L.oracle_euler2dte_dt.argtypes = [INT, V, V, I64, PP]
        L.oracle_euler2dte_dt.restype = D
        for name in ("oracle_euler3d_facerhs", "oracle_euler2d_facerhs", "oracle_swe_fluxrhs"):
            getattr(L, name).argtypes = [INT, V, V, V, I64, V, V, V, PP, INT]
        L.oracle_updatesolution.argtypes = [INT, INT, INT, INT, V, V, V, I64, D, D, D, INT]
        L.oracle_swe_updatesolution.argtypes = [INT, V, V, V, I64, D, D, D, INT, INT]
        L.oracle_lsrk.argtypes = [INT, INT, INT, V, V, V, V, V, I64, V, V, V, D, INT, PP, INT, INT]
        L.oracle_l2energysquared.argtypes = [INT, INT, INT, INT, V, V, I64]
        L.oracle_l2energysquared.restype = D
        L.oracle_euler_cfl.argtypes = [INT, INT, V, V, I64, PP]
        L.oracle_euler_cfl.restype = D
        L.oracle_lsrk54_coefs.argtypes = [V, V, V]
        L.oracle_set_pow_mode.argtypes = [INT]
        L.oracle_get_pow_mode.restype = INT
        L.oracle_pow.argtypes = [D, D]
        L.oracle_pow.restype = D
        _lib = L
    return _lib


def set_pow_mode(mode):
    """1 (default): the pressure's pow correctly rounded (quad precision
    rounded once); 0: the C library's pow (CPU-baseline timing)."""
    lib().oracle_set_pow_mode(int(mode))


def pow_mode():
    return lib().oracle_get_pow_mode()


def opow(x, y):
    """The oracle's pow in its current mode."""
    return lib().oracle_pow(float(x), float(y))


def _p(a):
    assert a.flags["F_CONTIGUOUS"] or a.ndim == 1
    return a.ctypes.data_as(C.c_void_p)


def ophys(phys):
    return OPhys(*(getattr(phys, n) for n, _ in OPhys._fields_))


def _tag(s):
    return {(0, 3): "euler3d", (0, 2): "euler2d", (1, 2): "swe", (2, 2): "euler2dte"}[(s.eqn, s.dim)]


def volumerhs(s, rhs, Q, nthreads=1, nelem=None):
    n = s.nelem if nelem is None else nelem
    f = getattr(lib(), "oracle_%s_volumerhs" % _tag(s))
    return f(s.N, _p(rhs), _p(Q), _p(s.vgeo), _p(s.D), n, C.byref(ophys(s.phys)), nthreads)


def facerhs(s, rhs, Q, nthreads=1, nelem=None, elemtobndy=None, vmapP=None):
    n = s.nelem if nelem is None else nelem
    bnd = s.elemtobndy if elemtobndy is None else elemtobndy
    vP = s.vmapP if vmapP is None else vmapP
    if s.eqn == 2:  # total energy: fluxrhs! reads vgeo's y (euler2d_set3c.jl:282)
        return lib().oracle_euler2dte_fluxrhs(s.N, _p(rhs), _p(Q), _p(s.sgeo), _p(s.vgeo), n,
                                              _p(s.vmapM), _p(vP), _p(bnd),
                                              C.byref(ophys(s.phys)), nthreads)
    name = "oracle_swe_fluxrhs" if s.eqn == 1 else "oracle_%s_facerhs" % _tag(s)
    return getattr(lib(), name)(s.N, _p(rhs), _p(Q), _p(s.sgeo), n, _p(s.vmapM), _p(vP),
                                _p(bnd), C.byref(ophys(s.phys)), nthreads)


def updatesolution(s, rhs, Q, rka, rkb, dt, nthreads=1, advection=False):
    if s.eqn == 1:
        return lib().oracle_swe_updatesolution(s.N, _p(rhs), _p(Q), _p(s.vgeo), s.nelem,
                                               rka, rkb, dt, 1 if advection else 0, nthreads)
    imji = 10 if s.dim == 3 else 5
    return lib().oracle_updatesolution(s.Np, s.nstate, s.nvgeo, imji, _p(rhs), _p(Q),
                                       _p(s.vgeo), s.nelem, rka, rkb, dt, nthreads)


def lsrk(s, Q, rhs, dt, nsteps, nthreads=1, advection=False):
    """nsteps LSRK54 steps of the reference stage loop, in place on Q and rhs."""
    return lib().oracle_lsrk(s.eqn, s.dim, s.N, _p(Q), _p(rhs), _p(s.vgeo), _p(s.sgeo),
                             _p(s.D), s.nelem, _p(s.vmapM), _p(s.vmapP), _p(s.elemtobndy),
                             float(dt), int(nsteps), C.byref(ophys(s.phys)),
                             1 if advection else 0, nthreads)


def rhs_eval(s, Q, rhs0=None, nthreads=1):
    """One RHS evaluation: rhs0 + volume - face (the reference's two calls)."""
    rhs = np.zeros_like(Q, order="F") if rhs0 is None else np.array(rhs0, order="F", copy=True)
    volumerhs(s, rhs, Q, nthreads)
    rc = facerhs(s, rhs, Q, nthreads)
    if rc:
        raise RuntimeError("oracle facerhs error %d" % rc)
    return rhs


def lsrk54_coefs():
    a, b, c = (C.c_double * 5)(), (C.c_double * 5)(), (C.c_double * 5)()
    lib().oracle_lsrk54_coefs(a, b, c)
    return tuple(a), tuple(b), tuple(c)


def l2energysquared(s, Q):
    imj = 9 if s.dim == 3 else 4
    return lib().oracle_l2energysquared(s.Np, s.nstate, s.nvgeo, imj, _p(Q), _p(s.vgeo), s.nelem)


def euler2dte_dt(s, Q):
    """courantnumber's time step of the total-energy example."""
    return lib().oracle_euler2dte_dt(s.N, _p(s.vgeo), _p(Q), s.nelem, C.byref(ophys(s.phys)))


def euler_cfl(s, Q):
    return lib().oracle_euler_cfl(s.dim, s.N, _p(s.vgeo), _p(Q), s.nelem, C.byref(ophys(s.phys)))
