"""The contracted volume term (FP64_FMA) and the matrix-core form
(FP64_MFMA: the volume term's derivative contractions as FP64 16x16x4 MFMA
tiles, 3D Euler) against the oracle: volumerhs! on bricks with walls and
periodic directions and on warped elements for N = 1..7 (every tile shape the
form builds: 1 to 10 tiles per direction, 2 to 8 line products per tile);
the fused stage over the reference and the prepared arrays against the
oracle's stage; the library loop on its paths against the oracle's
lowstorageRK.

Tolerances are the contracted form's: 1e-12 relative L2 per volume term and
per stage, 1e-10 after 40 steps.  The matrix core sums the NQ products of a
line in its own order and rounding, so neither form is held to the exact
order's bits."""
import numpy as np
import pytest

import canary_amd.kernels as K
import canary_amd.problems as P
from canary_amd._lib import (ERR_UNSUPPORTED, EULER, FP64, FP64_FMA, FP64_MFMA,
                             PATH_FACE_UPDATE, PATH_FUSED)
from oracle import oracle as O

from gpu_util import DeviceSetup, dev, host
from states import rel_l2, small_setup, smooth_euler_state, warped_setup

pytestmark = pytest.mark.gpu


def volume(s, ds, Q, rhs0, prec):
    r = dev(rhs0)
    K.volumerhs(s.dim, s.N, r, dev(Q), ds.vgeo, ds.D, s.nelem, eqn=s.eqn, phys=s.phys, prec=prec)
    K.check_errors()
    return host(r, Q.shape)


def oracle_volume(s, Q, rhs0):
    want = np.array(rhs0, order="F", copy=True)
    O.volumerhs(s, want, np.asfortranarray(Q))
    return want


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6, 7])
def test_volume_term_against_the_oracle(cuda, N):
    print("mfma volumerhs 3D N=%d" % N, flush=True)
    s = small_setup(EULER, 3, N, periodic=(True, False, False))
    ds = DeviceSetup(s)
    Q = np.asfortranarray(smooth_euler_state(s, seed=N))
    zero = s.zeros_state()
    rhs0 = np.asfortranarray(np.random.default_rng(N).standard_normal(Q.shape))
    for r0 in (zero, rhs0):
        want = oracle_volume(s, Q, r0)
        for prec in (FP64_FMA, FP64_MFMA):
            e = rel_l2(volume(s, ds, Q, r0, prec), want)
            print("  prec %d rhs0 %s: rel L2 %.3e" % (prec, "0" if r0 is zero else "random", e),
                  flush=True)
            assert e <= 1e-12


@pytest.mark.parametrize("N", [3, 4, 7])
def test_volume_term_on_warped_elements(cuda, N):
    print("mfma volumerhs warped 3D N=%d" % N, flush=True)
    s = warped_setup(EULER, 3, N, Ne=3, periodic=(True,) * 3, seed=N)
    ds = DeviceSetup(s)
    Q = np.asfortranarray(smooth_euler_state(s, seed=5))
    want = oracle_volume(s, Q, s.zeros_state())
    for prec in (FP64_FMA, FP64_MFMA):
        e = rel_l2(volume(s, ds, Q, s.zeros_state(), prec), want)
        print("  prec %d: rel L2 %.3e" % (prec, e), flush=True)
        assert e <= 1e-12


def test_volume_term_with_partial_workgroups(cuda):
    """Element counts that leave workgroup slots empty (N = 1, 2: several
    elements per workgroup) and a count smaller than one workgroup."""
    for N in (1, 2):
        s = small_setup(EULER, 3, N, periodic=(True, True, False))
        ds = DeviceSetup(s)
        Q = np.asfortranarray(smooth_euler_state(s, seed=11))
        for n in (1, 7, s.nelem - 1):
            print("mfma volumerhs N=%d nelem=%d" % (N, n), flush=True)
            r = dev(s.zeros_state())
            K.volumerhs(3, N, r, dev(Q), ds.vgeo, ds.D, n, eqn=EULER, phys=s.phys, prec=FP64_MFMA)
            K.check_errors()
            got = host(r, Q.shape)
            want = s.zeros_state()
            O.volumerhs(s, want, Q, nelem=n)
            assert rel_l2(got[:, :, :n], want[:, :, :n]) <= 1e-12
            assert not got[:, :, n:].any()  # elements past nreal untouched


@pytest.mark.parametrize("N", [2, 4, 6])
def test_fused_stages_against_the_oracle(cuda, N):
    print("mfma fused stage 3D N=%d" % N, flush=True)
    s = small_setup(EULER, 3, N, periodic=(True, False, False))
    ds = DeviceSetup(s)
    prep = K.prepare_geometry(3, N, ds.vgeo, ds.sgeo, ds.vmapM, ds.vmapP, ds.bndy, s.nelem)
    K.check_errors()
    Q = np.asfortranarray(smooth_euler_state(s, seed=3 + N))
    rhs0 = np.asfortranarray(np.random.default_rng(N).standard_normal(Q.shape))
    A, B, _ = P.rk_coefs()
    want_r, want_q = rhs0.copy(order="F"), Q.copy(order="F")
    O.volumerhs(s, want_r, want_q)
    O.facerhs(s, want_r, want_q)
    O.updatesolution(s, want_r, want_q, A[2], B[1], 0.013)
    for layout in ("reference", "prepared"):
        qin, qout, r = dev(Q), dev(np.zeros_like(Q)), dev(rhs0)
        if layout == "reference":
            K.stage_fused(3, N, r, qin, qout, ds.vgeo, ds.sgeo, ds.D, s.nelem, ds.vmapM,
                          ds.vmapP, ds.bndy, A[2], B[1], 0.013, eqn=EULER, phys=s.phys,
                          prec=FP64_MFMA)
        else:
            K.stage_prepared(3, N, r, qin, qout, prep, ds.D, s.nelem, A[2], B[1], 0.013,
                             eqn=EULER, phys=s.phys, prec=FP64_MFMA)
        K.check_errors()
        eq, er = rel_l2(host(qout, Q.shape), want_q), rel_l2(host(r, Q.shape), want_r)
        print("  %s arrays: rel L2 Q %.3e rhs %.3e" % (layout, eq, er), flush=True)
        assert eq <= 1e-12 and er <= 1e-12


def test_library_loops_against_oracle_lowstoragerk(cuda):
    s = P.bubble3d(Ne=4, N=4, dtheta=2.0, jitter=1e-3, seed=3)
    ds = DeviceSetup(s)
    dt, nsteps = 0.02, 40
    Qo, ro = np.asfortranarray(s.Q.copy()), s.zeros_state()
    assert O.lsrk(s, Qo, ro, dt, nsteps, nthreads=8) == 0
    for path, graph in ((PATH_FUSED, False), (PATH_FUSED, True), (PATH_FACE_UPDATE, False)):
        q, q2, rhs = dev(s.Q), dev(s.zeros_state()), dev(s.zeros_state())
        out = K.lsrk54(3, 4, q, rhs, ds.vgeo, ds.sgeo, ds.D, s.nelem, ds.vmapM, ds.vmapP,
                       ds.bndy, dt, nsteps, eqn=EULER, phys=s.phys, path=path, prec=FP64_MFMA,
                       Qalt=q2, graph=graph)
        K.check_errors()
        e = rel_l2(host(out, Qo.shape), Qo)
        print("mfma loop path %d graph %s vs oracle after %d steps: %.3e" % (path, graph, nsteps, e),
              flush=True)
        assert e <= 1e-10
    prep = K.prepare_geometry(3, 4, ds.vgeo, ds.sgeo, ds.vmapM, ds.vmapP, ds.bndy, s.nelem)
    q, q2, rhs = dev(s.Q), dev(s.zeros_state()), dev(s.zeros_state())
    out = K.lsrk54_prepared(3, 4, q, q2, rhs, prep, ds.D, s.nelem, dt, nsteps, eqn=EULER,
                            phys=s.phys, prec=FP64_MFMA, graph=True)
    K.check_errors()
    e = rel_l2(host(out, Qo.shape), Qo)
    print("mfma prepared loop vs oracle after %d steps: %.3e" % (nsteps, e), flush=True)
    assert e <= 1e-10


def test_matrix_core_form_refuses_what_it_does_not_hold(cuda):
    """facerhs_update has no contraction and no matrix-core form: the unit
    refuses it (the library loop's face + update path runs it contracted)."""
    import ctypes as C
    from canary_amd._lib import default_phys, lib
    s = small_setup(EULER, 3, 2)
    ds = DeviceSetup(s)
    q, q2, r = dev(s.zeros_state()), dev(s.zeros_state()), dev(s.zeros_state())
    p = default_phys()
    ptr = lambda t: C.c_void_p(t.data_ptr())
    rc = lib().canary_facerhs_update(EULER, 3, 2, FP64_MFMA, ptr(r), ptr(q), ptr(q2),
                                     ptr(ds.vgeo), ptr(ds.sgeo), s.nelem, s.nelem,
                                     ptr(ds.vmapM), ptr(ds.vmapP), ptr(ds.bndy), 0.0, 1.0, 0.01,
                                     0, C.byref(p), None)
    assert rc == ERR_UNSUPPORTED
