"""The 2D Euler equations in total-energy form (euler2d_set3c.jl; eqn
CANARY_EULER_TE): the HIP kernels through the C ABI against the CPU oracle.

The pressure P = (R/c_v)(E - (U^2 + V^2)/(2 rho) - rho g y) needs no pow, so
in the exact order every path gives the oracle's bits: volumerhs!,
fluxrhs! (which reads vgeo's y at the M and P nodes), updatesolution!, the
fused stage and the face + update fusion, and the library's LSRK54 loop on
each path; the contracted stage (FP64_FMA) is held to 1e-12 relative L2 per
stage and 1e-10 after 100 steps of the example's bubble.
"""
import numpy as np
import pytest
import torch

import canary_amd.kernels as K
import canary_amd.problems as P
from canary_amd import _lib
from canary_amd._lib import (EULER_TE, ERR_ARG, ERR_UNSUPPORTED, FP64, FP64_FMA, PATH_FACE_UPDATE,
                             PATH_FUSED, PATH_KERNELS, CanaryError)
from oracle import oracle as O

from gpu_util import DeviceSetup, dev, host
from states import max_ulp, rel_l2, small_setup, smooth_te_state, warped_setup

pytestmark = pytest.mark.gpu


def setup(N, periodic, warped=False):
    if warped:
        return warped_setup(EULER_TE, 2, N, Ne=3, periodic=periodic, seed=N)
    return small_setup(EULER_TE, 2, N, Ne=4, periodic=periodic)


def gpu_kernels(s, ds, Q, rhs0, rka, rkb, dt):
    q, r = dev(Q), dev(rhs0)
    K.volumerhs(2, s.N, r, q, ds.vgeo, ds.D, s.nelem, eqn=EULER_TE, phys=s.phys)
    K.facerhs(2, s.N, r, q, ds.sgeo, s.nelem, ds.vmapM, ds.vmapP, ds.bndy, eqn=EULER_TE,
              phys=s.phys, vgeo=ds.vgeo)
    K.check_errors()
    rhs_after = host(r, Q.shape)
    K.updatesolution(2, s.N, r, q, ds.vgeo, s.nelem, rka, rkb, dt, eqn=EULER_TE)
    K.check_errors()
    return rhs_after, host(q, Q.shape), host(r, Q.shape)


@pytest.mark.parametrize("N,periodic,warped", [(2, (True, False), False), (3, (False, False), False),
                                               (4, (True, True), False), (7, (True, False), False),
                                               (4, (True, False), True), (5, (True, True), True)])
def test_kernels_and_fused_match_oracle(cuda, N, periodic, warped):
    print("total energy N=%d periodic=%s warped=%s" % (N, periodic, warped), flush=True)
    s = setup(N, periodic, warped)
    ds = DeviceSetup(s)
    Q = np.asfortranarray(smooth_te_state(s, seed=N))
    rhs0 = np.asfortranarray(np.random.default_rng(N).standard_normal(Q.shape) * 1e2)
    A, B, _ = P.rk_coefs()
    rka, rkb, dt = A[2], B[1], 0.013
    want_r = O.rhs_eval(s, Q, rhs0)
    want_q = Q.copy(order="F")
    want_rr = want_r.copy(order="F")
    O.updatesolution(s, want_rr, want_q, rka, rkb, dt)
    got_r, got_q, got_rr = gpu_kernels(s, ds, Q, rhs0, rka, rkb, dt)
    assert max_ulp(got_r, want_r) == 0
    assert max_ulp(got_q, want_q) == 0 and max_ulp(got_rr, want_rr) == 0
    # the fused stage and the face + update fusion give the same bits
    for which in ("fused", "face_update"):
        qin, qout, r = dev(Q), dev(np.zeros_like(Q)), dev(rhs0)
        if which == "fused":
            K.stage_fused(2, N, r, qin, qout, ds.vgeo, ds.sgeo, ds.D, s.nelem, ds.vmapM, ds.vmapP,
                          ds.bndy, rka, rkb, dt, eqn=EULER_TE, phys=s.phys)
        else:
            K.volumerhs(2, N, r, qin, ds.vgeo, ds.D, s.nelem, eqn=EULER_TE, phys=s.phys)
            K.facerhs_update(2, N, r, qin, qout, ds.vgeo, ds.sgeo, s.nelem, ds.vmapM, ds.vmapP,
                             ds.bndy, rka, rkb, dt, eqn=EULER_TE, phys=s.phys)
        K.check_errors()
        assert max_ulp(host(qout, Q.shape), want_q) == 0, which
        assert max_ulp(host(r, Q.shape), want_rr) == 0, which
    # the contracted fused stage: within 1e-12 of the reference order
    qin, qout, r = dev(Q), dev(np.zeros_like(Q)), dev(rhs0)
    K.stage_fused(2, N, r, qin, qout, ds.vgeo, ds.sgeo, ds.D, s.nelem, ds.vmapM, ds.vmapP,
                  ds.bndy, rka, rkb, dt, eqn=EULER_TE, phys=s.phys, prec=FP64_FMA)
    K.check_errors()
    assert rel_l2(host(r, Q.shape), want_rr) <= 1e-12
    assert rel_l2(host(qout, Q.shape), want_q) <= 1e-12


@pytest.mark.parametrize("path", [PATH_KERNELS, PATH_FUSED, PATH_FACE_UPDATE])
def test_lsrk_loop_matches_oracle(cuda, path):
    """The example's bubble (periodic in x, walls in y), 10 LSRK54 steps of
    its courantnumber dt through the library loop, against the oracle's."""
    s = P.bubble2d_te(Ne=4, N=4)
    ds = DeviceSetup(s)
    dt = P.te_dt(s, s.Q)
    q, q2, r = dev(s.Q), dev(s.Q), dev(s.zeros_state())
    fin = K.lsrk54(2, 4, q, r, ds.vgeo, ds.sgeo, ds.D, s.nelem, ds.vmapM, ds.vmapP, ds.bndy,
                   dt, 10, eqn=EULER_TE, phys=s.phys, path=path, prec=FP64,
                   Qalt=None if path == PATH_KERNELS else q2)
    K.check_errors()
    Q = s.Q.copy(order="F")
    rhs = s.zeros_state()
    assert O.lsrk(s, Q, rhs, dt, 10) == 0
    got = host(fin, s.Q.shape)
    assert np.isfinite(got).all()
    assert max_ulp(got, Q) == 0
    assert max_ulp(host(r, s.Q.shape), rhs) == 0


def test_contracted_loop_after_100_steps(cuda):
    s = P.bubble2d_te(Ne=6, N=4)
    ds = DeviceSetup(s)
    dt = P.te_dt(s, s.Q)
    out = {}
    for prec in (FP64, FP64_FMA):
        q, q2, r = dev(s.Q), dev(s.Q), dev(s.zeros_state())
        fin = K.lsrk54(2, 4, q, r, ds.vgeo, ds.sgeo, ds.D, s.nelem, ds.vmapM, ds.vmapP, ds.bndy,
                       dt, 100, eqn=EULER_TE, phys=s.phys, path=PATH_FUSED, prec=prec, Qalt=q2)
        K.check_errors()
        out[prec] = host(fin, s.Q.shape)
    Q = s.Q.copy(order="F")
    assert O.lsrk(s, Q, s.zeros_state(), dt, 100, nthreads=8) == 0
    assert max_ulp(out[FP64], Q) == 0
    e = rel_l2(out[FP64_FMA], Q)
    print("contracted vs oracle after 100 steps: %.3e" % e, flush=True)
    assert e <= 1e-10
    # the bubble rises: vertical momentum appears where the state was at rest
    assert np.max(np.abs(Q[:, 1, :])) > 0


@pytest.mark.parametrize("N", [2, 4, 7])
def test_contracted_volume_term(cuda, N):
    """canary_volumerhs and canary_facerhs_geo in the contracted form
    (CANARY_FP64_FMA) read the node heights too: within 1e-12 of the
    oracle's volumerhs! and of its RHS."""
    s = setup(N, (True, False), warped=N == 4)
    ds = DeviceSetup(s)
    Q = np.asfortranarray(smooth_te_state(s, seed=N))
    want = s.zeros_state()
    O.volumerhs(s, want, Q)
    r = dev(s.zeros_state())
    K.volumerhs(2, N, r, dev(Q), ds.vgeo, ds.D, s.nelem, eqn=EULER_TE, phys=s.phys,
                prec=FP64_FMA)
    K.check_errors()
    e = rel_l2(host(r, Q.shape), want)
    print("contracted total-energy volume term N=%d: rel L2 %.3e" % (N, e), flush=True)
    assert e <= 1e-12
    # and the face term (it reads the heights too), contracted
    want = O.rhs_eval(s, Q, s.zeros_state())
    q = dev(Q)
    K.facerhs(2, N, r, q, ds.sgeo, s.nelem, ds.vmapM, ds.vmapP, ds.bndy, eqn=EULER_TE,
              phys=s.phys, vgeo=ds.vgeo, prec=FP64_FMA)
    K.check_errors()
    e = rel_l2(host(r, Q.shape), want)
    print("contracted total-energy RHS N=%d: rel L2 %.3e" % (N, e), flush=True)
    assert e <= 1e-12


def test_face_term_needs_the_heights(cuda):
    """canary_facerhs has no vgeo: the total-energy face term is refused
    there, and the prepared arrays (which hold no heights) have no
    total-energy stage."""
    s = small_setup(EULER_TE, 2, 3, Ne=3)
    ds = DeviceSetup(s)
    q, r = dev(smooth_te_state(s)), dev(s.zeros_state())
    lib = _lib.lib()
    rc = lib.canary_facerhs(EULER_TE, 2, 3, FP64, K._ptr(r), K._ptr(q), K._ptr(ds.sgeo),
                            s.nelem, s.nelem, K._ptr(ds.vmapM), K._ptr(ds.vmapP),
                            K._ptr(ds.bndy), None, None)
    assert rc == ERR_ARG
    with pytest.raises(ValueError):
        K.facerhs(2, 3, r, q, ds.sgeo, s.nelem, ds.vmapM, ds.vmapP, ds.bndy, eqn=EULER_TE,
                  phys=s.phys)
    prep = K.prepare_geometry(2, 3, ds.vgeo, ds.sgeo, ds.vmapM, ds.vmapP, ds.bndy, s.nelem)
    K.check_errors()
    with pytest.raises(CanaryError) as ei:
        K.stage_prepared(2, 3, r, q, dev(s.zeros_state()), prep, ds.D, s.nelem, 0.0, 1.0, 0.01,
                         eqn=EULER_TE, phys=s.phys)
    assert ei.value.code == ERR_UNSUPPORTED
