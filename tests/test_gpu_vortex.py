"""Isentropic vortex on the GPU: the CPU-runnable configuration (periodic
[-5,5]^2, 10x10, N=4, 100 LSRK54 steps) against the oracle, through both the
fused stage and the three drop-in kernels, and the h-convergence of the GPU
path (the same orders as the oracle's in test_convergence.py).  The
contracted stage runs here too: the vortex's pressure arguments leave the
range of its series pow, so its device log/exp fallback is exercised."""
import numpy as np
import pytest

import canary_amd.kernels as K
import canary_amd.problems as P
from canary_amd._lib import FP64, FP64_FMA
from oracle import oracle as O

from gpu_util import DeviceSetup, dev, host
from states import max_ulp, rel_l2

pytestmark = pytest.mark.gpu


def gpu_run(s, dt, nsteps, fused=True):
    A, B, _ = P.rk_coefs()
    ds = DeviceSetup(s)
    q, q2, rhs = dev(s.Q), dev(s.Q), dev(np.zeros_like(s.Q))
    for _ in range(nsteps):
        for st in range(5):
            if fused:
                K.stage_fused(2, s.N, rhs, q, q2, ds.vgeo, ds.sgeo, ds.D, s.nelem, ds.vmapM,
                              ds.vmapP, ds.bndy, A[(st + 1) % 5], B[st], dt, phys=s.phys,
                              prec=FP64_FMA if fused == "fma" else FP64)
                q, q2 = q2, q
            else:
                K.volumerhs(2, s.N, rhs, q, ds.vgeo, ds.D, s.nelem, phys=s.phys)
                K.facerhs(2, s.N, rhs, q, ds.sgeo, s.nelem, ds.vmapM, ds.vmapP, ds.bndy,
                          phys=s.phys)
                K.updatesolution(2, s.N, rhs, q, ds.vgeo, s.nelem, A[(st + 1) % 5], B[st], dt)
    K.check_errors()
    return host(q, s.Q.shape)


@pytest.mark.parametrize("fused", [True, False, "fma"])
def test_config1_vortex_matches_oracle(cuda, fused):
    s = P.vortex2d(Ne=10, N=4)
    dt = 0.5 * P.cfl(s, s.Q) / s.N ** np.sqrt(2)
    want = np.array(s.Q, order="F", copy=True)
    O.lsrk(s, want, np.zeros_like(want, order="F"), dt, 100, nthreads=8)
    got = gpu_run(s, dt, 100, fused)
    e = rel_l2(got, want)
    kind = {True: "fused", False: "3 kernels", "fma": "fused, contracted"}[fused]
    print("config 1 GPU (%s) vs oracle after 100 steps: %.3e" % (kind, e), flush=True)
    assert e <= 1e-10
    if fused != "fma":
        assert max_ulp(got, want) == 0  # exact order, correctly rounded pow


def test_vortex_h_convergence_gpu(cuda):
    N, tend = 4, 1.0
    errs = []
    for Ne in (5, 10, 20):
        s = P.vortex2d(Ne=Ne, N=N)
        nsteps = int(np.ceil(tend / P.euler_courant(s, s.Q)))
        Q = gpu_run(s, tend / nsteps, nsteps)
        ex = P.vortex_state(s, tend, s.vortex)
        MJ = s.vgeo[:, 4, :]
        d = Q[:, 2, :] - ex[:, 2, :]
        errs.append(float(np.sqrt(np.sum(MJ * d * d) / np.sum(MJ * ex[:, 2, :] ** 2))))
    orders = [np.log(errs[i] / errs[i + 1]) / np.log(2) for i in range(2)]
    print("GPU vortex N=4 errors", ["%.3e" % e for e in errs], "orders",
          ["%.2f" % o for o in orders], flush=True)
    assert orders[-1] >= 4.0 and orders[0] > 3.0
