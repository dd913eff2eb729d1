"""Precision form (FP32 perturbation about an FP64 background at rest)
against the FP64 GPU path, itself bit-compared with the CPU oracle in
test_gpu_parity.py.  Bar: relative L2 error of rho, the momenta and
rho*theta <= 1e-5 after 1, 10 and 100 LSRK54 steps, for the 3D bubble
(hydrostatic background) and the 2D isentropic vortex (uniform background).
The batched precision form is bit-identical to the single-solve one."""
import numpy as np
import pytest
import torch

import canary_amd.kernels as K
import canary_amd.problems as P
from canary_amd._lib import CanaryError
from canary_amd.perturb import Perturb32
from canary_amd.batch import Batch32

from gpu_util import DeviceSetup, dev, host
from states import max_ulp

pytestmark = pytest.mark.gpu


def fp64_run(s, ds, Q, dt, nsteps):
    A, B, _ = P.rk_coefs()
    q, q2, rhs = dev(Q), torch.empty(Q.size, dtype=torch.float64, device="cuda"), dev(np.zeros_like(Q))
    for _ in range(nsteps):
        for st in range(5):
            K.stage_fused(s.dim, s.N, rhs, q, q2, ds.vgeo, ds.sgeo, ds.D, s.nelem, ds.vmapM,
                          ds.vmapP, ds.bndy, A[(st + 1) % 5], B[st], dt, phys=s.phys)
            q, q2 = q2, q
    K.check_errors()
    return host(q, Q.shape)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("N", [3, 4])
def test_perturb32_tracks_fp64(cuda, N, mixed):
    """Both variants: compact (FP32 geometry and LSRK register) and mixed
    (FP64 geometry, normals and register)."""
    s = P.bubble3d(Ne=5, N=N, dtheta=2.0, jitter=1e-3, seed=4)
    Q0 = P.bubble_state(s, dtheta=0.0, jitter=0.0)
    dt = P.cfl(s, s.Q) / s.N ** np.sqrt(2)
    ds = DeviceSetup(s)
    pf = Perturb32(s, Q0, s.Q, mixed=mixed)
    # round trip: reconstruct(prepare(Q)) == Q up to FP32 rounding of q'
    Qr = host(pf.full_state(), s.Q.shape)
    assert np.max(np.abs(Qr - s.Q)) <= 1e-6 * np.max(np.abs(s.Q - Q0)) + 1e-300
    done = 0
    for n in (1, 10, 100):
        pf.run(dt, n - done)
        done = n
        pf.check()
        got = host(pf.full_state(), s.Q.shape)
        want = fp64_run(s, ds, s.Q, dt, n)
        e = {nm: rel(got[:, k, :], want[:, k, :]) for k, nm in ((3, "rho"), (0, "rhou"), (2, "rhow"),
                                                               (4, "rhotheta"))}
        print("N=%d mixed=%s steps=%d" % (N, mixed, n), {k: "%.2e" % v for k, v in e.items()},
              flush=True)
        assert all(v <= 1e-5 for v in e.values()), e


@pytest.mark.parametrize("mixed", [False, True])
def test_perturb32_refuses_what_the_compact_arrays_drop(cuda, mixed):
    """prepare checks every property the compact arrays rely on: vmapM the
    canonical face mask, each vmapP column the canonical mask of one
    neighbour face (one int32 code per face is kept), elemtobndy 0 or 1,
    MJI == 1 ./ MJ (the stage divides)."""
    from canary_amd._lib import ERR_ARG, ERR_BC, ERR_ORIENTATION, ERR_VMAP
    base = P.bubble3d(Ne=2, N=3, dtheta=2.0, jitter=0.0)
    Q0 = P.bubble_state(base, dtheta=0.0, jitter=0.0)

    def expect(code, **change):
        s = P.bubble3d(Ne=2, N=3, dtheta=2.0, jitter=0.0)
        for name, fn in change.items():
            a = getattr(s, name).copy(order="F")
            fn(a, s)
            setattr(s, name, a)
        with pytest.raises(CanaryError) as ei:
            Perturb32(s, Q0, s.Q, mixed=mixed)
        assert ei.value.code == code, (change, ei.value.code)

    def bump_M(a, s):
        a[0, 0, 0] += 1

    def flip_P(a, s):
        e, f = next((e, f) for e in range(s.nelem) for f in range(6) if s.elemtobndy[f, e] == 0)
        a[:, f, e] = a[::-1, f, e]

    def bad_bc(a, s):
        a[1, 0] = 2

    def bad_MJI(a, s):
        a[7, 10, 1] = np.nextafter(a[7, 10, 1], 0.0)  # field 10 = MJI

    expect(ERR_VMAP, vmapM=bump_M)
    expect(ERR_ORIENTATION, vmapP=flip_P)
    expect(ERR_BC, elemtobndy=bad_bc)
    expect(ERR_ARG, vgeo=bad_MJI)
    Perturb32(base, Q0, base.Q, mixed=mixed).check()  # the untouched arrays prepare cleanly


def state_errors(got, want, dim):
    names = ["rhou", "rhov", "rhow"][:dim] + ["rho", "rhotheta"]
    return {nm: rel(got[:, k, :], want[:, k, :]) for k, nm in enumerate(names)}


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("N", [4, 7])
def test_perturb32_vortex2d_tracks_fp64(cuda, N, mixed):
    s = P.vortex2d(Ne=8, N=N)
    dt = P.euler_courant(s, s.Q)
    ds = DeviceSetup(s)
    pf = Perturb32(s, s.Q0, s.Q, mixed=mixed)
    done = 0
    for n in (1, 10, 100):
        pf.run(dt, n - done)
        done = n
        pf.check()
        got = host(pf.full_state(), s.Q.shape)
        want = fp64_run(s, ds, s.Q, dt, n)
        e = state_errors(got, want, 2)
        print("vortex N=%d mixed=%s steps=%d" % (N, mixed, n),
              {k: "%.2e" % v for k, v in e.items()}, flush=True)
        assert all(v <= 1e-5 for v in e.values()), e


def p32_mix():
    out = []
    for Ne, N, seed in ((3, 3, 1), (2, 4, 2), (4, 2, 3), (2, 5, 4), (2, 6, 5), (1, 7, 6)):
        s = P.bubble3d(Ne=Ne, N=N, dtheta=2.0, jitter=1e-3, seed=seed)
        out.append((s, P.cfl(s, s.Q) / s.N ** np.sqrt(2)))
    return out


def vortex_mix():
    out = []
    for Ne, N in ((4, 3), (3, 5), (2, 7), (5, 4)):
        s = P.vortex2d(Ne=Ne, N=N)
        out.append((s, P.euler_courant(s, s.Q)))
    return out


@pytest.mark.parametrize("dim,mixf", [(3, p32_mix), (2, vortex_mix)])
def test_batch32_equals_single(cuda, dim, mixf):
    mix = mixf()
    b = Batch32(dim)
    ids = [b.add(s, dt) for s, dt in mix]
    b.finalize(autotune=True)
    print("autotune", b.epw, flush=True)
    b.run(3)
    b.check()
    for (s, dt), i in zip(mix, ids):
        pf = Perturb32(s, s.Q0, s.Q)
        pf.run(dt, 3)
        pf.check()
        want = host(pf.full_state(), s.Q.shape)
        assert np.isfinite(want).all()
        assert max_ulp(b.state(i), want) == 0, "solve N=%d" % s.N


def test_batch32_join_and_leave(cuda):
    mix = vortex_mix()
    b = Batch32(2)
    a = b.add(*mix[0])
    b.finalize()
    b.run(2)
    c = b.add(*mix[2])
    b.finalize()
    b.run(3)
    Qa = b.remove(a)
    b.finalize()
    b.run(1)
    for (s, dt), n, got in ((mix[0], 5, Qa), (mix[2], 4, b.state(c))):
        pf = Perturb32(s, s.Q0, s.Q)
        pf.run(dt, n)
        assert max_ulp(got, host(pf.full_state(), s.Q.shape)) == 0
