"""Device geometry, ghost-element buffers and reductions against the host
restatements (themselves pinned to the reference's fixtures in
test_host_operators_metric.py / test_host_mesh.py).

Bars: geometry bit-identical to the host (same operation order, no FMA
contraction, IEEE division and sqrt on both sides), on brick meshes and on
randomly perturbed (non-affine) elements; buffer copies exact; the energy
within 1e-13 relative (summation order differs); cfl bit-identical for the
shallow-water formula (no pow) and within 4 ulp for Euler (pow)."""
import numpy as np
import pytest
import torch

import canary_amd.kernels as K
import canary_amd.mesh as M
import canary_amd.problems as P
from canary_amd._lib import CanaryError, EULER, SWE, ERR_ARG, ERR_VMAP

from gpu_util import dev, host
from states import max_ulp, small_setup, smooth_euler_state, smooth_swe_state

pytestmark = pytest.mark.gpu


def perturbed(s, amp, seed):
    """elemtocoord of the setup with every vertex moved by up to amp of the
    element size (shared vertices moved alike, so the mesh stays conforming)."""
    etc = np.array(s.mesh["elemtocoord"], order="F", copy=True)
    rng = np.random.default_rng(seed)
    keys = {}
    h = np.ptp(etc[:, :, 0], axis=1).max()
    for e in range(etc.shape[2]):
        for v in range(etc.shape[1]):
            k = tuple(np.round(etc[:, v, e], 9))
            if k not in keys:
                keys[k] = rng.uniform(-amp, amp, size=etc.shape[0]) * h
            etc[:, v, e] += keys[k]
    return etc


@pytest.mark.parametrize("dim,N", [(2, 1), (2, 3), (2, 4), (2, 7), (3, 1), (3, 2), (3, 4),
                                   (3, 5), (3, 7)])
@pytest.mark.parametrize("warp", [0.0, 0.1])
def test_geometry_bit_identical(cuda, dim, N, warp):
    s = small_setup(EULER, dim, N, Ne=3 if dim == 3 else 4)
    etc = perturbed(s, warp, seed=N) if warp else np.asfortranarray(s.mesh["elemtocoord"])
    vg_h, sg_h = M.computegeometry(dim, N, etc, s.vmapM)
    vg_d, sg_d = K.computegeometry(dim, N, dev(etc), dev(s.vmapM))
    K.check_errors()
    assert max_ulp(host(vg_d, vg_h.shape), vg_h) == 0
    assert max_ulp(host(sg_d, sg_h.shape), sg_h) == 0


def test_geometry_bad_vmapM_reported(cuda):
    s = small_setup(EULER, 2, 3, Ne=3)
    vM = np.array(s.vmapM, order="F", copy=True)
    vM[1, 2, 0] = s.nelem * s.Np + 5
    K.computegeometry(2, 3, dev(s.mesh["elemtocoord"]), dev(vM))
    with pytest.raises(CanaryError) as ei:
        K.check_errors()
    assert ei.value.code == ERR_VMAP


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_send_and_receive_buffers(cuda, dtype):
    dim, N, ns, nreal, nghost = 3, 3, 5, 40, 7
    w = (N + 1) ** dim * ns
    rng = np.random.default_rng(3)
    Q = torch.from_numpy(rng.standard_normal(w * (nreal + nghost))).to(dtype).cuda()
    send = rng.choice(nreal, size=11, replace=False) + 1
    sendQ = torch.empty(w * send.size, dtype=dtype, device="cuda")
    K.fillsendQ(dim, N, ns, sendQ, Q, torch.from_numpy(send).cuda(), nreal + nghost)
    Qh = Q.cpu().numpy().reshape(w, -1, order="F")
    want = Qh[:, send - 1]
    assert np.array_equal(sendQ.cpu().numpy().reshape(w, -1, order="F"), want)
    recv = torch.from_numpy(rng.standard_normal(w * nghost)).to(dtype).cuda()
    Q2 = Q.clone()
    K.transferrecvQ(dim, N, ns, Q2, recv, nreal)
    K.check_errors()
    got = Q2.cpu().numpy().reshape(w, -1, order="F")
    assert np.array_equal(got[:, :nreal], Qh[:, :nreal])
    assert np.array_equal(got[:, nreal:], recv.cpu().numpy().reshape(w, -1, order="F"))


def test_send_buffer_bad_index_reported(cuda):
    Q = torch.zeros(16 * 4 * 3, dtype=torch.float64, device="cuda")
    sendQ = torch.empty(16 * 4 * 2, dtype=torch.float64, device="cuda")
    K.fillsendQ(2, 3, 4, sendQ, Q, torch.tensor([1, 9], dtype=torch.int64, device="cuda"), 3)
    with pytest.raises(CanaryError) as ei:
        K.check_errors()
    assert ei.value.code == ERR_ARG


@pytest.mark.parametrize("eqn,dim,N", [(EULER, 3, 4), (EULER, 2, 5), (SWE, 2, 3)])
def test_reductions(cuda, eqn, dim, N):
    s = small_setup(eqn, dim, N, Ne=4 if dim == 3 else 6)
    Q = np.asfortranarray(smooth_swe_state(s, seed=2) if eqn == SWE else
                          smooth_euler_state(s, seed=2))
    qd, vd = dev(Q), dev(s.vgeo)
    e_h = P.l2energysquared(s, Q)
    e_d = float(K.l2energysquared(dim, N, qd, vd, s.nelem, nstate=s.nstate))
    assert abs(e_d - e_h) <= 1e-13 * abs(e_h)
    c_h = P.cfl(s, Q)
    c_d = float(K.cfl(dim, N, vd, qd, s.nelem, eqn=eqn, phys=s.phys))
    K.check_errors()
    if eqn == SWE:
        assert c_d == c_h
    else:
        assert max_ulp(np.array([c_d]), np.array([c_h])) <= 4
    # run-to-run deterministic
    assert float(K.l2energysquared(dim, N, qd, vd, s.nelem, nstate=s.nstate)) == e_d
