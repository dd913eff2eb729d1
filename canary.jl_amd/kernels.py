"""Device kernels with the reference call surface (euler3d_set2c.jl:713-734):

    volumerhs(dim, N, rhs, Q, vgeo, D, elems)
    facerhs(dim, N, rhs, Q, sgeo, elems, vmapM, vmapP, elemtobndy)
    updatesolution(dim, N, rhs, Q, vgeo, elems, rka, rkb, dt)

plus the fused single-pass stage.  Arrays are torch tensors on the GPU in
the reference layout (a C-order tensor of shape (nelem, nstate, Np) is
Julia's (Np, nstate, nelem)).  Every call goes through the C ABI of
libcanary_dg.so; there is no other code path.  Shapes are checked on the
host before any launch.
"""
import ctypes as C

import torch

from . import _lib
from ._lib import EULER, EULER_TE, SWE, FP64, FP64_FMA, FP64_MFMA, check, lib

_NS = {(EULER, 3): 5, (EULER, 2): 4, (SWE, 2): 4, (EULER_TE, 2): 4}
_NV = {3: 14, 2: 8}
_NSG = {3: 5, 2: 4}


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _stream(stream):
    s = torch.cuda.current_stream() if stream is None else stream
    return C.c_void_p(s.cuda_stream)


def _nreal(elems):
    if isinstance(elems, range):
        if elems.start != 1 or elems.step != 1:
            raise ValueError("elems must be 1:nreal (the device methods use length(elems))")
        return len(elems)
    return int(elems)


def _need(t, name, numel, dtype=torch.float64):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("%s must be a CUDA tensor" % name)
    if t.dtype != dtype:
        raise TypeError("%s must be %s, got %s" % (name, dtype, t.dtype))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)
    if t.numel() < numel:
        raise ValueError("%s has %d elements, needs at least %d" % (name, t.numel(), numel))


def _mapM(vmapM, numel):
    """vmapM, or None for the canonical face mask (the map is then not read)."""
    if vmapM is None:
        return C.c_void_p(None)
    _need(vmapM, "vmapM", numel, torch.int64)
    return _ptr(vmapM)


def _phys(phys):
    return C.byref(phys) if phys is not None else C.byref(_lib.default_phys())


def volumerhs(dim, N, rhs, Q, vgeo, D, elems, eqn=EULER, phys=None, stream=None, prec=FP64):
    """volumerhs!(Val(dim), Val(N), rhs, Q, vgeo, D, elems): rhs += volume term.

    prec: FP64 (the reference evaluation order), FP64_FMA (contracted) or
    FP64_MFMA (contracted, the derivative contractions on the matrix cores;
    3D Euler only)."""
    if prec not in (FP64, FP64_FMA, FP64_MFMA):
        raise ValueError("volumerhs: prec must be FP64, FP64_FMA or FP64_MFMA")
    n = _nreal(elems)
    Np = (N + 1) ** dim
    ns = _NS[(eqn, dim)]
    _need(rhs, "rhs", Np * ns * n)
    _need(Q, "Q", Np * ns * n)
    _need(vgeo, "vgeo", Np * _NV[dim] * n)
    _need(D, "D", (N + 1) ** 2)
    rc = lib().canary_volumerhs(eqn, dim, N, prec, _ptr(rhs), _ptr(Q), _ptr(vgeo), _ptr(D), n,
                                _phys(phys), _stream(stream))
    check(rc, "volumerhs")


def facerhs(dim, N, rhs, Q, sgeo, elems, vmapM, vmapP, elemtobndy, eqn=EULER, phys=None,
            nelem_total=None, stream=None, vgeo=None, prec=FP64):
    """facerhs!(Val(dim), Val(N), rhs, Q, sgeo, elems, vmapM, vmapP, elemtobndy);
    the total-energy equations' fluxrhs! also takes vgeo (its pressure reads
    the node height), of all nelem_total elements.  prec: FP64 (the
    reference evaluation order) or FP64_FMA (contracted)."""
    if prec not in (FP64, FP64_FMA):
        raise ValueError("facerhs: prec must be FP64 or FP64_FMA")
    n = _nreal(elems)
    Np, Nfp, nf = (N + 1) ** dim, (N + 1) ** (dim - 1), 2 * dim
    ns = _NS[(eqn, dim)]
    ntot = n if nelem_total is None else int(nelem_total)
    _need(rhs, "rhs", Np * ns * n)
    _need(Q, "Q", Np * ns * ntot)
    _need(sgeo, "sgeo", _NSG[dim] * Nfp * nf * n)
    pM = _mapM(vmapM, Nfp * nf * n)
    _need(vmapP, "vmapP", Nfp * nf * n, torch.int64)
    _need(elemtobndy, "elemtobndy", nf * n, torch.int64)
    if eqn == EULER_TE and vgeo is None:
        raise ValueError("the total-energy face term needs vgeo (node heights)")
    if vgeo is not None:
        _need(vgeo, "vgeo", Np * _NV[dim] * ntot)
    rc = lib().canary_facerhs_geo(eqn, dim, N, prec, _ptr(rhs), _ptr(Q), _ptr(sgeo),
                                  None if vgeo is None else _ptr(vgeo), n, ntot, pM,
                                  _ptr(vmapP), _ptr(elemtobndy), _phys(phys), _stream(stream))
    check(rc, "facerhs")


fluxrhs = facerhs  # the shallow-water example's name for the face kernel


def updatesolution(dim, N, rhs, Q, vgeo, elems, rka, rkb, dt, eqn=EULER, advection=False,
                   stream=None, prec=FP64):
    """updatesolution!(Val(dim), Val(N), rhs, Q, vgeo, elems, rka, rkb, dt); prec
    FP64 (the reference evaluation order) or FP64_FMA (contracted)."""
    if prec not in (FP64, FP64_FMA):
        raise ValueError("updatesolution: prec must be FP64 or FP64_FMA")
    n = _nreal(elems)
    Np = (N + 1) ** dim
    ns = _NS[(eqn, dim)]
    _need(rhs, "rhs", Np * ns * n)
    _need(Q, "Q", Np * ns * n)
    _need(vgeo, "vgeo", Np * _NV[dim] * n)
    rc = lib().canary_updatesolution(eqn, dim, N, prec, _ptr(rhs), _ptr(Q), _ptr(vgeo), n,
                                     float(rka), float(rkb), float(dt), 1 if advection else 0,
                                     _stream(stream))
    check(rc, "updatesolution")


def stage_fused(dim, N, rhs, Qin, Qout, vgeo, sgeo, D, elems, vmapM, vmapP, elemtobndy,
                rka, rkb, dt, eqn=EULER, phys=None, nelem_total=None, stream=None,
                prec=FP64, advection=False):
    """One LSRK stage (volume + faces + update) in one pass; Qin -> Qout.

    prec: FP64 (the reference evaluation order, bit-identical to volumerhs +
    facerhs + updatesolution), FP64_FMA (multiply-adds fused, pressure via
    exp/log; within 1e-12 relative L2 per stage, faster) or FP64_MFMA
    (FP64_FMA with the derivative contractions on the matrix cores; 3D Euler
    only).  advection (shallow water): the velocity is kept through the
    update, as in updatesolution."""
    if prec not in (FP64, FP64_FMA, FP64_MFMA):
        raise ValueError("stage_fused: prec must be FP64, FP64_FMA or FP64_MFMA")
    n = _nreal(elems)
    Np, Nfp, nf = (N + 1) ** dim, (N + 1) ** (dim - 1), 2 * dim
    ns = _NS[(eqn, dim)]
    ntot = n if nelem_total is None else int(nelem_total)
    _need(rhs, "rhs", Np * ns * n)
    _need(Qin, "Qin", Np * ns * ntot)
    _need(Qout, "Qout", Np * ns * n)
    _need(vgeo, "vgeo", Np * _NV[dim] * n)
    _need(sgeo, "sgeo", _NSG[dim] * Nfp * nf * n)
    _need(D, "D", (N + 1) ** 2)
    pM = _mapM(vmapM, Nfp * nf * n)
    _need(vmapP, "vmapP", Nfp * nf * n, torch.int64)
    _need(elemtobndy, "elemtobndy", nf * n, torch.int64)
    if Qin.data_ptr() == Qout.data_ptr():
        raise ValueError("stage_fused needs distinct Qin and Qout")
    rc = lib().canary_stage_fused(eqn, dim, N, prec, _ptr(rhs), _ptr(Qin), _ptr(Qout),
                                  _ptr(vgeo), _ptr(sgeo), _ptr(D), n, ntot, pM,
                                  _ptr(vmapP), _ptr(elemtobndy), float(rka), float(rkb),
                                  float(dt), 1 if advection else 0, _phys(phys), _stream(stream))
    check(rc, "stage_fused")


def facerhs_update(dim, N, rhs, Qin, Qout, vgeo, sgeo, elems, vmapM, vmapP, elemtobndy,
                   rka, rkb, dt, eqn=EULER, phys=None, nelem_total=None, stream=None,
                   prec=FP64, advection=False):
    """facerhs! then updatesolution! of one stage in one pass (after a
    separate volumerhs); Qin -> Qout, rhs in place.  prec and advection as
    for stage_fused."""
    if prec not in (FP64, FP64_FMA):
        raise ValueError("facerhs_update: prec must be FP64 or FP64_FMA")
    n = _nreal(elems)
    Np, Nfp, nf = (N + 1) ** dim, (N + 1) ** (dim - 1), 2 * dim
    ns = _NS[(eqn, dim)]
    ntot = n if nelem_total is None else int(nelem_total)
    _need(rhs, "rhs", Np * ns * n)
    _need(Qin, "Qin", Np * ns * ntot)
    _need(Qout, "Qout", Np * ns * n)
    _need(vgeo, "vgeo", Np * _NV[dim] * n)
    _need(sgeo, "sgeo", _NSG[dim] * Nfp * nf * n)
    pM = _mapM(vmapM, Nfp * nf * n)
    _need(vmapP, "vmapP", Nfp * nf * n, torch.int64)
    _need(elemtobndy, "elemtobndy", nf * n, torch.int64)
    if Qin.data_ptr() == Qout.data_ptr():
        raise ValueError("facerhs_update needs distinct Qin and Qout")
    rc = lib().canary_facerhs_update(eqn, dim, N, prec, _ptr(rhs), _ptr(Qin), _ptr(Qout),
                                     _ptr(vgeo), _ptr(sgeo), n, ntot, pM, _ptr(vmapP),
                                     _ptr(elemtobndy), float(rka), float(rkb), float(dt),
                                     1 if advection else 0, _phys(phys), _stream(stream))
    check(rc, "facerhs_update")


def lsrk54(dim, N, Q, rhs, vgeo, sgeo, D, elems, vmapM, vmapP, elemtobndy, dt, nsteps,
           eqn=EULER, phys=None, nelem_total=None, path=_lib.PATH_FUSED, prec=FP64, Qalt=None,
           advection=False, graph=False, stream=None):
    """nsteps LSRK54 steps of one solve in the library's native time loop
    (lowstorageRK).  path: PATH_KERNELS (three kernels per stage, Q in place),
    PATH_FUSED or PATH_FACE_UPDATE (Q and Qalt ping-pong).  Returns the tensor
    holding the final state (Q or Qalt)."""
    n = _nreal(elems)
    Np, Nfp, nf = (N + 1) ** dim, (N + 1) ** (dim - 1), 2 * dim
    ns = _NS[(eqn, dim)]
    ntot = n if nelem_total is None else int(nelem_total)
    _need(Q, "Q", Np * ns * ntot)
    _need(rhs, "rhs", Np * ns * n)
    _need(vgeo, "vgeo", Np * _NV[dim] * n)
    _need(sgeo, "sgeo", _NSG[dim] * Nfp * nf * n)
    _need(D, "D", (N + 1) ** 2)
    pM = _mapM(vmapM, Nfp * nf * n)
    _need(vmapP, "vmapP", Nfp * nf * n, torch.int64)
    _need(elemtobndy, "elemtobndy", nf * n, torch.int64)
    if path != _lib.PATH_KERNELS:
        if Qalt is None:
            raise ValueError("lsrk54: the ping-pong paths need Qalt")
        _need(Qalt, "Qalt", Np * ns * ntot)
    alt = C.c_int(0)
    rc = lib().canary_lsrk54(eqn, dim, N, prec, path, int(nsteps), float(dt), _ptr(Q),
                             _ptr(Qalt) if Qalt is not None else None, _ptr(rhs), _ptr(vgeo),
                             _ptr(sgeo), _ptr(D), n, ntot, pM, _ptr(vmapP),
                             _ptr(elemtobndy), _phys(phys), 1 if advection else 0,
                             1 if graph else 0, C.byref(alt), _stream(stream))
    check(rc, "lsrk54")
    return Qalt if alt.value else Q


def prepare_geometry(dim, N, vgeo, sgeo, vmapM, vmapP, elemtobndy, nelem, nelem_total=None,
                     stream=None):
    """The prepared arrays of the fused stage (pvgeo, psgeo, pface tensors),
    built on the device from the reference ones; the checks it makes are
    reported by check_errors()."""
    Np, Nfp, nf = (N + 1) ** dim, (N + 1) ** (dim - 1), 2 * dim
    ntot = nelem if nelem_total is None else int(nelem_total)
    _need(vgeo, "vgeo", Np * _NV[dim] * nelem)
    _need(sgeo, "sgeo", _NSG[dim] * Nfp * nf * nelem)
    _need(vmapM, "vmapM", Nfp * nf * nelem, torch.int64)
    _need(vmapP, "vmapP", Nfp * nf * nelem, torch.int64)
    _need(elemtobndy, "elemtobndy", nf * nelem, torch.int64)
    dv = vgeo.device
    out = {"pvgeo": torch.empty(Np * (dim * dim + 1) * nelem, dtype=torch.float64, device=dv),
           "psgeo": torch.empty((dim + 1) * Nfp * nf * nelem, dtype=torch.float64, device=dv),
           "pface": torch.empty(nf * nelem, dtype=torch.int32, device=dv)}
    rc = lib().canary_prepare_geometry(dim, N, nelem, ntot, _ptr(vgeo), _ptr(sgeo), _ptr(vmapM),
                                       _ptr(vmapP), _ptr(elemtobndy), _ptr(out["pvgeo"]),
                                       _ptr(out["psgeo"]), _ptr(out["pface"]), _stream(stream))
    check(rc, "prepare_geometry")
    return out


def stage_prepared(dim, N, rhs, Qin, Qout, prep, D, elems, rka, rkb, dt, eqn=EULER, phys=None,
                   nelem_total=None, stream=None, prec=FP64, advection=False):
    """stage_fused over the prepared arrays `prep` (prepare_geometry); prec as
    for stage_fused."""
    if prec not in (FP64, FP64_FMA, FP64_MFMA):
        raise ValueError("stage_prepared: prec must be FP64, FP64_FMA or FP64_MFMA")
    n = _nreal(elems)
    Np, Nfp, nf = (N + 1) ** dim, (N + 1) ** (dim - 1), 2 * dim
    ns = _NS[(eqn, dim)]
    ntot = n if nelem_total is None else int(nelem_total)
    _need(rhs, "rhs", Np * ns * n)
    _need(Qin, "Qin", Np * ns * ntot)
    _need(Qout, "Qout", Np * ns * n)
    _need(prep["pvgeo"], "pvgeo", Np * (dim * dim + 1) * n)
    _need(prep["psgeo"], "psgeo", (dim + 1) * Nfp * nf * n)
    _need(prep["pface"], "pface", nf * n, torch.int32)
    _need(D, "D", (N + 1) ** 2)
    if Qin.data_ptr() == Qout.data_ptr():
        raise ValueError("stage_prepared needs distinct Qin and Qout")
    rc = lib().canary_stage_prepared(eqn, dim, N, prec, _ptr(rhs), _ptr(Qin), _ptr(Qout),
                                     _ptr(prep["pvgeo"]), _ptr(prep["psgeo"]),
                                     _ptr(prep["pface"]), _ptr(D), n, ntot, float(rka),
                                     float(rkb), float(dt), 1 if advection else 0, _phys(phys),
                                     _stream(stream))
    check(rc, "stage_prepared")


def lsrk54_prepared(dim, N, Q, Qalt, rhs, prep, D, elems, dt, nsteps, eqn=EULER, phys=None,
                    prec=FP64, advection=False, graph=False, stream=None):
    """lsrk54 over the prepared arrays; returns the tensor holding the final
    state (Q or Qalt)."""
    n = _nreal(elems)
    Np, Nfp, nf = (N + 1) ** dim, (N + 1) ** (dim - 1), 2 * dim
    ns = _NS[(eqn, dim)]
    for t, nm in ((Q, "Q"), (Qalt, "Qalt"), (rhs, "rhs")):
        _need(t, nm, Np * ns * n)
    _need(prep["pvgeo"], "pvgeo", Np * (dim * dim + 1) * n)
    _need(prep["psgeo"], "psgeo", (dim + 1) * Nfp * nf * n)
    _need(prep["pface"], "pface", nf * n, torch.int32)
    alt = C.c_int(0)
    rc = lib().canary_lsrk54_prepared(eqn, dim, N, prec, int(nsteps), float(dt), _ptr(Q),
                                      _ptr(Qalt), _ptr(rhs), _ptr(prep["pvgeo"]),
                                      _ptr(prep["psgeo"]), _ptr(prep["pface"]), _ptr(D), n,
                                      _phys(phys), 1 if advection else 0, 1 if graph else 0,
                                      C.byref(alt), _stream(stream))
    check(rc, "lsrk54_prepared")
    return Qalt if alt.value else Q


def lsrk54_coefs():
    """(RKA, RKB, RKC) of the library's LSRK54 tables."""
    a, b, c = (C.c_double * 5)(), (C.c_double * 5)(), (C.c_double * 5)()
    lib().canary_lsrk54_coefs(a, b, c)
    return tuple(a), tuple(b), tuple(c)


def check_errors(stream=None):
    """Synchronise and raise if a kernel recorded a bad bc / vmap index."""
    info = (C.c_int64 * 4)()
    rc = lib().canary_check_errors(_stream(stream), info)
    if rc:
        err = _lib.CanaryError(rc, "device kernel (face %d, element %d, value %d)"
                               % (info[1], info[2], info[3]))
        err.info = tuple(info)
        raise err


# ------------------------------------------------ geometry, halos, reductions
def computegeometry(dim, N, elemtocoord, vmapM, stream=None):
    """computegeometry on the GPU: elemtocoord (CUDA float64, the reference's
    (d, 2^d, nelem) column-major array flattened), vmapM (CUDA int64) ->
    (vgeo, sgeo) flat float64 tensors in the reference layouts."""
    nelem = elemtocoord.numel() // (dim * 2 ** dim)
    Np, Nfp, nf = (N + 1) ** dim, (N + 1) ** (dim - 1), 2 * dim
    _need(elemtocoord, "elemtocoord", dim * 2 ** dim * nelem)
    _need(vmapM, "vmapM", Nfp * nf * nelem, torch.int64)
    vgeo = torch.empty(Np * _NV[dim] * nelem, dtype=torch.float64, device=elemtocoord.device)
    sgeo = torch.empty(_NSG[dim] * Nfp * nf * nelem, dtype=torch.float64,
                       device=elemtocoord.device)
    rc = lib().canary_computegeometry_device(dim, N, nelem, _ptr(elemtocoord), _ptr(vmapM),
                                             _ptr(vgeo), _ptr(sgeo), _stream(stream))
    check(rc, "computegeometry_device")
    return vgeo, sgeo


def _word(t):
    if t.dtype == torch.float64:
        return 8
    if t.dtype == torch.float32:
        return 4
    raise TypeError("states must be float64 or float32")


def fillsendQ(dim, N, nstate, sendQ, Q, sendelems, nelem_total, stream=None):
    """fillsendQ!: sendQ[:, :, e] = Q[:, :, sendelems[e]] (1-based sendelems)."""
    nsend = sendelems.numel()
    w = (N + 1) ** dim * nstate
    _need(sendelems, "sendelems", nsend, torch.int64)
    _need(sendQ, "sendQ", w * nsend, Q.dtype)
    _need(Q, "Q", w * nelem_total, Q.dtype)
    rc = lib().canary_fillsendQ(dim, N, nstate, _ptr(sendQ), _ptr(Q), _ptr(sendelems), nsend,
                                int(nelem_total), _word(Q), _stream(stream))
    check(rc, "fillsendQ")


def transferrecvQ(dim, N, nstate, Q, recvQ, nrealelem, stream=None):
    """transferrecvQ!: Q[:, :, nrealelem+1:end] = recvQ."""
    w = (N + 1) ** dim * nstate
    nrecv = recvQ.numel() // w
    _need(recvQ, "recvQ", w * nrecv, Q.dtype)
    _need(Q, "Q", w * (nrealelem + nrecv), Q.dtype)
    rc = lib().canary_transferrecvQ(dim, N, nstate, _ptr(Q), _ptr(recvQ), nrecv, int(nrealelem),
                                    _word(Q), _stream(stream))
    check(rc, "transferrecvQ")


def pack_nodes(dim, N, nstate, buf, Q, nodes, nelem_total, stream=None):
    """Face-trace halo pack: buf[k + nnode*s] = Q[node k (1-based id), s]."""
    nnode = nodes.numel()
    _need(nodes, "nodes", nnode, torch.int64)
    _need(buf, "buf", nnode * nstate, Q.dtype)
    _need(Q, "Q", (N + 1) ** dim * nstate * nelem_total, Q.dtype)
    rc = lib().canary_pack_nodes(dim, N, nstate, _ptr(buf), _ptr(Q), _ptr(nodes), nnode,
                                 int(nelem_total), _word(Q), _stream(stream))
    check(rc, "pack_nodes")


def unpack_nodes(dim, N, nstate, Q, buf, nodes, nelem_total, stream=None):
    """Face-trace halo unpack: Q[node k, s] = buf[k + nnode*s]."""
    nnode = nodes.numel()
    _need(nodes, "nodes", nnode, torch.int64)
    _need(buf, "buf", nnode * nstate, Q.dtype)
    _need(Q, "Q", (N + 1) ** dim * nstate * nelem_total, Q.dtype)
    rc = lib().canary_unpack_nodes(dim, N, nstate, _ptr(Q), _ptr(buf), _ptr(nodes), nnode,
                                   int(nelem_total), _word(Q), _stream(stream))
    check(rc, "unpack_nodes")


def l2energysquared(dim, N, Q, vgeo, nelem, nstate=None, stream=None):
    """L2energysquared on the GPU (deterministic tree sum); returns a
    0-dimensional CUDA tensor."""
    Np = (N + 1) ** dim
    ns = nstate if nstate is not None else dim + 2
    _need(Q, "Q", Np * ns * nelem)
    _need(vgeo, "vgeo", Np * _NV[dim] * nelem)
    work = torch.empty(_lib.REDUCE_WORK, dtype=torch.float64, device=Q.device)
    out = torch.empty((), dtype=torch.float64, device=Q.device)
    imj = 9 if dim == 3 else 4
    rc = lib().canary_l2energysquared_device(Np, ns, _NV[dim], imj, _ptr(Q), _ptr(vgeo), nelem,
                                             _ptr(work), _ptr(out), _stream(stream))
    check(rc, "l2energysquared_device")
    return out


def cfl(dim, N, vgeo, Q, nelem, eqn=EULER, phys=None, stream=None):
    """The drivers' cfl() on the GPU (min over nodes); 0-dimensional tensor."""
    Np = (N + 1) ** dim
    _need(Q, "Q", Np * _NS[(eqn, dim)] * nelem)
    _need(vgeo, "vgeo", Np * _NV[dim] * nelem)
    work = torch.empty(_lib.REDUCE_WORK, dtype=torch.float64, device=Q.device)
    out = torch.empty((), dtype=torch.float64, device=Q.device)
    rc = lib().canary_cfl_device(eqn, dim, N, _ptr(vgeo), _ptr(Q), nelem, _phys(phys), _ptr(work),
                                 _ptr(out), _stream(stream))
    check(rc, "cfl_device")
    return out


def cfl_noalloc(dim, N, vgeo, Q, nelem, out, eqn=EULER, phys=None, stream=None):
    """cfl() into the caller's 0-dimensional CUDA tensor `out` through the
    entry point that allocates nothing (partials in the library's static
    device array; calls overlapping on one device race)."""
    Np = (N + 1) ** dim
    _need(Q, "Q", Np * _NS[(eqn, dim)] * nelem)
    _need(vgeo, "vgeo", Np * _NV[dim] * nelem)
    _need(out, "out", 1)
    rc = lib().canary_cfl_device_noalloc(eqn, dim, N, _ptr(vgeo), _ptr(Q), nelem, _phys(phys),
                                         _ptr(out), _stream(stream))
    check(rc, "cfl_device_noalloc")
    return out
