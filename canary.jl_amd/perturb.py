"""Precision form of the 2D / 3D Euler stage: FP32 perturbation about an
FP64 background at rest (hydrostatic, or uniform without gravity); no
reference counterpart: the reference evolves Q in one precision
(euler3d_set2c.jl:754-891, euler2d_set2c.jl:639-775).

    pf = Perturb32(setup, Q_background, Q_initial)
    pf.run(dt, nsteps)          # LSRK54 steps, one fused FP32 launch per stage
    Q = pf.full_state()         # FP64 (Np, dim+2, nelem) tensor: Q0 + q'

Two variants share the FP32 state q' and the FP32 RHS arithmetic:
  compact (default): metric terms, normals and the LSRK register dq stored
    in FP32 (the fewest bytes);
  mixed (mixed=True): metric terms, MJ, MJI, normals, sMJ and dq kept in
    FP64 on the GPU, as the background is in both.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from .problems import device_arrays, rk_coefs


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _dev(a, device):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).reshape(-1, order="F"))).to(device)


class Perturb32:
    def __init__(self, s, Q0, Q, device="cuda", stream=None, mixed=False):
        if s.eqn != _lib.EULER or s.dim not in (2, 3):
            raise ValueError("the precision form is implemented for 2D / 3D Euler")
        if not 1 <= s.N <= 7:
            raise ValueError("precision form compiled for N = 1..7")
        self.s = s
        self.device = torch.device(device)
        self.stream = stream
        self.mixed = bool(mixed)
        d = s.dim
        Np, Nfp, ne, nf, ns = s.Np, (s.N + 1) ** (d - 1), s.nelem, 2 * d, d + 2
        f32, f64 = torch.float32, torch.float64
        gw = f64 if self.mixed else f32  # geometry and LSRK register word
        dv = self.device
        self.vg = torch.empty(Np * (d * d + 1) * ne, dtype=gw, device=dv)
        self.sg = torch.empty(4 * Nfp * nf * ne, dtype=gw, device=dv)
        self.nbr = torch.empty(nf * ne, dtype=torch.int32, device=dv)
        self.bg = torch.empty(Np * 3 * ne, dtype=f64, device=dv)
        self.bg32 = torch.empty(Np * 3 * ne, dtype=f32, device=dv)
        self.q = torch.empty(Np * ns * ne, dtype=f32, device=dv)
        self.q2 = torch.empty_like(self.q)
        self.rhs = torch.zeros(Np * ns * ne, dtype=gw, device=dv)
        # FP64 drop-in arrays: the device ones of a device-prepared setup, else uploads
        arr = device_arrays(s, dv)
        self.D = arr["D"]
        vgeo, sgeo, vM, vP, bd = arr["vgeo"], arr["sgeo"], arr["vmapM"], arr["vmapP"], arr["bndy"]
        q0 = Q0 if isinstance(Q0, torch.Tensor) else _dev(Q0, dv)
        qf = Q if isinstance(Q, torch.Tensor) else _dev(Q, dv)
        prep = lib().canary_perturb_mixed_prepare if self.mixed else lib().canary_perturb32_prepare
        rc = prep(d, s.N, ne, ne, _ptr(vgeo), _ptr(sgeo), _ptr(vM), _ptr(vP), _ptr(bd), _ptr(q0),
                  _ptr(qf), C.byref(s.phys), _ptr(self.vg), _ptr(self.sg), _ptr(self.nbr),
                  _ptr(self.bg), _ptr(self.bg32), _ptr(self.q), self._st())
        check(rc, "perturb prepare")
        self.check()
        del vgeo, sgeo, vM, vP, bd, q0, qf

    def _st(self):
        s = torch.cuda.current_stream(self.device) if self.stream is None else self.stream
        return C.c_void_p(s.cuda_stream)

    def check(self):
        info = (C.c_int64 * 4)()
        rc = lib().canary_check_errors(self._st(), info)
        if rc:
            raise _lib.CanaryError(rc, "perturb32 (face %d, element %d, value %d)"
                                   % (info[1], info[2], info[3]))

    def stage(self, rka, rkb, dt):
        s = self.s
        fn = lib().canary_stage_perturb_mixed if self.mixed else lib().canary_stage_perturb32
        rc = fn(s.dim, s.N, _ptr(self.rhs), _ptr(self.q), _ptr(self.q2), _ptr(self.bg32),
                _ptr(self.vg), _ptr(self.sg), _ptr(self.nbr), _ptr(self.D),
                s.nelem, float(rka), float(rkb), float(dt), C.byref(s.phys), self._st())
        check(rc, "perturb stage")
        self.q, self.q2 = self.q2, self.q

    def run(self, dt, nsteps):
        A, B, _ = rk_coefs()
        for _ in range(nsteps):
            for st in range(5):
                self.stage(A[(st + 1) % 5], B[st], dt)

    def full_state(self):
        s = self.s
        out = torch.empty(s.Np * s.nstate * s.nelem, dtype=torch.float64, device=self.device)
        check(lib().canary_perturb32_reconstruct(s.dim, s.N, s.nelem, _ptr(self.q), _ptr(self.bg),
                                                 _ptr(out), self._st()), "perturb32_reconstruct")
        return out

    @staticmethod
    def bytes_per_element_stage(N, dim=3, mixed=False):
        """Bytes one element moves per stage, every array counted once: q' r+w
        (FP32), dq r+w, dim^2 + 1 metric fields (metric terms, MJ) and 4 face
        fields (FP32, or FP64 in the mixed variant), the background's 3
        fields in FP32 and one int32 face code per face."""
        Nq = N + 1
        Np, Nfp, nf, ns = Nq ** dim, Nq ** (dim - 1), 2 * dim, dim + 2
        g = 8 if mixed else 4
        return (4 * 2 * ns * Np + g * (2 * ns * Np + (dim * dim + 1) * Np + nf * Nfp * 4)
                + 4 * 3 * Np + 4 * nf)
