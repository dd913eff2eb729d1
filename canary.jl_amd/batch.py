"""Batched form: many independent solves advanced together (no reference
counterpart; the reference runs one solve per process, euler3d_set2c.jl:1049
and the shallow-water driver).

    b = Batch(SWE, 2, prec=FP64)  # drop-in arrays (FP64; FP32 copies for SWE)
    b = Batch32(dim=3)            # Euler precision form (FP32 perturbation)
    i = b.add(setup, dt)
    b.finalize(autotune=True)     # work table; elements per workgroup chosen per N by timing
    b.run(nsteps)                 # LSRK54 steps, one hipGraph replay per step
    b.run(nsteps, graph=False)    # the same steps from the library's own loop
    Q = b.state(i)                # float64 (Np, nstate, nelem), reference layout

Solves may differ in N, mesh, element count, dt and physical constants.  A
stage launches one kernel per (N, elements-per-workgroup) class over that
class's slice of the device work table; inside a captured step the classes
run on forked streams, so a step is a graph of 5 stages x (classes in
parallel).  finalize(single_launch=True) (drop-in form) makes each stage
one launch over the whole table instead, the kernel switching on each
entry's N.  All solves share one stage parity: buffer [parity] holds the
current state and a stage writes buffer [1 - parity].
"""
import ctypes as C
import gc

import numpy as np
import torch

from . import _lib
from ._lib import EULER, SWE, FP64, FP32, FP64_FMA, check, lib
from .perturb import Perturb32
from .problems import device_arrays, rk_coefs
from .streams import OwnedStream, reap

_NS = {(EULER, 3): 5, (EULER, 2): 4, (SWE, 2): 4}
_NSR = {(EULER, 3): 5, (EULER, 2): 4, (SWE, 2): 3}

ITEM_DTYPE = np.dtype([("solve", "<i4"), ("N", "<i2"), ("epw", "<i2"), ("e0", "<i8")])
assert ITEM_DTYPE.itemsize == C.sizeof(_lib.BatchItem) == 16


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _dev(a, device, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(np.asarray(a).reshape(-1, order="F")))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device)


def _query(fn, *args):
    out = []
    for which in range(3):
        e, t = C.c_int(), C.c_int()
        rc = fn(*args, which, C.byref(e), C.byref(t))
        if rc == 0 and e.value not in [v[0] for v in out]:
            out.append((e.value, t.value))
    return out


def variants(eqn, dim, prec, N):
    """[(epw, threads)] candidates of the batched drop-in kernel for order N."""
    return _query(lib().canary_batch_variant, eqn, dim, prec, N)


def variants32(dim, N):
    """[(epw, threads)] candidates of the batched precision-form kernel."""
    return _query(lib().canary_p32_batch_variant, dim, N)


def bytes_per_element_stage(eqn, dim, N, word=8, prepared=False):
    """Bytes one element moves per fused stage, every array counted once:
    Q read + write (all states), the LSRK register read + write (evolved
    states), metric terms + MJ + MJI, normals + sMJ at `word` bytes, Int64
    vmapM + vmapP and elemtobndy; over the prepared arrays: metric terms +
    MJ, normals + sMJ and one int32 face code per face."""
    Nq = N + 1
    Np, Nfp, nface = Nq ** dim, Nq ** (dim - 1), 2 * dim
    ns, nsr = _NS[(eqn, dim)], _NSR[(eqn, dim)]
    nsurf = dim + 1
    if prepared:
        return (word * (2 * ns * Np + 2 * nsr * Np + (dim * dim + 1) * Np + nface * Nfp * nsurf)
                + 4 * nface)
    nvol = dim * dim + 2
    return (word * (2 * ns * Np + 2 * nsr * Np + nvol * Np + nface * Nfp * nsurf)
            + 2 * nface * Nfp * 8 + nface * 8)


class _Solve:
    pass


class _BatchBase:
    """Work table, stage launches, graphs and autotuning shared by the
    batched drop-in and precision forms."""

    _rowtype = None
    _has_single_launch = False

    def __init__(self, device="cuda"):
        self.device = torch.device(device)
        self.single_launch = False
        self.solves = {}          # id -> _Solve
        self._next_id = 0
        self.parity = 0
        self.epw = {}             # N -> chosen elements per workgroup
        self.tuning = {}          # N -> {epw: microseconds per stage}
        self.groups = []          # [(N, epw, first item, item count)]
        self._table = None
        self._items = None
        self.dirty = True         # solves added / removed since the tables were built
        self._graphs = {}
        self._streams = []        # OwnedStream per N class (forks of a captured stage)
        self._capture = None      # OwnedStream the step graphs are captured on
        self._version = 0

    # -- to be provided by the concrete forms
    def _variants(self, N):
        raise NotImplementedError

    def _launch_raw(self, N, epw, table, items_ptr, count, parity, rka, rkb, stream):
        raise NotImplementedError

    def _fill_row(self, row, v):
        raise NotImplementedError

    def _word(self):
        raise NotImplementedError

    def _bytes_per_element_stage(self, v):
        raise NotImplementedError

    # ------------------------------------------------------------ solves
    def _register(self, v):
        sid = self._next_id
        self._next_id += 1
        v.steps = 0
        self.solves[sid] = v
        self.dirty = True
        return sid

    def remove(self, sid):
        """Drop a solve (its device arrays are freed); returns its final state."""
        Q = self.state(sid)
        del self.solves[sid]
        self.dirty = True
        return Q

    # --------------------------------------------------------- work table
    def finalize(self, epw=None, autotune=False, single_launch=None):
        """Build the solve table and the work table (entries grouped by N).
        epw: {N: elements per workgroup} overrides; autotune: time every
        candidate per N on this batch and keep the fastest.  single_launch
        (drop-in form): every stage is ONE launch over the entries of all
        orders (canary_batch_stage_all, each order at its fullest 256-lane
        variant) instead of one launch per order; None keeps the current
        choice."""
        if not self.solves:
            raise ValueError("empty batch")
        if single_launch is not None:
            if single_launch and not self._has_single_launch:
                raise ValueError("this batched form has no single-launch kernel")
            if bool(single_launch) != self.single_launch:
                self.single_launch = bool(single_launch)
                self.dirty = True
        epw = dict(epw or {})
        Ns = sorted({v.N for v in self.solves.values()})
        if self.single_launch:
            if epw or autotune:
                raise ValueError("single_launch fixes elements per workgroup per order")
            for N in Ns:
                self.epw[N] = self._all_variant(N)
            self._build_tables()
            return self
        for N in Ns:
            cands = [e for e, _ in self._variants(N)]
            if N in epw:
                if epw[N] not in cands:
                    raise ValueError("epw %d is not a candidate for N=%d (%s)" % (epw[N], N, cands))
                self.epw[N] = epw[N]
            elif N not in self.epw:
                self.epw[N] = cands[0]
        self._build_tables()
        if autotune:
            self.autotune(only=[N for N in Ns if N not in epw])
        return self

    def _solve_table(self):
        ids = sorted(self.solves)
        tab = (self._rowtype * len(ids))()
        self._row = {}
        for r, sid in enumerate(ids):
            self._fill_row(tab[r], self.solves[sid])
            self._row[sid] = r
        raw = np.frombuffer(bytes(tab), dtype=np.uint8).copy()
        return torch.from_numpy(raw).to(self.device)

    def _items_for(self, N, epw):
        rows = []
        for sid in sorted(self.solves):
            v = self.solves[sid]
            if v.N != N:
                continue
            e0 = np.arange(0, v.nelem, epw, dtype=np.int64)
            blk = np.zeros(e0.size, dtype=ITEM_DTYPE)
            blk["solve"], blk["N"], blk["epw"], blk["e0"] = self._row[sid], N, epw, e0
            rows.append(blk)
        return np.concatenate(rows) if rows else np.zeros(0, dtype=ITEM_DTYPE)

    def _build_tables(self):
        if self._table is not None or self._graphs:
            # the step graphs and work tables about to be dropped may still be
            # executing on the runtime's streams: destroying a graph (or
            # handing its table memory back) under a running launch is a
            # use after free
            torch.cuda.synchronize(self.device)
        self._graphs = {}
        reap()  # streams of collected runtimes, destroyed here (no capture in flight)
        self.dirty = False
        self._table = self._solve_table()
        parts, self.groups, off = [], [], 0
        for N in sorted({v.N for v in self.solves.values()}):
            it = self._items_for(N, self.epw[N])
            parts.append(it)
            self.groups.append((N, self.epw[N], off, it.size))
            off += it.size
        items = np.concatenate(parts)
        if self.single_launch:  # one class of every entry (N = epw = 0 marks it)
            self.groups = [(0, 0, 0, off)]
        self._items = torch.from_numpy(items.view(np.uint8).copy()).to(self.device)
        self._graphs = {}
        self._version += 1
        while len(self._streams) < len(self.groups):
            self._streams.append(OwnedStream(self.device))

    # ------------------------------------------------------------ stages
    def _st(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _launch(self, N, epw, items_ptr, count, parity, rka, rkb, stream):
        rc = self._launch_raw(N, epw, self._table, C.c_void_p(items_ptr), count, parity,
                              float(rka), float(rkb), stream)
        check(rc, "batch stage")

    def stage(self, rka, rkb, concurrent=False):
        """One LSRK stage of every solve: buffer [parity] -> [1 - parity]."""
        if self.dirty:
            self.finalize()
        base = self._items.data_ptr()
        if concurrent and len(self.groups) > 1:
            cur = torch.cuda.current_stream(self.device)
            for g, (N, epw, off, cnt) in enumerate(self.groups):
                s = self._streams[g].stream
                s.wait_stream(cur)
                with torch.cuda.stream(s):
                    self._launch(N, epw, base + 16 * off, cnt, self.parity, rka, rkb,
                                 C.c_void_p(s.cuda_stream))
            for g in range(len(self.groups)):
                cur.wait_stream(self._streams[g].stream)
        else:
            st = self._st()
            for N, epw, off, cnt in self.groups:
                self._launch(N, epw, base + 16 * off, cnt, self.parity, rka, rkb, st)
        self.parity ^= 1

    def step(self, concurrent=False):
        A, B, _ = rk_coefs()
        for s in range(5):
            self.stage(A[(s + 1) % 5], B[s], concurrent=concurrent)
        for v in self.solves.values():
            v.steps += 1

    def _graph(self):
        key = (self._version, self.parity)
        g = self._graphs.get(key)
        if g is None:
            p0 = self.parity
            if self._capture is None:
                self._capture = OwnedStream(self.device)
            side = self._capture.stream
            # the capture protocol of torch.cuda.graph (idle device, no garbage
            # left to be freed during the capture), without its cache flush
            torch.cuda.synchronize(self.device)
            gc.collect()
            g = torch.cuda.CUDAGraph()
            A, B, _ = rk_coefs()
            with torch.cuda.stream(side):
                g.capture_begin()
                try:
                    for s in range(5):
                        self.stage(A[(s + 1) % 5], B[s], concurrent=True)
                finally:
                    g.capture_end()
            torch.cuda.current_stream(self.device).wait_stream(side)
            self.parity = p0  # capture recorded the launches, it did not run them
            self._graphs[key] = g
        return g

    def run(self, nsteps, graph=True):
        """nsteps LSRK54 steps of every solve; graph=True replays one captured
        hipGraph per step (one per starting parity), graph=False launches the
        stages from the library's own loop (canary_batch_run) on the current
        stream."""
        if self.dirty:
            self.finalize()
        if not graph:
            self._run_native(nsteps)
            return
        for _ in range(nsteps):
            self._graph().replay()
            self.parity ^= 1  # five stages flip it an odd number of times
            for v in self.solves.values():
                v.steps += 1

    def _run_native(self, nsteps):
        if nsteps <= 0 or not self.groups:
            return
        base = self._items.data_ptr()
        cls = (_lib.BatchClass * len(self.groups))()
        for i, (N, epw, off, cnt) in enumerate(self.groups):
            cls[i].items, cls[i].nitems, cls[i].N, cls[i].epw = base + 16 * off, cnt, N, epw
        kind, eqn, dim, prec = self._run_kind()
        rc = lib().canary_batch_run(kind, eqn, dim, prec, _ptr(self._table), cls, len(self.groups),
                                    int(nsteps), self.parity, self._st())
        check(rc, "batch run")
        self.parity ^= nsteps & 1  # five stages per step flip it an odd number of times
        for v in self.solves.values():
            v.steps += nsteps

    def check(self):
        info = (C.c_int64 * 4)()
        rc = lib().canary_check_errors(self._st(), info)
        if rc:
            raise _lib.CanaryError(rc, "batch stage (solve %d, element %d, value %d)"
                                   % (info[1], info[2], info[3]))

    # ----------------------------------------------------------- autotune
    def _time_class(self, N, epw, reps, warmup):
        items = self._items_for(N, epw)
        dev = torch.from_numpy(items.view(np.uint8).copy()).to(self.device)
        A, B, _ = rk_coefs()
        st = self._st()
        for _ in range(warmup):
            self._launch(N, epw, dev.data_ptr(), items.size, self.parity, A[1], B[0], st)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            self._launch(N, epw, dev.data_ptr(), items.size, self.parity, A[1], B[0], st)
        e1.record()
        e1.synchronize()
        return 1e3 * e0.elapsed_time(e1) / reps

    def _registers(self):
        """The LSRK registers (restored after autotuning)."""
        return {sid: v.rhs for sid, v in self.solves.items()}

    def autotune(self, reps=20, warmup=3, only=None):
        """Time each elements-per-workgroup candidate per N on this batch's
        solves and keep the fastest.  The LSRK registers are restored after;
        buffer [parity] is only read."""
        if self.dirty:
            self._build_tables()
        regs = self._registers()
        saved = {sid: r.clone() for sid, r in regs.items()}
        Ns = sorted({v.N for v in self.solves.values()}) if only is None else only
        for N in Ns:
            res = {}
            for e, _ in self._variants(N):
                res[e] = self._time_class(N, e, reps, warmup)
            self.tuning[N] = res
            self.epw[N] = min(res, key=res.get)
        for sid, r in regs.items():
            r.copy_(saved[sid])
        self.check()
        self._build_tables()
        return dict(self.epw)

    # ------------------------------------------------------------ figures
    def ndof(self):
        return sum(v.setup.ndof() for v in self.solves.values())

    def bytes_per_stage(self):
        return sum(v.nelem * self._bytes_per_element_stage(v) for v in self.solves.values())


class Batch(_BatchBase):
    """Batched drop-in form: each solve keeps the reference arrays (FP64), or
    FP32 copies of Q, rhs, vgeo and sgeo (prec = FP32, shallow water)."""

    _rowtype = _lib.BatchSolve
    _has_single_launch = True

    def __init__(self, eqn, dim, prec=FP64, device="cuda", prepared=False):
        """prepared: every solve runs over the prepared arrays of
        canary_prepare_geometry (metric terms and MJ, normal and sMJ, one
        face code per face) instead of the reference arrays: fewer bytes per
        element-stage, and in the exact order the same bits."""
        if (eqn, dim) not in _NS:
            raise ValueError("unknown (eqn, dim) = (%r, %r)" % (eqn, dim))
        if prec not in (FP64, FP64_FMA, FP32):
            raise ValueError("prec must be FP64, FP64_FMA or FP32")
        super().__init__(device)
        self.eqn, self.dim, self.prec = eqn, dim, prec
        self.prepared = bool(prepared)
        self.rdtype = torch.float32 if prec == FP32 else torch.float64

    def _variants(self, N):
        return variants(self.eqn, self.dim, self.prec, N)

    def _launch_raw(self, N, epw, table, items_ptr, count, parity, rka, rkb, stream):
        if N == 0:  # the single-launch class
            fn = (lib().canary_batch_stage_all_prepared if self.prepared
                  else lib().canary_batch_stage_all)
            return fn(self.eqn, self.dim, self.prec, _ptr(table), items_ptr, count, parity, rka,
                      rkb, stream)
        fn = lib().canary_batch_stage_prepared if self.prepared else lib().canary_batch_stage
        return fn(self.eqn, self.dim, self.prec, N, epw, _ptr(table), items_ptr, count, parity,
                  rka, rkb, stream)

    def _all_variant(self, N):
        e, t = C.c_int(), C.c_int()
        rc = lib().canary_batch_all_variant(self.eqn, self.dim, self.prec, N, C.byref(e),
                                            C.byref(t))
        if rc:
            raise ValueError("order N=%d has no single-launch variant (error %d)" % (N, rc))
        return e.value

    def _run_kind(self):
        if self.prepared:
            kind = _lib.BATCH_PREPARED_ALL if self.single_launch else _lib.BATCH_PREPARED
        else:
            kind = _lib.BATCH_DROPIN_ALL if self.single_launch else _lib.BATCH_DROPIN
        return kind, self.eqn, self.dim, self.prec

    def _word(self):
        return 4 if self.prec == FP32 else 8

    def _bytes_per_element_stage(self, v):
        return bytes_per_element_stage(self.eqn, self.dim, v.N, self._word(),
                                       prepared=self.prepared)

    def add(self, s, dt, Q=None, advection=False):
        """Upload one solve (a problems.Setup; Q defaults to s.Q).  Returns its
        id.  May be called between steps of a running batch: the state goes
        into the current parity buffer and the work table is rebuilt by the
        next finalize().  advection (shallow water): the solve keeps its
        velocity through each update, as updatesolution's flag does."""
        if s.eqn != self.eqn or s.dim != self.dim:
            raise ValueError("solve (eqn %d, dim %d) does not match the batch" % (s.eqn, s.dim))
        if advection and s.eqn != _lib.SWE:
            raise ValueError("advection is a shallow-water mode")
        if not self._variants(s.N):
            raise _lib.CanaryError(_lib.ERR_UNSUPPORTED, "batched form at N=%d" % s.N)
        if Q is None:
            Q = s.Q if s.Q is not None else getattr(s, "Qd", None)
        nq = None if Q is None else (Q.numel() if isinstance(Q, torch.Tensor) else Q.size)
        if nq != s.Np * s.nstate * s.nelem:
            raise ValueError("Q must hold (Np, nstate, nelem) values")
        v = _Solve()
        v.setup, v.dt, v.N, v.nelem = s, float(dt), s.N, s.nelem
        dv, rd = self.device, self.rdtype
        q = Q.to(dv, rd).clone() if isinstance(Q, torch.Tensor) else _dev(Q, dv, rd)
        v.Q = [q, torch.empty_like(q)] if self.parity == 0 else [torch.empty_like(q), q]
        v.rhs = torch.zeros_like(q)
        arr = device_arrays(s, dv)
        v.D = arr["D"]
        if self.prepared:
            from . import kernels as K
            prep = K.prepare_geometry(s.dim, s.N, arr["vgeo"], arr["sgeo"], arr["vmapM"],
                                      arr["vmapP"], arr["bndy"], s.nelem)
            K.check_errors()
            v.vgeo, v.sgeo = prep["pvgeo"].to(rd), prep["psgeo"].to(rd)
            v.pface = prep["pface"]
            v.vmapM = v.vmapP = v.bndy = None
            del prep
        else:
            v.vgeo = arr["vgeo"].to(rd)
            v.sgeo = arr["sgeo"].to(rd)
            v.vmapM, v.vmapP, v.bndy = arr["vmapM"], arr["vmapP"], arr["bndy"]
            v.pface = None
        v.phys = s.phys
        v.advection = bool(advection)
        if s.eqn == _lib.EULER:  # the pressure's pow table, before any capture
            lib().canary_pow_register(s.phys.c_p / s.phys.c_v, None)
        return self._register(v)

    def _fill_row(self, t, v):
        t.Q[0], t.Q[1] = v.Q[0].data_ptr(), v.Q[1].data_ptr()
        t.rhs = v.rhs.data_ptr()
        t.vgeo, t.sgeo, t.D = v.vgeo.data_ptr(), v.sgeo.data_ptr(), v.D.data_ptr()
        if v.pface is not None:
            t.vmapM = t.vmapP = t.elemtobndy = None
            t.pface = v.pface.data_ptr()
        else:
            t.vmapM, t.vmapP = v.vmapM.data_ptr(), v.vmapP.data_ptr()
            t.elemtobndy = v.bndy.data_ptr()
            t.pface = None
        t.nreal = t.nelem_total = v.nelem
        t.dt = v.dt
        t.phys = v.phys
        t.N = v.N
        t.advection = 1 if v.advection else 0

    def footprint(self, s):
        """Device bytes a solve of this setup holds in the batch."""
        w = self._word()
        d = s.dim
        Np, Nfp, nf, ne = s.Np, (s.N + 1) ** (d - 1), 2 * d, s.nelem
        nq = Np * s.nstate * ne
        if self.prepared:
            return (w * (3 * nq + (d * d + 1) * Np * ne + (d + 1) * Nfp * nf * ne)
                    + 4 * nf * ne + 8 * (s.N + 1) ** 2)
        nvg, nsg = (14, 5) if d == 3 else (8, 4)
        return (w * (3 * nq + nvg * Np * ne + nsg * Nfp * nf * ne)
                + 8 * (2 * Nfp * nf * ne + nf * ne + (s.N + 1) ** 2))

    def state_tensor(self, sid):
        """Current state of a solve on the device (float64, flat)."""
        return self.solves[sid].Q[self.parity].to(torch.float64)

    def state(self, sid):
        s = self.solves[sid].setup
        q = self.state_tensor(sid).cpu().numpy()
        return q.reshape((s.Np, s.nstate, s.nelem), order="F")


class Batch32(_BatchBase):
    """Batched precision form of the Euler equations (2D or 3D): each solve
    is an FP32 perturbation about its FP64 background at rest (setup.Q0)."""

    _rowtype = _lib.P32Solve

    def __init__(self, dim, device="cuda"):
        if dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        super().__init__(device)
        self.dim = dim
        self.eqn = EULER

    def _variants(self, N):
        return variants32(self.dim, N)

    def _launch_raw(self, N, epw, table, items_ptr, count, parity, rka, rkb, stream):
        return lib().canary_p32_batch_stage(self.dim, N, epw, _ptr(table), items_ptr, count,
                                            parity, rka, rkb, stream)

    def _run_kind(self):
        return _lib.BATCH_P32, EULER, self.dim, 0

    def _word(self):
        return 4

    def _bytes_per_element_stage(self, v):
        return Perturb32.bytes_per_element_stage(v.N, self.dim)

    def add(self, s, dt, Q=None, Q0=None):
        """Upload one solve: background Q0 (default s.Q0) and state Q (default
        s.Q).  Returns its id."""
        if s.eqn != EULER or s.dim != self.dim:
            raise ValueError("solve (eqn %d, dim %d) does not match the batch" % (s.eqn, s.dim))
        if not self._variants(s.N):
            raise _lib.CanaryError(_lib.ERR_UNSUPPORTED, "batched precision form at N=%d" % s.N)
        if Q is None:
            Q = s.Q if s.Q is not None else getattr(s, "Qd", None)
        if Q0 is None:
            Q0 = s.Q0 if s.Q0 is not None else getattr(s, "Q0d", None)
        if Q0 is None:
            raise ValueError("the precision form needs a background state (setup.Q0)")
        pf = Perturb32(s, Q0, Q, device=self.device)
        v = _Solve()
        v.setup, v.dt, v.N, v.nelem = s, float(dt), s.N, s.nelem
        v.pf = pf
        v.Q = [pf.q, pf.q2] if self.parity == 0 else [pf.q2, pf.q]
        v.rhs = pf.rhs
        v.phys = s.phys
        return self._register(v)

    def _fill_row(self, t, v):
        pf = v.pf
        t.q[0], t.q[1] = v.Q[0].data_ptr(), v.Q[1].data_ptr()
        t.rhs = v.rhs.data_ptr()
        t.bg32, t.vg, t.sg = pf.bg32.data_ptr(), pf.vg.data_ptr(), pf.sg.data_ptr()
        t.nbr, t.D = pf.nbr.data_ptr(), pf.D.data_ptr()
        t.nreal = v.nelem
        t.dt = v.dt
        t.phys = v.phys
        t.N = v.N
        t.dim = self.dim

    def footprint(self, s):
        """Device bytes a solve of this setup holds in the batch: q', its
        ping-pong twin and dq, the metric terms and MJ, the face records
        (FP32), the background in FP64 and FP32, the face codes and D."""
        d = s.dim
        Np, Nfp, nf, ns, ne = s.Np, (s.N + 1) ** (d - 1), 2 * d, d + 2, s.nelem
        return ne * (4 * (3 * ns * Np + (d * d + 1) * Np + 4 * nf * Nfp + 3 * Np + nf)
                     + 8 * 3 * Np) + 8 * s.D.size

    def state_tensor(self, sid):
        v = self.solves[sid]
        v.pf.q = v.Q[self.parity]
        return v.pf.full_state()

    def state(self, sid):
        s = self.solves[sid].setup
        return self.state_tensor(sid).cpu().numpy().reshape((s.Np, s.nstate, s.nelem), order="F")
