"""Benchmark: LSRK54 time stepping of the 3D rising-bubble Euler problem,
32^3 hexahedra, N = 4 (4.1M nodes x 5 states), FP64, one MI355X.

A step is one LSRK54 step (5 stages).  Each stage is one launch of the fused
single-pass kernel (volume + faces + update, Q ping-pong) in contracted FP64
(multiply-adds fused); the same kernel in the reference evaluation order and
the three-kernel drop-in path (volumerhs, facerhs, updatesolution) are timed
beside it.

Metric: GDOF-stage/s = nelem * Np * nstate * stages / seconds / 1e9, with
all arrays resident in HBM before the timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def load_package():
    if "canary_amd" in sys.modules:
        return sys.modules["canary_amd"]
    pkg = os.path.join(ROOT, "canary.jl_amd")
    spec = importlib.util.spec_from_file_location("canary_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["canary_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def algorithmic_bytes(dim, N, nstate=None, word=8):
    """Bytes one element moves per stage when every array the three reference
    kernels touch is counted once: Q r+w, dQ r+w, the metric terms + MJ + MJI,
    normals + sMJ, vmapM + vmapP (Int64) and elemtobndy."""
    Nq = N + 1
    Np, Nfp, nface = Nq ** dim, Nq ** (dim - 1), 2 * dim
    ns = nstate if nstate is not None else dim + 2
    nvol = dim * dim + 2
    nsurf = dim + 1
    return word * (4 * ns * Np + nvol * Np + nface * Nfp * nsurf) + 2 * nface * Nfp * 8 + nface * 8


def prepared_bytes(dim, N, word=8):
    """Bytes one element moves per stage over the prepared arrays: Q r+w,
    dQ r+w, the metric terms + MJ, normals + sMJ, one int32 face code per
    face (MJI, vmapM, vmapP and elemtobndy are not read)."""
    Nq = N + 1
    Np, Nfp, nface, ns = Nq ** dim, Nq ** (dim - 1), 2 * dim, dim + 2
    return word * (4 * ns * Np + (dim * dim + 1) * Np + nface * Nfp * (dim + 1)) + 4 * nface


def swe_bytes(N):
    """Shallow-water bytes per element-stage by the same rule: Q read with b
    (4 fields), written without it (3), dQ r+w (3 fields), 4 metric terms +
    MJ + MJI, normals + sMJ, vmapM + vmapP and elemtobndy."""
    Np, Nfp = (N + 1) ** 2, N + 1
    return 8 * ((4 + 3 + 2 * 3) * Np + 6 * Np + 4 * Nfp * 3) + 2 * 4 * Nfp * 8 + 4 * 8


def total_energy_bytes(N):
    """Bytes per element-stage of the 2D total-energy equations by the B3
    rule: Q and dQ r+w (4 states), the metric terms + MJ + MJI + y (7
    vgeo fields), normals + sMJ, vmapM + vmapP and elemtobndy."""
    Np, Nfp = (N + 1) ** 2, N + 1
    return 8 * (16 * Np + 7 * Np + 4 * Nfp * 3) + 2 * 4 * Nfp * 8 + 4 * 8


def matrix_core_form(s, K, pnative, timed_own, Q0, q_measured, vgeo, D, steps, warmup,
                     bytes_launch):
    """The derivative contractions on the matrix cores (FP64_MFMA) against
    the same loop and kernel with them on the vector ALUs (FP64_FMA): the
    prepared-array loop of the measured value, and the volume term alone
    (canary_volumerhs, µs per launch by HIP events over 20 launches)."""
    import torch
    from canary_amd._lib import FP64_FMA, FP64_MFMA
    stages = 5 * steps
    dof = s.nelem * s.Np * 5
    out = {}
    qm, qm2, rm = Q0.clone(), torch.empty_like(Q0), torch.zeros_like(Q0)
    torch.cuda.synchronize()
    qm, qm2 = pnative(qm, qm2, rm, warmup, prec=FP64_MFMA)
    hold = {}

    def go():
        hold["q"] = pnative(qm, qm2, rm, steps, prec=FP64_MFMA)

    t, _ = timed_own(go)
    K.check_errors()
    qf = hold["q"][0]
    out["prepared_loop_gdof_stage_s"] = round(dof * stages / t / 1e9, 3)
    out["prepared_loop_us_per_stage"] = round(t / stages * 1e6, 2)
    out["prepared_loop_frac"] = round(bytes_launch / (t / stages) / 1e9 / HBM_PEAK_GBS, 4)
    out["vs_vector_alu_state_rel_l2"] = float(
        "%.3e" % (torch.linalg.vector_norm(qf - q_measured)
                  / torch.linalg.vector_norm(q_measured)).item())
    del qm, qm2, rm, hold, qf
    rv = torch.zeros_like(Q0)
    for name, prec in (("vector_alu", FP64_FMA), ("matrix_core", FP64_MFMA)):
        for _ in range(3):
            K.volumerhs(3, s.N, rv, Q0, vgeo, D, s.nelem, phys=s.phys, prec=prec)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(20):
            K.volumerhs(3, s.N, rv, Q0, vgeo, D, s.nelem, phys=s.phys, prec=prec)
        e1.record()
        torch.cuda.synchronize()
        K.check_errors()
        out["volumerhs_%s_us" % name] = round(e0.elapsed_time(e1) / 20 * 1e3, 2)
    del rv
    return out


def two_d_form(ne, N, steps, warmup):
    """Single large 2D solves through the library's LSRK54 loop (fused
    stage, launched from C): the Euler set2c isentropic vortex, shallow
    water with a bump and the total-energy example's rising bubble, ne^2
    quads, contracted and exact-order arithmetic, over the reference arrays
    and (not the total-energy form, whose pressure reads node heights the
    prepared arrays do not hold) over the prepared ones."""
    import numpy as np
    import torch
    import canary_amd.kernels as K
    import canary_amd.problems as P
    from canary_amd._lib import FP64, FP64_FMA, PATH_FUSED, SWE, EULER, EULER_TE
    from canary_amd.batch import bytes_per_element_stage
    from canary_amd.streams import OwnedStream

    dev = torch.device("cuda:0")

    def d(a):
        return torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order="F"))).to(dev)

    own = OwnedStream(dev)
    out = {}
    for name, make in (("euler2d_vortex", lambda: P.vortex2d(Ne=ne, N=N)),
                       ("swe2d_bump", lambda: P.swe2d(Ne=ne, N=N)),
                       ("euler2d_total_energy_bubble", lambda: P.bubble2d_te(Ne=ne, N=N))):
        s = make()
        eqn = s.eqn
        if eqn == EULER:
            dt = 0.5 * P.cfl(s, s.Q) / N ** np.sqrt(2)
        elif eqn == SWE:
            dt = P.swe_courant(s, s.Q)
        else:
            dt = P.te_dt(s, s.Q)
        vgeo, sgeo, D = d(s.vgeo), d(s.sgeo), d(s.D)
        vM, vP, bnd = d(s.vmapM), d(s.vmapP), d(s.elemtobndy)
        Q0 = d(s.Q)
        dof = s.nelem * s.Np * (3 if eqn == SWE else 4)
        nbytes = {SWE: swe_bytes(N), EULER: algorithmic_bytes(2, N),
                  EULER_TE: total_energy_bytes(N)}[eqn] * s.nelem
        res = {"workload": "%s, %d^2 quads, N=%d" % (s.label, ne, N), "dof": dof,
               "bytes_per_element_stage": nbytes // s.nelem}
        forms = [("fp64_fma", FP64_FMA, False), ("fp64", FP64, False)]
        if eqn != EULER_TE:
            prep = K.prepare_geometry(2, N, vgeo, sgeo, vM, vP, bnd, s.nelem)
            K.check_errors()
            pbytes = bytes_per_element_stage(eqn, 2, N, prepared=True)
            forms += [("fp64_fma_prepared", FP64_FMA, True), ("fp64_prepared", FP64, True)]
        for tag, prec, prepared in forms:
            q, q2, r = Q0.clone(), torch.empty_like(Q0), torch.zeros_like(Q0)
            torch.cuda.synchronize()

            def loop(n):
                with torch.cuda.stream(own.stream):
                    if prepared:
                        return K.lsrk54_prepared(2, N, q, q2, r, prep, D, s.nelem, dt, n, eqn=eqn,
                                                 phys=s.phys, prec=prec)
                    return K.lsrk54(2, N, q, r, vgeo, sgeo, D, s.nelem, vM, vP, bnd, dt, n,
                                    eqn=eqn, phys=s.phys, path=PATH_FUSED, prec=prec, Qalt=q2)
            loop(warmup)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(own.stream)
            fin = loop(steps)
            e1.record(own.stream)
            torch.cuda.synchronize()
            K.check_errors()
            if not bool(torch.isfinite(fin).all()):
                raise SystemExit("non-finite 2D state")
            t = e0.elapsed_time(e1) / 1e3 / (5 * steps)
            res[tag] = {"us_per_stage": round(t * 1e6, 2),
                        "gdof_stage_s": round(dof / t / 1e9, 3),
                        "hbm_frac": round(nbytes / t / 1e9 / HBM_PEAK_GBS, 4)}
            if prepared:
                res[tag]["hbm_frac_on_own_bytes"] = round(pbytes * s.nelem / t / 1e9
                                                          / HBM_PEAK_GBS, 4)
            del q, q2, r, fin
        if eqn != EULER_TE:
            res["bytes_per_element_stage_prepared"] = pbytes
            del prep
        out[name] = res
        print("2D form %s: %s" % (name, res["fp64_fma"]), file=sys.stderr, flush=True)
        del vgeo, sgeo, D, vM, vP, bnd, Q0, s
        torch.cuda.empty_cache()
    return out


def sweep_form(steps=10, warmup=10):
    """The library LSRK54 loop over the contracted fused stage on the 3D
    bubble for N = 2..7, each at about 20 M DOF, over the reference arrays
    and over the prepared ones (the measured path), with the roofline
    fraction of both on the reference-array byte rule B3(N); and the prepared
    loop with the derivative contractions on the matrix cores (FP64_MFMA)."""
    import numpy as np
    import torch
    import canary_amd.kernels as K
    import canary_amd.problems as P
    from canary_amd._lib import FP64_FMA, FP64_MFMA, PATH_FUSED
    from canary_amd.streams import OwnedStream

    dev = torch.device("cuda:0")

    def d(a):
        return torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order="F"))).to(dev)

    own = OwnedStream(dev)
    out = []
    for N, ne in ((2, 53), (3, 40), (4, 32), (5, 27), (6, 23), (7, 20)):
        s = P.bubble3d(Ne=ne, N=N, gravity=9.81, dtheta=2.0, jitter=1e-3, seed=1)
        dt = P.cfl(s, s.Q) / N ** np.sqrt(2)
        vgeo, sgeo, D = d(s.vgeo), d(s.sgeo), d(s.D)
        vM, vP, bnd = d(s.vmapM), d(s.vmapP), d(s.elemtobndy)
        q, q2, r = d(s.Q), torch.empty(s.Q.size, dtype=torch.float64, device=dev), \
            torch.zeros(s.Q.size, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()

        prep = K.prepare_geometry(3, N, vgeo, sgeo, vM, vP, bnd, s.nelem)
        K.check_errors()

        def loop(n, prepared, prec=FP64_FMA):
            with torch.cuda.stream(own.stream):
                if prepared:
                    return K.lsrk54_prepared(3, N, q, q2, r, prep, D, s.nelem, dt, n,
                                             phys=s.phys, prec=prec)
                return K.lsrk54(3, N, q, r, vgeo, sgeo, D, s.nelem, vM, vP, bnd, dt, n,
                                phys=s.phys, path=PATH_FUSED, prec=FP64_FMA, Qalt=q2)
        dof = s.nelem * s.Np * 5
        row = {"N": N, "elements": s.nelem, "dof": dof,
               "bytes_per_element_stage": algorithmic_bytes(3, N),
               "bytes_per_element_stage_prepared": prepared_bytes(3, N)}
        for tag, prepared, prec in (("", False, FP64_FMA), ("prepared_", True, FP64_FMA),
                                    ("prepared_mfma_", True, FP64_MFMA)):
            loop(warmup, prepared, prec)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(own.stream)
            fin = loop(steps, prepared, prec)
            e1.record(own.stream)
            torch.cuda.synchronize()
            K.check_errors()
            if not bool(torch.isfinite(fin).all()):
                raise SystemExit("non-finite state in the sweep at N=%d" % N)
            t = e0.elapsed_time(e1) / 1e3 / (5 * steps)
            row[tag + "us_per_stage"] = round(t * 1e6, 2)
            row[tag + "gdof_stage_s"] = round(dof / t / 1e9, 3)
            row[tag + "hbm_frac"] = round(algorithmic_bytes(3, N) * s.nelem / t / 1e9
                                          / HBM_PEAK_GBS, 4)
            del fin
        out.append(row)
        print("sweep N=%d: %s" % (N, out[-1]), file=sys.stderr, flush=True)
        del vgeo, sgeo, D, vM, vP, bnd, q, q2, r, s, prep
        torch.cuda.empty_cache()
    return {"workload": "3D Euler set2c rising bubble, about 20 M DOF per N, contracted fused "
                        "stage through canary_lsrk54", "steps": steps, "warmup": warmup,
            "runs": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--ne", type=int, default=32)
    ap.add_argument("--N", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-elems", type=int, default=32768)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--precision-ne", type=int, default=48,
                    help="elements per direction of the FP32-vs-FP64 precision-form run (0: skip)")
    ap.add_argument("--precision-steps", type=str, default="1,100,1000",
                    help="LSRK steps after which the FP32 form's error is reported")
    ap.add_argument("--two-d-ne", type=int, default=512,
                    help="quads per direction of the single-solve 2D Euler / shallow-water "
                         "runs (0: skip)")
    ap.add_argument("--batched-solves", type=int, default=64,
                    help="shallow-water solves of the batched-form run (0: skip)")
    ap.add_argument("--batched-steps", type=int, default=20)
    ap.add_argument("--served-requests", type=int, default=64,
                    help="requests of the served-form run (0: skip)")
    ap.add_argument("--served-rate", type=float, default=200.0)
    ap.add_argument("--served-graph", action="store_true",
                    help="served form: also serve the stream with step-graph replay")
    ap.add_argument("--form", choices=["batched", "served", "sweep"], default=None,
                    help="run only that added form and print its JSON (used by the main run, "
                         "which measures each added form in a child process)")
    ap.add_argument("--copy-gib", type=float, default=0.0,
                    help="also run one device-to-device copy of this many GiB (untimed), the "
                         "known-size reference for HBM byte counters")
    ap.add_argument("--calib", action="store_true",
                    help="also run two more known-size kernels for the byte counters (untimed): "
                         "a 1 GiB read-only sum and a gather of 2^20 random 128-byte rows")
    ap.add_argument("--summarize-trace", metavar="DIR", default=None,
                    help="print the kernel table of a rocprofv3 --kernel-trace csv run of this "
                         "script (with the same --steps / --ne / --N) and exit")
    ap.add_argument("--summarize-pmc", metavar="DIR", default=None,
                    help="print the per-kernel counter means of a rocprofv3 --pmc csv run of "
                         "this script and exit")
    args = ap.parse_args()
    if args.summarize_trace:
        summarize_trace(args.summarize_trace, args.steps, args.ne, args.N)
        return
    if args.summarize_pmc:
        summarize_pmc(args.summarize_pmc, args.ne, args.N)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build measures the single-GPU kernel; "
              "multi-GPU sharding is described, not run (got --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    if args.form is not None:
        load_package()
        if args.form == "batched":
            res = batched_form(args.batched_solves, args.batched_steps)
        elif args.form == "sweep":
            res = sweep_form()
        else:
            res = served_form(args.served_requests, args.served_rate, graph=args.served_graph)
        print(json.dumps(res), flush=True)
        return

    import numpy as np
    import torch

    pkg = load_package()
    if not pkg.built():
        raise SystemExit("libcanary_dg.so is not built; run python canary.jl_amd/build.py")
    import canary_amd.kernels as K
    import canary_amd.problems as P
    from canary_amd._lib import FP64, FP64_FMA

    t0 = time.time()
    s = P.bubble3d(Ne=args.ne, N=args.N, gravity=9.81, dtheta=2.0, jitter=1e-3, seed=1)
    setup_s = time.time() - t0
    print("setup %s: %d elements, %.1f s" % (s.label, s.nelem, setup_s), file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")

    def d(a):
        return torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order="F"))).to(dev)

    vgeo, sgeo, D = d(s.vgeo), d(s.sgeo), d(s.D)
    vM, vP, bnd = d(s.vmapM), d(s.vmapP), d(s.elemtobndy)
    Q0 = d(s.Q)
    A, B, _ = P.rk_coefs()
    # the reference's time step: cfl(...) / N^sqrt(2) (euler3d_set2c.jl:1104)
    dt = P.cfl(s, s.Q) / s.N ** np.sqrt(2)
    stream = torch.cuda.current_stream()

    def run_fused(q, q2, rhs, nsteps, prec=FP64_FMA, mapM=vM):
        for _ in range(nsteps):
            for st in range(5):
                K.stage_fused(3, s.N, rhs, q, q2, vgeo, sgeo, D, s.nelem, mapM, vP, bnd,
                              A[(st + 1) % 5], B[st], dt, phys=s.phys, prec=prec)
                q, q2 = q2, q
        return q, q2

    def run_plain(q, rhs, nsteps, prec=FP64):
        for _ in range(nsteps):
            for st in range(5):
                K.volumerhs(3, s.N, rhs, q, vgeo, D, s.nelem, phys=s.phys, prec=prec)
                K.facerhs(3, s.N, rhs, q, sgeo, s.nelem, vM, vP, bnd, phys=s.phys, prec=prec)
                K.updatesolution(3, s.N, rhs, q, vgeo, s.nelem, A[(st + 1) % 5], B[st], dt,
                                 prec=prec)

    def timed(fn):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t_wall = time.perf_counter()
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3, time.perf_counter() - t_wall

    if args.copy_gib > 0:
        nwords = int(args.copy_gib * (1 << 30)) // 8
        src = torch.ones(nwords, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    if args.calib:
        # known-size references for the byte counters: a streaming copy made
        # by an elementwise kernel of its own (1 GiB negated into another:
        # 1 GiB read, 1 GiB written), a read-only stream (1 GiB summed:
        # 1 GiB read) and a gather of whole 128-byte rows in random order
        # (2^20 rows of a 128 MiB array, each once, plus the 8 MiB index:
        # 136 MiB read, 128 MiB written)
        x = torch.ones((1 << 30) // 8, dtype=torch.float64, device=dev)
        y = torch.empty_like(x)
        torch.neg(x, out=y)
        float(x.sum())
        del x, y
        rows = torch.ones(1 << 20, 16, dtype=torch.float64, device=dev)
        perm = torch.randperm(1 << 20, device=dev)
        g = rows.index_select(0, perm)
        torch.cuda.synchronize()
        del rows, perm, g

    # measured copy bandwidth of the box (1 GiB device-to-device, read + write)
    cw = (1 << 30) // 8
    csrc = torch.ones(cw, dtype=torch.float64, device=dev)
    cdst = torch.empty_like(csrc)
    cdst.copy_(csrc)
    t_copy, _ = timed(lambda: [cdst.copy_(csrc) for _ in range(10)])
    copy_gbs = 2 * (1 << 30) * 10 / t_copy / 1e9
    del csrc, cdst

    # the measured value: the library's LSRK54 time loop (the reference's
    # lowstorageRK) over the contracted fused stage reading the prepared
    # arrays (canary_prepare_geometry, built once from the reference arrays
    # before the timed region), launched from C on a stream of its own:
    # `warmup` untimed steps, then exactly `steps` timed steps
    from canary_amd.streams import OwnedStream
    from canary_amd._lib import PATH_FUSED
    own = OwnedStream(dev)
    prep = K.prepare_geometry(3, s.N, vgeo, sgeo, vM, vP, bnd, s.nelem)
    K.check_errors()

    def pnative(qa, qb, r, nsteps, prec=FP64_FMA, graph=False):
        """nsteps steps over the prepared arrays from qa; returns (state, spare)."""
        with torch.cuda.stream(own.stream):
            out = K.lsrk54_prepared(3, s.N, qa, qb, r, prep, D, s.nelem, dt, nsteps,
                                    phys=s.phys, prec=prec, graph=graph)
        return (out, qb) if out is qa else (out, qa)

    def native(qa, qb, r, nsteps, mapM=vM, graph=False):
        """nsteps steps from qa; returns (state, spare buffer)."""
        with torch.cuda.stream(own.stream):
            out = K.lsrk54(3, s.N, qa, r, vgeo, sgeo, D, s.nelem, mapM, vP, bnd, dt, nsteps,
                           phys=s.phys, path=PATH_FUSED, prec=FP64_FMA, Qalt=qb, graph=graph)
        return (out, qb) if out is qa else (out, qa)

    def timed_own(fn):
        """HIP events on the loop's own stream around fn."""
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t_wall = time.perf_counter()
        e0.record(own.stream)
        fn()
        e1.record(own.stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3, time.perf_counter() - t_wall

    stages = 5 * args.steps
    dof = s.nelem * s.Np * 5
    bytes_launch = algorithmic_bytes(3, s.N) * s.nelem
    own_bytes = prepared_bytes(3, s.N)
    extra = {}
    # The loops below run in this order: the reference-array loop, the
    # exact-order prepared loop, then the measured one.  The first second of
    # sustained load on a box runs through a clock / power transient (stage
    # times rise by up to a third for some 20 launches, then settle); the
    # measured loop's own warmup and timed steps come after it.
    # The same loop over the reference arrays themselves (canary_lsrk54,
    # contracted fused stage): the call surface the Julia shim binds
    q, q2, rhs = Q0.clone(), torch.empty_like(Q0), torch.zeros_like(Q0)
    torch.cuda.synchronize()
    q, q2 = native(q, q2, rhs, args.warmup)
    K.check_errors()
    holder = {}

    def go():
        holder["q"] = native(q, q2, rhs, args.steps)

    t_ref, _ = timed_own(go)
    K.check_errors()
    qf = holder["q"][0]
    extra["reference_arrays_loop_gdof_stage_s"] = round(dof * stages / t_ref / 1e9, 3)
    extra["reference_arrays_loop_us_per_stage"] = round(t_ref / stages * 1e6, 2)
    extra["reference_arrays_loop_frac"] = round(bytes_launch / (t_ref / stages) / 1e9
                                                / HBM_PEAK_GBS, 4)
    # the exact-order loop over the prepared arrays (bit-identical to the
    # oracle), from the same initial state
    qx_, qx2_, rx_ = Q0.clone(), torch.empty_like(Q0), torch.zeros_like(Q0)
    torch.cuda.synchronize()
    qx_, qx2_ = pnative(qx_, qx2_, rx_, args.warmup, prec=FP64)
    hold_x = {}

    def go_x():
        hold_x["q"] = pnative(qx_, qx2_, rx_, args.steps, prec=FP64)

    t_px, _ = timed_own(go_x)
    K.check_errors()
    extra["prepared_exact_order_gdof_stage_s"] = round(dof * stages / t_px / 1e9, 3)
    extra["prepared_exact_order_us_per_stage"] = round(t_px / stages * 1e6, 2)
    extra["prepared_exact_order_frac"] = round(bytes_launch / (t_px / stages) / 1e9
                                               / HBM_PEAK_GBS, 4)
    del qx_, qx2_, rx_, hold_x
    # the measured value
    qp_, qp2_, rp_ = Q0.clone(), torch.empty_like(Q0), torch.zeros_like(Q0)
    torch.cuda.synchronize()  # made on the current stream, used on the loop's own
    qp_, qp2_ = pnative(qp_, qp2_, rp_, args.warmup)
    K.check_errors()
    hold_p = {}

    def go_p():
        hold_p["q"] = pnative(qp_, qp2_, rp_, args.steps)

    t_fused, wall_fused = timed_own(go_p)
    K.check_errors()
    qpf = hold_p["q"][0]
    if not bool(torch.isfinite(qpf).all()):
        raise SystemExit("non-finite state after the timed steps")
    value = dof * stages / t_fused / 1e9
    launch_s = t_fused / stages
    achieved = bytes_launch / launch_s / 1e9

    extra["prepared_vs_reference_arrays_rel_l2"] = float(
        "%.3e" % (torch.linalg.vector_norm(qpf - qf) / torch.linalg.vector_norm(qf)).item())
    del qp_, qp2_, rp_
    extra["matrix_core_form"] = matrix_core_form(s, K, pnative, timed_own, Q0, qpf, vgeo, D,
                                                 args.steps, args.warmup, bytes_launch)

    # the same loop replaying two captured steps per hipGraph launch (capture
    # and instantiation inside the timed call)
    qg, qg2, rg = Q0.clone(), torch.empty_like(Q0), torch.zeros_like(Q0)
    torch.cuda.synchronize()
    qg, qg2 = native(qg, qg2, rg, args.warmup, graph=True)
    hold_g = {}

    def go_graph():
        hold_g["q"] = native(qg, qg2, rg, args.steps, graph=True)

    t_graph, _ = timed_own(go_graph)
    K.check_errors()
    extra["native_loop_graph_gdof_stage_s"] = round(dof * stages / t_graph / 1e9, 3)
    extra["native_loop_graph_same_bits"] = bool(torch.equal(hold_g["q"][0], qf))
    del qg, qg2, rg, hold_g
    # the same stages launched one by one from Python, from the same initial
    # state: the stage loop the reference writes
    qy, qy2, ry = Q0.clone(), torch.empty_like(Q0), torch.zeros_like(Q0)
    qy, qy2 = run_fused(qy, qy2, ry, args.warmup)
    hold_y = {}

    def go_py():
        hold_y["q"] = run_fused(qy, qy2, ry, args.steps)

    t_py, _ = timed(go_py)
    K.check_errors()
    extra["python_stage_loop_gdof_stage_s"] = round(dof * stages / t_py / 1e9, 3)
    extra["python_stage_loop_same_bits"] = bool(torch.equal(hold_y["q"][0], qf))
    del qy, qy2, ry, hold_y
    # the same stages in the reference evaluation order (bit-identical to the
    # three-kernel path), from the same initial state: its rate, and how far
    # the contracted state is from it after warmup + steps LSRK54 steps
    qe, qe2, re = Q0.clone(), torch.empty_like(Q0), torch.zeros_like(Q0)
    qe, qe2 = run_fused(qe, qe2, re, args.warmup, FP64)
    hold_e = {}

    def go_exact():
        hold_e["q"] = run_fused(qe, qe2, re, args.steps, FP64)

    t_exact, _ = timed(go_exact)
    K.check_errors()
    qx = hold_e["q"][0]
    extra["fused_exact_order_gdof_stage_s"] = round(dof * stages / t_exact / 1e9, 3)
    extra["fused_exact_order_ms_per_stage"] = round(t_exact / stages * 1e3, 4)
    extra["fma_vs_exact_rel_l2_after_%d_steps" % (args.warmup + args.steps)] = float(
        "%.3e" % (torch.linalg.vector_norm(qf - qx) / torch.linalg.vector_norm(qx)).item())
    del qe, qe2, re, hold_e, qx
    # the measured loop with vmapM = NULL (M-side ids from the canonical face
    # mask, the map not read); same bits as with the map
    qc, qc2, rc_ = Q0.clone(), torch.empty_like(Q0), torch.zeros_like(Q0)
    torch.cuda.synchronize()  # made on the current stream, used on the loop's own
    qc, qc2 = native(qc, qc2, rc_, args.warmup, None)
    hold_c = {}

    def go_canon():
        hold_c["q"] = native(qc, qc2, rc_, args.steps, None)

    t_canon, _ = timed_own(go_canon)
    K.check_errors()
    t_c = t_canon / stages
    extra["null_vmapM_gdof_stage_s"] = round(dof * stages / t_canon / 1e9, 3)
    extra["null_vmapM_ms_per_stage"] = round(t_c * 1e3, 4)
    extra["null_vmapM_same_bits"] = bool(torch.equal(hold_c["q"][0], qf))
    # its roofline fraction on the same algorithmic bytes, and on the bytes
    # without vmapM (the stricter minimum, which it does move)
    extra["null_vmapM_frac"] = round(bytes_launch / t_c / 1e9 / HBM_PEAK_GBS, 4)
    extra["null_vmapM_frac_without_vmapM_bytes"] = round(
        (algorithmic_bytes(3, s.N) - 6 * (s.N + 1) ** 2 * 8) * s.nelem / t_c / 1e9 / HBM_PEAK_GBS, 4)
    del qc, qc2, rc_, hold_c, own
    # one GPU's share of the mesh split over 8 ranks (contiguous Hilbert
    # chunks, ghosts appended): the stage at that real shape, and the halo
    # pack (fillsendQ) + unpack (transferrecvQ) a stage would add; the
    # exchange itself is not run (one GPU)
    extra["rank_share_of_8"] = rank_share_form(s, K, P, dev, dt, args.steps, FP64_FMA)
    if not args.no_plain:
        qp, rp = Q0.clone(), torch.zeros_like(Q0)
        run_plain(qp, rp, 1)
        t_plain, _ = timed(lambda: run_plain(qp, rp, max(1, args.steps // 2)))
        K.check_errors()
        st_plain = 5 * max(1, args.steps // 2)
        extra["plain_three_kernel_gdof_stage_s"] = round(dof * st_plain / t_plain / 1e9, 3)
        extra["plain_ms_per_stage"] = round(t_plain / st_plain * 1e3, 4)
        # the same three kernels in the contracted form (CANARY_FP64_FMA)
        qc3, rc3 = Q0.clone(), torch.zeros_like(Q0)
        run_plain(qc3, rc3, 1, FP64_FMA)
        t_c3, _ = timed(lambda: run_plain(qc3, rc3, max(1, args.steps // 2), FP64_FMA))
        K.check_errors()
        extra["plain_three_kernel_contracted_gdof_stage_s"] = round(dof * st_plain / t_c3 / 1e9, 3)
        del qc3, rc3
        # volumerhs, then facerhs + updatesolution fused (Q ping-pong)
        q2p = torch.empty_like(qp)

        def run_two(nsteps):
            a, b = qp, q2p
            for _ in range(nsteps):
                for st in range(5):
                    K.volumerhs(3, s.N, rp, a, vgeo, D, s.nelem, phys=s.phys)
                    K.facerhs_update(3, s.N, rp, a, b, vgeo, sgeo, s.nelem, vM, vP, bnd,
                                     A[(st + 1) % 5], B[st], dt, phys=s.phys)
                    a, b = b, a
        run_two(1)
        t_two, _ = timed(lambda: run_two(max(1, args.steps // 2)))
        K.check_errors()
        extra["volume_then_face_update_gdof_stage_s"] = round(dof * st_plain / t_two / 1e9, 3)
        del qp, q2p, rp

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(s, args.cpu_sample_elems)
    if args.precision_ne > 0:
        del q, q2, rhs, vgeo, sgeo, vM, vP, bnd, Q0, holder, qf, prep, hold_p, qpf
        torch.cuda.empty_cache()
        extra["precision_form"] = precision_form(args.precision_ne, s.N, args.precision_steps)
    torch.cuda.empty_cache()
    if args.two_d_ne > 0:
        extra["two_d_single_solves"] = two_d_form(args.two_d_ne, s.N, args.steps, args.warmup)
    if args.batched_solves > 0:
        extra["batched_form"] = form_child("batched", ["--batched-solves", str(args.batched_solves),
                                                       "--batched-steps", str(args.batched_steps)])
    if args.served_requests > 0:
        extra["served_form"] = form_child("served", ["--served-requests", str(args.served_requests),
                                                     "--served-rate", str(args.served_rate)])

    traffic = measured_traffic("prepared_euler3d_N%d_%d" % (s.N, args.ne))
    line = {
        "metric": "GDOF-stage/s 3D Euler N=4 FP64 (LSRK54 stages, fused single-pass kernel over "
                  "prepared geometry, multiply-adds fused, library time loop)",
        "value": round(value, 3),
        "unit": "GDOF-stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_fused / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded rising thermal bubble initial state",
        "config": {"workload": "3D Euler set2c rising bubble, [0,1000]^3 m, %d^3 hexes, N=%d, "
                               "no-flux walls, g=9.81, +2 K bubble, 1e-3 K jitter, LSRK54"
                               % (args.ne, s.N),
                   "nelem": s.nelem, "nodes": s.nelem * s.Np, "dof": dof,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": "stage_prep_kernel<Euler3DFma,5,1,1,false> (fused volume+face+update "
                               "over the prepared arrays)",
                     "bytes_per_element_stage": algorithmic_bytes(3, s.N),
                     "bytes_per_element_stage_own": own_bytes,
                     "frac_on_own_bytes": round(own_bytes * s.nelem / launch_s / 1e9
                                                / HBM_PEAK_GBS, 4),
                     "avg_launch_us": round(launch_s * 1e6, 2),
                     "measured_copy_gbs": round(copy_gbs, 1),
                     "frac_of_measured_copy": round(achieved / copy_gbs, 4)},
        "cpu_baseline": cpu,
        "wall_s_timed_region": round(wall_fused, 4),
        **extra,
    }
    print(json.dumps(line), flush=True)


def rank_share_form(s, K, P, dev, dt, steps, prec, nranks=8, rank=0):
    import numpy as np
    import torch
    sh = P.rank_share(s, nranks, rank)
    Np, ns = s.Np, s.nstate

    def d(a):
        return torch.from_numpy(np.ascontiguousarray(np.asarray(a).reshape(-1, order="F"))).to(dev)

    vg, sg, D = d(sh.vgeo), d(sh.sgeo), d(s.D)
    vM, vP, bd = d(sh.vmapM), d(sh.vmapP), d(sh.elemtobndy)
    q = d(sh.local_state(s.Q))
    q2, rhs = torch.empty_like(q), torch.zeros(Np * ns * sh.nreal, dtype=torch.float64, device=dev)
    se = d(sh.sendelems)
    sendQ = torch.empty(Np * ns * sh.sendelems.size, dtype=torch.float64, device=dev)
    recvQ = q.view(-1, Np * ns)[sh.nreal:].clone().reshape(-1)
    sn, rn = d(sh.send_nodes), d(sh.recv_nodes)
    sendT = torch.empty(sh.send_nodes.size * ns, dtype=torch.float64, device=dev)
    recvT = torch.zeros(sh.recv_nodes.size * ns, dtype=torch.float64, device=dev)
    A, B, _ = P.rk_coefs()

    def run(n, halo):
        a, b = q, q2
        for _ in range(n):
            for k in range(5):
                if halo == "elements":
                    K.fillsendQ(3, s.N, ns, sendQ, a, se, sh.nelem_total)
                    K.transferrecvQ(3, s.N, ns, a, recvQ, sh.nreal)
                elif halo == "traces":
                    K.pack_nodes(3, s.N, ns, sendT, a, sn, sh.nelem_total)
                    K.unpack_nodes(3, s.N, ns, a, recvT, rn, sh.nelem_total)
                K.stage_fused(3, s.N, rhs, a, b, vg, sg, D, sh.nreal, vM, vP, bd,
                              A[(k + 1) % 5], B[k], dt, phys=s.phys, nelem_total=sh.nelem_total,
                              prec=prec)
                a, b = b, a

    def timed(fn):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3

    run(1, None)
    t = timed(lambda: run(steps, None)) / (5 * steps)
    th = timed(lambda: run(steps, "elements")) / (5 * steps)
    tt = timed(lambda: run(steps, "traces")) / (5 * steps)
    K.check_errors()
    dof = sh.nreal * Np * ns
    return {"ranks": nranks, "rank": rank, "real_elements": int(sh.nreal),
            "ghost_elements": int(sh.nelem_total - sh.nreal),
            "send_elements": int(sh.sendelems.size),
            "stage_us": round(t * 1e6, 2), "gdof_stage_s": round(dof / t / 1e9, 3),
            "stage_with_pack_unpack_us": round(th * 1e6, 2),
            "halo_bytes_per_stage_whole_elements": int((sh.sendelems.size + sh.nelem_total
                                                        - sh.nreal) * Np * ns * 8),
            "stage_with_trace_pack_unpack_us": round(tt * 1e6, 2),
            "halo_bytes_per_stage_face_traces": int((sh.send_nodes.size + sh.recv_nodes.size)
                                                    * ns * 8)}


def precision_form(ne, N, steps_csv, timed_steps=10):
    """FP32 perturbation form against the FP64 fused path on the same bubble:
    time per stage of both, and the relative L2 error of rho, rho*u, rho*w and
    rho*theta of the FP32 form after the given numbers of LSRK54 steps."""
    import numpy as np
    import torch
    import canary_amd.kernels as K
    import canary_amd.problems as P
    from canary_amd._lib import FP64, FP64_FMA
    from canary_amd.perturb import Perturb32

    t0 = time.time()
    s = P.bubble3d(Ne=ne, N=N, gravity=9.81, dtheta=2.0, jitter=1e-3, seed=1)
    Q0 = P.bubble_state(s, dtheta=0.0, jitter=0.0)
    dt = P.cfl(s, s.Q) / N ** np.sqrt(2)
    print("setup precision form %s: %.1f s" % (s.label, time.time() - t0), file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")

    def d(a):
        return torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order="F"))).to(dev)

    vgeo, sgeo, D = d(s.vgeo), d(s.sgeo), d(s.D)
    vM, vP, bnd = d(s.vmapM), d(s.vmapP), d(s.elemtobndy)
    Qinit = d(s.Q)
    A, B, _ = P.rk_coefs()
    stream = torch.cuda.current_stream()
    st = {"q": Qinit.clone(), "q2": torch.empty_like(Qinit), "rhs": torch.zeros_like(Qinit)}

    def run64(nsteps, prec=FP64):
        for _ in range(nsteps):
            for k in range(5):
                K.stage_fused(3, N, st["rhs"], st["q"], st["q2"], vgeo, sgeo, D, s.nelem, vM, vP,
                              bnd, A[(k + 1) % 5], B[k], dt, phys=s.phys, prec=prec)
                st["q"], st["q2"] = st["q2"], st["q"]

    def time_fma():
        # the contracted FP64 stage on a scratch copy (the error reference
        # state st stays on the exact-order path), after one untimed step
        saved = {k: v.clone() for k, v in st.items()}
        run64(1, FP64_FMA)
        t = timed(lambda: run64(timed_steps, FP64_FMA)) / (5 * timed_steps)
        for k in st:
            st[k].copy_(saved[k])
        return t

    pf = Perturb32(s, Q0, s.Q)
    pfm = Perturb32(s, Q0, s.Q, mixed=True)

    def timed(fn):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3

    marks = sorted(int(v) for v in steps_csv.split(","))
    tp = {}
    errors, errors_m = {}, {}
    done = 0
    t64 = t32 = t32m = None

    def errs(p):
        a = p.full_state().cpu().numpy().reshape(s.Q.shape, order="F")
        b = st["q"].cpu().numpy().reshape(s.Q.shape, order="F")
        return {nm: float("%.3e" % (np.linalg.norm(a[:, k] - b[:, k]) / np.linalg.norm(b[:, k])))
                for k, nm in ((3, "rho"), (0, "rhou"), (2, "rhow"), (4, "rhotheta"))}

    prep = K.prepare_geometry(3, N, vgeo, sgeo, vM, vP, bnd, s.nelem)
    K.check_errors()

    def time_prepared(prec):
        # the library loop over the prepared arrays (the single-solve
        # benchmark's measured path) on scratch copies of the current state
        qa, qb, r = st["q"].clone(), torch.empty_like(st["q"]), st["rhs"].clone()
        K.lsrk54_prepared(3, N, qa, qb, r, prep, D, s.nelem, dt, 1, phys=s.phys, prec=prec)
        t = timed(lambda: K.lsrk54_prepared(3, N, qa, qb, r, prep, D, s.nelem, dt, timed_steps,
                                            phys=s.phys, prec=prec)) / (5 * timed_steps)
        K.check_errors()
        del qa, qb, r
        return t

    def time_all():
        t64f = time_fma()
        t64 = timed(lambda: run64(timed_steps)) / (5 * timed_steps)
        t32 = timed(lambda: pf.run(dt, timed_steps)) / (5 * timed_steps)
        t32m = timed(lambda: pfm.run(dt, timed_steps)) / (5 * timed_steps)
        tp["fma"] = time_prepared(FP64_FMA)
        tp["exact"] = time_prepared(FP64)
        return t64f, t64, t32, t32m

    for m in marks:
        n = m - done
        if t64 is None and n >= timed_steps:
            t64f, t64, t32, t32m = time_all()
            n -= timed_steps
        run64(n)
        pf.run(dt, n)
        pfm.run(dt, n)
        done = m
        K.check_errors()
        pf.check()
        pfm.check()
        errors[str(m)] = errs(pf)
        errors_m[str(m)] = errs(pfm)
    if t64 is None:
        run64(1)
        pf.run(dt, 1)
        pfm.run(dt, 1)
        t64f, t64, t32, t32m = time_all()
    dof = s.nelem * s.Np * 5
    b32 = Perturb32.bytes_per_element_stage(N)
    b32m = Perturb32.bytes_per_element_stage(N, mixed=True)
    b64 = algorithmic_bytes(3, N)
    return {"workload": "3D bubble %d^3 hexes N=%d, FP32 perturbation about an FP64 hydrostatic "
                        "background vs FP64 fused stage" % (ne, N),
            "fp64_ms_per_stage": round(t64 * 1e3, 4), "fp32_ms_per_stage": round(t32 * 1e3, 4),
            "fp64_fma_ms_per_stage": round(t64f * 1e3, 4),
            "speedup_fp32_vs_fp64": round(t64 / t32, 3),
            "speedup_fp32_vs_fp64_fma": round(t64f / t32, 3),
            "fp64_prepared_ms_per_stage": {"fma": round(tp["fma"] * 1e3, 4),
                                           "exact": round(tp["exact"] * 1e3, 4)},
            "speedup_fp32_vs_fp64_prepared": {"fma": round(tp["fma"] / t32, 3),
                                              "exact": round(tp["exact"] / t32, 3)},
            "fp32_gdof_stage_s": round(dof / t32 / 1e9, 3),
            "fp64_gdof_stage_s": round(dof / t64 / 1e9, 3),
            "fp32_bytes_per_element_stage": b32,
            "fp32_hbm_frac": round(b32 * s.nelem / t32 / 1e9 / HBM_PEAK_GBS, 4),
            "rel_l2_vs_fp64": errors,
            "mixed": {"what": "FP32 state q' and RHS arithmetic; metric terms, MJ, normals, "
                              "sMJ and the LSRK register dq in FP64 (MJI = 1/MJ in FP64)",
                      "ms_per_stage": round(t32m * 1e3, 4),
                      "speedup_vs_fp64": round(t64 / t32m, 3),
                      "gdof_stage_s": round(dof / t32m / 1e9, 3),
                      "bytes_per_element_stage": b32m,
                      "byte_ratio_fp64_over_mixed": round(b64 / b32m, 3),
                      "hbm_frac": round(b32m * s.nelem / t32m / 1e9 / HBM_PEAK_GBS, 4),
                      "rel_l2_vs_fp64": errors_m},
            "carries_the_speedup": "compact (FP32 metric terms, normals, LSRK register and "
                                   "FP32 copy of the background: %d B per element-stage against "
                                   "%d B for the FP64 stage, %d B over the prepared arrays); the "
                                   "mixed variant moves %d B, so bytes cap it at %.2fx"
                                   % (b32, b64, prepared_bytes(3, N), b32m, b64 / b32m),
            "dt": dt}


def form_child(name, extra_args, timeout=480):
    """Measure one added form in a child process of its own (fresh device
    state, as in the tests); a failure there is reported in the JSON line
    instead of ending the benchmark."""
    import subprocess
    cmd = [sys.executable, "-X", "faulthandler", os.path.abspath(__file__), "--form", name,
           *extra_args]
    sys.stderr.flush()
    try:  # the child's progress lines go straight to this process's stderr
        r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=timeout)
    except subprocess.TimeoutExpired:
        return {"error": "timed out after %d s" % timeout}
    if r.returncode != 0:
        return {"error": "exit status %d" % r.returncode}
    try:
        return json.loads(r.stdout.strip().splitlines()[-1])
    except (ValueError, IndexError):
        return {"error": "no JSON from the child", "stdout_tail": r.stdout[-500:]}


def _gpu_timer():
    import torch

    def timed(fn):
        stream = torch.cuda.current_stream()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3
    return timed


def batched_form(nsolves, steps, seed=42):
    """Ensemble of independent periodic shallow-water solves (Ne^2 quads, Ne
    uniform in 8..64, N in {3,4,5,7}, seeded bumps) advanced together:
    FP64 (reference evaluation order), contracted FP64 and FP32 batches
    (autotuned elements per workgroup, one graph per step) against the same
    solves run one after another with the single-solve fused stage; each
    precision also with every stage one launch over all orders, and both
    launch forms over the prepared arrays (hbm_frac on the reference-array
    byte rule throughout, hbm_frac_on_own_bytes on the bytes each form
    moves)."""
    import numpy as np
    import torch
    import canary_amd.kernels as K
    import canary_amd.problems as P
    from canary_amd._lib import SWE, FP32, FP64, FP64_FMA
    from canary_amd.batch import Batch, bytes_per_element_stage

    t0 = time.time()
    ens = P.swe_ensemble(n=nsolves, seed=seed)
    print("setup batched form: %d solves, %.1f s" % (nsolves, time.time() - t0), file=sys.stderr,
          flush=True)
    timed = _gpu_timer()
    dof = sum(s.ndof() for s, _ in ens)
    out = {"workload": "%d periodic shallow-water solves, Ne^2 quads (Ne uniform 8..64), "
                       "N in {3,4,5,7}, depth 10, 2 seeded Gaussian bumps (height 0.01-0.1, "
                       "width 0.05-0.2)" % nsolves,
           "solves": nsolves, "dof": dof, "elements": sum(s.nelem for s, _ in ens),
           "orders": sorted({s.N for s, _ in ens}), "steps": steps}
    states = {}
    for name, prec in (("fp64", FP64), ("fp64_fma", FP64_FMA), ("fp32", FP32)):
        b = Batch(SWE, 2, prec)
        ids = [b.add(s, dt) for s, dt in ens]
        b.finalize(autotune=True)
        b.run(2)
        b.check()
        t = timed(lambda: b.run(steps))
        b.check()
        out[name] = {"gdof_stage_s": round(dof * 5 * steps / t / 1e9, 3),
                     "ms_per_step": round(t / steps * 1e3, 4),
                     "hbm_frac": round(b.bytes_per_stage() * 5 * steps / t / 1e9 / HBM_PEAK_GBS, 4),
                     "epw": {str(k): v for k, v in sorted(b.epw.items())}}
        states[name] = [b.state(i) for i in ids]
        print("batched form %s: %.3f GDOF-stage/s" % (name, out[name]["gdof_stage_s"]),
              file=sys.stderr, flush=True)
        del b
        torch.cuda.empty_cache()
        # the same batch with every stage ONE launch over all orders, and
        # both launch forms over the prepared arrays; same steps, step graphs
        ref_bytes = None
        for tag, prepared, single in (("single_launch", False, True),
                                      ("prepared", True, False),
                                      ("prepared_single_launch", True, True)):
            b = Batch(SWE, 2, prec, prepared=prepared)
            ids = [b.add(s, dt) for s, dt in ens]
            if single:
                b.finalize(single_launch=True)
            else:
                b.finalize(autotune=True)
            b.run(2)
            b.check()
            t = timed(lambda: b.run(steps))
            b.check()
            got = [b.state(i) for i in ids]
            if prec == FP64:
                same = {"same_bits_as_reference_arrays_per_order":
                        all(np.array_equal(a, c) for a, c in zip(got, states[name]))}
            else:
                same = {"rel_l2_vs_reference_arrays_per_order": float("%.3e" % max(
                    float(np.linalg.norm(a[:, :3] - c[:, :3]) / np.linalg.norm(c[:, :3]))
                    for a, c in zip(got, states[name])))}
            if ref_bytes is None:
                ref_bytes = sum(v.nelem * bytes_per_element_stage(SWE, 2, v.N, 4 if prec == FP32
                                                                  else 8)
                                for v in b.solves.values())
            out[name][tag] = {
                "gdof_stage_s": round(dof * 5 * steps / t / 1e9, 3),
                "ms_per_step": round(t / steps * 1e3, 4),
                "hbm_frac": round(ref_bytes * 5 * steps / t / 1e9 / HBM_PEAK_GBS, 4),
                "hbm_frac_on_own_bytes": round(b.bytes_per_stage() * 5 * steps / t / 1e9
                                               / HBM_PEAK_GBS, 4),
                "epw": {str(k): v for k, v in sorted(b.epw.items())}, **same}
            print("batched form %s, %s: %.3f GDOF-stage/s"
                  % (name, tag, out[name][tag]["gdof_stage_s"]), file=sys.stderr, flush=True)
            del b, got
            torch.cuda.empty_cache()
    errs = [np.linalg.norm(a[:, :3] - c[:, :3]) / np.linalg.norm(c[:, :3])
            for a, c in zip(states["fp32"], states["fp64"])]
    out["fp32_rel_l2_vs_fp64_max"] = float("%.3e" % max(errs))
    errs = [np.linalg.norm(a[:, :3] - c[:, :3]) / np.linalg.norm(c[:, :3])
            for a, c in zip(states["fp64_fma"], states["fp64"])]
    out["fp64_fma_rel_l2_vs_fp64_max"] = float("%.3e" % max(errs))
    # the same solves one after another, single-solve fused stage, FP64
    A, B, _ = P.rk_coefs()
    dev = torch.device("cuda:0")

    def d(a):
        return torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order="F"))).to(dev)

    seq = []
    for s, dt in ens:
        arrs = [d(x) for x in (s.vgeo, s.sgeo, s.D, s.vmapM, s.vmapP, s.elemtobndy)]
        seq.append((s, dt, arrs, d(s.Q), torch.empty(s.Q.size, dtype=torch.float64, device=dev),
                    torch.zeros(s.Q.size, dtype=torch.float64, device=dev)))

    def run_seq(nsteps):
        for s, dt, (vg, sg, D, vM, vP, bd), q, q2, rhs in seq:
            for _ in range(nsteps):
                for k in range(5):
                    K.stage_fused(2, s.N, rhs, q, q2, vg, sg, D, s.nelem, vM, vP, bd,
                                  A[(k + 1) % 5], B[k], dt, eqn=SWE, phys=s.phys)
                    q, q2 = q2, q
    run_seq(1)
    t = timed(lambda: run_seq(steps))
    K.check_errors()
    out["sequential_fp64"] = {"gdof_stage_s": round(dof * 5 * steps / t / 1e9, 3),
                              "ms_per_step": round(t / steps * 1e3, 4)}
    out["speedup_batched_fp64_vs_sequential"] = round(
        out["fp64"]["gdof_stage_s"] / out["sequential_fp64"]["gdof_stage_s"], 3)
    out["speedup_batched_fp64_fma_vs_sequential"] = round(
        out["fp64_fma"]["gdof_stage_s"] / out["sequential_fp64"]["gdof_stage_s"], 3)
    return out


def served_form(nreq, rate, seed=2024, graph=False):
    """Poisson request stream (half 2D vortices, half 3D bubbles) through the
    continuous-batching server (FP32 perturbation form; each request's
    preparation -- host mesh, device geometry and states -- inside its
    latency), against the same requests solved one after another with the
    plain FP64 fused stage (preparation excluded)."""
    import numpy as np
    import torch
    import canary_amd.kernels as K
    import canary_amd.problems as P
    from canary_amd.serve import Server, make_setup_device, poisson_stream, summary

    reqs = poisson_stream(n=nreq, rate=rate, seed=seed)
    # an untimed tiny stream first (first launches, allocator); the measured
    # server starts with an empty elements-per-workgroup table and autotunes
    # each new order inside the stream
    Server("fp32").serve(poisson_stream(n=8, rate=1000.0, seed=1, vortex_ne=(2, 3),
                                        bubble_ne=(2, 2), steps=(2, 3)))
    # admission binds: an HBM budget of a quarter of the stream's summed
    # footprint (but at least its largest request), so requests wait for room
    from canary_amd.serve import make_setup_host
    probe = Server("fp32")
    fps = [probe.footprint(make_setup_host(r)) for r in reqs]
    del probe
    budget = max(max(fps), sum(fps) // 4)
    # work tiers: requests of at most 5e8 DOF-steps in a batch of their own,
    # the larger ones advanced every 8th iteration while small ones are
    # resident (swept on this stream: 2e8 / 5e8 x 4 / 8)
    srv = Server("fp32", keep_setups=True, tier_work=5e8, tier_ratio=8, hbm_budget=budget)
    t0 = time.time()
    out_reqs = srv.serve(reqs)
    wall = time.time() - t0
    st = summary(out_reqs, srv.makespan)
    print("served %d requests in %.1f s" % (nreq, wall), file=sys.stderr, flush=True)
    # the same stream in lock step (one batch per dimension); results must
    # not change
    srv_l = Server("fp32", hbm_budget=budget)
    out_l = srv_l.serve(poisson_stream(n=nreq, rate=rate, seed=seed))
    st_l = summary(out_l, srv_l.makespan)
    by_rid = {r.rid: r.result for r in out_l}
    lock = {k: round(st_l[k], 4) for k in ("p50_latency_s", "p99_latency_s", "max_latency_s",
                                           "makespan_s", "aggregate_gdof_stage_s")}
    lock["same_results"] = all(np.array_equal(r.result, by_rid[r.rid]) for r in out_reqs)
    del out_l, by_rid, srv_l
    print("served in lock step: p50 %.3f s p99 %.3f s" % (lock["p50_latency_s"],
                                                          lock["p99_latency_s"]),
          file=sys.stderr, flush=True)
    graphed = None
    if graph:
        # the tiered server again, each batch step replayed from its captured
        # hipGraph (re-captured as requests join and leave); same results
        srv_g = Server("fp32", tier_work=5e8, tier_ratio=8, hbm_budget=budget, graph=True)
        out_g = srv_g.serve(poisson_stream(n=nreq, rate=rate, seed=seed))
        st_g = summary(out_g, srv_g.makespan)
        by_rid = {r.rid: r.result for r in out_g}
        graphed = {k: round(st_g[k], 4) for k in ("p50_latency_s", "p99_latency_s",
                                                  "max_latency_s", "makespan_s",
                                                  "aggregate_gdof_stage_s")}
        graphed["same_results"] = all(np.array_equal(r.result, by_rid[r.rid]) for r in out_reqs)
        del out_g, by_rid, srv_g
        print("served with graph replay: p50 %.3f s p99 %.3f s" % (
            graphed["p50_latency_s"], graphed["p99_latency_s"]), file=sys.stderr, flush=True)
    # sequential plain FP64 solves of the same requests
    timed = _gpu_timer()
    A, B, _ = P.rk_coefs()
    dev = torch.device("cuda:0")

    def d(a):
        return torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order="F"))).to(dev)

    t_seq, errs = 0.0, []
    for r in out_reqs:
        s, dt = make_setup_device(r, r.setup)   # the same device preparation, FP64 arrays kept
        a = P.device_arrays(s)
        q, q2, rhs = s.Qd.clone(), torch.empty_like(s.Qd), torch.zeros_like(s.Qd)
        hold = {}

        def go():
            x, y = q, q2
            for _ in range(r.nsteps):
                for k in range(5):
                    K.stage_fused(s.dim, s.N, rhs, x, y, a["vgeo"], a["sgeo"], a["D"], s.nelem,
                                  a["vmapM"], a["vmapP"], a["bndy"], A[(k + 1) % 5], B[k], dt,
                                  phys=s.phys)
                    x, y = y, x
            hold["q"] = x
        t_seq += timed(go)
        if r.rid % 8 == 7:
            print("sequential FP64 baseline: request %d done" % r.rid, file=sys.stderr, flush=True)
        ref = hold["q"].cpu().numpy().reshape((s.Np, s.nstate, s.nelem), order="F")
        errs.append(np.linalg.norm(r.result - ref) / np.linalg.norm(ref))
        del a, q, q2, rhs, hold
        s.dev = s.Qd = s.Q0d = None
        r.setup = None
    K.check_errors()
    nk = {"vortex2d": sum(r.kind == "vortex2d" for r in out_reqs)}
    return {"workload": "%d requests, Poisson arrivals (mean %g/s, seed %d): %d 2D isentropic "
                        "vortices (16^2-64^2 quads, N 3..7), %d 3D rising bubbles (8^3-24^3 "
                        "hexes, N 3..5), 50-500 LSRK54 steps each; FP32 perturbation form, "
                        "work tiers, HBM budget a quarter of the summed footprint"
                        % (nreq, rate, seed, nk["vortex2d"], nreq - nk["vortex2d"]),
            **{k: (round(v, 4) if isinstance(v, float) else v) for k, v in st.items()},
            "work_tiers": {"tier_work_dof_steps": 5e8, "tier_ratio": 8},
            "hbm_budget_bytes": budget, "stream_footprint_bytes": sum(fps),
            "max_resident_bytes": max(res for _, _, res in srv.trace),
            "requests_that_waited_for_room": sum(1 for r in out_reqs
                                                 if r.admitted - r.prepared > 5e-3),
            "lock_step": lock,
            "graph_replay": graphed,
            "sequential_fp64_gpu_s": round(t_seq, 4),
            "sequential_fp64_gdof_stage_s": round(st["dof_stages"] / t_seq / 1e9, 3),
            "fp32_rel_l2_vs_fp64": {"max": float("%.3e" % max(errs)),
                                    "median": float("%.3e" % float(np.median(errs)))}}


HEADLINE_KERNEL = "stage_prep_kernel<canary::Euler3DFma, 5, 1, 1, false>"


def _csv_rows(d, pattern):
    import csv
    import glob
    out = []
    for f in sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True)):
        with open(f) as fh:
            out += list(csv.DictReader(fh))
    return out


def _short(name, n=96):
    return name if len(name) <= n else name[:n - 3] + "..."


def summarize_trace(d, steps, ne, N):
    """Per-kernel table of a `rocprofv3 --kernel-trace` run of this script
    (csv output under d), and the headline kernel's average over its last
    5 * steps dispatches -- the timed launches of the measured loop, which
    is the only loop that launches that kernel at that grid -- with the
    roofline fraction recomputed from it."""
    groups = {}
    for r in _csv_rows(d, "*kernel_trace.csv"):
        key = (r["Kernel_Name"], int(r.get("Grid_Size", r.get("Grid_Size_X"))),
               int(r.get("Workgroup_Size", r.get("Workgroup_Size_X"))))
        t = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        groups.setdefault(key, []).append((int(r["Dispatch_Id"]), t))
    print("%-96s %9s %4s %6s %9s %9s %9s" % ("kernel", "grid", "wg", "calls", "avg_us", "min_us",
                                             "max_us"))
    for (name, grid, wg), v in sorted(groups.items(), key=lambda kv: -sum(t for _, t in kv[1]))[:40]:
        ts = [t for _, t in v]
        print("%-96s %9d %4d %6d %9.2f %9.2f %9.2f" % (_short(name), grid, wg, len(ts),
                                                       sum(ts) / len(ts), min(ts), max(ts)))
    nelem = ne ** 3
    grid = nelem * 128  # one 128-lane workgroup per element
    sel = [v for (name, g, _), v in groups.items() if HEADLINE_KERNEL in name and g == grid]
    if len(sel) != 1:
        print("headline kernel %s at grid %d: not found once" % (HEADLINE_KERNEL, grid))
        return
    ts = [t for _, t in sorted(sel[0])][-5 * steps:]
    avg = sum(ts) / len(ts)
    nbytes = algorithmic_bytes(3, N) * nelem
    print("timed launches of %s (last %d of %d dispatches at grid %d): avg %.2f us, "
          "min %.2f, max %.2f" % (HEADLINE_KERNEL, len(ts), len(sel[0]), grid, avg, min(ts),
                                  max(ts)))
    print("recomputed: achieved %.1f GB/s, frac %.4f (%d B/element-stage x %d elements), "
          "5 x avg = %.4f ms per step"
          % (nbytes / avg / 1e3, nbytes / avg / 1e3 / HBM_PEAK_GBS, algorithmic_bytes(3, N), nelem,
             5 * avg / 1e3))


# request sizes of the TCC->memory counters (bytes per request)
_READ_SIZES = {"TCC_EA0_RDREQ_32B_sum": 32, "TCC_EA0_RDREQ_64B_sum": 64,
               "TCC_EA0_RDREQ_128B_sum": 128}


def summarize_pmc(d, ne, N):
    """Per-kernel means of every counter of a `rocprofv3 --pmc` run of this
    script (csv output under d).  With the request counters
    TCC_EA0_RDREQ_{32B,64B,128B}_sum the HBM read bytes are their sizes
    summed; with TCC_EA0_WRREQ_sum and TCC_EA0_WRREQ_64B_sum the write bytes
    are 64 per 64-byte request and 32 per other one.  The known-size kernels
    of --calib are listed with their expected bytes."""
    vals = {}
    for r in _csv_rows(d, "*counter_collection.csv"):
        key = (r["Kernel_Name"], int(r["Grid_Size"]))
        vals.setdefault(key, {}).setdefault(r["Counter_Name"], []).append(
            float(r["Counter_Value"]))
    rows = []
    for (name, grid), cs in vals.items():
        mean = {c: sum(v) / len(v) for c, v in cs.items()}
        calls = max(len(v) for v in cs.values())
        rd = wr = None
        if all(c in mean for c in _READ_SIZES):
            rd = sum(mean[c] * b for c, b in _READ_SIZES.items())
        if "TCC_EA0_WRREQ_sum" in mean and "TCC_EA0_WRREQ_64B_sum" in mean:
            w64 = mean["TCC_EA0_WRREQ_64B_sum"]
            wr = 64 * w64 + 32 * (mean["TCC_EA0_WRREQ_sum"] - w64)
        rows.append((name, grid, calls, mean, rd, wr))
    rows.sort(key=lambda r: -r[2] * max([v for v in r[3].values()] + [0]))
    counters = sorted({c for r in rows for c in r[3]})
    print("%-80s %9s %6s %s %14s %14s" % ("kernel", "grid", "calls",
                                         " ".join("%16s" % c[-16:] for c in counters),
                                         "read_MB", "write_MB"))
    for name, grid, calls, mean, rd, wr in rows[:30]:
        print("%-80s %9d %6d %s %14s %14s" % (
            _short(name, 80), grid, calls,
            " ".join("%16.0f" % mean.get(c, float("nan")) for c in counters),
            "-" if rd is None else "%.3f" % (rd / 1e6), "-" if wr is None else "%.3f" % (wr / 1e6)))
    nelem = ne ** 3
    head = [r for r in rows if HEADLINE_KERNEL in r[0] and r[1] == nelem * 128]
    if head:
        name, grid, calls, mean, rd, wr = head[0]
        own = prepared_bytes(3, N) * nelem
        print("headline %s: %d dispatches; read %s, write %s bytes per launch; own algorithmic "
              "bytes %d (read %d, write %d)" % (HEADLINE_KERNEL, calls, rd, wr, own,
                                               own - 2 * 5 * (N + 1) ** 3 * 8 * nelem,
                                               2 * 5 * (N + 1) ** 3 * 8 * nelem))


def measured_traffic(tag):
    """HBM bytes per launch from the committed rocprofv3 counter summary, if any."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % tag)
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    return None


def cpu_baseline(s, sample_elems):
    """The oracle (a C restatement of the reference's Julia CPU kernels) on the
    host: one LSRK stage (volumerhs!, facerhs!, updatesolution!) over the first
    `sample_elems` elements of the same mesh, single thread and on up to 16
    threads, median of 5 after a warm-up; the host's CPU model and the
    threads it offers this process are reported beside it.  The pressure's
    pow is the C library's here, as the reference calls its libm pow (the
    oracle's correctly rounded pow, its parity default, is about 6x slower
    per call and is not what the reference spends)."""
    import numpy as np
    from oracle import oracle as O
    O.set_pow_mode(0)
    n = min(sample_elems, s.nelem)
    Q = np.asfortranarray(s.Q.copy())
    rhs = s.zeros_state()
    A, B, _ = __import__("canary_amd.problems", fromlist=["rk_coefs"]).rk_coefs()
    def stage_time(threads):
        t0 = time.perf_counter()
        O.volumerhs(s, rhs, Q, nthreads=threads, nelem=n)
        O.facerhs(s, rhs, Q, nthreads=threads, nelem=n)
        O.lib().oracle_updatesolution(s.Np, 5, 14, 10, O._p(rhs), O._p(Q), O._p(s.vgeo), n,
                                      A[1], B[0], 0.01, threads)
        return time.perf_counter() - t0

    stage_time(1)  # warm-up (page-in, first touch)
    t1 = sorted(stage_time(1) for _ in range(5))[2]
    ncpu = min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 8)
    stage_time(ncpu)
    tn = sorted(stage_time(ncpu) for _ in range(5))[2]
    O.set_pow_mode(1)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"value": round(n * s.Np * 5 / t1 / 1e9, 5), "unit": "GDOF-stage/s", "cores": 1,
            "cpu_model": model, "threads_available": avail,
            "kind": "port",
            "sample": "LSRK stages (volumerhs!, facerhs!, updatesolution!) over %d of %d "
                      "elements, median of 5 after a warm-up, single thread; C restatement of "
                      "the Julia CPU kernels (gcc -O2 -ffp-contract=off, C library pow as the "
                      "reference's libm pow)" % (n, s.nelem),
            "all_cores": {"value": round(n * s.Np * 5 / tn / 1e9, 5), "cores": ncpu}}


if __name__ == "__main__":
    main()
