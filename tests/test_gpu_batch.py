"""Batched form on the GPU.

Every element of a batched stage runs the same exact-order FP64 code as the
single-solve fused stage, so the bar is bit-identity with per-solve
stage_fused runs (themselves bit-compared with the CPU oracle in
test_gpu_parity.py), for any mix of N, meshes, dt and constants, for every
elements-per-workgroup variant, eagerly and through captured graphs, and
when solves join and leave a running batch.  The FP32 shallow-water variant
is held to a relative L2 error against FP64 instead, and the contracted FP64
variant (FP64_FMA) to the single-solve contracted stage and the exact-order
batch.
"""
import numpy as np
import pytest
import torch

import canary_amd.kernels as K
import canary_amd.problems as P
from canary_amd._lib import CanaryError, EULER, SWE, FP32, FP64, FP64_FMA, ERR_ARG
from canary_amd.batch import Batch, ITEM_DTYPE, variants
from oracle import oracle as O

from gpu_util import DeviceSetup, dev, host
from states import max_ulp, rel_l2, smooth_euler_state

pytestmark = pytest.mark.gpu


def swe_mix():
    """Shallow-water solves of mixed N, size, boundary and gravity."""
    out = []
    spec = [(3, 4, True, 10.0, 1), (4, 3, False, 9.81, 2), (5, 5, True, 10.0, 3),
            (7, 2, True, 12.0, 4), (3, 6, False, 10.0, 5), (2, 3, True, 10.0, 6)]
    for N, Ne, per, g, seed in spec:
        rng = np.random.default_rng(seed)
        bumps = [(rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), 0.05, rng.uniform(0.05, 0.2))
                 for _ in range(2)]
        s = P.swe2d(Ne=Ne, N=N, bumps=bumps, periodic=per, gravity=g)
        out.append((s, P.swe_courant(s, s.Q)))
    return out


def euler2d_mix():
    out = []
    for N, Ne, seed in ((3, 4, 1), (4, 3, 2), (6, 2, 3)):
        s = P.bubble2d(Ne=Ne, N=N)
        s.Q = np.asfortranarray(smooth_euler_state(s, seed=seed))
        out.append((s, P.cfl(s, s.Q) / s.N ** np.sqrt(2)))
    v = P.vortex2d(Ne=4, N=4)  # other constants (p0 = R = 1, g = 0)
    out.append((v, P.euler_courant(v, v.Q)))
    return out


def euler3d_mix():
    out = []
    for N, Ne, seed in ((2, 3, 1), (3, 2, 2), (4, 2, 3), (5, 2, 4), (6, 2, 5), (7, 1, 6)):
        s = P.bubble3d(Ne=Ne, N=N, jitter=1e-3, seed=seed)
        s.Q = np.asfortranarray(smooth_euler_state(s, seed=seed))
        out.append((s, P.cfl(s, s.Q) / s.N ** np.sqrt(2)))
    return out


def single_run(s, dt, nsteps, prec=FP64):
    """nsteps LSRK54 steps of one solve through the single-solve fused stage."""
    A, B, _ = P.rk_coefs()
    ds = DeviceSetup(s)
    q, q2, rhs = dev(s.Q), dev(s.Q), dev(np.zeros_like(s.Q))
    for _ in range(nsteps):
        for st in range(5):
            K.stage_fused(s.dim, s.N, rhs, q, q2, ds.vgeo, ds.sgeo, ds.D, s.nelem, ds.vmapM,
                          ds.vmapP, ds.bndy, A[(st + 1) % 5], B[st], dt, eqn=s.eqn, phys=s.phys,
                          prec=prec)
            q, q2 = q2, q
    K.check_errors()
    out = host(q, s.Q.shape)
    assert np.isfinite(out).all()
    return out


def make_batch(eqn, dim, mix, prec=FP64, epw=None):
    b = Batch(eqn, dim, prec)
    ids = [b.add(s, dt) for s, dt in mix]
    b.finalize(epw=epw)
    return b, ids


@pytest.mark.parametrize("eqn,dim,mixf", [(SWE, 2, swe_mix), (EULER, 2, euler2d_mix),
                                          (EULER, 3, euler3d_mix)])
def test_batch_equals_single_solves(cuda, eqn, dim, mixf):
    mix = mixf()
    print("batch eqn=%d dim=%d N=%s" % (eqn, dim, [s.N for s, _ in mix]), flush=True)
    b, ids = make_batch(eqn, dim, mix)
    b.run(3, graph=False)
    b.check()
    for (s, dt), i in zip(mix, ids):
        want = single_run(s, dt, 3)
        assert max_ulp(b.state(i), want) == 0, "solve N=%d" % s.N


@pytest.mark.parametrize("eqn,dim,mixf", [(SWE, 2, swe_mix), (EULER, 2, euler2d_mix),
                                          (EULER, 3, euler3d_mix)])
def test_fma_batch_tracks_single_and_exact(cuda, eqn, dim, mixf):
    mix = mixf()
    print("fma batch eqn=%d dim=%d N=%s" % (eqn, dim, [s.N for s, _ in mix]), flush=True)
    b, ids = make_batch(eqn, dim, mix, prec=FP64_FMA)
    b.run(3, graph=False)
    b.check()
    ex, eids = make_batch(eqn, dim, mix)
    ex.run(3, graph=False)
    for (s, dt), i, j in zip(mix, ids, eids):
        got = b.state(i)
        want = single_run(s, dt, 3, prec=FP64_FMA)
        print("N=%d ulp vs single contracted %d, rel L2 vs exact %.2e"
              % (s.N, max_ulp(got, want), rel_l2(got, ex.state(j))), flush=True)
        assert rel_l2(got, want) <= 1e-13
        assert rel_l2(got, ex.state(j)) <= 1e-12


def test_batch_swe_matches_oracle(cuda):
    s, dt = swe_mix()[1]
    b, ids = make_batch(SWE, 2, [(s, dt)])
    b.run(2, graph=False)
    Q = np.array(s.Q, order="F", copy=True)
    rhs = np.zeros_like(Q, order="F")
    O.lsrk(s, Q, rhs, dt, 2)
    assert np.isfinite(Q).all()
    assert max_ulp(b.state(ids[0]), Q) == 0


@pytest.mark.parametrize("eqn,dim,mixf", [(SWE, 2, swe_mix), (EULER, 3, euler3d_mix)])
def test_every_variant_same_bits(cuda, eqn, dim, mixf):
    mix = mixf()
    ref, rids = make_batch(eqn, dim, mix)
    ref.run(2, graph=False)
    Ns = sorted({s.N for s, _ in mix})
    nvar = max(len(variants(eqn, dim, FP64, N)) for N in Ns)
    for k in range(1, nvar):
        epw = {}
        for N in Ns:
            c = variants(eqn, dim, FP64, N)
            epw[N] = c[min(k, len(c) - 1)][0]
        print("variant", epw, flush=True)
        b, ids = make_batch(eqn, dim, mix, epw=epw)
        b.run(2, graph=False)
        b.check()
        for i, j in zip(ids, rids):
            assert max_ulp(b.state(i), ref.state(j)) == 0


def test_graph_replays_match_eager(cuda):
    mix = swe_mix()
    eager, eids = make_batch(SWE, 2, mix)
    graph, gids = make_batch(SWE, 2, mix)
    for rep in range(4):
        graph.run(1, graph=True)       # replays alternate between the two parity graphs
        eager.run(1, graph=False)
        x = torch.ones(1000, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        assert float(x) == 1000.0
        for i, j in zip(gids, eids):
            assert max_ulp(graph.state(i), eager.state(j)) == 0, "replay %d" % rep
    graph.check()


def test_autotune_keeps_results(cuda):
    mix = swe_mix()
    ref, rids = make_batch(SWE, 2, mix)
    ref.run(2, graph=False)
    b = Batch(SWE, 2)
    ids = [b.add(s, dt) for s, dt in mix]
    chosen = b.finalize(autotune=True).epw
    print("autotune", chosen, b.tuning, flush=True)
    for N, res in b.tuning.items():
        assert chosen[N] == min(res, key=res.get)
        assert all(t > 0 for t in res.values())
    b.run(2)
    for i, j in zip(ids, rids):
        assert max_ulp(b.state(i), ref.state(j)) == 0


def test_join_and_leave(cuda):
    """Continuous batching: a solve joining a running batch, and one leaving
    it, get exactly the bits they get alone."""
    mix = swe_mix()
    b = Batch(SWE, 2)
    a = b.add(*mix[0])
    c = b.add(*mix[2])
    b.finalize()
    b.run(2)
    d = b.add(*mix[3])           # joins after 2 steps
    b.finalize()
    b.run(1)
    Qc = b.remove(c)             # leaves after 3 steps
    b.finalize()
    b.run(2)
    assert max_ulp(b.state(a), single_run(*mix[0], 5)) == 0
    assert max_ulp(Qc, single_run(*mix[2], 3)) == 0
    assert max_ulp(b.state(d), single_run(*mix[3], 3)) == 0
    b.check()


def test_fp32_swe_tracks_fp64(cuda):
    mix = swe_mix()
    b64, ids64 = make_batch(SWE, 2, mix)
    b32, ids32 = make_batch(SWE, 2, mix, prec=FP32)
    for n in (1, 10):
        b64.run(n)
        b32.run(n)
        b32.check()
        for (s, _), i, j in zip(mix, ids32, ids64):
            got, want = b32.state(i), b64.state(j)
            # states (U, V, h, b): momenta and the height perturbation h over
            # the depth b (b is copied through unchanged)
            eh = rel_l2(got[:, 2, :], want[:, 2, :])
            eU = rel_l2(got[:, 0:2, :], want[:, 0:2, :])
            print("N=%d steps %d: h %.2e U %.2e" % (s.N, n, eh, eU), flush=True)
            assert eh <= 1e-4 and eU <= 1e-4
            assert max_ulp(got[:, 3, :], want[:, 3, :].astype(np.float32).astype(np.float64)) == 0


def test_bad_table_entry_is_reported(cuda):
    """An entry whose N does not match its kernel is skipped and reported."""
    s, dt = swe_mix()[0]
    b, ids = make_batch(SWE, 2, [(s, dt)])
    items = np.zeros(1, dtype=ITEM_DTYPE)
    items["solve"], items["N"], items["epw"], items["e0"] = 0, s.N + 1, b.epw[s.N], 0
    t = torch.from_numpy(items.view(np.uint8).copy()).cuda()
    before = b.state(ids[0])
    b._launch(s.N, b.epw[s.N], t.data_ptr(), 1, b.parity, 0.0, 1.0, b._st())
    with pytest.raises(CanaryError) as ei:
        b.check()
    assert ei.value.code == ERR_ARG
    b.check()  # cleared
    assert max_ulp(b.state(ids[0]), before) == 0


def test_swe_ensemble_of_64(cuda):
    """The 64-solve shallow-water ensemble at its own size (periodic, Ne^2
    quads with Ne uniform in 8..64, N in {3,4,5,7}, still depth 10, two
    seeded Gaussian bumps of height 0.01-0.1 and width 0.05-0.2 domain
    widths; 102,836 elements, 16.3 M DOF), elements per workgroup
    autotuned per N, 3 LSRK54 steps:
      * FP64 (reference evaluation order), eager and replaying step graphs:
        every solve bit-identical to its own single-solve fused run;
      * one solve of each N bit-identical to the oracle's lowstorageRK;
      * contracted FP64 within 1e-13 of FP64, FP32 within 1e-5 (relative
        L2 of h, Hu, Hv per solve)."""
    ens = P.swe_ensemble(n=64, seed=42)
    assert len(ens) == 64 and sorted({s.N for s, _ in ens}) == [3, 4, 5, 7]
    assert all(8 <= int(round(np.sqrt(s.nelem))) <= 64 for s, _ in ens)
    print("ensemble: %d elements, %d DOF" % (sum(s.nelem for s, _ in ens),
                                            sum(s.ndof() for s, _ in ens)), flush=True)
    states = {}
    for name, prec, graph in (("fp64", FP64, False), ("fp64 graph", FP64, True),
                              ("fp64_fma", FP64_FMA, False), ("fp32", FP32, False)):
        b = Batch(SWE, 2, prec)
        ids = [b.add(s, dt) for s, dt in ens]
        b.finalize(autotune=True)
        b.run(3, graph=graph)
        b.check()
        states[name] = [b.state(i) for i in ids]
        print("%s: epw %s" % (name, dict(sorted(b.epw.items()))), flush=True)
        del b
        torch.cuda.empty_cache()
    for k, (s, dt) in enumerate(ens):
        assert max_ulp(states["fp64 graph"][k], states["fp64"][k]) == 0, k
        want = single_run(s, dt, 3)
        assert max_ulp(states["fp64"][k], want) == 0, "solve %d (N=%d, %d elements)" % (
            k, s.N, s.nelem)
        ref = states["fp64"][k][:, :3]
        assert rel_l2(states["fp64_fma"][k][:, :3], ref) <= 1e-13, k
        e32 = rel_l2(states["fp32"][k][:, :3], ref)
        assert e32 <= 1e-5, (k, e32)
    for N in (3, 4, 5, 7):
        k = min((k for k, (s, _) in enumerate(ens) if s.N == N), key=lambda k: ens[k][0].nelem)
        s, dt = ens[k]
        Q = np.array(s.Q, order="F", copy=True)
        assert O.lsrk(s, Q, np.zeros_like(Q, order="F"), dt, 3, nthreads=8) == 0
        assert max_ulp(states["fp64"][k], Q) == 0, "oracle, solve %d (N=%d)" % (k, N)


def test_batch_swe_advection_mode(cuda):
    """Solves flagged with the shallow-water advection mode (velocity kept
    through each update) beside unflagged ones in one batch: FP64 bit-identical
    to single-solve fused runs with the same flag, FP32 within 1e-5."""
    from states import smooth_swe_state
    mix = swe_mix()[:4]
    for s, _ in mix:
        s.Q = np.asfortranarray(smooth_swe_state(s, seed=s.N))
    flags = [True, False, True, True]
    A, B, _ = P.rk_coefs()
    out = {}
    for prec in (FP64, FP32):
        b = Batch(SWE, 2, prec)
        ids = [b.add(s, dt, advection=f) for (s, dt), f in zip(mix, flags)]
        b.finalize()
        b.run(3, graph=False)
        b.check()
        out[prec] = [b.state(i) for i in ids]
    for k, ((s, dt), f) in enumerate(zip(mix, flags)):
        ds = DeviceSetup(s)
        q, q2, rhs = dev(s.Q), dev(s.Q), dev(np.zeros_like(s.Q))
        for _ in range(3):
            for st in range(5):
                K.stage_fused(2, s.N, rhs, q, q2, ds.vgeo, ds.sgeo, ds.D, s.nelem, ds.vmapM,
                              ds.vmapP, ds.bndy, A[(st + 1) % 5], B[st], dt, eqn=SWE, phys=s.phys,
                              advection=f)
                q, q2 = q2, q
        K.check_errors()
        want = host(q, s.Q.shape)
        assert max_ulp(out[FP64][k], want) == 0, k
        assert rel_l2(out[FP32][k][:, :3], want[:, :3]) <= 1e-5, k
        if f:
            Qo = np.array(s.Q, order="F", copy=True)
            assert O.lsrk(s, Qo, np.zeros_like(Qo, order="F"), dt, 3, advection=True) == 0
            assert max_ulp(want, Qo) == 0, k


@pytest.mark.parametrize("eqn,dim,mixf", [(SWE, 2, swe_mix), (EULER, 2, euler2d_mix),
                                          (EULER, 3, euler3d_mix)])
def test_single_launch_same_bits(cuda, eqn, dim, mixf):
    """One launch per stage over the entries of every order
    (canary_batch_stage_all) gives the per-order launches' bits: in the
    exact order (FP64, and FP32 for shallow water) the same bits, from the
    library loop and from replayed step graphs; contracted FP64 within 1e-14
    (the two kernels may fuse different products).  3D orders 6 and 7 have
    no single-launch variant and are refused."""
    mix = mixf()
    if dim == 3:
        b = Batch(eqn, dim)
        for s, dt in mix:
            b.add(s, dt)
        with pytest.raises(ValueError):
            b.finalize(single_launch=True)
        mix = [(s, dt) for s, dt in mix if s.N <= 5]
    print("single launch eqn=%d dim=%d N=%s" % (eqn, dim, [s.N for s, _ in mix]), flush=True)
    precs = (FP64, FP64_FMA, FP32) if eqn == SWE else (FP64, FP64_FMA)
    for prec in precs:
        ref = Batch(eqn, dim, prec)
        rids = [ref.add(s, dt) for s, dt in mix]
        ref.finalize()
        ref.run(3, graph=False)
        ref.check()
        for graph in (False, True):
            b = Batch(eqn, dim, prec)
            ids = [b.add(s, dt) for s, dt in mix]
            b.finalize(single_launch=True)
            assert len(b.groups) == 1
            b.run(3, graph=graph)
            b.check()
            for (s, _), i, j in zip(mix, ids, rids):
                got, want = b.state(i), ref.state(j)
                if prec == FP64_FMA:
                    assert rel_l2(got, want) <= 1e-14, (s.N, graph)
                else:
                    assert max_ulp(got, want) == 0, (prec, s.N, graph)


def test_single_launch_bad_entry_is_reported(cuda):
    """An entry whose epw is not its order's single-launch variant is skipped
    and reported."""
    s, dt = swe_mix()[0]
    b = Batch(SWE, 2)
    i = b.add(s, dt)
    b.finalize(single_launch=True)
    items = np.zeros(1, dtype=ITEM_DTYPE)
    items["solve"], items["N"], items["epw"], items["e0"] = 0, s.N, b.epw[s.N] - 1, 0
    t = torch.from_numpy(items.view(np.uint8).copy()).cuda()
    before = b.state(i)
    b._launch(0, 0, t.data_ptr(), 1, b.parity, 0.0, 1.0, b._st())
    with pytest.raises(CanaryError) as ei:
        b.check()
    assert ei.value.code == ERR_ARG
    assert max_ulp(b.state(i), before) == 0


@pytest.mark.parametrize("eqn,dim,mixf", [(SWE, 2, swe_mix), (EULER, 2, euler2d_mix),
                                          (EULER, 3, euler3d_mix)])
def test_prepared_batch_same_bits(cuda, eqn, dim, mixf):
    """Solves in the prepared layout (canary_prepare_geometry's arrays):
    in the exact order the same bits as over the reference arrays, per order
    and in one launch, from the library loop and from step graphs;
    contracted within 1e-14 (another kernel instantiation may fuse other
    products); FP32 shallow water within 1e-5 of FP64 (its MJI is 1/MJ in
    FP32 here, the rounded FP64 MJI over the reference arrays)."""
    mix = mixf()
    print("prepared batch eqn=%d dim=%d N=%s" % (eqn, dim, [s.N for s, _ in mix]), flush=True)
    for prec in ((FP64, FP64_FMA, FP32) if eqn == SWE else (FP64, FP64_FMA)):
        ref = Batch(eqn, dim, prec)
        rids = [ref.add(s, dt) for s, dt in mix]
        ref.finalize()
        ref.run(3, graph=False)
        ref.check()
        want = [ref.state(j) for j in rids]
        del ref
        ex = None
        if prec == FP32:
            ex = Batch(eqn, dim, FP64)
            eids = [ex.add(s, dt) for s, dt in mix]
            ex.finalize()
            ex.run(3, graph=False)
            ex = [ex.state(j) for j in eids]
        for single in (False, True):
            sub = [(k, s, dt) for k, (s, dt) in enumerate(mix) if not single or dim == 2 or
                   s.N <= 5]
            for graph in (False, True):
                b = Batch(eqn, dim, prec, prepared=True)
                ids = [b.add(s, dt) for _, s, dt in sub]
                b.finalize(single_launch=single)
                b.run(3, graph=graph)
                b.check()
                for (k, s, _), i in zip(sub, ids):
                    got = b.state(i)
                    tag = (prec, s.N, single, graph)
                    if prec == FP64:
                        assert max_ulp(got, want[k]) == 0, tag
                    elif prec == FP64_FMA:
                        assert rel_l2(got, want[k]) <= 1e-14, tag
                    else:
                        assert rel_l2(got[:, :3], ex[k][:, :3]) <= 1e-5, tag
                del b
        torch.cuda.empty_cache()


def test_prepared_layout_must_match_the_entry(cuda):
    """A prepared solve's entry launched through the reference-array entry
    (and the other way round) is skipped and reported."""
    s, dt = swe_mix()[0]
    for prepared in (True, False):
        b = Batch(SWE, 2, prepared=prepared)
        i = b.add(s, dt)
        b.finalize()
        before = b.state(i)
        items = np.zeros(1, dtype=ITEM_DTYPE)
        items["solve"], items["N"], items["epw"], items["e0"] = 0, s.N, b.epw[s.N], 0
        t = torch.from_numpy(items.view(np.uint8).copy()).cuda()
        b.prepared = not prepared  # route through the other entry
        b._launch(s.N, b.epw[s.N], t.data_ptr(), 1, b.parity, 0.0, 1.0, b._st())
        b.prepared = prepared
        with pytest.raises(CanaryError) as ei:
            b.check()
        assert ei.value.code == ERR_ARG
        assert max_ulp(b.state(i), before) == 0
