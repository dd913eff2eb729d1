"""Served form on the GPU: continuous batching gives every request exactly
the result it gets alone (the batched kernels compute each element with the
same code whatever else is resident), admission keeps the resident HBM
footprint within the budget, and every request finishes."""
import numpy as np
import pytest

import canary_amd.kernels as K
import canary_amd.problems as P
from canary_amd.perturb import Perturb32
from canary_amd.serve import (Request, Server, make_setup_device, make_setup_host, poisson_stream,
                              summary)

from gpu_util import host
from states import max_ulp, rel_l2

pytestmark = pytest.mark.gpu


def small_stream(seed=7, n=8):
    return poisson_stream(n=n, rate=200.0, seed=seed, vortex_ne=(3, 6), vortex_n=(3, 4, 5, 7),
                          bubble_ne=(2, 3), bubble_n=(2, 3, 4), steps=(3, 12))


def fp32_alone(s, dt, nsteps):
    pf = Perturb32(s, s.Q0d, s.Qd)
    pf.run(dt, nsteps)
    pf.check()
    return host(pf.full_state(), (s.Np, s.nstate, s.nelem))


def fp64_alone(s, dt, nsteps):
    A, B, _ = P.rk_coefs()
    a = P.device_arrays(s)
    q, q2, rhs = s.Qd.clone(), s.Qd.clone(), s.Qd.new_zeros(s.Qd.shape)
    for _ in range(nsteps):
        for st in range(5):
            K.stage_fused(s.dim, s.N, rhs, q, q2, a["vgeo"], a["sgeo"], a["D"], s.nelem,
                          a["vmapM"], a["vmapP"], a["bndy"], A[(st + 1) % 5], B[st], dt,
                          eqn=s.eqn, phys=s.phys)
            q, q2 = q2, q
    K.check_errors()
    return host(q, (s.Np, s.nstate, s.nelem))


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_served_results_equal_alone(cuda, precision):
    reqs = small_stream()
    hosts = {r.rid: make_setup_host(r) for r in reqs}
    sizes = sorted(Server(precision).footprint(s) for s in hosts.values())
    budget = sizes[-1] + sizes[-2]  # at most about two of the largest resident at once
    srv = Server(precision, hbm_budget=budget, step_chunk=4)
    out = srv.serve(reqs)
    st = summary(out, srv.makespan)
    print(precision, {k: round(v, 4) if isinstance(v, float) else v for k, v in st.items()},
          flush=True)
    assert all(r.done is not None and r.latency > 0 for r in out)
    assert all(r.admitted >= r.prepared >= r.arrival for r in out)
    assert max(res for _, _, res in srv.trace) <= max(budget, sizes[-1])
    alone = fp32_alone if precision == "fp32" else fp64_alone
    for r in out:
        s, dt = make_setup_device(r, hosts[r.rid])
        assert dt == r.dt
        want = alone(s, dt, r.nsteps)
        assert np.isfinite(want).all()
        assert max_ulp(r.result, want) == 0, "request %d (%s N=%d)" % (r.rid, r.kind, r.N)
        if precision == "fp32":
            ref = fp64_alone(s, dt, r.nsteps)
            err = rel_l2(r.result, ref)
            assert err <= 1e-5, (r.rid, err)


def test_served_in_mix_equals_served_alone(cuda):
    reqs = small_stream(seed=11, n=6)
    hosts = {r.rid: make_setup_host(r) for r in reqs}
    mixed = Server("fp32", step_chunk=3).serve(reqs, setups=hosts)
    for r in mixed:
        solo = Request(rid=r.rid, kind=r.kind, Ne=r.Ne, N=r.N, nsteps=r.nsteps, arrival=0.0,
                       seed=r.seed)
        got = Server("fp32").serve([solo], setups=hosts)[0]
        assert max_ulp(got.result, r.result) == 0


def test_graph_replay_in_the_serving_loop(cuda):
    """Each batch step replayed from its captured hipGraph (graphs
    re-captured as requests join and leave, meshes prepared on worker
    threads meanwhile, work tiers): every result equals the launch-by-launch
    server's, bit for bit."""
    reqs = small_stream(seed=13, n=8)
    hosts = {r.rid: make_setup_host(r) for r in reqs}
    works = sorted(hosts[r.rid].ndof() * r.nsteps for r in reqs)
    plain = Server("fp32", step_chunk=2).serve(reqs, setups=hosts)
    by = {r.rid: r for r in plain}
    for tier in (None, works[len(works) // 2]):
        again = [Request(rid=r.rid, kind=r.kind, Ne=r.Ne, N=r.N, nsteps=r.nsteps,
                         arrival=r.arrival, seed=r.seed) for r in reqs]
        print("graph replay, tier_work %s" % tier, flush=True)
        got = Server("fp32", step_chunk=2, graph=True, tier_work=tier, tier_ratio=3).serve(
            again, setups=None if tier is None else hosts)
        for r in got:
            assert max_ulp(r.result, by[r.rid].result) == 0, r.rid


def test_work_tiers_give_the_same_results(cuda):
    """With work tiers (small requests in a batch of their own, the large
    ones advanced every third iteration) every result equals the untiered
    server's, bit for bit."""
    reqs = small_stream(seed=12, n=8)
    hosts = {r.rid: make_setup_host(r) for r in reqs}
    works = sorted(hosts[r.rid].ndof() * r.nsteps for r in reqs)
    flat = Server("fp32", step_chunk=2).serve(reqs, setups=hosts)
    again = [Request(rid=r.rid, kind=r.kind, Ne=r.Ne, N=r.N, nsteps=r.nsteps,
                     arrival=r.arrival, seed=r.seed) for r in reqs]
    srv = Server("fp32", step_chunk=2, tier_work=works[len(works) // 2], tier_ratio=3)
    tiered = srv.serve(again, setups=hosts)
    assert {t for (_, t), b in srv.batches.items()} == {0, 1}
    by = {r.rid: r for r in flat}
    for r in tiered:
        assert max_ulp(r.result, by[r.rid].result) == 0, r.rid


@pytest.mark.parametrize("kind", ["bubble3d", "vortex2d"])
def test_device_preparation_matches_host(cuda, kind):
    """Geometry bit-identical; initial states within a few ulps (device libm);
    the time-step reductions within a few ulps of the host formulas."""
    r = Request(rid=0, kind=kind, Ne=3, N=4, nsteps=1, arrival=0.0, seed=5)
    s, dt = make_setup_device(r, make_setup_host(r))
    if kind == "bubble3d":
        h = P.bubble3d(Ne=3, N=4, dtheta=2.0, jitter=1e-3, seed=5)
        dth = P.cfl(h, h.Q) / h.N ** np.sqrt(2)
    else:
        h = P.vortex2d(Ne=3, N=4)
        dth = P.euler_courant(h, h.Q)
    assert max_ulp(host(s.dev["vgeo"], h.vgeo.shape), h.vgeo) == 0
    assert max_ulp(host(s.dev["sgeo"], h.sgeo.shape), h.sgeo) == 0
    Q, Q0 = host(s.Qd, h.Q.shape), host(s.Q0d, h.Q.shape)
    assert rel_l2(Q, h.Q) <= 1e-15 and rel_l2(Q0, h.Q0) <= 1e-15
    assert max_ulp(np.array([dt]), np.array([dth])) <= 8


def test_stream_of_64_under_a_binding_budget(cuda):
    """The 64-request Poisson stream at its own sizes (mean 200 requests/s,
    seed 2024; 2D isentropic vortices of 16^2-64^2 quads with N 3..7 and 3D
    rising bubbles of 8^3-24^3 hexes with N 3..5, 50-500 LSRK54 steps
    each), FP32 perturbation form, under an HBM budget of a quarter of the
    stream's summed footprint:
      * the resident footprint never exceeds the budget, and requests wait
        for room (admission by footprint binds);
      * every result is bit-identical to the same request served alone;
      * every result is within 1e-5 relative L2 of the request's plain FP64
        solve."""
    reqs = poisson_stream(n=64, rate=200.0, seed=2024)
    hosts = {r.rid: make_setup_host(r) for r in reqs}
    probe = Server("fp32")
    fps = {rid: probe.footprint(s) for rid, s in hosts.items()}
    del probe
    budget = max(max(fps.values()), sum(fps.values()) // 4)
    srv = Server("fp32", hbm_budget=budget)
    out = srv.serve(reqs, setups=hosts)
    st = summary(out, srv.makespan)
    waited = [r for r in out if r.admitted - r.prepared > 5e-3]
    print({k: round(v, 4) if isinstance(v, float) else v for k, v in st.items()},
          "budget %.0f MB of %.0f MB, %d requests waited for room"
          % (budget / 1e6, sum(fps.values()) / 1e6, len(waited)), flush=True)
    assert all(r.done is not None and r.latency > 0 for r in out)
    assert max(res for _, _, res in srv.trace) <= budget
    assert len(waited) >= 8
    for r in out:
        solo = Request(rid=r.rid, kind=r.kind, Ne=r.Ne, N=r.N, nsteps=r.nsteps, arrival=0.0,
                       seed=r.seed)
        got = Server("fp32", epw=srv.epw).serve([solo], setups=hosts)[0]
        assert max_ulp(got.result, r.result) == 0, "request %d (%s N=%d)" % (r.rid, r.kind, r.N)
        s, dt = make_setup_device(r, hosts[r.rid])
        assert dt == r.dt
        ref = fp64_alone(s, dt, r.nsteps)
        s.dev = s.Qd = s.Q0d = None
        err = rel_l2(r.result, ref)
        assert err <= 1e-5, (r.rid, r.kind, r.N, err)
    print("max FP32 vs FP64 error checked for %d requests" % len(out), flush=True)
