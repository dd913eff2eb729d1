"""Host side of the served form: the seeded request stream and the latency
summary (no GPU)."""
import numpy as np
import pytest

from canary_amd.serve import Request, Server, poisson_stream, summary


def test_stream_is_seeded_and_in_range():
    a = poisson_stream(n=64, seed=5)
    b = poisson_stream(n=64, seed=5)
    assert [(r.kind, r.Ne, r.N, r.nsteps, r.arrival) for r in a] == \
        [(r.kind, r.Ne, r.N, r.nsteps, r.arrival) for r in b]
    assert poisson_stream(n=64, seed=6)[0].arrival != a[0].arrival
    arr = np.array([r.arrival for r in a])
    assert np.all(np.diff(arr) > 0) and arr[0] > 0
    for r in a:
        assert 50 <= r.nsteps <= 500
        if r.kind == "vortex2d":
            assert 16 <= r.Ne <= 64 and 3 <= r.N <= 7
        else:
            assert r.kind == "bubble3d" and 8 <= r.Ne <= 24 and r.N in (3, 4, 5)


def test_stream_rate_and_mix():
    r = poisson_stream(n=4000, rate=200.0, seed=1)
    gaps = np.diff([0.0] + [x.arrival for x in r])
    assert abs(gaps.mean() - 1 / 200.0) < 0.1 / 200.0
    frac = np.mean([x.kind == "vortex2d" for x in r])
    assert 0.45 < frac < 0.55


def test_summary_percentiles():
    reqs = []
    for i in range(100):
        q = Request(rid=i, kind="vortex2d", Ne=4, N=3, nsteps=10, arrival=float(i))
        q.admitted, q.done, q.dof = float(i), float(i) + (i + 1) * 0.01, 1000
        reqs.append(q)
    s = summary(reqs, makespan=200.0)
    assert abs(s["p50_latency_s"] - np.percentile([(i + 1) * 0.01 for i in range(100)], 50)) < 1e-12
    assert abs(s["p99_latency_s"] - np.percentile([(i + 1) * 0.01 for i in range(100)], 99)) < 1e-12
    assert abs(s["aggregate_gdof_stage_s"] - 100 * 5 * 1000 * 10 / 200.0 / 1e9) < 1e-18


def test_graph_replay_is_an_option_of_the_serving_loop():
    assert Server("fp32", graph=True).graph is True
    assert Server("fp32").graph is False


def test_collected_streams_are_queued_not_destroyed():
    """A finalizer runs in whichever thread triggers the collection, so an
    OwnedStream's finalizer makes no HIP call: it queues the stream for
    reap(), which the runtimes call from their launch thread."""
    import threading
    import canary_amd.streams as S
    before = list(S._graveyard)
    objs = []
    for p in (0x1234, 0x5678):
        o = object.__new__(S.OwnedStream)
        o.ptr = p
        objs.append(o)
    del o

    def drop():  # collected on another thread
        objs.clear()

    t = threading.Thread(target=drop)
    t.start()
    t.join()
    assert sorted(S._graveyard[len(before):]) == [0x1234, 0x5678]
    del S._graveyard[len(before):]  # never destroyed: they are not real streams
