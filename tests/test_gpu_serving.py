"""Serving (kestrel_amd.serving): requests admitted as slots and memory free
up run together with the others, and each result is the same bits as the
scenario run alone; admission keeps resident device memory within the
budget."""
import numpy as np
import pytest

from kestrel_amd import serving
from tests.helpers import FIELDS, stepper_for

pytestmark = pytest.mark.gpu


def small_trace(n=8, rate=8.0):
    return serving.poisson_trace(n, rate, side_range=(64, 256), t_range=(20.0, 60.0))


def test_serving_result_alone_equals_mix(gpu):
    """Eight requests arriving together through four slots: they queue, join
    running scenarios and reuse freed slots; every result equals its run
    alone."""
    reqs = small_trace(rate=1000.0)
    srv = serving.Server(4, 200e9, poll_iters=8, keep_state=True)
    out = srv.run(reqs)
    srv.close()
    assert out["completed"] == len(reqs), [r.status for r in srv.done]
    assert out["peak_resident"] == 4
    for r in srv.done:
        g = stepper_for(r.workload)
        st = g.integrate_to(r.tend)
        a = g.state()
        g.close()
        assert st["steps"] == r.stats["steps"] and st["cell_updates"] == r.stats["cell_updates"]
        assert a["t"] == r.state["t"] == r.tend
        for f in FIELDS:
            assert np.array_equal(a[f], r.state[f]), f"{r.workload.name}: {f}"
    lat = [r.finished - r.arrival for r in srv.done]
    assert out["latency_p50_s"] == pytest.approx(float(np.percentile(lat, 50)))


def test_serving_admission_respects_budget(gpu):
    """With a budget of two scenarios' memory at most two are resident."""
    reqs = small_trace(6)
    need = max(serving.device_bytes(r.workload.cfg) for r in reqs)
    srv = serving.Server(8, 2 * need, poll_iters=8)
    out = srv.run(reqs)
    srv.close()
    assert out["completed"] == len(reqs)
    assert out["peak_device_bytes"] <= 2 * need
    assert out["peak_resident"] <= 2


def test_serving_full_trace_sample_equals_alone(gpu):
    """The serving trace at its stated size: Poisson arrivals at 4 per second,
    sides log-uniform 64-2048, T ~ U[60, 600] s, closures sampled as in the
    ensemble, a 200 GB budget, mixed precision; every request completes and
    a seeded sample of six results (the largest grid among them) are the
    same bits as their runs alone."""
    reqs = serving.poisson_trace(32, 4.0)
    assert max(r.workload.cfg.NX for r in reqs) >= 1024
    srv = serving.Server(256, 200e9, keep_state=True)
    out = srv.run(reqs)
    srv.close()
    assert out["completed"] == len(reqs), [r.status for r in srv.done]
    assert out["latency_p99_s"] >= out["latency_p50_s"] > 0.0
    rng = np.random.default_rng(7)
    done = sorted(srv.done, key=lambda r: r.workload.cfg.NX)
    sample = [done[-1]] + [done[k] for k in rng.choice(len(done) - 1, 5, replace=False)]
    for r in sample:
        g = stepper_for(r.workload)
        st = g.integrate_to(r.tend)
        a = g.state()
        g.close()
        assert st["steps"] == r.stats["steps"] and st["cell_updates"] == r.stats["cell_updates"]
        assert a["t"] == r.state["t"] == r.tend
        for f in FIELDS:
            assert np.array_equal(a[f], r.state[f]), f"{r.workload.name}: {f}"
