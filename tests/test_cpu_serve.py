"""Serving scheduler on the CPU with a fake engine: admission by HBM budget, LRU eviction,
rejection of oversize requests, continuous batching and per-request result routing."""
import time

import numpy as np

from ddfmm_amd.serve import Request, Server, latency_stats, poisson_requests


class FakeEngine:
    def __init__(self, key, N, nrhs_slot, nbytes, delay=0.003):
        self.key, self.N, self.nrhs_slot, self.bytes, self.delay = key, N, nrhs_slot, nbytes, delay
        self.slots = [dict(busy=None), dict(busy=None)]
        self.launches = []
        self.closed = False

    def free_slot(self):
        return next((s for s in self.slots if s["busy"] is None), None)

    def launch(self, slot, den, token):
        # a linear map standing in for the FMM: column-wise scale by (1 + key length)
        slot["busy"] = (token, den * (1 + len(self.key)), time.perf_counter() + self.delay)
        self.launches.append(den.shape[0])

    def poll(self):
        out = []
        for s in self.slots:
            if s["busy"] is not None and time.perf_counter() >= s["busy"][2]:
                out.append((s["busy"][0], s["busy"][1]))
                s["busy"] = None
        return out

    def idle(self):
        return all(s["busy"] is None for s in self.slots)

    def close(self):
        self.closed = True


def make_server(budget, sizes, slot=16):
    made = {}

    def mk(key):
        e = FakeEngine(key, sizes[key][0], slot, sizes[key][1])
        made.setdefault(key, []).append(e)
        return e

    return Server(mk, lambda key: sizes[key][1], budget), made


def test_results_routed_and_batched():
    sizes = {"a": (50, 100), "bb": (80, 100)}
    srv, made = make_server(1000, sizes)
    reqs = poisson_requests({k: v[0] for k, v in sizes.items()}, qps=3000, count=60, seed=1, max_rhs=5)
    srv.run(reqs)
    for r in reqs:
        assert r.error is None
        assert np.allclose(r.result, r.den * (1 + len(r.geometry)))
    # continuous batching: more requests than launches
    launches = sum(len(e.launches) for es in made.values() for e in es)
    assert launches < len(reqs)
    assert all(n <= 16 for es in made.values() for e in es for n in e.launches)
    st = latency_stats(reqs)
    assert st["served"] == 60 and st["p99_ms"] >= st["p50_ms"] > 0


def test_admission_evicts_idle_lru_and_rejects_oversize():
    sizes = {"a": (10, 600), "b": (10, 600), "huge": (10, 5000)}
    srv, made = make_server(1000, sizes)
    den = np.ones((2, 10), complex)
    reqs = [Request(0, "a", den, 0.0), Request(1, "b", den, 0.05), Request(2, "huge", den, 0.06),
            Request(3, "a", den, 0.1)]
    srv.run(reqs)
    assert reqs[2].error and "HBM budget" in reqs[2].error
    for i in (0, 1, 3):
        assert reqs[i].error is None and np.allclose(reqs[i].result, den * (1 + len(reqs[i].geometry)))
    assert srv.evictions >= 2                     # a evicted for b, then b for a
    assert srv.resident_bytes() <= 1000
    assert made["a"][0].closed


def test_request_wider_than_slot_is_rejected_without_disturbing_others():
    sizes = {"a": (10, 100)}
    srv, _ = make_server(1000, sizes, slot=4)
    reqs = [Request(0, "a", np.ones((9, 10), complex), 0.0), Request(1, "a", np.ones((3, 10), complex), 0.0)]
    srv.run(reqs)
    assert reqs[0].error and "slot capacity" in reqs[0].error
    assert reqs[1].error is None and np.allclose(reqs[1].result, 2 * np.ones((3, 10)))
