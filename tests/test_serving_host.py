"""Host logic of the request loop (no GPU): the seeded request stream, the
footprint rule and the per-request chunk schedule."""
import math

import numpy as np


def test_request_stream_is_seeded_and_matches_the_mix():
    from qusim_amd.serving import make_requests
    a = make_requests(2000, rate=400.0)
    b = make_requests(2000, rate=400.0)
    assert [(r.kind, r.n, r.work, r.arrival) for r in a] == \
        [(r.kind, r.n, r.work, r.arrival) for r in b]
    kinds = [r.kind for r in a]
    share = {k: kinds.count(k) / len(a) for k in ("evo1d", "evo2d", "born2d")}
    # 2000 draws: the shares are within 4 binomial sigmas of 60/30/10
    for k, p in (("evo1d", 0.6), ("evo2d", 0.3), ("born2d", 0.1)):
        assert abs(share[k] - p) < 4 * math.sqrt(p * (1 - p) / 2000)
    t = np.array([r.arrival for r in a])
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    # exponential gaps: mean 1/rate, coefficient of variation 1
    assert abs(gaps.mean() * 400.0 - 1) < 0.1
    assert abs(gaps.std() / gaps.mean() - 1) < 0.1
    for r in a:
        if r.kind == "evo1d":
            assert r.n in (1024, 2048, 4096, 8192, 16384) and 100 <= r.work <= 2000
        elif r.kind == "evo2d":
            assert r.n in (256, 512, 1024) and 50 <= r.work <= 500
            assert len(r.params["bumps"]) == 16
        else:
            assert r.n in (256, 512, 1024) and r.work == 200
            assert 0.5 <= r.params["energy"] <= 4.5


def test_footprint_rule():
    from qusim_amd.serving import Request, footprint
    r1 = Request(0, "evo2d", 1024, 100, 0.0, {})
    assert footprint(r1, 32) == 72 * 1024 * 1024 + (1 << 20)
    assert footprint(r1, 64) == 64 * 1024 * 1024 + (1 << 20)
    r2 = Request(1, "evo1d", 16384, 100, 0.0, {})
    assert footprint(r2, 64) == 192 * 16384 + (1 << 20)   # no fused FP64 kernel at 16384
    r3 = Request(2, "born2d", 512, 200, 0.0, {})
    assert footprint(r3, 32) < footprint(r3, 64)


def test_chunk_schedule_depends_only_on_the_request():
    from qusim_amd.serving import Request, default_quantum
    r = Request(0, "evo2d", 512, 173, 0.0, {})
    q = default_quantum(r)
    done, chunks = 0, []
    while done < r.work:
        c = min(q, r.work - done)
        chunks.append(c)
        done += c
    assert chunks == [50, 50, 50, 23]


class _FakeBorn:
    """Stands in for a device Born solve: request 1 fails on its first turn."""

    def __init__(self, req, precision):
        self.req, self.done = req, 0

    def advance(self, q, stream):
        if self.req.rid == 1:
            raise RuntimeError("planted failure")
        self.done += q

    def finished(self):
        return self.done >= self.req.work

    def result(self, keep, stream=None):
        return {"iterations": self.done, "dsigma": [1.0]}

    def close(self):
        pass


def _host_only(monkeypatch):
    from qusim_amd import serving
    monkeypatch.setattr(serving.N, "init", lambda *a, **k: None)
    monkeypatch.setattr(serving.N, "call", lambda *a, **k: 0)
    monkeypatch.setattr(serving, "BornJob", _FakeBorn)
    return serving


def test_failed_born_request_is_reported_and_the_rest_complete(monkeypatch):
    """a solve that raises inside a worker thread fails alone: its error is
    recorded, the other requests finish, and run() returns"""
    serving = _host_only(monkeypatch)
    reqs = [serving.Request(i, "born2d", 256, 200, 0.0, {}) for i in range(4)]
    s = serving.Server(precision=32, realtime=False, born_workers=2)
    st = s.run(reqs)
    assert set(s.errors) == {1} and "planted failure" in s.errors[1]
    assert set(s.results) == {0, 2, 3}
    assert st["completed"] == 3 and st["errors"] == 1


def test_dead_born_worker_stops_the_run(monkeypatch):
    """a worker that cannot start (no stream) ends run() with an error
    instead of leaving its requests waiting forever"""
    import pytest
    serving = _host_only(monkeypatch)

    def no_stream(name, *a, **k):
        if name == "qsim_hip_stream_create":
            raise RuntimeError("no device")
        return 0

    monkeypatch.setattr(serving.N, "call", no_stream)
    reqs = [serving.Request(i, "born2d", 256, 200, 0.0, {}) for i in range(2)]
    with pytest.raises(RuntimeError, match="Born worker stopped"):
        serving.Server(precision=32, realtime=False, born_workers=1).run(reqs)


def test_request_beyond_the_budget_is_refused_and_the_rest_complete(monkeypatch):
    """a request whose footprint exceeds the whole HBM budget is refused
    with an error at admission (it would otherwise wait forever); the
    others are admitted and finish"""
    serving = _host_only(monkeypatch)
    reqs = [serving.Request(0, "born2d", 256, 200, 0.0, {}),
            serving.Request(1, "born2d", 1024, 200, 0.0, {}),
            serving.Request(2, "born2d", 256, 200, 0.0, {})]
    budget = serving.footprint(reqs[0], 32) * 2
    assert serving.footprint(reqs[1], 32) > budget
    s = serving.Server(precision=32, hbm_budget=budget, realtime=False, born_workers=2)
    st = s.run(reqs)
    assert set(s.errors) == {1} and "exceeds the HBM budget" in s.errors[1]
    assert set(s.results) == {0, 2}
    assert st["completed"] == 2 and st["errors"] == 1
    assert s.peak_used <= budget
