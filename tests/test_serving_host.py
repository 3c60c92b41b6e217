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
