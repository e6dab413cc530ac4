"""Serving loop host logic (no GPU): the request stream, the footprint bound
and admission under slot and HBM-budget limits, with a stand-in runner."""
import threading
import time

import numpy as np
import pytest

from pybrown_amd.serving import (MODES, OPEN_BOX, Request, Server, footprint_bytes, make_request, stream_params,
                                 summary)


def test_stream_is_deterministic_and_follows_the_mix():
    a = stream_params(4096, rate=10.0)
    b = stream_params(4096, rate=10.0)
    assert a == b
    Ns = np.array([p["N"] for p in a])
    assert Ns.min() >= 32 and Ns.max() <= 3000
    # log-uniform: the median of N is near sqrt(32 * 3000) = 310
    assert 250 < np.median(Ns) < 380
    phis = np.array([p["phi"] for p in a])
    assert phis.min() >= 0.02 and phis.max() <= 0.30
    frac = {m: sum(p["mode"] == m for p in a) / len(a) for m in MODES}
    assert abs(frac["rpy_smith"] - 0.60) < 0.03 and abs(frac["rpy"] - 0.25) < 0.03
    assert abs(frac["rpy_smith_lub"] - 0.15) < 0.03
    t = np.array([p["arrival"] for p in a])
    assert np.all(np.diff(t) >= 0)
    # Poisson arrivals at rate 10/s: mean gap 0.1 s
    assert abs(np.diff(t).mean() - 0.1) < 0.01
    assert [p["seed"] for p in a[:5]] == [0, 1, 2, 3, 4]
    # without a rate every request is there at the start, same parameters
    c = stream_params(16)
    assert all(p["arrival"] == 0.0 for p in c)
    assert [p["N"] for p in c] == [p["N"] for p in a[:16]]


def test_make_request_boxes():
    ps = stream_params(64)
    for mode in MODES:
        p = dict(next(q for q in ps if q["mode"] == mode), N=40)
        r = make_request(p, steps=5)
        assert r.N == 40 and r.xyz.shape == (40, 3) and r.steps == 5
        assert r.inp["check_overlaps"] is (mode == "rpy_smith_lub") and r.inp["seed"] == p["seed"]
        if mode == "rpy":
            assert r.inp["box_length"] == OPEN_BOX
        else:
            L = r.inp["box_length"]
            assert np.all(np.abs(r.xyz) <= L / 2)
            assert abs(40 * 4 / 3 * np.pi * 1000.0 / L ** 3 - p["phi"]) < 1e-9
        assert ("ewald_alpha" in r.inp) == (mode != "rpy")
        assert ("lubrication_cutoff" in r.inp) == (mode == "rpy_smith_lub")
        assert r.inp.get("propagation_scheme", "ermak") == ("midpoint" if mode == "rpy_smith_lub" else "ermak")
        if mode == "rpy_smith_lub":
            from oracle.orc import Oracle
            P = Oracle().pointer_pbc_matrix(r.xyz, r.inp["box_length"])
            d = np.linalg.norm(P, axis=2) + 1e9 * np.eye(r.N)
            assert d.min() >= 20.1 - 1e-9


def test_footprint_grows_with_size_and_mode():
    f = [footprint_bytes(N, "rpy_smith") for N in (32, 100, 1000, 3000)]
    assert all(x < y for x, y in zip(f, f[1:]))
    # lubrication holds four dense 3008 x 3008 FP64 matrices at N = 1000 (D,
    # B, the far-field resistance, an inverse scratch), the other matrix
    # modes three (D, B, the look-ahead copy of D)
    assert footprint_bytes(1000, "rpy_smith_lub") - footprint_bytes(1000, "rpy_smith") >= 8 * 3008 ** 2
    assert footprint_bytes(1000, "rpy_smith", check_overlaps=True) > footprint_bytes(1000, "rpy_smith")
    assert footprint_bytes(1000, "rpy_smith") > 3 * 8 * 3008 ** 2


class _Probe:
    """Stand-in runner: sleeps in proportion to N and records how many
    requests and how many footprint bytes are resident at once."""

    def __init__(self):
        self.lock = threading.Lock()
        self.active = 0
        self.mem = 0
        self.max_active = 0
        self.max_mem = 0
        self.order = []

    def __call__(self, req):
        fp = footprint_bytes(req.N, req.mode)
        with self.lock:
            self.active += 1
            self.mem += fp
            self.max_active = max(self.max_active, self.active)
            self.max_mem = max(self.max_mem, self.mem)
            self.order.append(req.rid)
        time.sleep(0.002 + req.N * 1e-5)
        with self.lock:
            self.active -= 1
            self.mem -= fp
        return np.zeros((req.steps, 1), dtype=np.int64), req.xyz.copy(), fp


def _fake_requests(params):
    return [Request(p["rid"], np.zeros((p["N"], 3)), np.full(p["N"], 10.0), {"hydrodynamics": p["mode"]}, 1.0, 3,
                    p["arrival"]) for p in params]


def test_admission_respects_slots_and_budget():
    reqs = _fake_requests(stream_params(60, rate=2000.0))
    budget = 3 * footprint_bytes(3000, "rpy_smith_lub")
    probe = _Probe()
    res = Server(slots=5, hbm_budget=budget, runner=probe).run(reqs)
    assert [r.rid for r in res] == [q.rid for q in reqs]
    assert all(r.status == "ok" for r in res)
    assert probe.max_active <= 5
    assert probe.max_mem <= budget
    # first come, first served
    assert probe.order == sorted(probe.order)
    for r in res:
        assert r.arrival <= r.admit + 1e-9 <= r.start + 1e-6 and r.start <= r.end
    s = summary(res)
    assert s["completed"] == 60 and s["failed"] == 0 and s["p50_s"] <= s["p99_s"]


def test_admission_waits_for_memory():
    # budget for one large simulation at a time: they must not overlap
    reqs = _fake_requests([{"rid": k, "N": 3000, "mode": "rpy_smith", "arrival": 0.0} for k in range(4)])
    probe = _Probe()
    Server(slots=4, hbm_budget=1.5 * footprint_bytes(3000, "rpy_smith"), runner=probe).run(reqs)
    assert probe.max_active == 1


def test_oversized_request_is_rejected_not_stalled():
    reqs = _fake_requests([{"rid": 0, "N": 3000, "mode": "rpy_smith_lub", "arrival": 0.0},
                           {"rid": 1, "N": 50, "mode": "rpy", "arrival": 0.0}])
    res = Server(slots=2, hbm_budget=footprint_bytes(1000, "rpy_smith"), runner=_Probe()).run(reqs)
    assert res[0].status.startswith("rejected")
    assert res[1].status == "ok"


def test_runner_errors_are_reported_per_request():
    def runner(req):
        if req.rid == 1:
            raise RuntimeError("boom")
        return np.zeros((1, 1), dtype=np.int64), req.xyz, 0
    reqs = _fake_requests(stream_params(3))
    res = Server(slots=2, runner=runner).run(reqs)
    assert [r.status for r in res] == ["ok", "error: boom", "ok"]
    assert summary(res)["completed"] == 2


def test_realtime_arrivals_are_honoured():
    reqs = _fake_requests([{"rid": k, "N": 32, "mode": "rpy", "arrival": 0.02 * k} for k in range(4)])
    res = Server(slots=4, runner=_Probe()).run(reqs)
    for r in res:
        assert r.admit >= r.arrival - 1e-9
    assert res[-1].admit >= 0.06 - 1e-3


def test_bad_slots():
    with pytest.raises(ValueError):
        Server(slots=0)
