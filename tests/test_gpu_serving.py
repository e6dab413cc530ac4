"""Serving path on the GPU: the in-library step loop equals step-by-step
propagation, the footprint bound holds for what the library allocates, and a
request's result is the same alone and beside other simulations."""
import numpy as np
import pytest

from pybrown_amd.serving import Server, make_request, request_footprint, run_one, stream_params, summary

pytestmark = pytest.mark.gpu


def _box(req):
    from pybrown_amd.bead import Bead
    from pybrown_amd.box import Box
    return Box([Bead(req.xyz[i].copy(), float(req.radii[i]), label="XXX") for i in range(req.N)], req.inp)


def _request(mode, N, rid=0, steps=12):
    p = next(q for q in stream_params(64) if q["mode"] == mode)
    return make_request(dict(p, N=N, rid=rid, seed=100 + rid), steps=steps)


@pytest.mark.parametrize("mode", ["rpy_smith", "rpy", "rpy_smith_lub"])
def test_run_equals_stepwise_propagation(mode):
    req = _request(mode, 70)
    freqs = dict(diff_freq=2, lub_freq=3, chol_freq=2)
    a = _box(req)
    b = _box(req)
    try:
        flux, frames = a.run(req.dt, req.steps, j0=0, xyz_every=3, **freqs)
        fl, fr = [], []
        for j in range(req.steps):
            if j % 3 == 0:
                fr.append(b._pos.copy())
            b.propagate(req.dt, j % freqs["diff_freq"] == 0, j % freqs["lub_freq"] == 0, j % freqs["chol_freq"] == 0)
            fl.append([b.net_flux["XXX"]])
        assert np.array_equal(a._pos, b._pos)
        assert np.array_equal(flux, np.array(fl, dtype=np.int64))
        assert np.array_equal(frames, np.array(fr))
        assert a.net_flux == b.net_flux
        # a second call continues the same trajectory (j0 carries the cadence)
        flux2, _ = a.run(req.dt, 5, j0=req.steps, **freqs)
        for j in range(req.steps, req.steps + 5):
            b.propagate(req.dt, j % freqs["diff_freq"] == 0, j % freqs["lub_freq"] == 0, j % freqs["chol_freq"] == 0)
        assert np.array_equal(a._pos, b._pos)
    finally:
        a.close()
        b.close()


def test_run_reports_a_failing_factorisation():
    # free-space RPY with minimum-image vectors in a dense periodic box is not
    # positive definite: the first step's factorisation fails
    req = _request("rpy", 300, steps=5)
    req.inp["box_length"] = 150.0
    req.xyz = req.xyz - 150.0 * np.round(req.xyz / 150.0)
    box = _box(req)
    try:
        with pytest.raises(np.linalg.LinAlgError, match="not positive definite"):
            box.run(req.dt, req.steps)
    finally:
        box.close()


@pytest.mark.parametrize("mode,N", [("rpy_smith", 40), ("rpy_smith", 700), ("rpy", 333), ("rpy_smith_lub", 250)])
def test_footprint_bounds_the_allocation(mode, N):
    req = _request(mode, N, steps=1)
    box = _box(req)
    try:
        held = box.device_bytes
    finally:
        box.close()
    est = request_footprint(req)
    print("%s N=%d held %d estimate %d" % (mode, N, held, est))
    assert held <= est <= held * 1.05 + (2 << 20)


def test_results_are_the_same_alone_and_in_the_mix():
    stream = stream_params(400)
    params = []
    for mode in ("rpy_smith", "rpy", "rpy_smith_lub"):
        params += [p for p in stream if p["N"] <= 160 and p["mode"] == mode][:3]
    # lubrication requests with 16 and more tile columns (3N > 960): their
    # factorisations take the one-workgroup-per-CU build with helper jobs
    params += [p for p in stream if 320 <= p["N"] <= 450 and p["mode"] == "rpy_smith_lub"][:2]
    params.sort(key=lambda p: p["rid"])
    assert len(params) == 11
    reqs = [make_request(p, steps=10) for p in params]
    alone = Server(slots=1).run(reqs, realtime=False)
    mixed = Server(slots=4).run(reqs, realtime=False)
    for a, m in zip(alone, mixed):
        assert a.status == "ok" and m.status == "ok", (a.status, m.status)
        assert np.array_equal(a.xyz, m.xyz), a.rid
        assert np.array_equal(a.flux, m.flux), a.rid
    # and the same as a direct run
    f, x, _ = run_one(reqs[0])
    assert np.array_equal(x, alone[0].xyz) and np.array_equal(f, alone[0].flux)
    s = summary(mixed)
    assert s["completed"] == len(reqs)


def test_c5_sample_served_against_oracle():
    """A stratified sample of the served request stream (request k: seed k),
    served concurrently on 8 slots with 20 steps each (the stream's 200 cut
    to 20 so the CPU oracle stays within the test's time): every request
    with N <= 200 follows the oracle trajectory (oracle/box_oracle.py, the
    reference's noise stream) with its per-step flux exact and final
    positions within 1e-10 relative; the sample's largest lubrication
    request (3N above 960, 16+ tile columns) and a large Ewald request give
    the same bits served in the mix as run alone."""
    from oracle.box_oracle import OracleBox
    stream = stream_params(4096)
    small = []
    for mode in ("rpy_smith", "rpy", "rpy_smith_lub"):
        small += [p for p in stream if p["N"] <= 200 and p["mode"] == mode][:2]
    big_lub = next(p for p in stream if 330 <= p["N"] <= 600 and p["mode"] == "rpy_smith_lub")
    big_ewald = next(p for p in stream if 700 <= p["N"] <= 1000 and p["mode"] == "rpy_smith")
    params = sorted(small + [big_lub, big_ewald], key=lambda p: p["rid"])
    reqs = [make_request(p, steps=20) for p in params]
    served = Server(slots=8).run(reqs, realtime=False)
    for r in served:
        assert r.status == "ok", (r.rid, r.status)
    by_rid = {r.rid: r for r in served}
    for req in reqs:
        res = by_rid[req.rid]
        if req.N <= 200:
            o = OracleBox(req.xyz, req.radii, ["XXX"] * req.N, req.inp)
            flux = []
            for _ in range(req.steps):
                o.propagate(req.dt)
                flux.append(o.net_flux["XXX"])
            err = np.abs(res.xyz - o.r).max() / np.abs(o.r).max()
            print("rid %d %s N=%d: position error %.3e after %d steps" % (req.rid, req.mode, req.N, err, req.steps))
            assert err <= 1e-10, req.rid
            assert np.array_equal(res.flux[:, 0], np.array(flux)), req.rid
        if req.rid in (big_lub["rid"], big_ewald["rid"]):
            f, x, _ = run_one(req)
            assert np.array_equal(x, res.xyz) and np.array_equal(f, res.flux), req.rid
