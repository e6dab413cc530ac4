"""Headline benchmark: steps/s of the periodic-RPY (Smith-Ewald) Brownian
dynamics step at N = 1000 beads (a = 10 A, phi = 0.05, L = 437.6 A,
alpha = sqrt(pi), |m|_1, |n|_1 <= 2, Ermak-McCammon, dt = 1 ps, flux plane
x = 0), FP64, on one MI355X.

A step is one Box.propagate: Ewald mobility assembly, Cholesky of the
3000 x 3000 diffusion matrix, B.xi propagation with overlap check, flux and
wrap; positions and matrices are resident in HBM, the host supplies the
seeded N(0,1) stream.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def cpu_baseline(name, steps=3, threads=1):
    """Reference CPU path on the host cores: the reference's clib kernels
    (oracle/_ref, built from its sources; the restatement oracle/liborc.so if
    that build is absent) driven by the box.py step restatement with
    numpy.linalg.cholesky.  threads = 1: one BLAS thread, as the reference's
    default build runs (its OpenMP pragma is inert); threads > 1: the OpenMP
    build of the same sources (the reference's parallel loop over bead rows)
    and that many BLAS threads.  Bounded sample: 1 warm-up step + `steps`
    timed steps."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from oracle.box_oracle import OracleBox
    from oracle.orc import ref_available
    from pybrown_amd.workloads import workload

    omp = threads > 1
    if omp:
        os.environ["OMP_NUM_THREADS"] = str(threads)
    ref = ref_available(omp)
    lib = ("ref_omp" if omp else "ref") if ref else "orc"
    # the step itself is this repository's restatement of box.py (the
    # reference's Python is read, not run), so the baseline is a port even
    # where its kernels are the reference's own C
    kind = "port"

    xyz, a, inp, dt, _ = workload(name, seed=0)
    with threadpool_limits(limits=threads):
        box = OracleBox(xyz, a, ["XXX"] * len(a), inp, lib=lib)
        box.propagate(dt)
        times = []
        for _ in range(steps):
            t0 = time.perf_counter()
            box.propagate(dt)
            times.append(time.perf_counter() - t0)
    per_step = float(np.median(times))
    kernels = ("the reference's clib kernels (diff_tensor.c, bead.c) compiled from its sources at -O2 (its own "
               "default build is unoptimised)" if ref else "the oracle's C restatement of those kernels (oracle/orc.c)")
    how = (kernels + ", driven by a numpy restatement of box.py's step (oracle/box_oracle.py) "
           "with numpy.linalg.cholesky; 1 thread" if not omp else
           "the same, the clib sources compiled -O2 -fopenmp (the reference's own parallel loop over bead rows), "
           "%d OpenMP and BLAS threads" % threads)
    return {"value": 1.0 / per_step, "unit": "steps/s", "cores": threads, "kind": kind,
            "sample": "%d timed steps (after 1 warm-up) of the same N=%d workload; median %.3f s/step; %s"
                      % (steps, len(a), per_step, how),
            "host": host_info()}


def host_info():
    """os.cpu_count() and the CPU model of this host (/proc/cpuinfo)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_count": os.cpu_count(), "cpu_model": model}


def host_threads():
    """Threads for the multi-core baseline: the host's cores, at most 16 (one
    GPU's share of the host on the machines this runs on)."""
    return max(1, min(16, os.cpu_count() or 1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--potrf", type=int, default=0, help="Cholesky variant (1, 2, 4, 6; 0 = library default)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed", "mixed_refined", "mixed_factor"],
                    help="dense algebra: FP64; mixed = FP32 matrix-core bulk of the Cholesky and of the SPD "
                         "inverses (FP64 diagonal tiles); mixed_refined = mixed plus one Newton-Schulz step per "
                         "inverse; mixed_factor = the Cholesky's bulk only (c5: the served mode, lubrication "
                         "requests as mixed_factor)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (HBM counter calibration)")
    ap.add_argument("--requests", type=int, default=64, help="c5: requests taken from the head of the stream")
    ap.add_argument("--slots", type=int, default=16, help="c5: concurrent simulations")
    ap.add_argument("--request-steps", type=int, default=200, help="c5: steps per request")
    ap.add_argument("--graphs", action="store_true",
                    help="c5: HIP graphs: whole steps of the in-library runs (Ermak without overlap checks, RPY "
                         "and Ewald), and the factorisation and SPD-inverse sequences elsewhere")
    ap.add_argument("--no-c1-leg", action="store_true", help="c2: skip the small C1 leg")
    ap.add_argument("--no-served-leg", action="store_true", help="c2: skip the served request-stream sample")
    ap.add_argument("--max-n", type=int, default=0, help="c5: keep only the requests with N <= this")
    ap.add_argument("--request", type=int, default=None, metavar="RID",
                    help="c5: time stream request RID alone (--request-steps steps after 2 untimed ones)")
    ap.add_argument("--tune", action="store_true",
                    help="c5: time the factorisation's launch shapes per order first (pybrown_amd.autotune) and "
                         "serve with the tuned table")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it (got --gpus %d)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload == "c5":
        serve_bench(args)
        return
    # the served-sample leg runs 16 simulations at once: one hardware queue
    # each (read when the HIP runtime starts; one stream alone uses one).
    # Set, not defaulted: an inherited smaller value (4 is HIP's default)
    # would put four simulations on each queue.
    os.environ["GPU_MAX_HW_QUEUES"] = os.environ.get("PBX_BENCH_HWQ") or "16"

    import ctypes

    import numpy as np
    import torch

    from pybrown_amd.bead import Bead
    from pybrown_amd.box import Box
    from pybrown_amd.workloads import workload

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_per_copy": nbytes, "copies": 3}))
        return
    from pybrown_amd._lib import lib
    if args.potrf:
        assert lib.pbx_set_potrf_variant(args.potrf) == 0
    xyz, a, inp, dt, _ = workload(args.workload, seed=0)
    N = len(a)
    n = 3 * N
    box = Box([Bead(p, r) for p, r in zip(xyz, a)], inp, xi_pool_attempts=64, precision=args.precision)
    for _ in range(args.warmup):
        box.propagate(dt)
    torch.cuda.synchronize()
    snap = box.snapshot()  # the stage-timing pass below replays the timed steps from here
    # HIP events on the simulation's own stream around the factorisation of
    # every timed step (two marker packets per step; the roofline's time)
    lib.pbx_sim_set_timing(box._sim, 2)
    k0 = lib.pbx_kernel_launches()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        box.propagate(dt)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kernels_per_step = (lib.pbx_kernel_launches() - k0) / args.steps
    ms = (t1 - t0) * 1e3 / args.steps
    st = np.zeros(3)
    nt = lib.pbx_sim_get_timing(box._sim, st.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    chol_timed = st[1] / max(nt, 1)
    # every stage bound (six marker packets per step): the same steps again,
    # replayed from the snapshot (restores are bit-exact, so the trajectory
    # and its attempt batches are the timed steps')
    box.restore(snap)
    lib.pbx_sim_set_timing(box._sim, 1)
    for _ in range(args.steps):
        box.propagate(dt)
    torch.cuda.synchronize()
    nt2 = lib.pbx_sim_get_timing(box._sim, st.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    lib.pbx_sim_set_timing(box._sim, 0)
    stage = st / max(nt2, 1)
    stage[1] = chol_timed
    print("stage ms/step: assembly %.3f  cholesky %.3f (timed steps)  propagation %.3f  (assembly and propagation "
          "over the %d timed steps replayed; wall %.3f)" % (stage[0], stage[1], stage[2], nt2, ms),
          file=sys.stderr)

    accuracy = None
    if args.precision != "fp64":
        accuracy = mixed_accuracy(box, xyz, a, inp, dt, precision=args.precision)
        print("mixed precision errors: %s" % accuracy, file=sys.stderr)
    out = {
        "metric": METRIC[args.workload],
        "value": 1e3 / ms,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE_TEXT[args.precision],
        "data": "synthetic",
        "config": {"workload": WORKLOAD_TEXT[args.workload], "N": N, "box_length": round(inp["box_length"], 4),
                   "precision": args.precision},
        "stage_ms": {"assembly": stage[0], "cholesky": stage[1], "propagation": stage[2]},
        "kernels_per_step": kernels_per_step,
        "roofline": roofline(args.workload, N, stage, args.precision),
    }
    if args.workload == "c2" and not args.no_c1_leg:
        out["c1_leg"] = c1_leg()
    if args.workload == "c2" and not args.no_served_leg:
        out["served_sample"] = served_legs(64, 100, 16, log=lambda m: print(m, file=sys.stderr))
    if accuracy is not None:
        out["accuracy"] = accuracy
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.workload)
        # the same reference path on all host cores (at most 16), beside it
        out["cpu_baseline_multicore"] = cpu_baseline(args.workload, steps=5, threads=host_threads())
    print(json.dumps(out))


def c1_leg(warmup=20, steps=500):
    """The launch-bound small case beside the headline: free-space RPY,
    N = 100 (a = 5 A, L = 150 A), Ermak with overlap checks, flux x = 0;
    steps/s and kernels launched per step."""
    import torch

    from pybrown_amd._lib import lib
    from pybrown_amd.bead import Bead
    from pybrown_amd.box import Box
    from pybrown_amd.workloads import workload
    xyz, a, inp, dt, _ = workload("c1", seed=0)
    box = Box([Bead(p, r) for p, r in zip(xyz, a)], inp)
    try:
        for _ in range(warmup):
            box.propagate(dt)
        torch.cuda.synchronize()
        k0 = lib.pbx_kernel_launches()
        t0 = time.perf_counter()
        for _ in range(steps):
            box.propagate(dt)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        return {"workload": WORKLOAD_TEXT["c1"], "steps": steps, "steps_per_s": steps / (t1 - t0),
                "ms_per_step": (t1 - t0) * 1e3 / steps,
                "kernels_per_step": (lib.pbx_kernel_launches() - k0) / steps}
    finally:
        box.close()


def mixed_accuracy(box, xyz, a, inp, dt, steps=3, precision="mixed"):
    """Errors of the mixed-precision engine against an FP64 twin stepped from
    the same positions with the same noise, each the largest over `steps`
    steps: the noise covariance residual ||B B^T - D||_F / ||D||_F of both
    factors (on the device); the max-norm relative error of the diffusion
    tensor D and (lubrication) of the far-field resistance Rff = inv(Mff);
    the step displacement (2-norm relative); and the per-step diffusion
    coefficient <|d|^2> / (6 dt).  The mobility is assembled in FP64 on
    both paths (identical)."""
    from pybrown_amd.bead import Bead
    from pybrown_amd.box import Box
    ref = Box([Bead(p, r) for p, r in zip(box._pos.copy(), a)], dict(inp, seed=inp["seed"] + 1))
    mix = Box([Bead(p, r) for p, r in zip(box._pos.copy(), a)], dict(inp, seed=inp["seed"] + 1), precision=precision)
    L = inp["box_length"]
    lub = inp["hydrodynamics"].endswith("_lub")
    try:
        errs, derrs, cerrs, rerrs = [], [], [], []
        for _ in range(steps):
            start = ref._pos.copy()
            mix._pos[:] = start
            ref.propagate(dt)
            mix.propagate(dt)
            d_ref = ref._pos - start
            d_ref -= L * np.round(d_ref / L)
            d_mix = mix._pos - start
            d_mix -= L * np.round(d_mix / L)
            errs.append(float(np.linalg.norm(d_mix - d_ref) / np.linalg.norm(d_ref)))
            c_ref = (d_ref ** 2).sum(axis=1).mean() / (6 * dt)
            c_mix = (d_mix ** 2).sum(axis=1).mean() / (6 * dt)
            cerrs.append(float(abs(c_mix - c_ref) / c_ref))
            D_ref = ref.D
            derrs.append(float(np.abs(mix.D - D_ref).max() / np.abs(D_ref).max()))
            if lub:
                R_ref = ref.Rff
                rerrs.append(float(np.abs(mix.Rff - R_ref).max() / np.abs(R_ref).max()))
        out = {"precision": precision, "factor_residual": mix.factor_residual(),
               "fp64_factor_residual": ref.factor_residual(), "D_rel_error": max(derrs),
               "step_displacement_rel_error": max(errs), "diffusion_coefficient_rel_error": max(cerrs),
               "steps_compared": steps}
        if lub:
            out["Rff_rel_error"] = max(rerrs)
        return out
    finally:
        ref.close()
        mix.close()


def served_legs(requests, steps, slots, precision="fp64", graphs=False, log=None, order="cost", max_n=None):
    """The served request stream on the first `requests` requests (request
    k: seed k; pybrown_amd.serving), `steps` steps each:
      plain   one at a time on the plain FP64 path (Box.propagate per step,
              one synchronised step at a time, as BD.py drives a Box);
      run     one at a time at the served precision, all steps inside the
              library (Box.run);
      burst   served on `slots` concurrent simulations, all arrived at the
              start, admitted in `order` (throughput X);
      poisson served with Poisson arrivals at 1.5 X (the reported sims/s and
              p50/p99 time to trajectory).
    Served results are checked against the one-at-a-time ones (bit for bit)
    and, in a mixed precision mode, against the plain FP64 ones (relative
    error of each request's final displacement)."""
    import functools

    import numpy as np

    from pybrown_amd.serving import Server, make_requests_parallel, run_one, run_plain, stream_params, summary
    log = log or (lambda *a: None)
    params = stream_params(requests)
    if max_n:  # the small-request regime: the head of the stream with N <= max_n
        params = [p for p in params if p["N"] <= max_n]
    t = time.perf_counter()
    reqs = make_requests_parallel(params, steps=steps, workers=min(16, os.cpu_count() or 1))
    log("c5: %d requests placed in %.1f s (N %d..%d, %s)" % (
        len(reqs), time.perf_counter() - t, min(r.N for r in reqs), max(r.N for r in reqs),
        {m: sum(r.mode == m for r in reqs) for m in ("rpy_smith", "rpy", "rpy_smith_lub")}))
    budget = 250e9
    runner = functools.partial(run_one, graphs=graphs, precision=precision)
    small = sorted(reqs, key=lambda r: r.N)[:2]
    Server(slots=2, hbm_budget=budget, runner=runner).run(small, realtime=False)  # code objects, allocator
    plain = Server(slots=1, hbm_budget=budget, runner=run_plain).run(reqs, realtime=False)
    log("c5 plain one at a time: %s" % summary(plain))
    one = Server(slots=1, hbm_budget=budget, runner=runner, precision=precision).run(reqs, realtime=False)
    log("c5 run one at a time: %s" % summary(one))
    burst_res = Server(slots=slots, hbm_budget=budget, runner=runner, precision=precision,
                       order=order).run(reqs, realtime=False)
    burst = summary(burst_res)
    log("c5 served, all at start (%s order): %s" % (order, burst))
    # the requests that set the window: the longest one at a time, and when
    # each of them started and ended in the served batch
    longest = sorted(range(len(reqs)), key=lambda i: -(plain[i].end - plain[i].start))[:5]
    log("c5 longest alone (rid N mode: alone s | served start-end s): %s" % "; ".join(
        "%d %d %s: %.2f | %.2f-%.2f" % (reqs[i].rid, reqs[i].N, reqs[i].mode, plain[i].end - plain[i].start,
                                        burst_res[i].start, burst_res[i].end) for i in longest))
    rate = 1.5 * burst["sims_per_s"]
    timed = {p["rid"]: p["arrival"] for p in stream_params(requests, rate=rate)}
    for r in reqs:
        r.arrival = timed[r.rid]
    res = Server(slots=slots, hbm_budget=budget, runner=runner, precision=precision).run(reqs, realtime=True)
    served = summary(res)
    log("c5 served, Poisson arrivals at %.3f/s: %s" % (rate, served))
    failed = [(r.rid, r.N, r.mode, r.status) for r in res + burst_res if r.status != "ok"]
    if failed:
        log("c5 failed requests: %s" % failed)
    # every served result against the one-at-a-time one (bit for bit in FP64)
    same = all(a.status == b.status == c.status == "ok" and np.array_equal(a.xyz, b.xyz) and
               np.array_equal(a.xyz, c.xyz) and np.array_equal(a.flux, b.flux) and np.array_equal(a.flux, c.flux)
               for a, b, c in zip(one, burst_res, res))
    plain_same = all(a.status == p_.status == "ok" and np.array_equal(a.xyz, p_.xyz) and np.array_equal(a.flux, p_.flux)
                     for a, p_ in zip(one, plain))
    base = summary(plain)
    mixed_err = None
    if precision != "fp64":
        # final displacement of each request (min image) against plain FP64;
        # a request whose overlap test decides one attempt differently takes
        # another trajectory from there (the largest errors name them)
        errs, worst = [], []
        for q, a, p_ in zip(reqs, res, plain):
            if a.status == p_.status == "ok":
                L = q.inp["box_length"]
                d64 = p_.xyz - q.xyz
                d64 -= L * np.round(d64 / L)
                dmx = a.xyz - q.xyz
                dmx -= L * np.round(dmx / L)
                errs.append(float(np.linalg.norm(dmx - d64) / np.linalg.norm(d64)))
                worst.append((errs[-1], q.rid, q.N, q.mode))
        worst.sort(reverse=True)
        mixed_err = {"final_displacement_rel_error_max": max(errs), "median": float(np.median(errs)),
                     "largest": [{"rid": r, "N": n_, "mode": m, "error": e} for e, r, n_, m in worst[:3]],
                     "flux_equal": sum(bool(np.array_equal(a.flux, p_.flux)) for a, p_ in zip(res, plain)),
                     "compared": len(errs)}
        log("c5 served %s against plain FP64: %s" % (precision, mixed_err))
    return {
        "requests": len(reqs), "stream_head": requests, "max_n": max_n, "steps_per_request": steps, "slots": slots,
        "precision": precision, "graphs": graphs,
        "sims_per_s": served["sims_per_s"], "p50_s": served["p50_s"], "p99_s": served["p99_s"],
        "completed": served["completed"], "failed": served["failed"], "arrival_rate_per_s": rate,
        "plain_one_at_a_time_sims_per_s": base["sims_per_s"],
        "run_one_at_a_time_sims_per_s": summary(one)["sims_per_s"],
        "burst_sims_per_s": burst["sims_per_s"],
        "speedup_vs_plain_one_at_a_time": served["sims_per_s"] / base["sims_per_s"],
        "served_equal_to_one_at_a_time": bool(same),
        "run_equal_to_plain": bool(plain_same),
        "burst_order": order,
        "against_plain_fp64": mixed_err,
        "gpu_max_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
    }


def serve_bench(args):
    """--workload c5: the served request stream (served_legs) as the bench
    line."""
    # hardware queues the HIP runtime spreads the simulations' streams over;
    # read once when the runtime starts, so set before anything touches it
    os.environ["GPU_MAX_HW_QUEUES"] = os.environ.get("PBX_BENCH_HWQ") or str(min(max(args.slots, 4), 32))
    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    tuned = None
    if args.tune:
        from pybrown_amd import autotune
        res = autotune.tune(grid=[t for t in autotune.DEFAULT_GRID if t <= 160], reps=3,
                            log=lambda m: print("tune " + m, file=sys.stderr))
        autotune.apply(res)
        tuned = {"grid": res["grid"], "codes": res["codes"], "seconds": res["seconds"], "device": res["device"]}
    if args.request is not None:
        return one_request_bench(args)
    leg = served_legs(args.requests, args.request_steps, args.slots, args.precision, args.graphs,
                      log=lambda m: print(m, file=sys.stderr), max_n=args.max_n or None)
    leg["potrf_table"] = tuned
    out = {
        "metric": "served sims/s + p99 latency (request stream)",
        "value": leg["sims_per_s"],
        "unit": "sims/s",
        "n_gpus": 1,
        "steps": args.request_steps,
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE_TEXT[args.precision],
        "data": "synthetic",
        "config": {"workload": "request stream: first %d requests, N log-uniform 32..3000, phi 0.02..0.30, "
                               "60%% rpy_smith / 25%% rpy (open box) ermak / 15%% rpy_smith_lub midpoint, dt=1ps, "
                               "%d steps each, flux x=0, Poisson arrivals at 1.5x the burst throughput"
                               % (args.requests, args.request_steps),
                   "slots": args.slots, "hbm_budget_bytes": 250e9, "graphs": args.graphs},
        "served": leg,
        "roofline": None,
    }
    print(json.dumps(out))


def one_request_bench(args):
    """--workload c5 --request RID: one request of the stream alone, in the
    library's queued runs (Box.run), its step time after two untimed steps."""
    from pybrown_amd import serving
    from pybrown_amd.bead import Bead
    from pybrown_amd.box import Box
    p = serving.stream_params(max(args.requests, args.request + 1))[args.request]
    req = serving.make_request(p, steps=args.request_steps + 2)
    beads = [Bead(req.xyz[i].copy(), float(req.radii[i]), label="XXX") for i in range(req.N)]
    box = Box(beads, req.inp, graphs=args.graphs, precision=serving.request_precision(req, args.precision))
    try:
        box.run(req.dt, 2)
        t0 = time.perf_counter()
        box.run(req.dt, args.request_steps)
        dt = time.perf_counter() - t0
    finally:
        box.close()
    ms = dt * 1e3 / args.request_steps
    print(json.dumps({"metric": "ms/step of one stream request", "value": ms, "unit": "ms/step",
                      "higher_is_better": False, "request": p, "n": 3 * req.N, "steps": args.request_steps,
                      "dtype": DTYPE_TEXT[args.precision], "data": "synthetic"}))


METRIC = {"c2": "steps/s at N=1000 RPY-Ewald", "c1": "steps/s at N=100 RPY", "c3": "steps/s at N=500 RPY-Ewald+JO midpoint",
          "c4": "steps/s at N=4000 RPY-Ewald"}
DTYPE_TEXT = {"fp64": "f64",
              "mixed": "f64 + f32 (matrix-core bulk of the Cholesky and of the SPD inverses)",
              "mixed_refined": "f64 + f32 (matrix-core bulk; SPD inverses refined by one FP64-residual "
                               "Newton-Schulz step)",
              "mixed_factor": "f64 + f32 (matrix-core bulk of the Cholesky; SPD inverses f64)"}
WORKLOAD_TEXT = {
    "c1": "rpy N=100 a=5 L=150 ermak dt=1ps flux x=0",
    "c2": "rpy_smith N=1000 a=10 phi=0.05 L=437.6 alpha=sqrt(pi) m=n=2 ermak dt=1ps flux x=0",
    "c3": "rpy_smith_lub N=500 a=20/10 phi=0.25 cutoff=0.5 midpoint m=100 dt=0.1ps",
    "c4": "rpy_smith N=4000 a=10 phi=0.10 L=551.3 alpha=sqrt(pi) m=n=2 ermak dt=1ps",
}

FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 (vector and matrix)


def smith_flops(N, m=2, nrec=2):
    """Algorithmic flops of one Smith-Ewald assembly: per pair 63 per
    real-space image (sqrt/erfc/exp counted as 1) and 10 per reciprocal
    vector; self blocks skip the m = 0 image."""
    def l1_ball(k):
        return sum(1 for x in range(-k, k + 1) for y in range(-k, k + 1) for z in range(-k, k + 1)
                   if abs(x) + abs(y) + abs(z) <= k)
    real, recip = l1_ball(m), l1_ball(nrec) - 1
    pair = real * 63 + recip * 10
    self_ = (real - 1) * 63 + recip * 10
    return pair * N * (N - 1) / 2 + self_ * N


def measured_traffic(kernel_key):
    """HBM bytes per launch of `kernel_key` from the committed counter summary
    (profiles/hbm_traffic.json, written from rocprofv3 FETCH_SIZE and
    WRITE_SIZE runs scaled by a plain-copy calibration), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            entry = json.load(f).get(kernel_key)
    except (OSError, ValueError):
        return None
    return None if entry is None else entry.get("hbm_bytes")


def roofline(name, N, stage, precision="fp64"):
    """Roofline entry for the dominant stage of the step (by measured time).
    The mixed factorisation's bulk runs on the FP32 matrix cores, so its
    fraction is against the FP32 peak."""
    n = 3 * N
    k = int(np.argmax(stage))
    if k == 1:
        from pybrown_amd._lib import lib
        flops = n ** 3 / 3.0
        ach = flops / (stage[1] * 1e-3) / 1e12
        peak = FP64_PEAK_TFLOPS if precision == "fp64" else FP32_PEAK_TFLOPS
        key = "cholesky_%s_n%d_v%d" % (name, n, lib.pbx_get_potrf_variant()) + ("" if precision == "fp64" else "_mixed")
        return {"kernel": "cholesky (potrf, n=%d, %s)" % (n, precision), "bound": "mfma", "achieved": ach,
                "peak": peak, "unit": "TFLOP/s", "frac": ach / peak, "traffic": measured_traffic(key),
                "algorithmic": "n^3/3 = %.3e flop per factorisation" % flops}
    if k == 0:
        flops = smith_flops(N) if name in ("c2", "c4", "c3") else 30.0 * N * (N - 1) / 2
        ach = flops / (stage[0] * 1e-3) / 1e12
        return {"kernel": "assemble_tile_kernel (mobility)", "bound": "mfma", "achieved": ach,
                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS, "traffic": None,
                "algorithmic": "%.3e flop per assembly" % flops}
    byts = 4.0 * n * (n + 1)
    ach = byts / (stage[2] * 1e-3) / 1e9
    return {"kernel": "propagation (B xi)", "bound": "hbm", "achieved": ach, "peak": 8000.0, "unit": "GB/s",
            "frac": ach / 8000.0, "traffic": None, "algorithmic": "4n(n+1) = %.3e B per step" % byts}


if __name__ == "__main__":
    main()
