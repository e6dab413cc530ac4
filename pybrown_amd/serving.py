"""Serving loop: a stream of independent simulations (replicas with their own
N, box, seed and hydrodynamics mode), admitted continuously while others run.

A request is one simulation run for a fixed number of Ermak steps with the
plane flux measured; its result is the per-step flux trajectory and the final
positions.  The loop admits a request (first come, first served) when

  1. it has arrived,
  2. one of `slots` worker slots is free, and
  3. its device footprint (`footprint_bytes`, an upper bound of what the
     library allocates for it) fits the HBM budget beside the simulations
     already resident.

Each slot is a host thread that creates the simulation on the device
(positions, matrices, factors and the generator state stay resident until
it finishes), runs all its steps inside the library (`Box.run`, one call)
and frees it.  Every simulation has its own HIP stream, so the steps of all
resident simulations overlap on the device; the kernels of one simulation
read only its own buffers, so a result does not depend on what runs beside
it (tests/test_gpu_serving.py checks that results are bit-identical alone and
in the mix).  The number of hardware queues the HIP runtime spreads those
streams over is set by GPU_MAX_HW_QUEUES before the runtime starts.

The request stream follows the served benchmark: request k has seed k;
N = round(exp(U(ln 32, ln 3000))), volume fraction U(0.02, 0.30), radius
10 A; 60 % periodic RPY with Ewald sums (rpy_smith, alpha = sqrt(pi),
m = n = 2), 25 % free-space RPY, 15 % Ewald RPY with Jeffrey-Onishi
lubrication (cutoff 0.5, Cichocki projection); dt = 1 ps, flux through the
plane x = 0; Poisson arrivals.  Three choices keep every request
runnable, as the reference itself would not be:
  - free-space RPY runs in an open box (box_length far beyond any
    displacement: no images, no wrap).  With minimum-image pair vectors in the
    periodic box the RPY matrix has negative eigenvalues at every (N, phi) of
    the stream (N = 100..3000, phi = 0.05..0.3), so the Cholesky would fail
    on the first step;
  - without lubrication, overlap checks are off: random sequential addition
    leaves near-contact pairs, and at these densities almost no move of all
    beads is overlap-free (the same reason as for the N = 4000 workload).
    RPY and its Ewald sum stay positive definite for overlapping beads;
  - lubrication requests use the midpoint scheme of the N = 500
    lubrication workload (whose Ewald-plus-JO definition "lubrication" takes
    here) with overlap checks on (placement with a 0.1 A minimum gap,
    overlap_treshold 0, max_move_attempts 10000 so a request that cannot
    move fails instead of stalling a slot).  The Jeffrey-Onishi series is
    undefined for overlapping pairs (an overlapped pair turns D into NaN, in
    the reference as here), and with Ermak steps, which lack the divergence
    term, lubricated pairs drift into contact: measured on requests 2, 19 and
    23 of the stream, rejections grow to hundreds of attempts per step and
    request 19 exhausts 10000 attempts, while midpoint steps need one
    attempt each.
"""
import math
import threading
import time

import numpy as np

from pybrown_amd.workloads import ETA_DEFAULT, T_DEFAULT, box_for, rsa_positions

MODES = ("rpy_smith", "rpy", "rpy_smith_lub")
OPEN_BOX = 1.0e6  # box length (A) of the open-box free-space runs
FLUX_X0 = {"normal": [1.0, 0.0, 0.0], "plane_point": [0.0, 0.0, 0.0]}


def stream_params(count, rate=None, seed=12345, n_min=32, n_max=3000, phi_min=0.02, phi_max=0.30):
    """Parameters of requests 0 .. count-1.  One RandomState(seed) is read in
    request order: inter-arrival time Exp(1) / rate (all at t = 0 when rate
    is None), then u_N, u_phi, u_mode.  Request k has simulation seed k."""
    rng = np.random.RandomState(seed)
    t = 0.0
    out = []
    for k in range(count):
        gap = rng.exponential(1.0)
        if rate:
            t += gap / rate
        N = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        phi = rng.uniform(phi_min, phi_max)
        u = rng.uniform()
        mode = MODES[0] if u < 0.60 else (MODES[1] if u < 0.85 else MODES[2])
        out.append({"rid": k, "seed": k, "N": N, "phi": phi, "mode": mode, "arrival": t})
    return out


class Request:
    """One simulation: positions (N, 3), radii (N,), input keywords of
    BD.py:37-64, time step, step count and arrival time (s, relative)."""

    def __init__(self, rid, xyz, radii, inp, dt, steps, arrival=0.0):
        self.rid = rid
        self.xyz = np.asarray(xyz, dtype=np.float64)
        self.radii = np.asarray(radii, dtype=np.float64)
        self.inp = inp
        self.dt = float(dt)
        self.steps = int(steps)
        self.arrival = float(arrival)

    @property
    def N(self):
        return self.radii.size

    @property
    def mode(self):
        return self.inp["hydrodynamics"]


def make_request(p, steps=200, dt=1.0, a=10.0):
    """Request from stream parameters: RSA placement with the request's
    seed in the box of volume fraction phi."""
    lub = p["mode"] == "rpy_smith_lub"
    radii = np.full(p["N"], a)
    L = box_for(radii, p["phi"])
    xyz = rsa_positions(radii, L, p["seed"], min_gap=0.1 if lub else 0.0)
    inp = {"T": T_DEFAULT, "viscosity": ETA_DEFAULT, "seed": p["seed"], "hydrodynamics": p["mode"],
           "check_overlaps": False, "measure_flux": FLUX_X0, "box_length": L}
    if p["mode"] in ("rpy_smith", "rpy_smith_lub"):
        inp.update(ewald_alpha=math.sqrt(math.pi), ewald_real=2, ewald_imag=2)
    if lub:
        inp.update(lubrication_cutoff=0.5, cichocki_correction=True, check_overlaps=True, overlap_treshold=0.0,
                   max_move_attempts=10000, propagation_scheme="midpoint", m_midpoint=100)
    if p["mode"] == "rpy":
        inp["box_length"] = OPEN_BOX
    return Request(p["rid"], xyz, radii, inp, dt, steps, p["arrival"])


def make_requests_parallel(params, steps=200, dt=1.0, workers=8):
    """make_request over a process pool (placement is host Python; the pool
    uses 'spawn' so no child inherits a GPU context)."""
    import concurrent.futures as cf
    import multiprocessing as mp
    if workers <= 1 or len(params) < 2:
        return [make_request(p, steps, dt) for p in params]
    with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(make_request, params, [steps] * len(params), [dt] * len(params)))


def footprint_bytes(N, mode, scheme="ermak", check_overlaps=False, xi_pool_attempts=64, precision="fp64"):
    """Upper bound of the device memory one simulation holds: the dense
    n x n matrices (D and B; without lubrication the look-ahead copy of D;
    with lubrication the far-field resistance and an inverse scratch, and
    with refined inverses a copy of the matrix inverted; midpoint one more
    each), the factorisation work
    areas (diagonal inverses and state), the noise ring (two chunks of
    xi_pool_attempts x 3N draws) and O(N) vectors."""
    n = 3 * N
    np_ = -(-n // 64) * 64
    T = np_ // 64
    lub = mode in ("rpy_lub", "rpy_smith_lub")
    mid = scheme == "midpoint"
    if mode == "nohi":
        mats, works = 0, 0
    else:
        mats = 2 + (1 if mid else 0) + ((2 + (1 if mid else 0)) if lub else 1)
        if lub and precision == "mixed_refined":
            mats += 1
        works = 2 if lub else 1
    work = 8 * (np_ * 64 + 8 + 64 * 64)
    ring = 8 * 2 * xi_pool_attempts * n
    vecs = 8 * (6 * np_ + 2 * N) + 4 * N + 4096
    batch = 0
    if check_overlaps and scheme == "ermak":
        S = max(2, (T + 11) // 12)  # column slab width of the batched attempt products
        batch = 8 * (33 * n + N) + 4 * 32 + (8 * (-(-T // S)) * np_ * 32 if mode != "nohi" else 0)
    return 8 * np_ * np_ * mats + works * work + ring + vecs + batch + (1 << 20)


def request_footprint(req, precision="fp64"):
    """footprint_bytes for a request's own keywords in a served precision
    mode."""
    return footprint_bytes(req.N, req.mode, req.inp.get("propagation_scheme", "ermak"),
                           bool(req.inp.get("check_overlaps", True)),
                           precision=request_precision(req, precision))


class Result:
    __slots__ = ("rid", "N", "mode", "status", "arrival", "admit", "start", "end", "flux", "xyz", "device_bytes")

    def __init__(self, rid, N, mode):
        self.rid, self.N, self.mode = rid, N, mode
        self.status = "pending"
        self.arrival = self.admit = self.start = self.end = None
        self.flux = self.xyz = None
        self.device_bytes = 0

    @property
    def latency(self):
        """Time to trajectory: arrival to the last step's results on the host."""
        return None if self.end is None else self.end - self.arrival


def request_precision(req, precision):
    """The Box precision a request runs at in a served precision mode:
    "fp64" throughout, or "mixed" (FP32 matrix-core bulk of the noise
    factor).  Lubrication requests keep their SPD inverses in FP64
    ("mixed_factor"): their midpoint drift takes the difference of D at two
    positions times m/2 = 50, so FP32-bulk inverses (D to a few 1e-6) put a
    few 1e-5 of relative error into the displacement, and refining them
    (one Newton-Schulz step, 2 n^3 FP64 flops per inverse against about n^3
    for the inverse itself) makes those requests slower than FP64."""
    if precision == "mixed" and req.mode in ("rpy_lub", "rpy_smith_lub"):
        return "mixed_factor"
    return precision


def run_one(req, graphs=False, precision="fp64"):
    """Create the simulation on the device, run all steps inside the library
    (Box.run: steps that need no host decision are queued back to back), free
    it.  Returns (per-step flux (steps, labels), final positions (N, 3),
    device bytes held).  graphs: replay the factorisation and SPD-inverse
    launch sequences from HIP graphs; precision: the served precision mode
    (request_precision)."""
    from pybrown_amd.bead import Bead
    from pybrown_amd.box import Box
    beads = [Bead(req.xyz[i].copy(), float(req.radii[i]), label="XXX") for i in range(req.N)]
    box = Box(beads, req.inp, graphs=graphs, precision=request_precision(req, precision))
    try:
        held = box.device_bytes
        flux, _ = box.run(req.dt, req.steps)
        return flux, box._pos.copy(), held
    finally:
        box.close()


def run_plain(req):
    """The plain one-at-a-time FP64 path: the simulation stepped from Python
    with Box.propagate, one synchronised step at a time, as BD.py:164-185
    drives the reference's Box.  Same results as run_one."""
    from pybrown_amd.bead import Bead
    from pybrown_amd.box import Box
    beads = [Bead(req.xyz[i].copy(), float(req.radii[i]), label="XXX") for i in range(req.N)]
    box = Box(beads, req.inp)
    try:
        held = box.device_bytes
        flux = np.zeros((req.steps, 1), dtype=np.int64)
        for j in range(req.steps):
            box.propagate(req.dt)
            flux[j, 0] = box.net_flux["XXX"]
        return flux, box._pos.copy(), held
    finally:
        box.close()


def estimated_cost(req):
    """Relative device work of a request, for ordering a batch: its steps
    times the dense algebra of one step (n^3/3 for the noise factor; with
    lubrication two SPD inverses of about n^3 each per D, and the midpoint's
    second D), plus a per-step floor for the launch-bound small ones."""
    n = 3.0 * req.N
    lub = req.mode in ("rpy_lub", "rpy_smith_lub")
    mid = req.inp.get("propagation_scheme", "ermak") == "midpoint"
    per_step = n ** 3 / 3.0 + (2.0 * n ** 3 * (2 if mid else 1) if lub else 0.0)
    return req.steps * (per_step + 2e9)


class Server:
    """Continuous admission of a request stream onto `slots` concurrent
    device-resident simulations within an HBM budget (bytes).  precision:
    the served precision mode (request_precision; it also sets the
    footprints).  order: "arrival" admits in arrival order; "cost" admits a
    batch that has all arrived (realtime=False) largest estimated_cost
    first, so the longest simulations start first and the short ones fill
    the slots around them (longest-processing-time order)."""

    def __init__(self, slots=8, hbm_budget=250e9, runner=run_one, precision="fp64", order="arrival"):
        if slots < 1:
            raise ValueError("slots must be >= 1")
        if order not in ("arrival", "cost"):
            raise ValueError("order must be 'arrival' or 'cost'")
        self.slots = int(slots)
        self.budget = float(hbm_budget)
        self.runner = runner
        self.precision = precision
        self.order = order

    def run(self, requests, realtime=True):
        """Serve `requests` (arrival times in seconds from the start; with
        realtime=False all have arrived at the start).  Returns the results
        in request order.  A request larger than the whole budget fails with
        status 'rejected: footprint'."""
        if realtime or self.order == "arrival":
            order = sorted(range(len(requests)), key=lambda i: (requests[i].arrival if realtime else 0.0, i))
        else:
            order = sorted(range(len(requests)), key=lambda i: (-estimated_cost(requests[i]), i))
        results = [Result(r.rid, r.N, r.mode) for r in requests]
        fps = [request_footprint(r, self.precision) for r in requests]
        cv = threading.Condition()
        state = {"free": self.slots, "mem": 0.0, "done": 0}
        work = []
        t0 = time.perf_counter()

        def worker():
            while True:
                with cv:
                    while not work:
                        cv.wait()
                    i = work.pop(0)
                if i is None:
                    return
                res, req = results[i], requests[i]
                res.start = time.perf_counter() - t0
                try:
                    res.flux, res.xyz, res.device_bytes = self.runner(req)
                    res.status = "ok"
                except Exception as e:  # reported per request, the server keeps going
                    res.status = "error: %s" % e
                res.end = time.perf_counter() - t0
                with cv:
                    state["free"] += 1
                    state["mem"] -= fps[i]
                    state["done"] += 1
                    cv.notify_all()

        threads = [threading.Thread(target=worker, daemon=True) for _ in range(self.slots)]
        for th in threads:
            th.start()
        k = 0
        with cv:
            while k < len(order):
                i = order[k]
                now = time.perf_counter() - t0
                arr = requests[i].arrival if realtime else 0.0
                results[i].arrival = arr
                if fps[i] > self.budget:
                    results[i].status = "rejected: footprint %.3g B over the budget" % fps[i]
                    results[i].admit = results[i].start = results[i].end = now
                    state["done"] += 1
                    k += 1
                    continue
                if arr > now:
                    cv.wait(timeout=arr - now)
                    continue
                if state["free"] == 0 or state["mem"] + fps[i] > self.budget:
                    cv.wait()
                    continue
                state["free"] -= 1
                state["mem"] += fps[i]
                results[i].admit = now
                work.append(i)
                cv.notify_all()
                k += 1
            while state["done"] < len(requests):
                cv.wait()
            work.extend([None] * self.slots)
            cv.notify_all()
        for th in threads:
            th.join()
        return results


def summary(results):
    """sims/s over the served window, latency percentiles (s) of the
    completed requests, and counts."""
    ok = [r for r in results if r.status == "ok"]
    if not ok:
        return {"completed": 0, "failed": len(results)}
    t_first = min(r.arrival for r in results)
    t_last = max(r.end for r in ok)
    lat = np.array([r.latency for r in ok])
    return {"completed": len(ok), "failed": len(results) - len(ok),
            "window_s": t_last - t_first, "sims_per_s": len(ok) / (t_last - t_first),
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean())}
