"""Trajectories of the device-resident Box against the oracle step
(oracle/box_oracle.py) with the same seeded RandomState stream.

Bars, after every step:
  - positions within 1e-10 relative (max|dr| / max|r|);
  - the step's displacement d = r_j - r_{j-1} (minimum image) within 1e-11
    relative (max|d_gpu - d_oracle| / max|d_oracle|), so an error in D, B or
    the lubrication correction is measured against the move, not against
    box-sized coordinates; lubrication cases take 1e-10 (the near-contact
    resistance is ill-conditioned);
  - flux counts, move-attempt counts and draw counts identical."""
import math

import numpy as np
import pytest

from oracle.box_oracle import OracleBox

pytestmark = pytest.mark.gpu

FLUX = {"normal": [1.0, 0.0, 0.0], "plane_point": [0.0, 0.0, 0.0]}


def _gpu_box(xyz, a, inp, labels=None, **kw):
    from pybrown_amd.bead import Bead
    from pybrown_amd.box import Box
    labels = labels or ["XXX"] * len(a)
    beads = [Bead(p, r, label=lab) for p, r, lab in zip(xyz, a, labels)]
    return Box(beads, inp, **kw)


def disp_rel_err(g0, g1, o0, o1, L):
    """max|d_gpu - d_oracle| / max|d_oracle| for one step's displacements,
    each reduced to the minimum image of the box (length L)."""
    dg, do = g1 - g0, o1 - o0
    if L:
        dg = dg - L * np.round(dg / L)
        do = do - L * np.round(do / L)
    den = np.abs(do).max()
    num = np.abs(dg - do).max()
    return num / den if den > 0 else num


def _run_pair(xyz, a, inp, steps, dt, labels=None, tol=1e-10, flags=None, disp_tol=None, **kw):
    labels = labels or ["XXX"] * len(a)
    if disp_tol is None:
        lub = "lub" in inp.get("hydrodynamics", "")
        disp_tol = 0.0 if tol == 0.0 else (1e-10 if lub else 1e-11)
    g = _gpu_box(xyz, a, inp, labels, **kw)
    o = OracleBox(xyz, a, labels, inp)
    L = inp.get("box_length")
    worst = worst_d = 0.0
    prev_g = np.array([b.r for b in g.beads])
    prev_o = o.r.copy()
    for j in range(steps):
        f = flags(j) if flags else (True, True, True)
        g.propagate(dt, *f)
        o.propagate(dt, *f)
        pos_g = np.array([b.r for b in g.beads])
        err = np.abs(pos_g - o.r).max() / np.abs(o.r).max()
        worst = max(worst, err)
        assert err <= tol, "step %d: position error %.3e" % (j, err)
        derr = disp_rel_err(prev_g, pos_g, prev_o, o.r, L)
        worst_d = max(worst_d, derr)
        assert derr <= disp_tol, "step %d: displacement error %.3e" % (j, derr)
        prev_g, prev_o = pos_g, o.r.copy()
        if "measure_flux" in inp:
            assert g.net_flux == o.net_flux, "step %d flux %s vs %s" % (j, g.net_flux, o.net_flux)
        assert g.last_move_attempts == o.attempt_log[-1]
    assert g.draw_count == o.draw_count
    print("worst position error %.3e, worst displacement error %.3e over %d steps" % (worst, worst_d, steps))
    return g, o, worst


def _grid(coords, a):
    xyz = np.array([[x, y, z] for x in coords for y in coords for z in coords], float)
    return xyz, np.full(len(xyz), a)


BASE = {"box_length": 35.0, "T": 298.0, "viscosity": 0.01, "seed": 0}


def test_ermak_nohi_bit_exact():
    xyz, a = _grid([-6.0, -3.0, 0.0, 3.0, 6.0], 1.0)
    inp = dict(BASE, hydrodynamics="nohi", measure_flux=FLUX)
    g, o, worst = _run_pair(xyz, a, inp, 100, 1e-10, tol=0.0)
    assert worst == 0.0


# dt = 0.01 ps moves a 1 A bead about 0.07 A per step (D0 = 0.22 A^2/ps at
# 298 K, 0.01 P): displacements of physical size, so the per-step
# displacement bar sees errors of D and B instead of rounding of positions
def test_ermak_rpy_grid():
    xyz, a = _grid([-3.0, 0.0, 3.0], 1.0)
    _run_pair(xyz, a, dict(BASE, hydrodynamics="rpy"), 100, 0.01)


@pytest.mark.parametrize("mn", [0, 1, 2])
def test_ermak_rpy_smith_grid(mn):
    xyz, a = _grid([-3.0, 0.0, 3.0], 1.0)
    inp = dict(BASE, hydrodynamics="rpy_smith", ewald_alpha=math.sqrt(math.pi), ewald_real=mn, ewald_imag=mn)
    _run_pair(xyz, a, inp, 100, 0.01)


def test_c1_rpy_flux_first_200_steps():
    from pybrown_amd.workloads import workload
    xyz, a, inp, dt, _ = workload("c1")
    _run_pair(xyz, a, inp, 200, dt)


def test_rpy_smith_c2_density_with_flux():
    from pybrown_amd.workloads import rsa_positions, box_for
    a = np.full(120, 10.0)
    Lb = box_for(a, 0.05)
    xyz = rsa_positions(a, Lb, 2)
    inp = {"box_length": Lb, "T": 293.15, "viscosity": 0.01005, "seed": 0, "hydrodynamics": "rpy_smith",
           "ewald_alpha": math.sqrt(math.pi), "ewald_real": 2, "ewald_imag": 2, "measure_flux": FLUX}
    _run_pair(xyz, a, inp, 60, 1.0)


def test_rejections_match():
    # beads 0.3 A apart: overlap rejections and redraws on most steps
    xyz, a = _grid([-2.3, 0.0, 2.3], 1.0)
    inp = dict(BASE, hydrodynamics="rpy", seed=3, measure_flux=FLUX)
    g, o, _ = _run_pair(xyz, a, inp, 30, 0.02)
    print("move attempts per step:", o.attempt_log)
    assert sum(o.attempt_log) > len(o.attempt_log), "workload produced no rejections"


@pytest.mark.parametrize("mn,m_mid,dt", [(1, 2.0, 1e-10), (2, 100.0, 1e-10), (1, 2.0, 0.002), (2, 100.0, 0.002)])
def test_midpoint_rpy_smith_lub_grid(mn, m_mid, dt):
    # test_box.py:409-511 design (cutoff 10, grid spacing 2.2), at the
    # reference test's dt = 1e-10 ps and at dt = 0.002 ps, where a step moves
    # a bead about 0.02 A across the 0.2 A gaps (rejections included).  At
    # 1e-10 ps a step moves about 7e-6 A, so rounding of the 2.2 A positions
    # alone is 6e-11 of a move: the displacement bar there is 1e-9.
    xyz, a = _grid([-2.2, 0.0, 2.2], 1.0)
    inp = dict(BASE, hydrodynamics="rpy_smith_lub", propagation_scheme="midpoint", lubrication_cutoff=10,
               ewald_alpha=math.sqrt(math.pi), ewald_real=mn, ewald_imag=mn, m_midpoint=m_mid)
    _run_pair(xyz, a, inp, 30, dt, disp_tol=1e-9 if dt < 1e-6 else 1e-10)


def test_midpoint_c3_like_bidisperse():
    from pybrown_amd.workloads import rsa_positions, box_for
    a = np.array([20.0] * 30 + [10.0] * 30)
    Lb = box_for(a, 0.25)
    xyz = rsa_positions(a, Lb, 4, min_gap=0.1)
    inp = {"box_length": Lb, "T": 293.15, "viscosity": 0.01005, "seed": 1, "hydrodynamics": "rpy_smith_lub",
           "ewald_alpha": math.sqrt(math.pi), "ewald_real": 2, "ewald_imag": 2, "lubrication_cutoff": 0.5,
           "propagation_scheme": "midpoint", "m_midpoint": 100, "overlap_treshold": 0.1, "measure_flux": FLUX}
    _run_pair(xyz, a, inp, 20, 0.1)


def test_midpoint_many_tile_rows():
    """N = 400 (19 tile rows of 64): the B0^T solve of the midpoint runs as
    one workgroup per tile row chained through completion flags."""
    from pybrown_amd.workloads import rsa_positions, box_for
    a = np.full(400, 5.0)
    Lb = box_for(a, 0.05)
    xyz = rsa_positions(a, Lb, 9, min_gap=0.5)
    inp = {"box_length": Lb, "T": 293.15, "viscosity": 0.01005, "seed": 2, "hydrodynamics": "rpy_smith",
           "ewald_alpha": math.sqrt(math.pi), "ewald_real": 1, "ewald_imag": 1, "propagation_scheme": "midpoint",
           "m_midpoint": 100, "measure_flux": FLUX, "check_overlaps": False}
    _run_pair(xyz, a, inp, 6, 1.0)


def test_midpoint_lub_many_tile_rows():
    """N = 400 with lubrication (19 tile rows): D_mid w from R_mid's factor
    and the two chained solves, the B0^T solve on the same chain."""
    from pybrown_amd.workloads import rsa_positions, box_for
    a = np.full(400, 5.0)
    Lb = box_for(a, 0.15)
    xyz = rsa_positions(a, Lb, 10, min_gap=0.2)
    inp = {"box_length": Lb, "T": 293.15, "viscosity": 0.01005, "seed": 3, "hydrodynamics": "rpy_smith_lub",
           "ewald_alpha": math.sqrt(math.pi), "ewald_real": 1, "ewald_imag": 1, "lubrication_cutoff": 1.0,
           "propagation_scheme": "midpoint", "m_midpoint": 100, "measure_flux": FLUX, "check_overlaps": False}
    _run_pair(xyz, a, inp, 5, 1.0)


@pytest.mark.parametrize("hyd", ["rpy_lub", "rpy_smith_lub"])
def test_midpoint_lub_with_external_force(hyd):
    """With forces the midpoint reads D and D_mid themselves (explicit
    inverses; D formed from the one-factorisation build when read), also on
    steps that reuse the last step's D and B (cadence flags)."""
    from pybrown_amd.workloads import rsa_positions, box_for
    a = np.full(45, 1.0)
    Lb = box_for(a, 0.1)
    xyz = rsa_positions(a, Lb, 7, min_gap=0.05)
    inp = {"box_length": Lb, "T": 298.0, "viscosity": 0.01, "seed": 8, "hydrodynamics": hyd,
           "ewald_real": 1, "ewald_imag": 1, "lubrication_cutoff": 1.0, "propagation_scheme": "midpoint",
           "m_midpoint": 100, "external_force": [2e-21, 0.0, -3e-21], "measure_flux": FLUX}
    _run_pair(xyz, a, inp, 16, 0.05, flags=lambda j: (j % 3 == 0, j % 2 == 0, j % 2 == 0))


@pytest.mark.parametrize("hyd", ["rpy_lub", "rpy_smith_lub"])
def test_ermak_lub_with_cadence_flags(hyd):
    from pybrown_amd.workloads import rsa_positions, box_for
    a = np.full(45, 1.0)
    Lb = box_for(a, 0.1)
    xyz = rsa_positions(a, Lb, 6, min_gap=0.05)
    inp = {"box_length": Lb, "T": 298.0, "viscosity": 0.01, "seed": 5, "hydrodynamics": hyd,
           "ewald_real": 1, "ewald_imag": 1, "lubrication_cutoff": 1.0, "measure_flux": FLUX}
    # diff_freq 3, lub_freq 2, chol_freq 2 (BD.py:177)
    _run_pair(xyz, a, inp, 24, 0.05, flags=lambda j: (j % 3 == 0, j % 2 == 0, j % 2 == 0))


def test_external_force_drift():
    xyz, a = _grid([-3.0, 0.0, 3.0], 1.0)
    inp = dict(BASE, hydrodynamics="rpy", external_force=[2e-21, 0.0, -3e-21], measure_flux=FLUX)
    _run_pair(xyz, a, inp, 30, 0.01)


def test_labels_split_flux():
    xyz, a = _grid([-6.0, -3.0, 0.0, 3.0, 6.0], 1.0)
    labels = ["A" if i % 2 else "B" for i in range(len(a))]
    inp = dict(BASE, hydrodynamics="rpy", seed=3, measure_flux=FLUX)
    g, o, _ = _run_pair(xyz, a, inp, 20, 0.05, labels=labels)
    assert set(g.net_flux) == {"A", "B"}


def test_panel_launch_cholesky_trajectory():
    # the one-launch-per-panel Cholesky with the folded solve inside the step
    # (variant 4: its tile-column workgroups wait for the published inverse)
    from pybrown_amd._lib import lib
    from pybrown_amd.workloads import rsa_positions, box_for
    a = np.full(200, 10.0)
    Lb = box_for(a, 0.05)
    xyz = rsa_positions(a, Lb, 12)
    inp = {"box_length": Lb, "T": 293.15, "viscosity": 0.01005, "seed": 2, "hydrodynamics": "rpy_smith",
           "ewald_alpha": math.sqrt(math.pi), "ewald_real": 2, "ewald_imag": 2, "measure_flux": FLUX}
    assert lib.pbx_set_potrf_variant(4) == 0
    try:
        _run_pair(xyz, a, inp, 10, 1.0)
    finally:
        lib.pbx_set_potrf_variant(0)


def _trajectory(xyz, a, inp, steps, dt, **kw):
    g = _gpu_box(xyz, a, inp, **kw)
    out = []
    for _ in range(steps):
        g.propagate(dt)
        out.append((np.array([b.r for b in g.beads]), dict(getattr(g, "net_flux", {})), g.last_move_attempts))
    return out, g.draw_count


def test_attempt_batch_size_does_not_change_the_trajectory():
    # rejection-heavy: batches of 1, 3 and 32 attempts and the adaptive size
    # give bit-identical positions, flux, attempt and draw counts
    xyz, a = _grid([-2.3, 0.0, 2.3], 1.0)
    inp = dict(BASE, hydrodynamics="rpy", seed=3, measure_flux=FLUX)
    ref, ref_draws = _trajectory(xyz, a, inp, 30, 0.02, attempt_batch=1)
    assert sum(t[2] for t in ref) > 30
    for K in (3, 32, 0):
        got, draws = _trajectory(xyz, a, inp, 30, 0.02, attempt_batch=K)
        assert draws == ref_draws
        for (p0, f0, n0), (p1, f1, n1) in zip(ref, got):
            assert np.array_equal(p0, p1) and f0 == f1 and n0 == n1


def test_nohi_batched_rejections_bit_exact():
    xyz, a = _grid([-2.2, 0.0, 2.2], 1.0)
    inp = dict(BASE, hydrodynamics="nohi", seed=4, measure_flux=FLUX)
    g, o, worst = _run_pair(xyz, a, inp, 40, 0.02, tol=0.0)
    assert sum(o.attempt_log) > len(o.attempt_log)


def test_callback_noise_matches_native():
    xyz, a = _grid([-2.3, 0.0, 2.3], 1.0)
    inp = dict(BASE, hydrodynamics="rpy_smith", ewald_real=1, ewald_imag=1, seed=11, measure_flux=FLUX)
    nat, d0 = _trajectory(xyz, a, inp, 20, 0.02, noise="native", xi_pool_attempts=5)
    cb, d1 = _trajectory(xyz, a, inp, 20, 0.02, noise="callback", xi_pool_attempts=7)
    assert d0 == d1
    for (p0, f0, n0), (p1, f1, n1) in zip(nat, cb):
        assert np.array_equal(p0, p1) and f0 == f1 and n0 == n1


def test_max_move_attempts_exhausted_matches_oracle():
    # gaps of 0.1 A: the first step of most seeds needs more than 4 attempts;
    # pick one on the CPU oracle
    xyz, a = _grid([-2.1, 0.0, 2.1], 1.0)
    inp = None
    for seed in range(200):
        trial = dict(BASE, hydrodynamics="rpy", seed=seed, max_move_attempts=4, measure_flux=FLUX)
        o = OracleBox(xyz, a, ["XXX"] * len(a), trial)
        try:
            o.propagate(0.02)
        except AssertionError:
            inp = trial
            break
    assert inp is not None
    g = _gpu_box(xyz, a, inp)
    with pytest.raises(AssertionError, match="move attepts"):
        g.propagate(0.02)
    assert g.last_move_attempts == 4 and g.draw_count == o.draw_count
    pos = np.array([b.r for b in g.beads])
    from pybrown_amd._lib import lib, ptr
    lib.pbx_sim_get_positions(g._sim, ptr(pos.reshape(-1)))
    # positions are the chain end before the wrap (no wrap after a failed step)
    assert np.abs(pos - o.r).max() <= 1e-10 * np.abs(o.r).max()


def test_c2_full_size_first_steps():
    # the benchmark workload (N = 1000, Smith-Ewald, ~20 attempts per step)
    from pybrown_amd.workloads import workload
    xyz, a, inp, dt, _ = workload("c2", seed=0)
    g, o, _ = _run_pair(xyz, a, inp, 3, dt)
    print("attempts:", o.attempt_log)


def _restart_pair(xyz, a, inp, M, dt, flags=None, **kw):
    """test_box.py:53-117: pickle after M steps, run M//2 more; unpickle the
    copy and run the same M//2 steps; positions, flux and counters agree bit
    for bit."""
    import pickle
    g = _gpu_box(xyz, a, inp, **kw)
    f = flags or (lambda j: (True, True, True))
    for j in range(M):
        g.propagate(dt, *f(j))
    saved = pickle.dumps(g)
    for j in range(M, M + M // 2):
        g.propagate(dt, *f(j))
    original = np.array([b.r for b in g.beads])
    flux0, draws0 = dict(getattr(g, "net_flux", {})), g.draw_count
    g.close()
    h = pickle.loads(saved)
    for j in range(M, M + M // 2):
        h.propagate(dt, *f(j))
    restarted = np.array([b.r for b in h.beads])
    assert np.array_equal(original, restarted)
    assert dict(getattr(h, "net_flux", {})) == flux0 and h.draw_count == draws0


@pytest.mark.parametrize("M", [4, 7, 11])
def test_restart_nohi_seed_sync(M):
    # test_box.py:53-82 design: 125 beads of radius 0.1 on a unit grid, seed 1
    xyz = np.array([[i, j, k] for i in range(5) for j in range(5) for k in range(5)], float)
    a = np.full(125, 0.1)
    inp = dict(BASE, hydrodynamics="nohi", seed=1, measure_flux=FLUX)
    _restart_pair(xyz, a, inp, M, 0.001)


@pytest.mark.parametrize("noise", ["native", "callback"])
def test_restart_smith_cadence(noise):
    # Cholesky every second step: the restored B must be the saved one
    xyz, a = _grid([-2.3, 0.0, 2.3], 1.0)
    inp = dict(BASE, hydrodynamics="rpy_smith", ewald_real=1, ewald_imag=1, seed=9, measure_flux=FLUX)
    _restart_pair(xyz, a, inp, 9, 0.02, flags=lambda j: (j % 2 == 0, True, j % 2 == 0), noise=noise,
                  xi_pool_attempts=5)


def test_restart_midpoint_lub():
    xyz, a = _grid([-2.2, 0.0, 2.2], 1.0)
    inp = dict(BASE, hydrodynamics="rpy_smith_lub", propagation_scheme="midpoint", lubrication_cutoff=10,
               ewald_real=1, ewald_imag=1, m_midpoint=100.0, seed=2)
    _restart_pair(xyz, a, inp, 6, 1e-10, flags=lambda j: (j % 3 == 0, j % 2 == 0, True))


def test_restore_rejects_a_different_simulation():
    xyz, a = _grid([-3.0, 0.0, 3.0], 1.0)
    g = _gpu_box(xyz, a, dict(BASE, hydrodynamics="rpy"))
    h = _gpu_box(xyz, a, dict(BASE, hydrodynamics="nohi"))
    with pytest.raises(RuntimeError):
        h.restore(g.snapshot())


@pytest.mark.parametrize("case", ["rpy_smith_ermak", "lub_midpoint", "mixed"])
def test_graph_replays_match_eager(case):
    """The factorisation and SPD-inverse sequences replayed from HIP graphs
    (captured on the first step, replayed on every later one with the step's
    eager launches and synchronisations in between) give the eager
    trajectory bit for bit."""
    from pybrown_amd.workloads import rsa_positions, box_for
    if case == "lub_midpoint":
        a = np.array([20.0] * 30 + [10.0] * 30)
        Lb = box_for(a, 0.25)
        xyz = rsa_positions(a, Lb, 4, min_gap=0.1)
        inp = {"box_length": Lb, "T": 293.15, "viscosity": 0.01005, "seed": 1, "hydrodynamics": "rpy_smith_lub",
               "ewald_alpha": math.sqrt(math.pi), "ewald_real": 2, "ewald_imag": 2, "lubrication_cutoff": 0.5,
               "propagation_scheme": "midpoint", "m_midpoint": 100, "overlap_treshold": 0.1, "measure_flux": FLUX}
        steps, dt, kw = 12, 0.1, {}
    else:
        a = np.full(300, 10.0)
        Lb = box_for(a, 0.05)
        xyz = rsa_positions(a, Lb, 3)
        inp = {"box_length": Lb, "T": 293.15, "viscosity": 0.01005, "seed": 0, "hydrodynamics": "rpy_smith",
               "ewald_alpha": math.sqrt(math.pi), "ewald_real": 2, "ewald_imag": 2, "measure_flux": FLUX}
        steps, dt, kw = 15, 1.0, ({"precision": "mixed"} if case == "mixed" else {})
    eager, d0 = _trajectory(xyz, a, inp, steps, dt, graphs=False, **kw)
    graph, d1 = _trajectory(xyz, a, inp, steps, dt, graphs=True, **kw)
    assert d0 == d1
    for j, ((p0, f0, n0), (p1, f1, n1)) in enumerate(zip(eager, graph)):
        assert np.array_equal(p0, p1) and f0 == f1 and n0 == n1, "step %d" % j


@pytest.mark.parametrize("case", ["rpy_smith", "rpy_open_box_force"])
def test_whole_step_graphs_match_eager_queued_runs(case):
    """In-library runs (Box.run) whose steps need no host decision replay
    whole steps from HIP graphs when graphs are on: one graph per (builds D,
    factors) pair, the step's noise slot and flux slot read from a device
    cursor the graph advances.  Against the same runs with graphs off,
    positions, per-step flux and draws agree to the bit, across build and
    factor cadences, noise chunk uploads every 5 steps, a change of time
    step, an eager step between runs and (second case) an external force
    (the deterministic sub-step in the graph).  Replays launch no kernels from
    the host, so the launch counter shows the graphs were used."""
    from pybrown_amd._lib import lib
    from pybrown_amd.serving import OPEN_BOX
    from pybrown_amd.workloads import rsa_positions, box_for
    a = np.full(120, 10.0)
    Lb = box_for(a, 0.05)
    xyz = rsa_positions(a, Lb, 6)
    inp = {"box_length": Lb, "T": 293.15, "viscosity": 0.01005, "seed": 4, "hydrodynamics": "rpy_smith",
           "ewald_alpha": math.sqrt(math.pi), "ewald_real": 2, "ewald_imag": 2, "measure_flux": FLUX,
           "check_overlaps": False}
    if case == "rpy_open_box_force":
        inp = dict(inp, hydrodynamics="rpy", box_length=OPEN_BOX, external_force=[0.5, -0.25, 0.0])

    def runs(graphs):
        box = _gpu_box(xyz, a, inp, graphs=graphs, xi_pool_attempts=5)
        out = []
        k0 = lib.pbx_kernel_launches()
        f, _ = box.run(1.0, 23, diff_freq=2, chol_freq=3)
        out.append((box._pos.copy(), f.copy(), box.draw_count))
        f, _ = box.run(1.0, 11, j0=23)
        out.append((box._pos.copy(), f.copy(), box.draw_count))
        f, _ = box.run(0.5, 7, j0=34, diff_freq=2, chol_freq=2)
        out.append((box._pos.copy(), f.copy(), box.draw_count))
        launches = lib.pbx_kernel_launches() - k0
        box.propagate(1.0)
        out.append((box._pos.copy(), np.array([box.net_flux["XXX"]]), box.draw_count))
        f, _ = box.run(1.0, 9, j0=42, diff_freq=3, chol_freq=3)
        out.append((box._pos.copy(), f.copy(), box.draw_count))
        box.close()
        return out, launches

    eager, n_eager = runs(False)
    graph, n_graph = runs(True)
    print("%s: launches over 41 steps: eager %d, graphs %d" % (case, n_eager, n_graph))
    for k, ((p0, f0, d0), (p1, f1, d1)) in enumerate(zip(eager, graph)):
        assert np.array_equal(p0, p1) and np.array_equal(f0, f1) and d0 == d1, "call %d" % k
    assert n_graph < n_eager / 3


@pytest.mark.parametrize("hydro", ["rpy", "rpy_smith"])
def test_lookahead_assembly_matches_plain_steps(hydro):
    """Steps with the look-ahead assembly of D (on by default: a step that
    built D assembles the next one's after its result copies) against the
    same steps with it off: positions, flux and the D of each step to the
    bit, over D built every step, then every second step, a bead moved from
    the host between steps, and the driver loop with a build cadence."""
    from pybrown_amd._lib import lib
    from pybrown_amd.workloads import rsa_positions, box_for
    a = np.full(200, 10.0)
    Lb = box_for(a, 0.05)
    xyz = rsa_positions(a, Lb, 5)
    # free-space RPY in an open box (with minimum-image vectors in the
    # periodic box its matrix is not positive definite at this density)
    from pybrown_amd.serving import OPEN_BOX
    inp = {"box_length": OPEN_BOX if hydro == "rpy" else Lb, "T": 293.15, "viscosity": 0.01005, "seed": 3,
           "hydrodynamics": hydro, "ewald_alpha": math.sqrt(math.pi), "ewald_real": 2, "ewald_imag": 2,
           "measure_flux": FLUX}

    def trajectory(look):
        box = _gpu_box(xyz, a, inp)
        assert lib.pbx_sim_set_lookahead(box._sim, 1 if look else 0) == 0
        out = []
        for j in range(10):
            if j == 6:
                box.beads[7].r[0] += 0.5  # moved from the host: the look-ahead D is stale
            every = j < 4
            box.propagate(1.0, build_Dff=every or j % 2 == 0, cholesky=every or j % 2 == 0)
            out.append((box._pos.copy(), dict(box.net_flux), box.D.copy()))
        flux, _ = box.run(1.0, 6, j0=10, diff_freq=3, chol_freq=3)
        out.append((box._pos.copy(), flux.copy(), box.D.copy()))
        return out

    plain, ahead = trajectory(False), trajectory(True)
    for j, ((p0, f0, d0), (p1, f1, d1)) in enumerate(zip(plain, ahead)):
        assert np.array_equal(p0, p1), "positions, step %d" % j
        assert np.array_equal(np.asarray(f0 if not isinstance(f0, dict) else list(f0.values())),
                              np.asarray(f1 if not isinstance(f1, dict) else list(f1.values()))), "flux, step %d" % j
        assert np.array_equal(d0, d1), "D, step %d" % j
