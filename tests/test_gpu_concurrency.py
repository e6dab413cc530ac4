"""Results do not depend on what runs beside a simulation.

Regression test for a race inside the wave-specialised 64 x 64 tile
factorisation (pbx_tile.h, tile_factor_ws): the two waves that stage the
next panels' columns once reported through one shared LDS counter, which
could reach the chain wave's target while one of them was a panel ahead and
the other still staging; the chain then read (and its stores could be
overwritten by) unstaged columns.  Waves drift apart only when other
workgroups share the compute unit, so it showed up only with many
simulations on the GPU at once (a wrong factor in about 1 of 1000
factorisations with 32 simulations on 32 hardware queues).

Here 24 identical simulations (same positions, same seed) run on 24 host
threads, each refactoring the same D every step (build_Dff off, Cholesky
on): every simulation's positions must equal the others' after every step,
since B xi depends on every entry of B."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant", [6, 4])
def test_concurrent_refactorisations_are_identical(variant):
    from pybrown_amd._lib import lib
    from pybrown_amd.bead import Bead
    from pybrown_amd.box import Box
    from pybrown_amd.workloads import box_for, rsa_positions
    S, N, reps = 24, 333, 100  # 3N = 999: 16 tile columns
    a = np.full(N, 10.0)
    L = box_for(a, 0.1)
    xyz = rsa_positions(a, L, 3)
    inp = {"box_length": L, "T": 293.15, "viscosity": 0.01005, "seed": 1, "hydrodynamics": "rpy_smith",
           "ewald_alpha": np.sqrt(np.pi), "ewald_real": 2, "ewald_imag": 2, "check_overlaps": False}
    assert lib.pbx_set_potrf_variant(variant) == 0
    boxes = [Box([Bead(p.copy(), r) for p, r in zip(xyz, a)], inp) for _ in range(S)]
    try:
        for b in boxes:
            b.propagate(1.0)  # D at the start positions and its factor
        traj = [np.empty((reps, N, 3)) for _ in range(S)]
        errors = []

        def work(s):
            try:
                for r in range(reps):
                    boxes[s].propagate(1.0, build_Dff=False, build_Dnf=False, cholesky=True)
                    traj[s][r] = boxes[s]._pos
            except Exception as e:  # reported below
                errors.append(e)

        th = [threading.Thread(target=work, args=(s,)) for s in range(S)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errors, errors
        differing = [s for s in range(1, S) if not np.array_equal(traj[s], traj[0])]
        assert not differing, "simulations %s left simulation 0's trajectory" % differing
    finally:
        lib.pbx_set_potrf_variant(0)
        for b in boxes:
            b.close()
