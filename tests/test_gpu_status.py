"""A failure status raised by a launch sequence is reported by the step that
ran it.  The failure is forced from the host (pbx_debug_inject_wait_failure:
the k-th factorisation or B0^T solve from now ends with its status word set
as an exhausted bounded wait sets it); every kernel runs to completion.

Status-reporting sequences of one step, in queue order (a lubrication step
that rebuilds D and factors it takes D and B from one factorisation, the
R inverse-and-factor):
  lubrication Ermak with the divergence term:  Rff inverse, R
      inverse-and-factor, then two inverses per coordinate in the divergence
      (box.py:560-610);
  midpoint:  the Cholesky, or with lubrication the Rff inverse and the R
      inverse-and-factor; [the midpoint Rff and R inverses,] the B0^T solve
      (box.py:254-349).  With lubrication and no forces: the Rff inverse,
      the factorisation of J R J; per attempt the B0 xi solve, the midpoint
      Rff inverse, the factorisation of R_mid and the two solves with it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

XYZ = np.array([[0.0, 0.0, 0.0], [-0.955, -1.779, 1.347], [0.518, -2.31, -0.105], [-0.511, 1.809, 1.133],
                [0.531, 0.272, 2.335], [1.211, -2.273, -2.492], [1.22, 2.251, -0.518]])


def _box(inp):
    from pybrown_amd.bead import Bead
    from pybrown_amd.box import Box
    return Box([Bead(p.copy(), 1.0) for p in XYZ], inp)


def _failing_step_raises(inp, k, dt=0.01):
    from pybrown_amd._lib import lib
    box = _box(inp)
    try:
        box.propagate(dt)  # a clean step first
        assert lib.pbx_debug_inject_wait_failure(k) == 0
        try:
            with pytest.raises(RuntimeError, match="wait"):
                box.propagate(dt)
        finally:
            lib.pbx_debug_inject_wait_failure(0)
        # the status word was cleared when the failure was reported
        box.propagate(dt)
    finally:
        box.close()


@pytest.mark.parametrize("overlaps", [False, True])
def test_divergence_inverse_failure_fails_its_step(overlaps):
    inp = {"box_length": 1.0e6, "T": 298.0, "viscosity": 0.01, "seed": 4, "hydrodynamics": "rpy_lub",
           "lubrication_cutoff": 1.0, "cichocki_correction": True, "divergence_term": True,
           "check_overlaps": overlaps}
    # sequences 1-2 are the step's own Rff inverse and R inverse-and-factor;
    # 3 and 4 the first coordinate's inverses in the divergence's rebuilds
    _failing_step_raises(inp, 4)


@pytest.mark.parametrize("hydro,k", [("rpy_smith", 2), ("rpy_smith_lub", 3), ("rpy_smith_lub", 5),
                                     ("rpy_smith_lub", 6), ("rpy_smith_lub", 7)])
def test_midpoint_solve_failure_fails_its_step(hydro, k):
    inp = {"box_length": 12.0, "T": 298.0, "viscosity": 0.01, "seed": 5, "hydrodynamics": hydro,
           "ewald_alpha": np.sqrt(np.pi), "ewald_real": 1, "ewald_imag": 1, "lubrication_cutoff": 1.0,
           "propagation_scheme": "midpoint", "m_midpoint": 100.0, "check_overlaps": False}
    # the k-th sequence: the B0^T solve; with lubrication the B0 xi solve (3),
    # R_mid's factorisation (5) and the solves with it (6, 7)
    _failing_step_raises(inp, k)
