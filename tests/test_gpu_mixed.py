"""Mixed-precision noise factor: trailing updates on FP32 matrix cores
(operands rounded to FP32, FP32 accumulation added into the FP64 matrix),
diagonal tiles and solves in FP64.  Checked against numpy's FP64 Cholesky and
the FP64 path of this library; the displacement tolerance is the one stated
for the mixed engine (1e-6 relative per step)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _smith_D(N, phi, seed):
    from oracle.orc import Oracle
    from pybrown_amd.workloads import box_for, rsa_positions
    orc = Oracle()
    a = np.full(N, 10.0)
    L = box_for(a, phi)
    xyz = rsa_positions(a, L, seed)
    return orc.rpy_smith_m_matrix(a, Oracle.packed(orc.pointer_pbc_matrix(xyz, L)), L, math.sqrt(math.pi), 2, 2)


@pytest.mark.parametrize("N", [21, 100, 333, 1000])
def test_mixed_cholesky_against_fp64(N):
    from pybrown_amd.diffusion import cholesky
    D = _smith_D(N, 0.1, N)
    Ln = np.linalg.cholesky(D)
    Lm = cholesky(D, precision="mixed")
    L64 = cholesky(D)
    rel = np.linalg.norm(Lm - Ln) / np.linalg.norm(Ln)
    res = np.linalg.norm(Lm @ Lm.T - D) / np.linalg.norm(D)
    res64 = np.linalg.norm(L64 @ L64.T - D) / np.linalg.norm(D)
    print("N=%d mixed: factor rel %.3e residual %.3e (fp64 residual %.3e)" % (N, rel, res, res64))
    assert np.all(np.triu(Lm, 1) == 0)
    T = -(-3 * N // 64)
    if T <= 2:
        # no trailing update: the mixed factor is the FP64 one
        assert np.array_equal(Lm, L64)
    else:
        # the FP32 bulk is in use (an FP64 result here would mean it is not)
        assert rel > 1e-12
    assert rel <= 1e-5
    assert res <= 1e-6


def test_mixed_cholesky_reports_non_spd():
    from pybrown_amd.diffusion import cholesky
    A = np.eye(300)
    A[200, 200] = -1.0
    with pytest.raises(np.linalg.LinAlgError):
        cholesky(A, precision="mixed")


def test_mixed_step_displacement_error():
    """Per-step displacement of the mixed simulation against the FP64 one
    from the same positions and noise: relative error <= 1e-6."""
    from pybrown_amd.bead import Bead
    from pybrown_amd.box import Box
    from pybrown_amd.workloads import box_for, rsa_positions
    N = 700
    a = np.full(N, 10.0)
    L = box_for(a, 0.1)
    xyz = rsa_positions(a, L, 5)
    inp = {"box_length": L, "T": 293.15, "viscosity": 0.01005, "seed": 7, "hydrodynamics": "rpy_smith",
           "ewald_alpha": math.sqrt(math.pi), "ewald_real": 2, "ewald_imag": 2, "check_overlaps": False}
    b64 = Box([Bead(p.copy(), r) for p, r in zip(xyz, a)], inp)
    bmx = Box([Bead(p.copy(), r) for p, r in zip(xyz, a)], inp, precision="mixed")
    try:
        errs = []
        for _ in range(4):
            start = b64._pos.copy()
            bmx._pos[:] = start
            b64.propagate(1.0)
            bmx.propagate(1.0)
            d64 = b64._pos - start
            dmx = bmx._pos - start
            errs.append(np.linalg.norm(dmx - d64) / np.linalg.norm(d64))
        r64, rmx = b64.factor_residual(), bmx.factor_residual()
        print("per-step displacement rel error %s; residual fp64 %.3e mixed %.3e" % (errs, r64, rmx))
        assert max(errs) <= 1e-6
        assert rmx <= 1e-6 and r64 <= 1e-14
        assert max(errs) > 0.0
    finally:
        b64.close()
        bmx.close()


def test_factor_residual_matches_numpy():
    from pybrown_amd.bead import Bead
    from pybrown_amd.box import Box
    from pybrown_amd.workloads import box_for, rsa_positions
    N = 150
    a = np.full(N, 10.0)
    L = box_for(a, 0.1)
    xyz = rsa_positions(a, L, 9)
    inp = {"box_length": L, "T": 293.15, "viscosity": 0.01005, "seed": 1, "hydrodynamics": "rpy_smith",
           "ewald_alpha": math.sqrt(math.pi), "ewald_real": 2, "ewald_imag": 2, "check_overlaps": False}
    box = Box([Bead(p.copy(), r) for p, r in zip(xyz, a)], inp, precision="mixed")
    try:
        box.propagate(1.0)
        D, B = box.D, box.B
        want = np.linalg.norm(B @ B.T - D) / np.linalg.norm(D)
        got = box.factor_residual()
        print("residual device %.6e numpy %.6e" % (got, want))
        assert abs(got - want) <= 1e-3 * want + 1e-18
    finally:
        box.close()


# ---------------------------------------------------------------------------
# Mixed-precision resistance inversion (box.py:514,524,537): the SPD inverse
# with its bulk on FP32 matrix cores, and with one Newton-Schulz refinement.

def _min_image(d, L):
    return d - L * np.round(d / L)


def _c3_subset(N, seed):
    """N beads of the C3 configuration (rpy_smith_lub, bidisperse 20/10,
    phi = 0.25 for all 500), half of each radius, in the C3 box."""
    from pybrown_amd.workloads import workload
    xyz, a, inp, dt, _ = workload("c3", seed=seed)
    idx = np.r_[0:N // 2, 250:250 + N - N // 2]
    return xyz[idx], a[idx], inp, dt


def _lub_mobility(N, seed):
    """The far-field mobility Mff and the resistance R of a lubrication
    configuration (the matrices box.py:514 and :537 invert)."""
    from oracle.box_oracle import OracleBox
    xyz, a, inp, dt = _c3_subset(N, seed)
    o = OracleBox(xyz, a, ["XXX"] * N, inp)
    o._compute_Dff_matrix()
    o._compute_Dtot_matrix()
    return o.Mff, o.R


@pytest.mark.parametrize("N", [60, 250, 500])
def test_spd_inverse_mixed_against_numpy(N):
    from pybrown_amd.diffusion import spd_inverse
    Mff, R = _lub_mobility(N, 0)
    for name, A in (("Mff", Mff), ("R", R)):
        Xn = np.linalg.inv(A)
        X64 = spd_inverse(A)
        Xm = spd_inverse(A, precision="mixed")
        Xr = spd_inverse(A, precision="mixed_refined")
        e64, em, er = (np.abs(X - Xn).max() / np.abs(Xn).max() for X in (X64, Xm, Xr))
        print("N=%d %s: cond %.3g; inverse rel error fp64 %.2e mixed %.2e refined %.2e"
              % (N, name, np.linalg.cond(A), e64, em, er))
        assert np.array_equal(Xm, Xm.T) and np.array_equal(Xr, Xr.T)
        assert e64 <= 1e-12
        assert em <= 1e-5
        if -(-3 * N // 64) > 2:
            assert em > 1e-12  # the FP32 bulk is in use
        # one Newton-Schulz step squares the residual I - A X: the refined
        # inverse is back at FP64 accuracy (times the condition number)
        assert er <= 1e-10 and er < 1e-3 * em


def test_spd_inverse_mixed_reports_non_spd():
    from pybrown_amd.diffusion import spd_inverse
    A = np.eye(300)
    A[250, 250] = -1.0
    for p in ("mixed", "mixed_refined"):
        with pytest.raises(np.linalg.LinAlgError):
            spd_inverse(A, precision=p)


@pytest.mark.parametrize("scheme", ["ermak", "midpoint"])
def test_lubrication_step_mixed_inverse_against_fp64(scheme):
    """A lubrication simulation (rpy_smith_lub, N = 200, phi = 0.25) with the
    SPD inverses and the noise factor in mixed precision, with and without
    refinement, against the FP64 simulation stepped from the same positions
    with the same noise.  Reported against FP64 for each quantity: the
    far-field resistance Rff = inv(Mff) (the mobility itself is assembled in
    FP64 on every path), the diffusion tensor D = kT inv(R), the noise
    covariance residual ||B B^T - D|| / ||D||, the displacement and the
    per-step diffusion coefficient <|d|^2> / (6 dt); bar: 1e-6 relative."""
    from pybrown_amd.bead import Bead
    from pybrown_amd.box import Box
    xyz, a, inp, dt = _c3_subset(200, 2)
    inp = dict(inp, propagation_scheme=scheme, check_overlaps=False)
    L = inp["box_length"]
    boxes = {p: Box([Bead(q.copy(), r) for q, r in zip(xyz, a)], inp, precision=p)
             for p in ("fp64", "mixed", "mixed_refined", "mixed_factor")}
    try:
        worst = {p: [0.0] * 5 for p in ("mixed", "mixed_refined", "mixed_factor")}
        for _ in range(3):
            start = boxes["fp64"]._pos.copy()
            d = {}
            for p, b in boxes.items():
                b._pos[:] = start
                b.propagate(dt)
                d[p] = _min_image(b._pos - start, L)
            D64, R64 = boxes["fp64"].D, boxes["fp64"].Rff
            c64 = (d["fp64"] ** 2).sum(axis=1).mean() / (6 * dt)
            for p in worst:
                b = boxes[p]
                e = [np.abs(b.Rff - R64).max() / np.abs(R64).max(),
                     np.abs(b.D - D64).max() / np.abs(D64).max(),
                     b.factor_residual(),
                     np.abs(d[p] - d["fp64"]).max() / np.abs(d["fp64"]).max(),
                     abs((d[p] ** 2).sum(axis=1).mean() / (6 * dt) - c64) / c64]
                worst[p] = [max(x, y) for x, y in zip(worst[p], e)]
        for p, w in worst.items():
            print("%s %s: Rff %.2e  D %.2e  covariance residual %.2e  displacement %.2e  diffusion coefficient %.2e"
                  % (scheme, p, *w))
        rff, dd, cov, disp, dcoef = worst["mixed"]
        # unrefined: the second inverse amplifies the FP32 rounding of the
        # first by cond(R) (about 2e3 here), so D carries a few 1e-6
        assert rff <= 2e-6 and dd <= 2e-5 and cov <= 1e-6
        assert dd > 1e-12  # FP32 bulk in use
        if scheme == "ermak":
            assert disp <= 5e-6 and dcoef <= 1e-6
        # (midpoint: its stochastic drift is m/2 (D_mid - D0) B0^{-T} xi with
        # m = 100, so unrefined D errors, which differ between the two
        # positions, reach the displacement times 50: a few 1e-5 here; the
        # mixed engine refines the inverses for lubrication midpoint runs)
        # refined: both inverses back at FP64 accuracy (times cond); the
        # noise factor stays mixed, so the displacement keeps its part
        rff, dd, cov, disp, dcoef = worst["mixed_refined"]
        assert rff <= 1e-11 and dd <= 1e-9 and cov <= 1e-6
        assert disp <= 1e-6 and dcoef <= 1e-6
        # FP64 inverses with the FP32-bulk noise factor (the served mixed
        # mode for lubrication): Rff is the FP64 one to the bit; D is FP64 as
        # well, by a different route (the FP64 path takes D and B from one
        # factorisation of the reversed R, this one inverts R and factors D)
        rff, dd, cov, disp, dcoef = worst["mixed_factor"]
        assert rff == 0.0 and dd <= 1e-13 and cov <= 1e-6
        assert disp <= 1e-6 and dcoef <= 1e-6
    finally:
        for b in boxes.values():
            b.close()
