"""The restatement against the reference suite's external known answers
(pecnut, HYDROLIB, BD_BOX; tests/golden/known_answers.json) and against
what the mathematics fixes (closed forms, symmetries, library routines)."""
import math

import numpy as np
import pytest

from oracle.orc import Oracle, block18_to_6x6

BIG = 1000000.0


def _P(orc, xyz, L=BIG):
    return Oracle.packed(orc.pointer_pbc_matrix(np.asarray(xyz, float), L))


def test_pecnut_rpy(orc, known_answers):
    for case in known_answers["pecnut_rpy"].values():
        xyz = [[x, 0.0, 0.0] for x in case["x"]]
        M = orc.rpy_m_matrix(np.full(3, case["radius"]), _P(orc, xyz))
        np.testing.assert_allclose(M, np.array(case["M"]), atol=0.5e-7, rtol=0)


def _R_inf(orc):
    # test_diffusion_vs_hydrolib.py:63-69: R of one sphere with a partner 3000 away
    M = orc.rpy_m_matrix(np.ones(2), _P(orc, [[0, 0, 0], [3000.0, 0, 0]]))
    return np.linalg.inv(M)[1][1]


@pytest.mark.parametrize("s", ["3.0", "2.5", "2.1", "2.01", "2.001"])
def test_hydrolib_two_sphere(orc, known_answers, s):
    case = known_answers["hydrolib_two_sphere"][s]
    R_h = np.array(case["R"]) / known_answers["R_hydrolib_infty"]
    Rinf = _R_inf(orc)
    R_jo = block18_to_6x6(orc.jo_2b_r_matrix(1.0, 1.0, [case["separation"], 0, 0])) / Rinf
    np.testing.assert_allclose(R_jo, R_h, atol=case["delta"], rtol=0)
    xyz = [[0, 0, 0], [case["separation"], 0, 0]]
    P = _P(orc, xyz)
    M = orc.rpy_m_matrix(np.ones(2), P)
    C = orc.jo_r_lubrication_correction_matrix(np.ones(2), P, BIG, False)
    np.testing.assert_allclose((C + np.linalg.inv(M)) / Rinf, R_h, atol=case["delta"], rtol=0)


def test_hydrolib_three_in_line(orc, known_answers):
    case = known_answers["hydrolib_3inline_3000"]
    xyz = [[x, 0, 0] for x in case["x"]]
    P = _P(orc, xyz)
    M = orc.rpy_m_matrix(np.ones(3), P)
    C = orc.jo_r_lubrication_correction_matrix(np.ones(3), P, BIG, False)
    R = (C + np.linalg.inv(M)) / _R_inf(orc)
    np.testing.assert_allclose(R, np.array(case["R"]) / known_answers["R_hydrolib_infty"], atol=1e-7, rtol=0)


def test_bdbox_smith_ratios(orc, known_answers):
    for case in known_answers["bdbox_smith"]:
        beads = case["beads"]
        xyz = [b[:3] for b in beads]
        a = np.array([b[3] for b in beads])
        L = case["box_length"]
        M = orc.rpy_smith_m_matrix(a, _P(orc, xyz, L), L, math.sqrt(math.pi), 2, 2)
        v = case["values"]
        d1 = v.get("D_bdbox_diag_e", v.get("D_bdbox_diag_1"))
        assert abs(M[3][0] / M[0][0] - v["D_bdbox_coupling_ee"] / d1) < case["delta"]
        assert abs(M[5][2] / M[0][0] - v["D_bdbox_coupling_rr"] / d1) < case["delta"]
        if "D_bdbox_diag_2" in v:
            assert abs(M[0][0] / M[3][3] - v["D_bdbox_diag_1"] / v["D_bdbox_diag_2"]) < case["delta"]


def test_rpy_closed_form_far_and_self(orc):
    # Rotne-Prager: M_ij = (1/8 pi r)[(1 + 2a^2/3r^2) I + (1 - 2a^2/r^2) rr/r^2] for equal a
    a, r = 1.3, 4.7
    M = orc.rpy_m_matrix(np.full(2, a), _P(orc, [[0, 0, 0], [0, 0, r]]))
    perp = (1 + 2 * a * a / (3 * r * r)) / (8 * math.pi * r)
    par = perp + (1 - 2 * a * a / (r * r)) / (8 * math.pi * r)
    assert M[3][0] == pytest.approx(perp, rel=1e-14)
    assert M[5][2] == pytest.approx(par, rel=1e-14)
    assert M[0][0] == pytest.approx(1 / (6 * math.pi * a), rel=1e-15)
    assert M[0][1] == 0.0


def test_rpy_continuity_at_branch_points(orc):
    # the three RPY branches agree at r = ai+aj and r = |ai-aj| (Wajnryb et al. 2013)
    ai, aj = 2.0, 0.7
    for r0 in (ai + aj, ai - aj):
        blocks = []
        for r in (r0 * (1 - 1e-9), r0 * (1 + 1e-9)):
            b = np.zeros(6)
            orc.lib.orc_rpy_pair(ai, aj, 0.0, 0.0, r, b.ctypes.data_as(__import__("ctypes").POINTER(__import__("ctypes").c_double)))
            blocks.append(b)
        np.testing.assert_allclose(blocks[0], blocks[1], rtol=1e-7, atol=1e-12)


def test_rpy_spd_with_overlaps(orc, ref_vectors):
    M = ref_vectors["rpy_M"]
    assert np.allclose(M, M.T, rtol=0, atol=0)
    assert np.linalg.eigvalsh(M).min() > 0


def test_smith_truncation_converges(orc):
    # At fixed alpha the L1-ball sums converge geometrically in (m, n); the
    # formulation is not alpha-independent (a property of the reference's
    # Smith sums, also visible in its BD_BOX diagonal), so the check is on
    # truncation, with a planted truncation (m = n = 1) that must fail it.
    rng = np.random.RandomState(3)
    N, L = 5, 30.0
    xyz = rng.uniform(-L / 2, L / 2, (N, 3))
    a = np.full(N, 1.0)
    P = Oracle.packed(orc.pointer_pbc_matrix(xyz, L))
    ref = orc.rpy_smith_m_matrix(a, P, L, math.sqrt(math.pi), 8, 8)
    scale = np.abs(ref).max()
    assert np.abs(orc.rpy_smith_m_matrix(a, P, L, math.sqrt(math.pi), 6, 6) - ref).max() < 1e-15 * scale
    assert np.abs(orc.rpy_smith_m_matrix(a, P, L, math.sqrt(math.pi), 4, 4) - ref).max() < 1e-8 * scale
    assert np.abs(orc.rpy_smith_m_matrix(a, P, L, math.sqrt(math.pi), 1, 1) - ref).max() > 1e-4 * scale


def test_smith_permutation_symmetry(orc):
    # test_diffusion.py:108-143 design: relabelling beads permutes the blocks
    rng = np.random.RandomState(11)
    N, L = 7, 20.0
    xyz = rng.normal(0, 4, (N, 3))
    a = rng.choice([1.0, 1.5], N)
    perm = rng.permutation(N)
    M = orc.rpy_smith_m_matrix(a, Oracle.packed(orc.pointer_pbc_matrix(xyz, L)), L, math.sqrt(math.pi), 2, 2)
    Mp = orc.rpy_smith_m_matrix(a[perm], Oracle.packed(orc.pointer_pbc_matrix(xyz[perm], L)), L,
                                math.sqrt(math.pi), 2, 2)
    idx = np.concatenate([[3 * p, 3 * p + 1, 3 * p + 2] for p in perm])
    np.testing.assert_allclose(Mp, M[np.ix_(idx, idx)], rtol=1e-12, atol=1e-15)


def test_rpy_2b_inverse_matches_numpy(orc):
    rng = np.random.RandomState(5)
    for _ in range(20):
        ai, aj = rng.choice([1.0, 2.0, 0.5], 2)
        r = rng.normal(size=3) * rng.uniform(0.1, 5)
        b = np.zeros(6)
        import ctypes
        orc.lib.orc_rpy_pair(ai, aj, r[0], r[1], r[2], b.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        B = np.array([[b[0], b[3], b[4]], [b[3], b[1], b[5]], [b[4], b[5], b[2]]])
        M6 = np.block([[np.eye(3) / (6 * math.pi * ai), B], [B, np.eye(3) / (6 * math.pi * aj)]])
        R = block18_to_6x6(orc.rpy_2b_r_matrix(ai, aj, r))
        np.testing.assert_allclose(R, np.linalg.inv(M6), rtol=1e-12, atol=1e-12 * np.abs(R).max())


def test_jo_lambda_inverse_identity(orc):
    # R22(lambda) = lambda R11(1/lambda) with the 6 pi a_i prefactor (diff_tensor.c:284-286)
    for s in (2.01, 2.3, 3.5):
        for l in (0.5, 1.0, 3.0):
            x11 = orc.lib.orc_jo_xa11(s, 1 / l) * l
            R = orc.jo_2b_r_matrix(1.0, l, [s * (1 + l) / 2, 0, 0])
            assert R[6] == pytest.approx(x11 * 6 * math.pi, rel=1e-13)


def test_cichocki_kills_rigid_translation(orc):
    rng = np.random.RandomState(9)
    for _ in range(10):
        t = rng.normal(size=18)
        R = block18_to_6x6(orc.cichocki(t))
        u = np.concatenate([np.ones(3), np.ones(3)])
        np.testing.assert_allclose(R @ u, 0, atol=1e-14)


def test_lubrication_permutation_symmetry(orc):
    # test_diffusion.py:30-104 design
    rng = np.random.RandomState(12)
    N = 8
    xyz = rng.uniform(-4, 4, (N, 3))
    a = rng.choice([0.8, 1.0], N)
    perm = rng.permutation(N)
    C = orc.jo_r_lubrication_correction_matrix(a, _P(orc, xyz), 10.0, True)
    Cp = orc.jo_r_lubrication_correction_matrix(a[perm], _P(orc, xyz[perm]), 10.0, True)
    idx = np.concatenate([[3 * p, 3 * p + 1, 3 * p + 2] for p in perm])
    np.testing.assert_allclose(Cp, C[np.ix_(idx, idx)], rtol=1e-11, atol=1e-11 * np.abs(C).max())


def test_planted_mistake_is_caught(orc, known_answers):
    # A dropped reciprocal term must break the BD_BOX check: emulate it by
    # truncating the reciprocal sum (n_max = 0) and expect a failure.
    case = known_answers["bdbox_smith"][0]
    xyz = [b[:3] for b in case["beads"]]
    a = np.array([b[3] for b in case["beads"]])
    L = case["box_length"]
    M = orc.rpy_smith_m_matrix(a, _P(orc, xyz, L), L, math.sqrt(math.pi), 2, 0)
    v = case["values"]
    assert abs(M[3][0] / M[0][0] - v["D_bdbox_coupling_ee"] / v["D_bdbox_diag_e"]) > case["delta"]
