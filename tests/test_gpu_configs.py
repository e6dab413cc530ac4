"""The benchmark workloads at full size on the device.

  C2  periodic RPY-Ewald, N = 1000: 100 steps against the oracle trajectory
      in tests/golden/c2_oracle_100.npz (tests/golden/make_c2_trajectory.py);
  C3  Ewald RPY with Jeffrey-Onishi lubrication, N = 500 bidisperse,
      midpoint: 10 steps against tests/golden/c3_oracle_10.npz
      (tests/golden/make_c3_trajectory.py);
  C4  N = 4000 (n = 12000): the mixed-precision noise factor against the
      FP64 one from the same positions and noise, through properties that do
      not need a CPU factorisation of this size.

Bars: positions 1e-10 relative to max|r| after every step; each step's
displacement 1e-11 relative to its largest component (C3: 1e-10, the
near-contact resistance is ill-conditioned); flux, move attempts and draws
exact.  C4: ||B B^T - D||_F / ||D||_F <= 1e-6 and the per-step displacement
of the mixed factor within 1e-6 of the FP64 one (the FP64 factor's
residual <= 1e-14)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _box(xyz, a, inp, **kw):
    from pybrown_amd.bead import Bead
    from pybrown_amd.box import Box
    return Box([Bead(p.copy(), r) for p, r in zip(xyz, a)], inp, **kw)


def _min_image(d, L):
    return d - L * np.round(d / L)


def test_c2_100_steps_against_oracle_fixture():
    from pybrown_amd.workloads import workload
    f = np.load(os.path.join(GOLDEN, "c2_oracle_100.npz"))
    xyz, a, inp, dt, _ = workload("c2")
    assert np.array_equal(xyz, f["xyz0"]) and inp["box_length"] == float(f["box_length"])
    L = inp["box_length"]
    sub, r_sub, flux, attempts = f["sub"], f["r_sub"], f["flux"], f["attempts"]
    full = {int(s): k for k, s in enumerate(f["full_steps"])}
    g = _box(xyz, a, inp)
    try:
        prev_g, prev_o = xyz[sub].copy(), xyz[sub].copy()
        worst = worst_d = 0.0
        for j in range(r_sub.shape[0]):
            g.propagate(dt)
            pos = g._pos
            err = np.abs(pos[sub] - r_sub[j]).max() / np.abs(r_sub[j]).max()
            dg, do = _min_image(pos[sub] - prev_g, L), _min_image(r_sub[j] - prev_o, L)
            derr = np.abs(dg - do).max() / np.abs(do).max()
            worst, worst_d = max(worst, err), max(worst_d, derr)
            assert err <= 1e-10, "step %d: position error %.3e" % (j + 1, err)
            assert derr <= 1e-11, "step %d: displacement error %.3e" % (j + 1, derr)
            assert g.net_flux["XXX"] == flux[j], "step %d flux" % (j + 1)
            assert g.last_move_attempts == attempts[j], "step %d attempts" % (j + 1)
            if j + 1 in full:
                r = f["r_full"][full[j + 1]]
                assert np.abs(pos - r).max() <= 1e-10 * np.abs(r).max(), "step %d, all beads" % (j + 1)
            prev_g, prev_o = pos[sub].copy(), r_sub[j].copy()
        assert g.draw_count == int(f["draw_count"])
        print("C2 100 steps: worst position error %.3e, worst displacement error %.3e, %d attempts"
              % (worst, worst_d, int(attempts.sum())))
    finally:
        g.close()


def test_c3_10_steps_against_oracle_fixture():
    from pybrown_amd.workloads import workload
    f = np.load(os.path.join(GOLDEN, "c3_oracle_10.npz"))
    xyz, a, inp, dt, _ = workload("c3")
    assert np.array_equal(xyz, f["xyz0"]) and np.array_equal(a, f["radii"])
    L = inp["box_length"]
    r_ref, flux, attempts = f["r"], f["flux"], f["attempts"]
    g = _box(xyz, a, inp)
    try:
        prev_g, prev_o = xyz.copy(), xyz.copy()
        worst = worst_d = 0.0
        for j in range(r_ref.shape[0]):
            g.propagate(dt)
            pos = g._pos
            err = np.abs(pos - r_ref[j]).max() / np.abs(r_ref[j]).max()
            dg, do = _min_image(pos - prev_g, L), _min_image(r_ref[j] - prev_o, L)
            derr = np.abs(dg - do).max() / np.abs(do).max()
            worst, worst_d = max(worst, err), max(worst_d, derr)
            assert err <= 1e-10, "step %d: position error %.3e" % (j + 1, err)
            assert derr <= 1e-10, "step %d: displacement error %.3e" % (j + 1, derr)
            assert g.net_flux["XXX"] == flux[j] and g.last_move_attempts == attempts[j], "step %d" % (j + 1)
            prev_g, prev_o = pos.copy(), r_ref[j].copy()
        assert g.draw_count == int(f["draw_count"])
        print("C3 10 steps: worst position error %.3e, worst displacement error %.3e" % (worst, worst_d))
    finally:
        g.close()


def test_c4_mixed_factor_against_fp64_full_size():
    """N = 4000 beads, n = 12000: the mixed-precision noise factor (FP32
    trailing updates) and the FP64 one, stepped from the same positions with
    the same noise.  Reports the error of each quantity against FP64: the
    mobility (assembled in FP64 on both paths: identical), the noise
    covariance residual, the displacement and the per-step diffusion
    coefficient <|d|^2> / (6 dt)."""
    from pybrown_amd.workloads import workload
    xyz, a, inp, dt, _ = workload("c4")
    L = inp["box_length"]
    b64 = _box(xyz, a, inp)
    bmx = _box(xyz, a, inp, precision="mixed")
    try:
        errs, derrs = [], []
        for _ in range(3):
            start = b64._pos.copy()
            bmx._pos[:] = start
            b64.propagate(dt)
            bmx.propagate(dt)
            d64 = _min_image(b64._pos - start, L)
            dmx = _min_image(bmx._pos - start, L)
            errs.append(np.abs(dmx - d64).max() / np.abs(d64).max())
            D64 = (d64 ** 2).sum(axis=1).mean() / (6 * dt)
            Dmx = (dmx ** 2).sum(axis=1).mean() / (6 * dt)
            derrs.append(abs(Dmx - D64) / D64)
        r64, rmx = b64.factor_residual(), bmx.factor_residual()
        print("C4: displacement rel error %s; diffusion coefficient rel error %s; residual fp64 %.3e mixed %.3e"
              % (["%.2e" % e for e in errs], ["%.2e" % e for e in derrs], r64, rmx))
        assert np.array_equal(b64.D, bmx.D)  # the mobility is FP64 on both paths
        assert rmx <= 1e-6 and r64 <= 1e-14
        assert max(errs) <= 1e-6 and max(derrs) <= 1e-6
        assert max(errs) > 0.0  # the FP32 bulk is in use
    finally:
        b64.close()
        bmx.close()
