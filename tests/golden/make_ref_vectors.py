"""Write tests/golden/ref_vectors.npz: inputs and outputs of the reference's
own C kernels, built from its sources by oracle/Makefile into
oracle/_ref/libpyBrown_ref.so.

These fixtures let the GPU parity tests compare against the reference itself
on a machine where the reference tree is absent.  Cases (seeded):

  pointers   N=16, L=20, Gaussian positions (sigma 5)     bead.c:27
  overlaps   200 configurations of 10 beads, thr 0 / 0.1   bead.c:169
  rpy        N=16 bidisperse (0.5, 1, 2) in L=20: far,
             overlapping and contained pairs                diff_tensor.c:73
  smith      N=16, L in {20, 150}, m = n in {0..3}          diff_tensor.c:170
  jo2b       24 (ai, aj, r) near-contact pairs              diff_tensor.c:270
  immobile   12 mobile x 7 immobile beads, L=20, positions
             up to 2.5 L out of the box (wrap loops)        bead.c:99

    make -C oracle && python tests/golden/make_ref_vectors.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle.orc import Oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_vectors.npz")


def main():
    ref = Oracle("ref")
    rng = np.random.RandomState(20240601)
    v = {}
    N, L = 16, 20.0
    xyz = rng.normal(0.0, 5.0, (N, 3))
    v["ptr_xyz"], v["ptr_L"] = xyz, L
    v["ptr_P3"] = ref.pointer_pbc_matrix(xyz, L)

    ov_xyz, ov_a, ov_thr, ov_out = [], [], [], []
    for k in range(200):
        x = rng.uniform(-5.0, 5.0, (10, 3))
        a = rng.choice([0.5, 1.0], 10)
        thr = 0.0 if k % 2 == 0 else 0.1
        ov_xyz.append(x)
        ov_a.append(a)
        ov_thr.append(thr)
        ov_out.append(ref.check_overlaps(x, a, 10.0, thr))
    v["ovl_xyz"], v["ovl_a"], v["ovl_thr"], v["ovl_L"] = np.array(ov_xyz), np.array(ov_a), np.array(ov_thr), 10.0
    v["ovl_out"] = np.array(ov_out, dtype=np.int32)

    a = rng.choice([0.5, 1.0, 2.0], N)
    xyz = rng.normal(0.0, 2.5, (N, 3))
    P = Oracle.packed(ref.pointer_pbc_matrix(xyz, L))
    v["rpy_a"], v["rpy_xyz"], v["rpy_L"], v["rpy_P"] = a, xyz, L, P
    v["rpy_M"] = ref.rpy_m_matrix(a, P)

    for L in (20.0, 150.0):
        a = rng.choice([1.0, 2.0], N) if L == 20.0 else rng.choice([5.0, 10.0], N)
        xyz = rng.normal(0.0, 5.0, (N, 3)) if L == 20.0 else rng.uniform(-L / 2, L / 2, (N, 3))
        P = Oracle.packed(ref.pointer_pbc_matrix(xyz, L))
        tag = "smith_L%d" % int(L)
        v[tag + "_a"], v[tag + "_xyz"], v[tag + "_P"] = a, xyz, P
        for mn in range(4):
            v[tag + "_M%d" % mn] = ref.rpy_smith_m_matrix(a, P, L, np.sqrt(np.pi), mn, mn)

    jo_in, jo_out = [], []
    for k in range(24):
        ai, aj = rng.choice([1.0, 2.0, 5.0], 2)
        gap = (ai + aj) * 10 ** rng.uniform(-3, 0.3)
        d = rng.normal(size=3)
        d *= (ai + aj + gap) / np.linalg.norm(d)
        jo_in.append([ai, aj, d[0], d[1], d[2]])
        jo_out.append(ref.jo_2b_r_matrix(ai, aj, d))
    v["jo_in"], v["jo_out"] = np.array(jo_in), np.array(jo_out)

    r2 = np.random.RandomState(20241203)  # own stream: earlier cases unchanged
    xm = r2.uniform(-50.0, 50.0, (12, 3))
    xi = r2.uniform(-50.0, 50.0, (7, 3))
    v["imm_xm"], v["imm_xi"], v["imm_L"] = xm, xi, 20.0
    v["imm_P3"] = ref.pointer_immobile_pbc_matrix(xm, xi, 20.0)
    np.savez_compressed(OUT, **v)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
