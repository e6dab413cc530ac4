"""Write tests/golden/c2_oracle_100.npz: 100 steps of the periodic RPY-Ewald
workload at full size (N = 1000 beads of radius 10 A, volume fraction 0.05,
alpha = sqrt(pi), m = n = 2, Ermak with overlap checks, dt = 1 ps, seed 0,
flux through the plane x = 0) run by the oracle step (oracle/box_oracle.py,
which restates box.py:150-250 over the C restatement of diff_tensor.c and
bead.c), with the RandomState(0) noise stream the reference draws.

Stored (float64 unless noted):
  xyz0        (1000, 3)  start positions (random sequential addition, seed 0)
  box_length  ()
  sub         (250,) int  beads 0, 4, 8, ...: the beads kept for every step
  r_sub       (100, 250, 3) their positions after each step
  full_steps  (3,) int    steps 10, 50, 100
  r_full      (3, 1000, 3) all positions after those steps
  flux        (100,) int  net flux of each step
  attempts    (100,) int  move attempts of each step
  draw_count  () int      normal draws consumed over the 100 steps

Run from the repository root:  python tests/golden/make_c2_trajectory.py
(about three minutes on one core)."""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.box_oracle import OracleBox  # noqa: E402
from pybrown_amd.workloads import workload  # noqa: E402

STEPS = 100
FULL = (10, 50, 100)


def main(out=os.path.join(ROOT, "tests", "golden", "c2_oracle_100.npz")):
    xyz, a, inp, dt, _ = workload("c2")
    ob = OracleBox(xyz, a, ["XXX"] * len(a), inp)
    sub = np.arange(0, len(a), 4)
    r_sub = np.empty((STEPS, sub.size, 3))
    r_full = np.empty((len(FULL), len(a), 3))
    flux = np.empty(STEPS, dtype=np.int64)
    t0 = time.time()
    for j in range(STEPS):
        ob.propagate(dt)
        r_sub[j] = ob.r[sub]
        flux[j] = ob.net_flux["XXX"]
        if j + 1 in FULL:
            r_full[FULL.index(j + 1)] = ob.r
        if (j + 1) % 10 == 0:
            print("step %d  %.1f s  attempts %d" % (j + 1, time.time() - t0, ob.attempt_log[-1]), flush=True)
    np.savez_compressed(out, xyz0=xyz, box_length=inp["box_length"], sub=sub, r_sub=r_sub,
                        full_steps=np.array(FULL), r_full=r_full, flux=flux,
                        attempts=np.array(ob.attempt_log, dtype=np.int64), draw_count=ob.draw_count)
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
