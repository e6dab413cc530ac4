"""Input and output of the simulation driver, in the reference's formats.

    read_str_file        pyBrown/input.py:113-172   `.str` bead list
    InputData            pyBrown/input.py:24-110    JSON keywords + defaults
    write_to_xyz_file    pyBrown/output.py:63-68    trajectory frames
    write_to_flux_file   pyBrown/output.py:122-138  per-label net flux
    write_to_restart_file, truncate_output_file_during_restart
                         pyBrown/output.py:44-59, 72-85

Text formats are reproduced field for field (Python's `format` of the same
numpy float64 values), so a run whose positions agree bit for bit writes
identical files.
"""
import json
import os
import pickle
import shutil
import time

import numpy as np

from pybrown_amd.bead import Bead


def timestamp(message, *variables):
    """output.py:23-27"""
    print("{}: ".format(time.asctime(time.localtime(time.time()))) + message.format(*variables) + "\n")


class InputData:
    """JSON keywords with defaults and keyword checks (input.py:24-96)."""

    def __init__(self, input_filename, obligatory_keywords=(), defaults=None, all_keywords=None):
        with open(input_filename, "r") as f:
            self.input_data = json.load(f)
        for k, v in (defaults or {}).items():
            self.input_data.setdefault(k, v)
        for k in obligatory_keywords:
            assert k in self.input_data, "Missing {} keyword in input JSON file.".format(k)
        if all_keywords is not None:
            for k in self.input_data:
                assert k in all_keywords, "Unrecognized {} keyword in input JSON file.".format(k)

    def __str__(self):
        return "".join("{}: {}\n".format(k, v) for k, v in self.input_data.items())

    __repr__ = __str__


def read_str_file(input_str_filename):
    """`.str` reader (input.py:113-172).

    sub   label id x y z a q 2r_LJ eps_LJ m   -> a Bead (hard-core radius r_LJ)
    bond  id1 id2 r_eq ? k                     -> bonded_with / bonded_how
    angle ? id1 id2 id3 theta_eq k             -> angled_with / angled_how
    """
    beads = []
    with open(input_str_filename) as f:
        for line in f:
            seg = line.split()
            if not seg:
                continue
            if seg[0] == "sub":
                beads.append(Bead(np.array([float(x) for x in seg[3:6]]), float(seg[6]), label=seg[1],
                                  hard_core_radius=float(seg[8]) / 2, epsilon_LJ=float(seg[9]),
                                  bead_id=int(seg[2])))
            elif seg[0] == "bond":
                id1, id2 = int(seg[1]), int(seg[2])
                b1 = [b for b in beads if b.bead_id == id1][-1]
                b1.bonded_with.append(id2)
                b1.bonded_how[id2] = [float(seg[3]), float(seg[5])]
            elif seg[0] == "angle":
                id1, id2, id3 = int(seg[2]), int(seg[3]), int(seg[4])
                b1 = [b for b in beads if b.bead_id == id1][-1]
                b1.angled_with.append([id2, id3])
                b1.angled_how[(id2, id3)] = [float(seg[5]), float(seg[6])]
    return beads


def write_to_xyz_file(xyz_file, xyz_filename, j, dt, beads):
    """output.py:63-68: count line, comment line with the time, one line per bead."""
    xyz_file.write("{}\n".format(len(beads)))
    xyz_file.write("{} time [ps] {}\n".format(xyz_filename, j * dt))
    for b in beads:
        xyz_file.write("{} {} {} {}\n".format(b.label, *b.r))


def write_to_flux_file(flux_file, j, dt, net_flux):
    """output.py:122-138: header at j = 0, then `time n_label1 n_label2 ...`
    (line j holds the flux of step j - 1)."""
    labels = list(net_flux.keys())
    if j == 0:
        flux_file.write(("time/ps" + " {}" * len(labels) + "\n").format(*labels))
    else:
        flux_file.write(("{}" + " {}" * len(labels) + "\n").format(j * dt, *[net_flux[k] for k in labels]))


def _file_length(filename):
    with open(filename) as f:
        return sum(1 for _ in f)


def write_to_restart_file(restart_filename, index, j, box, xyz_filename, extra_output_filenames=(), extra_data=None):
    """output.py:72-85: job index, step, the pickled Box (device state
    included) and the output files' line counts; a second copy `.rst2`."""
    with open(restart_filename, "wb", buffering=0) as f:
        pickle.dump(index, f)
        pickle.dump(j, f)
        pickle.dump(box, f)
        pickle.dump(_file_length(xyz_filename), f)
        for name in extra_output_filenames:
            pickle.dump(_file_length(name), f)
        if extra_data is not None:
            pickle.dump(extra_data, f)
    shutil.copy(restart_filename, restart_filename + "2")


def truncate_output_file_during_restart(filename, line):
    """output.py:44-59: keep the first `line` lines."""
    with open(filename) as f:
        keep = [f.readline() for _ in range(line)]
    with open(filename, "w") as f:
        f.writelines(keep)
    tmp = filename + ".tmp"
    if os.path.exists(tmp):
        os.remove(tmp)
