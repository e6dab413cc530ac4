"""The simulation driver (BD.py) and its restart (BD-restart.py) over the
device-resident Box:

    python -m pybrown_amd.bd input.json
    python -m pybrown_amd.bd --restart out.rst [--json-filename patch.json]

Keyword contract, defaults and output files follow BD.py:37-185 and
BD-restart.py: `.xyz` frames every xyz_write_freq steps, the `.flx` flux
file, the `.enr` energy file (E = 0: the hydrodynamic step has no force
field) and pickled `.rst` restart files every rst_write_freq steps, for each
job of `filename_range`.
"""
import argparse
import pickle
import sys
import time
from contextlib import ExitStack

import numpy as np

from pybrown_amd.box import Box
from pybrown_amd.io import (InputData, read_str_file, timestamp, truncate_output_file_during_restart,
                            write_to_flux_file, write_to_restart_file, write_to_xyz_file)

REQUIRED = ["output_xyz_filename", "input_str_filename", "box_length", "T", "viscosity", "dt", "number_of_steps"]

DEFAULTS = {"debug": False, "verbose": False, "hydrodynamics": "nohi",
            "ewald_alpha": np.sqrt(np.pi), "ewald_real": 0, "ewald_imag": 0,
            "diff_freq": 1, "lub_freq": 1, "chol_freq": 1, "xyz_write_freq": 1,
            "lubrication_cutoff": 1, "progress_bar": False, "seed": None, "immobile_labels": [],
            "propagation_scheme": "ermak", "m_midpoint": 100, "divergence_term": False,
            "check_overlaps": True, "external_force": [0.0, 0.0, 0.0],
            "lennard_jones_6": False, "lennard_jones_12": False, "lennard_jones_alpha": 4.0,
            "energy_unit": "joule", "custom_interactions": False, "cichocki_correction": True,
            "overlap_treshold": 0.0, "max_move_attempts": 1000000}

ALL_KEYWORDS = REQUIRED + list(DEFAULTS) + [
    "output_rst_filename", "rst_write_freq", "output_enr_filename", "enr_write_freq", "filename_range",
    "custom_interactions_filename", "auxiliary_custom_interactions_keywords", "external_force_region",
    "measure_flux", "measure_concentration", "walls"]


def write_to_enr_file(enr_file, j, dt, E, energy_unit_string):
    """output.py:108-118"""
    if j == 0:
        enr_file.write("time/ps energy/{}\n".format(energy_unit_string))
    enr_file.write("{} {}\n".format(j * dt, E))


def _names(inp, key, sub=None):
    name = inp[key][sub] if sub else inp[key]
    if "filename_range" in inp:
        return [name.format(j) for j in range(*inp["filename_range"])]
    return [name]


def _run_jobs(inp, first_index, box_rst=None, j_rst=None):
    flux = "measure_flux" in inp
    energy = "enr_write_freq" in inp
    restart = "rst_write_freq" in inp
    str_names = _names(inp, "input_str_filename")
    xyz_names = _names(inp, "output_xyz_filename")
    flx_names = _names(inp, "measure_flux", "output_flux_filename") if flux else None
    enr_names = _names(inp, "output_enr_filename") if energy else None
    rst_names = _names(inp, "output_rst_filename") if restart else None
    dt, n_steps = inp["dt"], inp["number_of_steps"]
    n_write, n_diff, n_lub, n_chol = inp["xyz_write_freq"], inp["diff_freq"], inp["lub_freq"], inp["chol_freq"]
    for index in range(first_index, len(str_names)):
        timestamp("{} job", str_names[index])
        extra = []
        if energy:
            extra.append(enr_names[index])
        if flux:
            extra.append(flx_names[index])
        start = time.time()
        if box_rst is not None and index == first_index:
            box, j0, mode = box_rst, j_rst + 1, "a"
        else:
            box, j0, mode = Box(read_str_file(str_names[index]), inp), 0, "w"
        timestamp("Random seed: {}", box.seed)
        with ExitStack() as stack:
            enr_file = stack.enter_context(open(enr_names[index], mode, buffering=1)) if energy else None
            flx_file = stack.enter_context(open(flx_names[index], mode, buffering=1)) if flux else None
            xyz_file = stack.enter_context(open(xyz_names[index], mode, buffering=1))
            for j in range(j0, n_steps):
                if inp["debug"]:
                    print("STEP {}\n".format(j))
                if j % n_write == 0:
                    write_to_xyz_file(xyz_file, xyz_names[index], j, dt, box.beads)
                if flux:
                    write_to_flux_file(flx_file, j, dt, box.net_flux)
                box.propagate(dt, j % n_diff == 0, j % n_lub == 0, j % n_chol == 0)
                if energy and j % inp["enr_write_freq"] == 0:
                    write_to_enr_file(enr_file, j, dt, box.E, inp["energy_unit"])
                if restart and j != 0 and j % inp["rst_write_freq"] == 0:
                    write_to_restart_file(rst_names[index], index, j, box, xyz_names[index], extra)
        box.close()
        print("{} seconds elapsed".format(time.time() - start))


def main_run(input_filename):
    """BD.py:main"""
    timestamp("Reading input from {} file", input_filename)
    i = InputData(input_filename, REQUIRED, DEFAULTS, ALL_KEYWORDS)
    timestamp("Input data:\n{}", i)
    _run_jobs(i.input_data, 0)


def main_restart(restart_filename, json_filename=None):
    """BD-restart.py:main: resume the pickled job, truncating the output
    files to their length at the checkpoint."""
    timestamp("Reading restart from {} file", restart_filename)
    patch = InputData(json_filename).input_data if json_filename else {}
    with open(restart_filename, "rb") as f:
        index_rst = pickle.load(f)
        j_rst = pickle.load(f)
        box_rst = pickle.load(f)
        xyz_len = pickle.load(f)
        enr_len = pickle.load(f) if box_rst.is_energy else None
        flx_len = pickle.load(f) if box_rst.is_flux else None
    inp = box_rst.inp
    for k, v in patch.items():
        inp[k] = v
    if box_rst.is_energy:
        truncate_output_file_during_restart(_names(inp, "output_enr_filename")[index_rst], enr_len)
    if box_rst.is_flux:
        truncate_output_file_during_restart(_names(inp, "measure_flux", "output_flux_filename")[index_rst], flx_len)
    truncate_output_file_during_restart(_names(inp, "output_xyz_filename")[index_rst], xyz_len)
    _run_jobs(inp, index_rst, box_rst, j_rst)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("input_filename", nargs="?")
    ap.add_argument("--restart", metavar="RST")
    ap.add_argument("-j", "--json-filename")
    a = ap.parse_args(argv)
    if a.restart:
        main_restart(a.restart, a.json_filename)
    elif a.input_filename:
        main_run(a.input_filename)
    else:
        ap.error("an input JSON file or --restart is required")


if __name__ == "__main__":
    sys.exit(main())
