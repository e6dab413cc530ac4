"""Autotuner for the noise factor's launch shape per matrix order.

The Cholesky (box.py:554, and the factorisations inside the SPD inverses of
box.py:514,524,537) runs in one of three shapes: variant 4 (one launch per
tile column, the solve of the next column folded in), or variant 6 (two tile
columns per launch, rank-128 trailing updates) with one or with two
workgroups per compute unit.  Which is fastest depends on the order: small
factorisations are bound by the diagonal chain, large ones by the trailing
updates.  The tuner times each shape on the device at a grid of orders
(`pbx_time_potrf`: the device part of pbx_cholesky on a synthetic SPD matrix)
and fills the library's per-order table (`pbx_set_potrf_table`) with the
fastest, each order between grid points taking its nearest grid point's
(in log order).  A shape replaces the default only when it is faster by more
than `margin`, so timing noise does not flip choices.

The shapes differ in the last bits of the factor (variant 4 accumulates
rank-64 trailing updates; variant 6 with two workgroups per CU alternates
rank-128 and rank-256 ones, with one per CU it applies rank-128 ones), all
within the parity tolerance.
The table is process-wide: simulations run alone and in the mix use the same
shapes, so served results stay identical to one-at-a-time ones under a fixed
table.

    python -m pybrown_amd.autotune --out potrf_table.json
    autotune.apply(autotune.load("potrf_table.json"))
"""
import argparse
import ctypes
import json
import math
import time

CODES = (4, 61, 62)
CODE_TEXT = {0: "library default", 4: "variant 4", 61: "variant 6, one workgroup per CU",
             62: "variant 6, two workgroups per CU"}
DEFAULT_GRID = (2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192)
TABLE_MAX = 1024


def default_code(T, own_cu_tiles=64):
    """The shape the library uses at T tile columns with no table."""
    return 61 if T <= own_cu_tiles else 62


def time_shape(T, code, reps=5, mixed=False):
    """Milliseconds per factorisation of order 64 T in shape `code`."""
    from pybrown_amd import _lib
    lib = _lib.lib
    table = (ctypes.c_int * (T + 1))()
    table[T] = code
    _lib.check_rc(lib.pbx_set_potrf_table(table, T + 1), "set_potrf_table")
    try:
        ms = ctypes.c_double()
        _lib.check_rc(lib.pbx_time_potrf(64 * T, reps, 1 if mixed else 0, ctypes.byref(ms)), "time_potrf")
        return ms.value
    finally:
        lib.pbx_set_potrf_table(None, 0)


def choose(times, T, margin=0.02, own_cu_tiles=64):
    """The fastest code of {code: ms}, or the default code unless another
    is faster than it by more than `margin` (relative)."""
    base = default_code(T, own_cu_tiles)
    best = min(times, key=times.get)
    if base in times and times[best] >= times[base] * (1.0 - margin):
        return base
    return best


def tune(grid=DEFAULT_GRID, reps=5, margin=0.02, log=None):
    """Time every shape at every order of `grid` (tile columns); returns a
    result dict {"grid", "codes", "ms", "reps", "margin", "device", "seconds"}
    (codes[k] is the choice at grid[k])."""
    from pybrown_amd import _lib
    lib = _lib.lib
    own = lib.pbx_get_potrf_own_cu_tiles()
    t0 = time.perf_counter()
    ms, codes = {}, []
    for T in grid:
        if T < 1 or T >= TABLE_MAX:
            raise ValueError("tile-column count %d outside [1, %d)" % (T, TABLE_MAX))
        t = {c: time_shape(T, c, reps) for c in CODES}
        # second pass in the opposite order: drifts in clock or temperature
        # then weigh on every shape alike; the smaller mean counts
        for c in reversed(CODES):
            t[c] = min(t[c], time_shape(T, c, reps))
        ms[T] = t
        codes.append(choose(t, T, margin, own))
        if log:
            log("T=%d (n=%d): %s -> %s" % (T, 64 * T, ", ".join("%d: %.3f ms" % (c, t[c]) for c in CODES),
                                          CODE_TEXT[codes[-1]]))
    return {"grid": list(grid), "codes": codes, "ms": {str(T): {str(c): v for c, v in t.items()} for T, t in ms.items()},
            "reps": reps, "margin": margin, "own_cu_tiles": own, "device": _device_name(),
            "seconds": time.perf_counter() - t0}


def table_from(result, t_max=TABLE_MAX - 1):
    """The per-order code list (index = tile columns, 0 .. t_max) from a tune
    result: each order takes the code of the grid point nearest to it in
    log(T)."""
    grid, codes = result["grid"], result["codes"]
    if len(grid) != len(codes) or not grid:
        raise ValueError("tune result without a grid")
    for c in codes:
        if c not in CODES and c != 0:
            raise ValueError("unknown shape code %r" % c)
    out = [0] * (t_max + 1)
    for T in range(1, t_max + 1):
        k = min(range(len(grid)), key=lambda i: (round(abs(math.log(grid[i] / T)), 9), grid[i]))
        out[T] = codes[k]
    return out


def apply(result, t_max=TABLE_MAX - 1):
    """Install a tune result as the library's per-order table."""
    from pybrown_amd import _lib
    codes = table_from(result, t_max)
    arr = (ctypes.c_int * len(codes))(*codes)
    _lib.check_rc(_lib.lib.pbx_set_potrf_table(arr, len(codes)), "set_potrf_table")
    return codes


def clear():
    from pybrown_amd import _lib
    _lib.check_rc(_lib.lib.pbx_set_potrf_table(None, 0), "set_potrf_table")


def save(result, path):
    with open(path, "w") as f:
        json.dump(result, f, indent=1)


def load(path):
    with open(path) as f:
        result = json.load(f)
    table_from(result)  # validates
    return result


def _device_name():
    """Marketing name and architecture of device 0, as the HIP runtime the
    library uses reports them (a table is valid for the device it was timed
    on)."""
    from pybrown_amd import _lib
    buf = ctypes.create_string_buffer(256)
    if _lib.lib.pbx_device_name(buf, len(buf)) != 0:
        return None
    return buf.value.decode(errors="replace")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default="potrf_table.json")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--margin", type=float, default=0.02)
    ap.add_argument("--grid", default=",".join(str(t) for t in DEFAULT_GRID),
                    help="tile-column counts to time (comma separated)")
    a = ap.parse_args(argv)
    grid = [int(x) for x in a.grid.split(",") if x]
    res = tune(grid, a.reps, a.margin, log=print)
    save(res, a.out)
    print("wrote %s (%.1f s)" % (a.out, res["seconds"]))


if __name__ == "__main__":
    main()
