"""Diagnostic: the Jacobian row update of one Kepler pair (n = 9: 12 rows x
63 columns, J + jac_ij * J with Kahan compensation) on the VALU and on the
FP64 matrix cores, timed on the same synthetic data (C ABI:
nbg_probe_row_update).  Prints one JSON line per mode and their agreement.

    python -m nbodygradient_amd.probe [--reps R] [--grid G] [--mode both|valu|mfma]
"""
import argparse
import ctypes as C
import json

import numpy as np

from . import _lib

ROWS, COLS, K = 12, 63, 14


def run(mode, reps, grid, ctx=None):
    ctx = ctx if ctx is not None else _lib.default_context()
    ms = C.c_double(0.0)
    out = np.zeros(ROWS * COLS)
    ctx.check(_lib.lib().nbg_probe_row_update(ctx.handle, int(mode), int(reps), int(grid), C.byref(ms),
                                               out.ctypes.data_as(_lib.D)), "nbg_probe_row_update")
    return ms.value, out.reshape(ROWS, COLS)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=4000)
    ap.add_argument("--grid", type=int, default=1024)
    ap.add_argument("--mode", default="both", choices=["both", "valu", "mfma"])
    a = ap.parse_args()
    modes = {"valu": [0], "mfma": [1], "both": [0, 1]}[a.mode]
    res = {}
    for m in modes:
        run(m, 10, a.grid)  # code loading
        ms, out = run(m, a.reps, a.grid)
        updates = a.reps * a.grid
        flops = 2.0 * ROWS * K * COLS * updates  # the 12 x 14 by 14 x 63 product
        rec = {"mode": ["valu", "mfma"][m], "reps": a.reps, "grid": a.grid, "ms": ms,
               "ns_per_update_per_wave": ms * 1e6 / a.reps, "updates_per_s": updates / (ms / 1e3),
               "product_tflops": flops / (ms / 1e3) / 1e12}
        res[m] = out
        print(json.dumps(rec), flush=True)
    if len(res) == 2:
        rel = float(np.max(np.abs(res[0] - res[1])) / np.max(np.abs(res[0])))
        print(json.dumps({"valu_vs_mfma_max_rel": rel}), flush=True)


if __name__ == "__main__":
    main()
