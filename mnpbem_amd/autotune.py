"""Shape-class autotuner for the tiles of the complex64 bf16x3 GEMM.

The mixed-precision factorisation launches C -= A B products of very
different shapes: large square updates near the top of the LU / TRSM
recursion and thin ones (M or N = 64, K = 64) at its bottom.  The library has
three tile variants of the bf16x3 kernel (0: 128 x 128 with 8 waves, one
workgroup per CU; 1: 64 x 64 with 4 waves, two workgroups per CU; 2: the
128 x 128 tile reading its C tile before the K loop) and routes
every launch through a 7 x 7 x 7 table of shape classes (class c of a
dimension x: x <= 64 * 2^c, the last class unbounded).  All variants
accumulate every output element in the same order, so the table changes
speed, never results.

``tune(run)`` executes ``run()`` once, reads back one launched shape per
class that the run touched, times every variant on that shape and keeps the
fastest one.  ``save_table`` / ``load_table`` keep the result as JSON
(``data/gemm_tuning_gfx950.json``, applied when the library loads).

    python -m mnpbem_amd.autotune             # tune on the 2884-face sweep, print, save
"""

import json
import os

import numpy as np

from . import _native as nat

NCLS = 7
NCELL = NCLS ** 3
VARIANTS = {0: "128x128 (8 waves)", 1: "64x64 (4 waves)", 2: "128x128 C prefetch"}
TABLE_PATH = nat.TUNING_PATH


def cell_of(M, N, K):
    def cls(x):
        c, t = 0, 64
        while x > t and c < NCLS - 1:
            c += 1
            t *= 2
        return c
    return (cls(M) * NCLS + cls(N)) * NCLS + cls(K)


def get_table():
    import ctypes
    buf = (ctypes.c_ubyte * NCELL)()
    nat.call("mnpbem_gemm_tuning_get", buf, NCELL)
    return np.frombuffer(bytes(buf), dtype=np.uint8).copy()


def set_table(table):
    import ctypes
    t = np.asarray(table, dtype=np.uint8)
    if t.shape != (NCELL,):
        raise ValueError("gemm tuning table must have %d cells" % NCELL)
    buf = (ctypes.c_ubyte * NCELL).from_buffer_copy(t.tobytes())
    nat.call("mnpbem_gemm_tuning_set", buf, NCELL)


def seen_shapes(clear=False):
    """{cell: (M, N, K, batch)} of the bf16x3 launches since the last clear."""
    import ctypes
    buf = (ctypes.c_int * (4 * NCELL))()
    nat.call("mnpbem_gemm_tuning_seen", buf, NCELL, 1 if clear else 0)
    a = np.frombuffer(bytes(buf), dtype=np.int32).reshape(NCELL, 4)
    return {int(c): tuple(int(v) for v in a[c]) for c in range(NCELL) if a[c, 0] > 0}


def time_variant(variant, M, N, K, batch, reps=3, device=None):
    """Best time (s) of one C -= A B launch of the given tile variant on random
    complex64 operands of the given shape."""
    import torch
    dev = torch.device("cuda") if device is None else torch.device(device)
    A = torch.randn((batch, M, K), dtype=torch.complex64, device=dev)
    B = torch.randn((batch, K, N), dtype=torch.complex64, device=dev)
    C = torch.zeros((batch, M, N), dtype=torch.complex64, device=dev)
    st = nat.stream_handle(dev)

    def run():
        nat.call("mnpbem_cgemm_sub_variant", variant, M, N, K, nat.ptr(A), K, M * K, nat.ptr(B), N, K * N,
                 nat.ptr(C), N, M * N, batch, st)

    run()
    best = float("inf")
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        run()
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1) * 1e-3)
    del A, B, C
    return best


def tune(run, reps=3, device=None, verbose=True):
    """Run the workload once, then time every tile variant on one recorded
    shape of every class it used and install the faster choice.  Returns the
    table and a list of (cell, shape, seconds per variant, choice)."""
    seen_shapes(clear=True)
    run()
    shapes = seen_shapes(clear=True)
    table = get_table()
    report = []
    for cell, (M, N, K, batch) in sorted(shapes.items()):
        t = [time_variant(v, M, N, K, batch, reps, device) for v in sorted(VARIANTS)]
        choice = int(np.argmin(t))
        table[cell] = choice
        report.append((cell, (M, N, K, batch), t, choice))
        if verbose:
            print("class %3d  M=%5d N=%5d K=%5d batch=%4d  %s  -> %s"
                  % (cell, M, N, K, batch, "  ".join("%s %.3f ms" % (VARIANTS[v], t[v] * 1e3) for v in sorted(VARIANTS)),
                     VARIANTS[choice]), flush=True)
    set_table(table)
    return table, report


def save_table(table, report=None, path=TABLE_PATH, workload=""):
    doc = {"arch": "gfx950", "kernel": "k_cgemm_sub_bf16x3", "variants": VARIANTS,
           "classes": "dimension class c: x <= 64 * 2**c (c = 0..5), 6: larger; cell = (cm * 7 + cn) * 7 + ck",
           "workload": workload, "table": [int(v) for v in table]}
    if report is not None:
        doc["measured"] = [{"cell": c, "shape": list(s), "ms": [round(x * 1e3, 4) for x in t], "choice": ch}
                           for c, s, t, ch in report]
    with open(path, "w") as fh:
        json.dump(doc, fh, indent=1)
    return path


def load_table(path=TABLE_PATH):
    """Install a saved table; returns False when there is none."""
    if not os.path.exists(path):
        return False
    with open(path) as fh:
        doc = json.load(fh)
    set_table(np.asarray(doc["table"], dtype=np.uint8))
    return True


def _sweep_workload(chunk=64):
    import mnpbem_amd as m
    lams = np.linspace(400.0, 800.0, 64)
    p = m.trisphere(1444, 40.0)
    cp = m.ComParticle([m.EpsConst(1.0), m.EpsDrude(10.0, 9.065, 0.0708)], [p], [[2, 1]], 1, interp="curv")
    bem = m.BEMRet(cp, precision="mixed")
    pw = m.PlaneWaveRet([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]], [[0.0, 0.0, 1.0], [0.0, 0.0, 1.0]])
    return lambda: bem.solve_sweep(lams, pw, chunk=chunk)


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=TABLE_PATH)
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args()
    tab, rep = tune(_sweep_workload(), reps=a.reps)
    print("saved", save_table(tab, rep, a.out, "mixed BEMRet sweep, trisphere(1444, 40) N=2884, 64 wavelengths, "
                                               "batch 64, 2 RHS"))
