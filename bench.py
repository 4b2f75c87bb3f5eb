"""Wavelength-sweep throughput of the retarded BEM solve on one MI355X.

Workload (the headline configuration): Drude gold sphere (eps_inf = 10,
wp = 9.065 eV, gamma = 0.0708 eV) of diameter 40 nm in vacuum, mesh
trisphere(1444, 40) = 2884 curved faces; plane waves polarised along x and y,
propagating along +z (2 right-hand sides); 64 wavelengths linspace(400, 800).
One step = the whole 64-wavelength sweep: Green assembly, factorisation,
solve and extinction for every wavelength, batched.

``value`` is the mixed-precision path (complex64 factorisation on MFMA +
FP64 iterative refinement) in wavelengths/s; the same sweep in complex128
("FP64 plain") is timed too and reported as ``fp64_value``.  The accuracy of
the mixed path is stated against the complex128 GPU path (all 64 wavelengths)
and against the FP64 CPU oracle (the baseline sample wavelengths).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_FP32_MFMA = 157.3e12   # MI355X dense FP32 matrix, FLOP/s
PEAK_FP64_MFMA = 78.6e12    # MI355X dense FP64 matrix, FLOP/s
PEAK_BF16_MFMA = 2.5e15     # MI355X dense bf16 matrix, FLOP/s
# the split-product GEMM executes three bf16 products per real product, so its
# ceiling in algorithmic (8 M N K per complex GEMM) flops is a third of bf16 peak
PEAK_BF16X3_ALG = PEAK_BF16_MFMA / 3.0
DRUDE = (10.0, 9.065, 0.0708)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(lams, pol, dr, seconds_budget=30.0):
    """Time the FP64 numpy/scipy oracle (the reference CPU algorithm) on the
    host on a bounded sample of the sweep's wavelengths."""
    import oracle as o
    nodes, tri6 = o.trisphere_mesh(1444, 40.0)
    comp = o.Composite([o.EpsConst(1.0), o.EpsDrude(*DRUDE)], [o.CurvedSurface(nodes, tri6)], [[2, 1]])
    t0 = time.time()
    setup = o.GreenSetup(comp)
    t_setup = time.time() - t0
    bem = o.BEMRetOracle(setup)
    done, ext = [], []
    t0 = time.time()
    for lam in lams:
        exc = o.planewave_potential(comp, pol, dr, lam)
        sig = bem.solve(exc)
        ext.append(o.extinction(comp, sig, pol, dr, lam))
        done.append(lam)
        if time.time() - t0 > seconds_budget:
            break
    dt = time.time() - t0
    return np.array(done), np.array(ext), len(done) / dt, t_setup


def blas_threads():
    for k in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS", "MKL_NUM_THREADS"):
        if os.environ.get(k):
            try:
                return int(os.environ[k])
            except ValueError:
                pass
    return len(os.sched_getaffinity(0))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--chunk", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the complex128 comparison sweep")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device with the library's plain copy kernel "
                         "(k_copy_cast<double>), the reference point that scales HBM byte counters")
    ap.add_argument("--traffic", type=str, default=os.path.join(ROOT, "profiles", "k_cgemm_sub_bf16x3_traffic_n2884.md"),
                    help="rocprofv3 summary holding the dominant kernel's measured HBM bytes per launch "
                         "(lines 'fetch_bytes_per_launch: F' and 'write_bytes_per_launch: W', from FETCH_SIZE / "
                         "WRITE_SIZE counter passes scaled by a plain copy) to report as roofline.traffic")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "mixed-fp32"],
                    help="mixed: complex64 GEMMs on bf16 matrix cores (3 split products); "
                         "mixed-fp32: on FP32 matrix cores")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; the sweep shards by wavelength across GPUs "
              "(see DESIGN.md) but no multi-GPU figure is produced", file=sys.stderr)
        sys.exit(2)

    import torch
    import mnpbem_amd as m
    from mnpbem_amd import _native as nat

    dev = torch.device("cuda:0")
    if args.copy_probe > 0:
        cnt = args.copy_probe * (1 << 20) // 8
        src = torch.ones(cnt, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        nat.call("mnpbem_convert", nat.DTYPE_C128, cnt, nat.ptr(src), nat.ptr(dst), nat.stream_handle(dev))
        torch.cuda.synchronize()
        log("copy probe: %d bytes read, %d bytes written by k_copy_cast<double>" % (cnt * 8, cnt * 8))
        del src, dst
    lams = np.linspace(400.0, 800.0, 64)
    pol = [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]
    dr = [[0.0, 0.0, 1.0], [0.0, 0.0, 1.0]]
    p = m.trisphere(1444, 40.0)
    cp = m.ComParticle([m.EpsConst(1.0), m.EpsDrude(*DRUDE)], [p], [[2, 1]], 1, interp="curv")
    N = cp.n
    pw = m.PlaneWaveRet(pol, dr)
    t0 = time.time()
    bem = m.BEMRet(cp, precision=args.precision)
    torch.cuda.synchronize()
    log("setup (mesh + GPU refinement coefficients): %.2f s, N = %d" % (time.time() - t0, N))

    def sweep(b):
        return b.solve_sweep(lams, pw, chunk=args.chunk)

    for w in range(args.warmup):
        sweep(bem)
        log("warmup %d done" % w)
    torch.cuda.synchronize()
    nat.call("mnpbem_prof_begin", 200000)
    t0 = time.perf_counter()
    for k in range(args.steps):
        ext_mixed = sweep(bem)
        log("step %d done (refinement iterations %s)" % (k, bem.last_refine_iters))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    import ctypes
    ms, fl, nl = ctypes.c_double(), ctypes.c_double(), ctypes.c_long()
    nat.call("mnpbem_prof_end", ctypes.byref(ms), ctypes.byref(fl), ctypes.byref(nl))
    by = ctypes.c_double()
    nat.call("mnpbem_prof_bytes", ctypes.byref(by))
    elapsed = t1 - t0
    value = len(lams) * args.steps / elapsed
    ms_per_step = 1000.0 * elapsed / args.steps
    gemm_achieved = fl.value / (ms.value * 1e-3) if ms.value > 0 else 0.0
    path_flops = 34.67 * float(N) ** 3 * len(lams) * args.steps
    log("mixed: %.2f wavelengths/s, %.1f ms/step; GEMM %d launches, %.3f ms avg, %.1f TFLOP/s"
        % (value, ms_per_step, nl.value, ms.value / max(nl.value, 1), gemm_achieved / 1e12))

    fp64 = None
    acc_vs_fp64 = None
    if not args.no_fp64:
        bem64 = m.BEMRet(cp, precision="double", green=bem.g)
        sweep(bem64)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ext64 = sweep(bem64)
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t0
        fp64 = len(lams) / t64
        acc_vs_fp64 = float(np.max(np.abs(ext_mixed / ext64 - 1)))
        log("fp64 plain: %.2f wavelengths/s; mixed vs fp64 max rel ext err %.2e" % (fp64, acc_vs_fp64))
        del bem64

    cpu = None
    acc_vs_oracle = None
    if not args.no_cpu:
        sample = lams[[0, 19, 38, 57]]
        done, ext_o, cpu_rate, t_setup = cpu_baseline(sample, pol, dr)
        idx = [int(np.argmin(np.abs(lams - l))) for l in done]
        acc_vs_oracle = float(np.max(np.abs(ext_mixed[idx] / ext_o - 1)))
        cores = blas_threads()
        cpu = {"value": round(cpu_rate, 4), "unit": "wavelengths/s", "cores": cores, "kind": "port",
               "sample": "%d of the 64 wavelengths (%s nm), N=%d, 2 RHS, complex128 numpy/scipy oracle "
                         "(reference CPU algorithm: explicit inverses + GEMMs), BLAS threads=%d; "
                         "one-off setup %.1f s not counted"
                         % (len(done), ",".join("%.1f" % l for l in done), N, cores, t_setup)}
        log("cpu oracle: %.3f wavelengths/s on %d threads; mixed vs oracle max rel ext err %.2e"
            % (cpu_rate, cores, acc_vs_oracle))

    traffic = None
    if args.traffic and args.precision == "mixed" and os.path.exists(args.traffic):
        tr = {}
        with open(args.traffic) as fh:
            for ln in fh:
                if ":" in ln and ln.split(":")[0].strip() in ("fetch_bytes_per_launch", "write_bytes_per_launch"):
                    tr[ln.split(":")[0].strip()] = int(ln.split(":")[1])
        if len(tr) == 2:
            traffic = tr["fetch_bytes_per_launch"] + tr["write_bytes_per_launch"]
    if args.precision == "mixed":
        kname = ("k_cgemm_sub_bf16x3 (complex64 GEMM on bf16 MFMA, 3 split products per real product; "
                 "achieved = 8MNK algorithmic flops / time, peak = bf16 dense peak / 3)")
        peak = PEAK_BF16X3_ALG
    else:
        kname = "k_cgemm_sub<float> (complex64 GEMM, FP32 MFMA)"
        peak = PEAK_FP32_MFMA
    line = {
        "metric": "wavelengths solved/s, 1444-face Au sphere, mixed vs FP64, <=1e-8 rel ext err",
        "value": round(value, 4),
        "unit": "wavelengths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64(bf16x3 GEMM)+f64-refine" if args.precision == "mixed" else "c64+f64-refine",
        "data": "synthetic",
        "config": {"workload": "Drude-gold sphere d=40nm, trisphere(1444) = %d curved faces, 64 wavelengths "
                               "400-800 nm, 2 plane waves (x,y pol, +z), mixed precision" % N,
                   "faces": N, "wavelengths": len(lams), "rhs": 2, "chunk": args.chunk},
        "fp64_value": None if fp64 is None else round(fp64, 4),
        "speedup_vs_fp64": None if fp64 is None else round(value / fp64, 3),
        "max_rel_ext_err_vs_fp64_gpu": acc_vs_fp64,
        "max_rel_ext_err_vs_cpu_oracle": acc_vs_oracle,
        "path_tflops_W": round(path_flops / elapsed / 1e12, 3),
        "roofline": {"bound": "mfma", "kernel": kname, "achieved": round(gemm_achieved / 1e12, 3),
                     "peak": round(peak / 1e12, 1), "unit": "TFLOP/s", "frac": round(gemm_achieved / peak, 4),
                     "traffic": traffic,
                     "compulsory_bytes_per_launch": round(by.value / max(nl.value, 1)),
                     "launches": nl.value,
                     "avg_launch_ms": round(ms.value / max(nl.value, 1), 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
