"""Wavelength-sweep throughput of the retarded BEM solve on one MI355X.

Workload (the headline configuration): Drude gold sphere (eps_inf = 10,
wp = 9.065 eV, gamma = 0.0708 eV) of diameter 40 nm in vacuum, mesh
trisphere(1444, 40) = 2884 curved faces; plane waves polarised along x and y,
propagating along +z (2 right-hand sides); 64 wavelengths linspace(400, 800).
One step = the whole 64-wavelength sweep: Green assembly, factorisation,
solve and extinction for every wavelength, batched.

``value`` is the mixed-precision path (complex64 factorisation on MFMA +
FP64 iterative refinement) in wavelengths/s, timed as replays of the
hipGraph-captured chunk step (``SweepGraph``; the eager launches of the same
step are timed beside it as ``eager_value`` and give the GEMM roofline
figures; ``--eager`` times them alone); the same sweep in complex128
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


def cpu_config1():
    """The FP64 numpy/scipy oracle on the reference's own CPU case in full:
    trisphere(144, 20) = 284 curved faces, Drude gold, x polarisation along
    +z, 41 wavelengths 400-800 nm (complex128).  Returns wavelengths/s and
    the seconds taken (one-off set-up not counted)."""
    import oracle as o
    nodes, tri6 = o.trisphere_mesh(144, 20.0)
    comp = o.Composite([o.EpsConst(1.0), o.EpsDrude(*DRUDE)], [o.CurvedSurface(nodes, tri6)], [[2, 1]])
    bem = o.BEMRetOracle(o.GreenSetup(comp))
    lams = np.linspace(400.0, 800.0, 41)
    t0 = time.time()
    for lam in lams:
        exc = o.planewave_potential(comp, [[1.0, 0.0, 0.0]], [[0.0, 0.0, 1.0]], lam)
        o.extinction(comp, bem.solve(exc), [[1.0, 0.0, 0.0]], [[0.0, 0.0, 1.0]], lam)
    dt = time.time() - t0
    return len(lams) / dt, dt


def blas_threads():
    for k in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS", "MKL_NUM_THREADS"):
        if os.environ.get(k):
            try:
                return int(os.environ[k])
            except ValueError:
                pass
    return len(os.sched_getaffinity(0))


def copy_probe(mib):
    """One plain device copy of known size through the library's copy kernel
    (k_copy_cast<double>), the calibration of the HBM byte counters."""
    import torch
    from mnpbem_amd import _native as nat
    dev = torch.device("cuda:0")
    cnt = mib * (1 << 20) // 8
    src = torch.ones(cnt, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    nat.call("mnpbem_convert", nat.DTYPE_C128, cnt, nat.ptr(src), nat.ptr(dst), nat.stream_handle(dev))
    torch.cuda.synchronize()
    log("copy probe: %d bytes read, %d bytes written by k_copy_cast<double>" % (cnt * 8, cnt * 8))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--chunk", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the complex128 comparison sweep")
    ap.add_argument("--no-fp32", action="store_true",
                    help="skip the mixed-fp32 sweep (complex64 GEMMs on FP32 matrix cores) timed beside the headline")
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
    ap.add_argument("--residual", choices=["auto", "stored", "matrix_free"], default="auto",
                    help="FP64 refinement residual of the mixed path: complex128 copies of G/H or entries "
                         "recomputed on the fly")
    ap.add_argument("--refine-tol", type=float, default=1e-12,
                    help="FP64 refinement stops when ||R - A X|| <= tol ||R|| for every wavelength")
    ap.add_argument("--workload", default="sphere", choices=["sphere", "uneven", "dimer", "adjoint"],
                    help="sphere: the headline 64-wavelength sweep of the 2884-face gold sphere (default); "
                         "uneven: 32 seeded spheres/ellipsoids of 284-2884 faces, 64 wavelengths each; "
                         "dimer: two 12500-face spheres, 2 nm gap, H-matrix GMRES at 500/600/700 nm; "
                         "adjoint: extinction and its vertex / Drude gradients of a 2880-face gold rod, "
                         "21 wavelengths 600-800 nm")
    ap.add_argument("--fd-vertices", type=int, default=20,
                    help="adjoint workload: random vertices checked against central differences")
    ap.add_argument("--seed", type=int, default=0, help="seed of the uneven-particle set")
    ap.add_argument("--dimer-precond", default="block", choices=["block", "whole"],
                    help="dimer workload: GMRES preconditioner (block: each sphere's direct solve, one "
                         "factorisation shared by the two translated spheres; whole: the direct solve of the "
                         "whole dimer, an n x n factorisation)")
    ap.add_argument("--eager", action="store_true",
                    help="time the headline sweep with eager launches only (default: value is the replayed "
                         "hipGraph of the 64-wavelength chunk step, eager steps timed beside it)")
    ap.add_argument("--graph-iters", type=int, default=3, help="refinement steps inside the captured step")
    args = ap.parse_args()
    if args.gpus == 1 and args.copy_probe > 0:
        copy_probe(args.copy_probe)
    if args.gpus == 1 and args.workload == "uneven":
        return bench_uneven(args)
    if args.gpus == 1 and args.workload == "dimer":
        return bench_dimer(args)
    if args.gpus == 1 and args.workload == "adjoint":
        return bench_adjoint(args)
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; the sweep would shard by wavelength across GPUs "
              "(one process per GPU, one gather of the spectra) but no multi-GPU figure is produced",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import mnpbem_amd as m
    from mnpbem_amd import _native as nat

    dev = torch.device("cuda:0")
    lams = np.linspace(400.0, 800.0, 64)
    pol = [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]
    dr = [[0.0, 0.0, 1.0], [0.0, 0.0, 1.0]]
    p = m.trisphere(1444, 40.0)
    cp = m.ComParticle([m.EpsConst(1.0), m.EpsDrude(*DRUDE)], [p], [[2, 1]], 1, interp="curv")
    N = cp.n
    pw = m.PlaneWaveRet(pol, dr)
    t0 = time.time()
    gemm_tiles = ("autotuned table %s" % os.path.relpath(nat.TUNING_PATH, ROOT)
                  if os.path.exists(nat.TUNING_PATH) else "built-in default table")
    bem = m.BEMRet(cp, precision=args.precision, refine_tol=args.refine_tol, residual=args.residual)
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
        log("step %d done (refinement solves %s, final residual %.2e)"
            % (k, bem.last_refine_iters, float(np.max(bem.last_residual))))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    import ctypes
    ms, fl, nl = ctypes.c_double(), ctypes.c_double(), ctypes.c_long()
    nat.call("mnpbem_prof_end", ctypes.byref(ms), ctypes.byref(fl), ctypes.byref(nl))
    by, cby = ctypes.c_double(), ctypes.c_double()
    nat.call("mnpbem_prof_bytes", ctypes.byref(by))
    nat.call("mnpbem_prof_cbytes", ctypes.byref(cby))
    elapsed = t1 - t0
    value = len(lams) * args.steps / elapsed
    ms_per_step = 1000.0 * elapsed / args.steps
    gemm_achieved = fl.value / (ms.value * 1e-3) if ms.value > 0 else 0.0
    path_flops = 32.0 * float(N) ** 3 * len(lams) * args.steps
    log("mixed: %.2f wavelengths/s, %.1f ms/step; GEMM %d launches, %.3f ms avg, %.1f TFLOP/s"
        % (value, ms_per_step, nl.value, ms.value / max(nl.value, 1), gemm_achieved / 1e12))

    graph = None
    if not args.eager and args.chunk >= len(lams):
        from mnpbem_amd.bem import SweepGraph
        sg = SweepGraph(bem, pw, len(lams), iters=args.graph_iters)
        sg.run(lams)  # capture
        torch.cuda.synchronize()
        tg0 = time.perf_counter()
        for _ in range(args.steps):
            ext_g = sg.run(lams)
        torch.cuda.synchronize()
        tg = (time.perf_counter() - tg0) / args.steps
        graph = {"graph_value": round(len(lams) / tg, 4), "graph_ms_per_step": round(1000.0 * tg, 1),
                 "graph_refine_steps": args.graph_iters,
                 "graph_max_rel_diff_vs_eager": float(np.max(np.abs(ext_g / ext_mixed - 1))),
                 "graph_final_residual": float(np.max(sg.last_residual)),
                 "graph_eager_fallbacks": sg.fallbacks}
        log("graph: %.2f wavelengths/s (%.1f ms/step), vs eager max rel diff %.2e, %d eager fallbacks"
            % (graph["graph_value"], graph["graph_ms_per_step"], graph["graph_max_rel_diff_vs_eager"],
               sg.fallbacks))
        del sg

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

    # the same mixed path with its complex64 GEMMs on FP32 matrix cores
    # instead of the bf16 split products, timed eagerly beside the headline
    fp32 = None
    if args.precision == "mixed" and not args.no_fp32:
        bem32 = m.BEMRet(cp, precision="mixed-fp32", refine_tol=args.refine_tol, residual=args.residual,
                         green=bem.g)
        sweep(bem32)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ext32 = sweep(bem32)
        torch.cuda.synchronize()
        t32 = time.perf_counter() - t0
        fp32 = {"value": round(len(lams) / t32, 4), "unit": "wavelengths/s", "execution": "eager launches",
                "refinement_solves": bem32.last_refine_iters,
                "max_rel_ext_err_vs_fp64_gpu": (None if fp64 is None
                                                else float(np.max(np.abs(ext32 / ext64 - 1))))}
        log("mixed-fp32: %.2f wavelengths/s (%s refinement solves)" % (fp32["value"], bem32.last_refine_iters))
        del bem32

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
        c1_rate, c1_s = cpu_config1()
        cpu["config1"] = {"value": round(c1_rate, 3), "unit": "wavelengths/s", "seconds": round(c1_s, 2),
                          "sample": "the reference's CPU case in full: trisphere(144, 20) = 284 faces, Drude gold, "
                                    "x-pol +z, 41 wavelengths, same oracle and threads (the reference publishes "
                                    "6.683 s for this mesh with tabulated gold on its EPYC host)"}
        log("cpu oracle, 284 faces x 41 wavelengths: %.2f s (%.2f wavelengths/s)" % (c1_s, c1_rate))

    # HBM bytes per launch of the dominant kernel: FETCH_SIZE from the counter
    # pass summarised in the traffic file; writes from its WRITE_SIZE line, or,
    # where the file has none, the compulsory writes of C of this run (the
    # kernel stores each output element once)
    traffic, traffic_write = None, None
    if args.traffic and args.precision == "mixed" and os.path.exists(args.traffic):
        tr = {}
        with open(args.traffic) as fh:
            for ln in fh:
                if ":" in ln and ln.split(":")[0].strip() in ("fetch_bytes_per_launch", "write_bytes_per_launch"):
                    tr[ln.split(":")[0].strip()] = int(ln.split(":")[1])
        if "fetch_bytes_per_launch" in tr:
            if "write_bytes_per_launch" in tr:
                traffic_write = "WRITE_SIZE"
                wr = tr["write_bytes_per_launch"]
            else:
                traffic_write = "compulsory C writes of this run (no WRITE_SIZE pass in the traffic file)"
                wr = round(cby.value / max(nl.value, 1))
            traffic = tr["fetch_bytes_per_launch"] + wr
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
                   "faces": N, "wavelengths": len(lams), "rhs": 2, "chunk": args.chunk,
                   "gemm_tiles": gemm_tiles, "residual": args.residual},
        "fp64_value": None if fp64 is None else round(fp64, 4),
        "mixed_fp32": fp32,
        "speedup_vs_fp64": None if fp64 is None else round(value / fp64, 3),
        "max_rel_ext_err_vs_fp64_gpu": acc_vs_fp64,
        "max_rel_ext_err_vs_cpu_oracle": acc_vs_oracle,
        "path_tflops_W": round(path_flops / elapsed / 1e12, 3),
        "roofline": {"bound": "mfma", "kernel": kname, "achieved": round(gemm_achieved / 1e12, 3),
                     "peak": round(peak / 1e12, 1), "unit": "TFLOP/s", "frac": round(gemm_achieved / peak, 4),
                     "traffic": traffic,
                     "traffic_write": traffic_write,
                     "compulsory_bytes_per_launch": round(by.value / max(nl.value, 1)),
                     "launches": nl.value,
                     "avg_launch_ms": round(ms.value / max(nl.value, 1), 4)},
        "cpu_baseline": cpu,
    }
    if graph is not None:
        # value: the same sweep as replays of the captured chunk step (all
        # kernels of the eager step, residual checked after every replay)
        line["eager_value"] = line["value"]
        line["eager_ms_per_step"] = line["ms_per_step"]
        line["value"] = graph["graph_value"]
        line["ms_per_step"] = graph["graph_ms_per_step"]
        line["speedup_vs_fp64"] = None if fp64 is None else round(line["value"] / fp64, 3)
        line["config"]["execution"] = ("hipGraph replay of the 64-wavelength chunk step (Green assembly, "
                                       "factorisations, %d refinement solves, extinction); roofline from the "
                                       "eager steps" % args.graph_iters)
        line["graph"] = graph
    else:
        line["config"]["execution"] = "eager launches"
    print(json.dumps(line), flush=True)


def bench_uneven(args):
    """32 particles of uneven size, 64 wavelengths each, solved together by
    BEMRetBatch: seeded kinds (sphere with curved faces / flat ellipsoid),
    sizes (144, 256, 400, 900, 1444 vertices -> 284 ... 2884 faces), sphere
    diameters U(20, 80) nm, ellipsoid semi-axes U(10, 40) nm, Drude gold with
    gamma U(0.05, 0.1) eV; mixed precision with FP64 refinement.  The
    particles are grouped by face count and every group's matrices of one
    wavelength block share each launch; the block step of all groups is one
    captured HIP graph, replayed 64 / L times per step (L from the free
    memory).  Set-up (GPU refinement data of all particles) is timed apart."""
    import torch
    import mnpbem_amd as m
    from mnpbem_amd.batch import BEMRetBatch, BatchSweepGraph
    rng = np.random.default_rng(args.seed)
    sizes = [144, 256, 400, 900, 1444]
    lams = np.linspace(400.0, 800.0, 64)
    pw = m.PlaneWaveRet([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]], [[0.0, 0.0, 1.0], [0.0, 0.0, 1.0]])
    specs = []
    for i in range(32):
        kind = "sphere" if rng.random() < 0.5 else "ellipsoid"
        nv = int(rng.choice(sizes))
        geo = float(rng.uniform(20.0, 80.0)) if kind == "sphere" else rng.uniform(10.0, 40.0, 3)
        gam = float(rng.uniform(0.05, 0.1))
        specs.append((kind, nv, geo, gam))

    def build(spec):
        kind, nv, geo, gam = spec
        part = m.trisphere(nv, geo) if kind == "sphere" else m.triellipsoid(nv, geo)
        return m.ComParticle([m.EpsConst(1.0), m.EpsDrude(10.0, 9.065, gam)], [part], [[2, 1]], 1)

    parts = [build(s) for s in specs]
    faces = [p.n for p in parts]
    t0 = time.perf_counter()
    bb = BEMRetBatch(parts, precision=args.precision)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    L = bb.auto_block(len(lams))
    groups = [(g.n, g.S) for g in bb.groups]
    log("uneven: %d particles, %d faces, groups (faces, particles) %s, block %d wavelengths, set-up %.2f s"
        % (len(parts), sum(faces), groups, L, t_setup))
    sg = BatchSweepGraph(bb, pw, L, iters=args.graph_iters)
    for w in range(max(args.warmup, 1)):
        sg.solve_sweep(lams)        # the first run captures
        log("warmup %d done" % w)
    torch.cuda.synchronize()
    fb0 = sg.fallbacks
    t1 = time.perf_counter()
    for k in range(args.steps):
        ext = sg.solve_sweep(lams)
        log("step %d done (final residual %.2e)" % (k, sg.last_residual))
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t1) / args.steps
    group_redo = (sg.fallbacks - fb0) / args.steps
    del sg
    torch.cuda.empty_cache()
    # the same batches with eager launches, and the per-particle loop of
    # BEMRet sweeps it replaces (64 wavelengths per particle in one chunk)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    ext_eager = bb.solve_sweep(lams, pw, block=L)
    torch.cuda.synchronize()
    dt_eager = time.perf_counter() - t2
    t3 = time.perf_counter()
    seq = []
    for g in bb.greens:
        seq.append(m.BEMRet(g.p1, precision=args.precision, green=g).solve_sweep(lams, pw, chunk=64))
    torch.cuda.synchronize()
    dt_seq = time.perf_counter() - t3
    diff_eager = max(float(np.max(np.abs(e / q - 1))) for e, q in zip(ext, ext_eager))
    diff_seq = max(float(np.max(np.abs(e / q - 1))) for e, q in zip(ext, seq))
    # accuracy: the largest particle in complex128
    big = int(np.argmax(faces))
    e64 = m.BEMRet(parts[big], precision="double", green=bb.greens[big]).solve_sweep(lams, pw, chunk=32)
    err = float(np.max(np.abs(ext[big] / e64 - 1)))
    units = len(specs) * len(lams)
    work = 32.0 * float(np.sum(np.asarray(faces, dtype=float) ** 3)) * len(lams)
    log("uneven: graph %.1f particle-wavelengths/s (%.3f s/step, %.1f groups per step solved again eagerly), "
        "eager %.3f s, per-particle loop %.3f s; largest vs complex128 %.2e"
        % (units / dt, dt, group_redo, dt_eager, dt_seq, err))
    line = {"metric": "particle-wavelengths solved/s, 32 uneven Drude-gold particles (config 3)",
            "value": round(units / dt, 3), "unit": "particle-wavelengths/s", "n_gpus": 1, "steps": args.steps,
            "warmup": max(args.warmup, 1), "ms_per_step": round(1000.0 * dt, 1), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None,
            "dtype": "c64(bf16x3 GEMM)+f64-refine" if args.precision == "mixed" else "c64+f64-refine",
            "data": "synthetic",
            "config": {"workload": "32 seeded spheres/ellipsoids, 64 wavelengths each, 2 plane waves, seed %d"
                                   % args.seed, "faces": faces, "total_faces": int(sum(faces)),
                       "groups_faces_particles": groups, "block_wavelengths": L,
                       "execution": "hipGraph replay of one wavelength block of all groups (%d replays per step, "
                                    "%d refinement solves)" % (len(lams) // L, args.graph_iters),
                       "setup_included": False},
            "setup_s": round(t_setup, 3),
            "graph_groups_resolved_per_step": group_redo,
            "path_tflops_W": round(work / dt / 1e12, 3),
            "eager_value": round(units / dt_eager, 3),
            "per_particle_loop_value": round(units / dt_seq, 3),
            "max_rel_diff_vs_eager": diff_eager, "max_rel_diff_vs_per_particle": diff_seq,
            "max_rel_ext_err_vs_fp64_largest": err}
    print(json.dumps(line), flush=True)


def bench_dimer(args):
    """Gold sphere dimer, two 12500-face golden-spiral spheres of 100 nm with a
    2 nm gap (25000 faces): BEMRetIter on ACA H-matrices whose entries the
    kernels generate in place (leaf 64, htol 1e-6, rank <= 30, blocks in
    complex64), GMRES restart 50, tol 1e-6, FP64 Arnoldi, preconditioned by
    each sphere's mixed-precision direct solve (--dimer-precond block,
    default) or by that of the whole dimer (whole); 500, 600, 700 nm; x
    polarisation along the dimer axis, propagation along z."""
    import torch
    import mnpbem_amd as m
    from mnpbem_amd import _native as nat
    t0 = time.perf_counter()
    p1 = m.trisphere_fibonacci(6252, 100.0)
    p2 = p1.shift([102.0, 0.0, 0.0])
    cp = m.ComParticle([m.EpsConst(1.0), m.EpsDrude(*DRUDE)], [p1, p2], [[2, 1], [2, 1]], 1, 2)
    lams = np.array([500.0, 600.0, 700.0])
    pw = m.PlaneWaveRet([[1.0, 0.0, 0.0]], [[0.0, 0.0, 1.0]])
    bem = m.BEMRetIter(cp, hmatrix=True, htol=1e-6, kmax=30, cleaf=64, tol=1e-6, restart=50, maxit=20,
                       precision="mixed", hmat_dtype="complex64",
                       preconditioner="block" if args.dimer_precond == "block" else "hlu_dense")
    torch.cuda.reset_peak_memory_stats()
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    log("dimer set-up %.1f s, N = %d" % (t_setup, cp.n))
    exts = []
    t1 = time.perf_counter()
    for lam in lams:
        exts.append(bem.solve_sweep([lam], pw, chunk=1)[0, 0])
        flag, rel, it = bem.info()
        st = [h.stat() for h in bem.last_hmats[0]]
        log("lambda %.0f: ext %.6e, GMRES flag %d relres %.2e it %s, compression %s, rank max %s" %
            (lam, exts[-1], flag[-1], rel[-1], list(it[-1]), [round(x["compression"], 3) for x in st],
             [x["rank_max"] for x in st]))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t1
    peak_gb = torch.cuda.max_memory_allocated() / 1e9
    # phases at 600 nm, outside the timed region: the preconditioner's
    # factorisation, the in-place ACA of the four Green functions, GMRES
    phase = {}
    lam = np.array([600.0])
    torch.cuda.synchronize()
    ta = time.perf_counter()
    if bem._block_bems is not None:
        for b in bem._block_bems:
            b.factor(lam)
    else:
        bem._direct.factor(lam)
    torch.cuda.synchronize()
    phase["precond_factor_s"] = round(time.perf_counter() - ta, 3)
    ta = time.perf_counter()
    kmat = bem.g.wavenumbers(lam)
    hms = [bem._aca.generate(s_, key, kmat[0], 600.0) for s_, key in ((0, "G"), (0, "H1"), (1, "G"), (1, "H2"))]
    torch.cuda.synchronize()
    phase["aca_in_place_s"] = round(time.perf_counter() - ta, 3)
    S = bem._prepare(lam)
    R = pw.rhs(cp, lam, bem.g)
    torch.cuda.synchronize()
    ta = time.perf_counter()
    bem.solve_batch(S, R)
    torch.cuda.synchronize()
    phase["gmres_s"] = round(time.perf_counter() - ta, 3)
    phase["gmres_iterations"] = int(bem.info()[2][-1][1])
    del S, R
    log("phases at 600 nm: %s; peak device memory of the timed solves %.1f GB" % (phase, peak_gb))
    hstat = [h.stat() for h in bem.last_hmats[0]]
    # H-matrix matvec as GMRES calls it (4 columns per side for one plane
    # wave), on the four H-matrices of the ACA phase above (GMRES's own have
    # released their block storage): algorithmic bytes = stored blocks read
    # once per product, timed with HIP events on the launching stream
    n = cp.n
    xv = torch.randn((n, 4), dtype=torch.complex128, device="cuda")
    yv = torch.empty_like(xv)
    for h in hms:
        h.matvec_device(nat.ptr(xv), 4, nat.ptr(yv), 4, 4)
    reps = 10
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        for h in hms:
            h.matvec_device(nat.ptr(xv), 4, nat.ptr(yv), 4, 4)
    e1.record()
    e1.synchronize()
    mv_ms = e0.elapsed_time(e1) / (reps * len(hms))
    # the same products on the block storage (one workgroup per leaf row
    # cluster walking its blocks), for comparison
    for h in hms:
        h.matvec_blocks_device(nat.ptr(xv), 4, nat.ptr(yv), 4, 4)
    e0.record()
    for _ in range(reps):
        for h in hms:
            h.matvec_blocks_device(nat.ptr(xv), 4, nat.ptr(yv), 4, 4)
    e1.record()
    e1.synchronize()
    mvb_ms = e0.elapsed_time(e1) / (reps * len(hms))
    mv_bytes = float(np.mean([h.bytes_per_matvec() for h in hms]))
    mv_gbs = mv_bytes / (mv_ms * 1e-3) / 1e9
    log("H-matrix matvec: %.3f ms per product, %.3e bytes -> %.0f GB/s (%.2f of 8 TB/s); on the block storage "
        "%.3f ms" % (mv_ms, mv_bytes, mv_gbs, mv_gbs / 8000.0, mvb_ms))
    del bem, hms, xv, yv
    torch.cuda.empty_cache()
    # reference: the dense mixed-precision direct solve of the same system
    t2 = time.perf_counter()
    ed = m.BEMRet(cp, precision="mixed").solve_sweep(lams, pw, chunk=1)[:, 0]
    torch.cuda.synchronize()
    dt_direct = time.perf_counter() - t2
    err = float(np.max(np.abs(np.array(exts) / ed - 1)))
    line = {"metric": "wavelengths solved/s, 25000-face gold dimer, H-matrix GMRES (config 5)",
            "value": round(len(lams) / dt, 4), "unit": "wavelengths/s", "n_gpus": 1, "steps": 1, "warmup": 0,
            "ms_per_step": round(1000.0 * dt, 1), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64 GMRES, c64 H-matrix blocks, c64 preconditioner", "data": "synthetic",
            "config": {"workload": "2 x trisphere_fibonacci(6252, 100 nm), 2 nm gap, 3 wavelengths",
                       "faces": cp.n, "cleaf": 64, "htol": 1e-6, "kmax": 30, "restart": 50, "tol": 1e-6,
                       "preconditioner": args.dimer_precond},
            "setup_s": round(t_setup, 2), "phases_600nm": phase, "peak_device_GB": round(peak_gb, 2),
            "direct_mixed_value": round(len(lams) / dt_direct, 4),
            "max_rel_ext_err_vs_direct": err,
            "hmatrix": [{"compression": round(x["compression"], 4), "rank_max": x["rank_max"],
                         "bytes": x["bytes"]} for x in hstat],
            "hmatvec": {"kernels": "k_hmat_vt_pk + k_hmat_rows_pk (packed panels, c64 blocks, 4 columns)",
                        "bound": "hbm", "bytes_per_product": round(mv_bytes), "avg_ms": round(mv_ms, 4),
                        "block_storage_avg_ms": round(mvb_ms, 4),
                        "achieved_GBps": round(mv_gbs, 1), "peak_GBps": 8000.0,
                        "frac": round(mv_gbs / 8000.0, 4)}}
    print(json.dumps(line), flush=True)


def bench_adjoint(args):
    """Gold nanorod (trirod(20, 60, [31, 13, 26]) = 2880 flat triangles,
    hemispherical caps) in eps = 1.77, plane wave polarised along the rod axis
    (z), propagating along x, 21 wavelengths 600-800 nm across the
    longitudinal resonance.  One step = extinction plus its gradient with
    respect to every vertex (forward solve, adjoint solve with the same
    factors, derivative kernels, chain rule), mixed precision.  Also timed:
    the forward sweep alone and the Drude-parameter gradient.  Outside the
    timed region the vertex gradient is checked against central differences
    (step 1e-4 nm, complex128 spectra) on --fd-vertices random vertices at
    the resonance wavelength."""
    import torch
    import mnpbem_amd as m
    from mnpbem_amd.adjoint import extinction_with_drude_gradient, extinction_with_shape_gradient
    rod = m.trirod(20.0, 60.0, [31, 13, 26], triangles=True)
    mk = lambda verts: m.ComParticle([m.EpsConst(1.77), m.EpsDrude(*DRUDE)],
                                     [m.geometry.Particle(verts, rod.faces, interp="flat")], [[2, 1]], 1)
    cp = mk(rod.verts)
    lams = np.linspace(600.0, 800.0, 21)
    pw = m.PlaneWaveRet([[0.0, 0.0, 1.0]], [[1.0, 0.0, 0.0]])
    bem = m.BEMRet(cp, precision="mixed")
    log("adjoint workload: rod N = %d, %d vertices" % (cp.n, len(rod.verts)))
    for _ in range(max(args.warmup, 1)):
        extinction_with_shape_gradient(bem, lams, pw, chunk=len(lams))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ext, grads = extinction_with_shape_gradient(bem, lams, pw, chunk=len(lams))
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    gv = grads[0]["verts"]
    t1 = time.perf_counter()
    for _ in range(args.steps):
        ext_f = bem.solve_sweep(lams, pw, chunk=len(lams))
    torch.cuda.synchronize()
    dt_fwd = (time.perf_counter() - t1) / args.steps
    t2 = time.perf_counter()
    _, gd = extinction_with_drude_gradient(bem, lams, pw, chunk=len(lams))
    torch.cuda.synchronize()
    dt_drude = time.perf_counter() - t2
    k_res = int(np.argmax(ext[:, 0]))
    lam_res = lams[k_res]
    log("sweep + vertex gradient %.3f s, forward only %.3f s, Drude gradient %.3f s; resonance %.0f nm"
        % (dt, dt_fwd, dt_drude, lam_res))
    # central differences on random vertices at the resonance (complex128)
    rng = np.random.default_rng(0)
    h = 1e-4
    scale = np.abs(gv[k_res, 0]).max()
    worst = 0.0
    t3 = time.perf_counter()
    for vi in rng.choice(len(rod.verts), args.fd_vertices, replace=False):
        for c in range(3):
            vals = []
            for sg in (1.0, -1.0):
                v = rod.verts.copy()
                v[vi, c] += sg * h
                vals.append(m.BEMRet(mk(v)).solve_sweep([lam_res], pw)[0, 0])
            fd = (vals[0] - vals[1]) / (2 * h)
            worst = max(worst, abs(gv[k_res, 0, vi, c] - fd) / scale)
    log("finite-difference check: %d vertices x 3, max |grad - fd| / max|grad| = %.2e (%.1f s)"
        % (args.fd_vertices, worst, time.perf_counter() - t3))
    line = {"metric": "wavelengths/s with d(ext)/d(vertex xyz), 2880-face gold rod (config 4)",
            "value": round(len(lams) / dt, 4), "unit": "wavelengths/s", "n_gpus": 1, "steps": args.steps,
            "warmup": max(args.warmup, 1), "ms_per_step": round(1000.0 * dt, 1), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "c64(bf16x3 GEMM)+f64-refine, f64 gradient kernels",
            "data": "synthetic",
            "config": {"workload": "trirod(20, 60, [31, 13, 26]) gold rod in eps 1.77, pol z, dir x, "
                                   "21 wavelengths 600-800 nm, mixed precision",
                       "faces": cp.n, "vertices": len(rod.verts), "wavelengths": len(lams)},
            "forward_only_value": round(len(lams) / dt_fwd, 4),
            "gradient_overhead": round(dt / dt_fwd, 3),
            "drude_gradient_value": round(len(lams) / dt_drude, 4),
            "resonance_nm": float(lam_res),
            "fd_check": {"vertices": args.fd_vertices, "step_nm": h, "wavelength_nm": float(lam_res),
                         "max_abs_err_over_max_grad": worst},
            "max_rel_ext_err_grad_vs_forward": float(np.max(np.abs(ext / ext_f - 1))),
            "drude_grad_at_resonance": [float(x) for x in gd[k_res, 0]]}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
