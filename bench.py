"""Benchmark: dielectric sphere (eps = 12+0.01i, radius 1 um) at omega = 1,
icosphere level 4, PMCHWT N = 15360 -- assemble + LU + solve per frequency.

A step is one frequency of the mixed-precision path: complex128 assembly
of the BEM matrix, complex64 LU on the matrix cores, complex128 iterative
refinement of the plane-wave solve, and the power.  The plain complex128
path (assembly, complex128 LU, solve) runs the same number of steps
beside it for the speedup and the power error, which must stay <= 1e-8.
The frequency-independent FIPPI table and the panel geometry are built
once, before the timed region, as the reference caches them.

Prints one JSON line (see --help for the knobs)."""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "scuff-em_amd"))
sys.path.insert(0, ROOT)

PEAK = {"mfma_f64": 78.6e12, "mfma_f32": 157.3e12, "hbm": 8.0e12}


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--level", type=int, default=4, help="icosphere level (4: N=15360)")
    ap.add_argument("--omega", type=float, default=1.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched-sweep leg")
    ap.add_argument("--no-far32", action="store_true", help="skip the single-precision far-pair leg")
    ap.add_argument("--cpu-lu-n", type=int, default=0,
                    help="order of the CPU LAPACK sample (0: the full N)")
    ap.add_argument("--cpu-rows", type=int, default=2, help="oracle assembly rows per host core")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (calibrates the HBM "
                         "byte counters of a profiler against a copy of known size) and exit")
    return ap.parse_args()


def copy_probe(mib):
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8, "copies": 5}))


def traffic_from_profiles(kernel):
    """HBM bytes per launch measured with rocprofv3 counters (committed
    summary), or None."""
    fn = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(fn):
        return None
    with open(fn) as f:
        d = json.load(f)
    v = d.get(kernel)
    return None if v is None else v.get("bytes_per_launch")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(level, omega, eps, rows_per_core, lu_n):
    """The oracle (a C restatement of the reference's path) on the host
    cores: assembly rows of a few edges per core, each core with its own
    oracle geometry and FIPPI cache (the calls release the GIL), timed
    warm and scaled by the number of edge pairs; plus LAPACK zgetrf
    (SciPy's, threaded by its BLAS) of a complex N x N matrix at the
    full N (lu_n = 0) or at lu_n, scaled by (N / lu_n)^3."""
    import threading
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    import scipy.linalg as sl
    from oracle import oracle as O
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    V, T = O.icosphere(level)
    NE = 30 * 4 ** level
    N = 2 * NE
    nrows = cores * rows_per_core
    edges = [int(round(i * (NE - 1) / max(1, nrows - 1))) for i in range(nrows)]
    parts = [edges[c::cores] for c in range(cores)]
    start = threading.Barrier(cores + 1)
    done = threading.Barrier(cores + 1)

    def work(mine):
        Go = O.Geometry()
        Go.add_surface(V, T, kind=1, eps=eps)
        for e in mine:                       # cold: fills this core's FIPPI cache
            Go.rows(omega, e, e + 1)
        start.wait()
        for e in mine:
            Go.rows(omega, e, e + 1)
        done.wait()

    # the Clenshaw-Curtis tables are set up once, before the threads start
    O.pcubature(lambda x: [x[0]], 1, [0.0], [1.0], maxeval=50)
    with ThreadPoolExecutor(cores) as ex:
        futs = [ex.submit(work, p) for p in parts]
        start.wait()
        t0 = time.perf_counter()
        done.wait()
        t_rows = time.perf_counter() - t0
        for f in futs:
            f.result()
    sampled_pairs = sum(NE - e for e in edges)
    all_pairs = NE * (NE + 1) / 2
    t_asm = t_rows * all_pairs / sampled_pairs
    n_lu = N if lu_n <= 0 else min(lu_n, N)
    rng = np.random.default_rng(20261202)
    A = rng.standard_normal((n_lu, n_lu)) + 1j * rng.standard_normal((n_lu, n_lu))
    A = np.asfortranarray(A)
    t0 = time.perf_counter()
    sl.lu_factor(A, overwrite_a=True, check_finite=False)
    t_lu = (time.perf_counter() - t0) * (N / n_lu) ** 3
    del A
    return {
        "value": t_asm + t_lu, "unit": "s/freq", "cores": cores, "kind": "port",
        "cpu_model": cpu_model(),
        "sample": f"{cores} host threads: oracle assembly rows of {nrows} of {NE} edges, one oracle geometry "
                  f"and FIPPI cache per thread (warm, {t_rows:.2f} s wall, scaled by edge pairs to "
                  f"{t_asm:.1f} s) + SciPy LAPACK zgetrf at N={n_lu} ({t_lu:.1f} s"
                  + ("" if n_lu == N else f", scaled by (N/{n_lu})^3") + ")",
    }


def sweep_leg():
    """The batch of 64 instances: 32 log-spaced frequencies in [0.05, 5] x
    two Drude spheres (eps = 1 - 100/(w(w + 0.1i)); level 2, N=960 and
    level 3, N=3840).  Batched: one scuffx_sweep_batched call (assembly,
    batched complex64 LU, complex128 refinement and power of all 64 in one
    captured graph).  The graph is recorded by a first sweep of the same
    shape over frequencies 0.3% higher and replayed for the timed sweep,
    since frequency data reach it through device tables; that first call,
    capture and instantiation included, is reported beside it.  Baseline:
    sequential plain solves (complex128 assembly, LU, solve, power, one
    instance after another, scuffx_batch_sweep with prec 64).  The FIPPI
    tables are built before either is timed."""
    import numpy as np
    import scuffx
    om = 0.05 * 100.0 ** (np.arange(32) / 31.0)
    pw = scuffx.PlaneWave([1.0, 0.0, 0.0], [0.0, 0.0, 1.0])
    Gs = []
    for level in (2, 3):
        V, T = scuffx.icosphere(level)
        G = scuffx.RWGGeometry()
        G.AddMesh(V, T, "DRUDE_10_0.1")
        G.BuildFIPPITable()
        Gs.append(G)
    # warm both paths once (first-use tables, workspaces, the autotuner);
    # the batched warm-up is a sweep of the same shape over other frequencies,
    # whose first call records and instantiates the graph ("cold")
    om_w = om * 1.003
    t0 = time.perf_counter()
    scuffx.SweepBatched(Gs, om_w, pw)
    t_cold = time.perf_counter() - t0
    for G in Gs:
        scuffx.BatchSweep(G, om[:1], pw, prec=64)
    phases = {}
    t0 = time.perf_counter()
    Pb, rel, info = scuffx.SweepBatched(Gs, om, pw, Timing=phases)
    t_batched = time.perf_counter() - t0
    print("# sweep: batched done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    Ps = [scuffx.BatchSweep(G, om, pw, prec=64) for G in Gs]
    t_seq = time.perf_counter() - t0
    print("# sweep: sequential done", file=sys.stderr, flush=True)
    diff = max(float(np.abs(Pb[g, k] - Ps[g][k]).max() / np.abs(Ps[g][k]).max())
               for g in range(2) for k in range(len(om)))
    return {"instances": 64, "workload": "32 log-spaced omega in [0.05,5] x Drude spheres (wp=10, gamma=0.1) "
                                         "icosphere level 2 (N=960) and level 3 (N=3840)",
            "batched_s": t_batched, "sequential_plain_s": t_seq, "speedup": t_seq / t_batched,
            "batched_first_call_s": t_cold, "speedup_first_call": t_seq / t_cold,
            "max_power_rel_diff": diff, "max_relres": float(rel.max()), "solved_again": info[0],
            "lu_block_width": info[1], "batched_phases_ms": phases}


def main():
    args = parse()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the sharded design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import scuffx
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    eps = 12 + 0.01j
    omega = args.omega
    V, T = scuffx.icosphere(args.level)
    G = scuffx.RWGGeometry()
    G.AddMesh(V, T, "CONST_EPS_12+0.01i")
    N = G.TotalBFs
    print(f"# N={N}: building FIPPI table", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    G.BuildFIPPITable()
    t_fippi = time.perf_counter() - t0
    M = G.AllocateBEMMatrix()
    LU32 = torch.empty((N, N), dtype=torch.complex64, device="cuda")
    ipiv = torch.zeros(N, dtype=torch.int32, device="cuda")
    pw = scuffx.PlaneWave([1.0, 0.0, 0.0], [0.0, 0.0, 1.0])
    RHS = G.AssembleRHSVector(omega, pw)
    KN = scuffx.HVector(N)
    x = scuffx.HVector(N)
    L = scuffx.lib()
    import ctypes as C

    phases = {"mixed": {}, "fp64": {}}

    def tick(path, name, t0):
        # every library call returns after its work is complete
        phases[path][name] = phases[path].get(name, 0.0) + time.perf_counter() - t0
        return time.perf_counter()

    def step_mixed():
        scuffx._sync_stream()
        t = time.perf_counter()
        G.AssembleBEMMatrix(omega, M)
        t = tick("mixed", "assemble", t)
        info = L.scuffx_lu32_from_c128(scuffx._ptr(M.t), N, N, scuffx._ptr(LU32), N, scuffx._ptr(ipiv))
        if info != 0:
            raise RuntimeError(f"complex64 LU failed ({info})")
        t = tick("mixed", "lu_c64", t)
        it, rel = C.c_int(), C.c_double()
        scuffx._check(L.scuffx_refine(scuffx._ptr(M.t), N, N, scuffx._ptr(LU32), N, scuffx._ptr(ipiv),
                                      scuffx._ptr(RHS.t), scuffx._ptr(x.t), 1e-12, 30, C.byref(it), C.byref(rel)))
        t = tick("mixed", "solve_refine", t)
        P = G.GetPower(omega, x, RHS)
        tick("mixed", "power", t)
        return P, it.value, rel.value

    def step_fp64():
        t = time.perf_counter()
        G.AssembleBEMMatrix(omega, M)
        t = tick("fp64", "assemble", t)
        KN.t.copy_(RHS.t)
        if M.LUFactorize() != 0:
            raise RuntimeError("complex128 LU reported a zero pivot")
        t = tick("fp64", "lu_c128", t)
        M.LUSolve(KN)
        t = tick("fp64", "solve", t)
        P = G.GetPower(omega, KN, RHS)
        tick("fp64", "power", t)
        return P

    def timed(fn, steps, label):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = None
        for s in range(steps):
            out = fn()
            print(f"# {label} step {s + 1}/{steps}", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / max(1, steps), out

    for w in range(args.warmup):
        step_mixed()
        step_fp64()
        print(f"# warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)

    # timed region: mixed path, every hot-kernel launch bracketed by HIP
    # events on its stream
    phases["mixed"].clear()
    phases["fp64"].clear()
    scuffx.profile(True)
    t_mixed, (Pm, iters, relres) = timed(step_mixed, args.steps, "mixed")
    prof_mixed = scuffx.profile_read()
    asm_stats = G.AssemblyStats()
    scuffx.profile(True)
    t_fp64, Pd = timed(step_fp64, args.steps, "fp64")
    prof_fp64 = scuffx.profile_read()
    scuffx.profile(False)

    perr = max(abs(Pm[k] - Pd[k]) / abs(Pd[k]) for k in ("PAbs", "PExt", "PScat"))

    # dominant kernel of the mixed step
    cand = {k: v for k, v in prof_mixed.items() if v[2] > 0}
    dom = max(cand, key=lambda k: cand[k][0])
    ms, work, launches = cand[dom]
    if dom.startswith("gemm") or dom.startswith("panel"):
        # the panel's flops are vector FP32/FP64 work, whose published
        # peaks equal the dense matrix-core ones on MI355X; the panel is
        # latency-bound (one global pivot exchange per column), so its
        # fraction is small by construction
        bound, unit = "mfma", "TFLOP/s"
        peak = PEAK["mfma_f32"] if dom.endswith("c64") else PEAK["mfma_f64"]
        achieved = (work / launches) / (ms / launches * 1e-3)
        scale = 1e12
    elif dom in ("resid", "stamp"):
        bound, unit, peak = "hbm", "GB/s", PEAK["hbm"]
        achieved = work / (ms * 1e-3)
        scale = 1e9
    else:
        bound, unit, peak = "mfma", "TFLOP/s", PEAK["mfma_f64"]
        achieved = work / (ms * 1e-3)
        scale = 1e12
    roofline = {
        "kernel": dom, "bound": bound, "achieved": achieved / scale, "peak": peak / scale, "unit": unit,
        "frac": achieved / peak, "traffic": traffic_from_profiles(dom),
        "launches": launches / args.steps, "avg_launch_ms": ms / launches,
        "work_per_launch": work / launches,
    }

    def kern_summary(p):
        return {k: {"ms_per_step": v[0] / args.steps, "launches_per_step": v[2] / args.steps,
                    "rate": (v[1] / (v[0] * 1e-3)) if v[0] > 0 and v[1] > 0 else None}
                for k, v in p.items() if v[2] > 0}

    # the mixed form with single-precision far pairs, beside the headline
    # (which keeps them in double: its power gate is 1e-8)
    far32 = None
    if not args.no_far32:
        saved = {p: dict(d) for p, d in phases.items()}
        G.SetFarPrecision(32)
        step_mixed()
        print("# far32 warmup done", file=sys.stderr, flush=True)
        phases["mixed"].clear()
        t32, (P32, it32, rel32) = timed(step_mixed, args.steps, "far32")
        asm32 = phases["mixed"]["assemble"] / args.steps
        far_ms32 = G.AssemblyStats()["far_ms"]
        M32 = M.t.clone()
        G.SetFarPrecision(64)
        G.AssembleBEMMatrix(omega, M)
        mdiff = float(torch.linalg.norm(M32 - M.t) / torch.linalg.norm(M.t))
        del M32
        phases = saved
        far32 = {"s_per_freq": t32, "speedup_vs_fp64": t_fp64 / t32, "assemble_s": asm32, "far_ms": far_ms32,
                 "far_ms_fp64": asm_stats["far_ms"], "matrix_rel_fro_vs_fp64": mdiff,
                 "power_rel_err_vs_fp64": max(abs(P32[k] - Pd[k]) / abs(Pd[k]) for k in ("PAbs", "PExt", "PScat")),
                 "refine_iters": it32, "refine_relres": rel32}

    sweep = None
    if not args.no_sweep:
        print("# batched sweep leg", file=sys.stderr, flush=True)
        sweep = sweep_leg()

    cpu = None
    if not args.no_cpu:
        print("# cpu baseline (oracle sample)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.level, omega, eps, args.cpu_rows, args.cpu_lu_n)

    W_lu = 8.0 / 3.0 * N ** 3
    line = {
        "metric": "N=15360 assemble+LU+solve s/freq, mixed vs FP64, at <=1e-8 power rel err",
        "value": t_mixed, "unit": "s/freq", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128 assembly + c64 LU + c128 refinement",
        "data": "synthetic: icosphere mesh generated in-process",
        "config": {"workload": f"dielectric sphere eps=12+0.01i R=1um omega={omega}, icosphere level "
                               f"{args.level}, PMCHWT N={N}, x-polarized plane wave along +z",
                   "N": N, "parallelism": "single GPU"},
        "fp64_s_per_freq": t_fp64, "speedup_vs_fp64": t_fp64 / t_mixed,
        "power_rel_err_vs_fp64": perr, "refine_iters": iters, "refine_relres": relres,
        "power_mixed": Pm, "power_fp64": Pd,
        "fippi_build_s": t_fippi, "assembly_ms": {k: asm_stats[k] for k in
                                                  ("self_ms", "far_ms", "near_ms", "stamp_ms", "mirror_ms")},
        "phase_s_per_freq": {p: {k: v / args.steps for k, v in d.items()} for p, d in phases.items()},
        "lu_tflops": {"c64": W_lu / (phases["mixed"]["lu_c64"] / args.steps) / 1e12,
                      "c128": W_lu / (phases["fp64"]["lu_c128"] / args.steps) / 1e12},
        "kernels_mixed": kern_summary(prof_mixed), "kernels_fp64": kern_summary(prof_fp64),
        "roofline": roofline, "cpu_baseline": cpu, "sweep": sweep, "far32": far32,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
