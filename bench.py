"""Benchmark of the MI355X pseudo-transient Newton-Krylov path.

Workload (the configuration the metric is quoted on, one GPU): the 144x48x48
Kuhn lattice, seed 1 (331,776 vertices, 2,258,399 edges), second-order
residual, plain FP64 GPU path: block ILU(0), GMRES(30) with modified
Gram-Schmidt, CFL ramp from 10, stop at a 1e-10 drop of the nonlinear
residual.  One "step" is one complete solve from the free stream; ``value``
is its wall-clock in seconds (mesh already resident in HBM).

The roofline object is for the kernel class that takes the most time in
the timed solves (HIP events on the context stream, kf_profile); its
algorithmic bytes per launch are the compulsory traffic of that kernel
(formulas in ALG_BYTES).  cpu_baseline runs the reference's own CPU path
(oracle/_ref/kref, built from the reference's sources) on the host cores
for a bounded sample -- per-call times of ComputeFlux, ComputeA,
ComputeNumericalILU, ComputeSparseTriangularSolve and one BLAS-1 pass on the
same mesh -- and extrapolates the solve with the GPU run's call counts.

mixed_precision repeats the solve with the FP32 and BF16-factor
preconditioners; learned_preconditioner runs flexible GMRES with ILU(0)
alone and with the learned correction trained online on the same mesh
(collection and training timed separately).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--prec fp64|fp32|bf16]
                    [--no-cpu] [--no-mixed] [--no-learned] [--learn-rhs R] [--learn-iters T]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "wall-clock to 1e-10 residual drop, mixed/learned vs plain FP64 GPU; flux GF/s"
LATTICE = (144, 48, 48)
SEED = 1
CONFIGS = {"c2": ((144, 48, 48), 1), "c3": ((216, 72, 72), 2), "c5": ((288, 96, 96), 4)}
HBM_PEAK_GBS = 8000.0
FP64_PEAK_GFLOPS = 78600.0


def alg_bytes(kind, N, E, S, Nf, nnzb, s):
    """Compulsory HBM bytes of one launch (each array moved once)."""
    return {
        # idx 8E, normal+area 32E, LSQ terms 48E; xyz 24N, q 32N, r 32N, grad written + read 2*96N
        "residual": 88 * E + 280 * N,
        "grad": 56 * E + 128 * N,                 # nbr 8E + terms 48E; q 32N, grad 96N
        "edge_flux": 72 * E + 152 * N,            # ends 8E, normal 32E, flux out 32E; xyz 24N, q 32N, grad 96N
        "residual_gather": 40 * E + 32 * N,       # ent 8E, flux in 32E; r 32N
        "timestep": 40 * E + 40 * N,
        "fill": 40 * E + 40 * N + 16 * s * nnzb,
        "ilu": 2 * 16 * s * nnzb + 4 * nnzb,
        "trsv_lower": E * (16 * s + 4) + 64 * N,  # lower blocks + cols; b in, x out
        "trsv_upper": (E + N) * (16 * s + 4) + 64 * N,
    }[kind]


# kernel classes of kf_profile -> the FP64 kernel names of rocprofv3 traces
PROFILE_TO_KERNEL = {"trsv_upper": "k_trsv_upper<double, double>", "trsv_lower": "k_trsv_lower<double, double>",
                     "ilu": "k_ilu_rows<double, true>"}


def measured_traffic(dom, prec, config="c2"):
    """HBM bytes per launch of the dominant kernel from the committed counter
    passes (profiles/traffic_c2_fp64.json: FETCH_SIZE and WRITE_SIZE runs,
    scaled by the reading of a copy of known size); None when not covered."""
    path = os.path.join(ROOT, "profiles", "traffic_c2_fp64.json")
    if config != "c2" or prec != "fp64" or dom not in PROFILE_TO_KERNEL or not os.path.exists(path):
        return None
    rec = json.load(open(path))["kernels"].get(PROFILE_TO_KERNEL[dom])
    return None if rec is None else rec["bytes"]


PROFILE_TO_KIND = {"grad": "grad", "edge_flux": "edge_flux", "residual": "residual_gather", "timestep": "timestep",
                   "fill": "fill", "ilu": "ilu", "trsv_lower": "trsv_lower", "trsv_upper": "trsv_upper"}


def cpu_baseline(mesh_path, sizes, counts, host_threads):
    kref = os.path.join(ROOT, "oracle", "_ref", "kref")
    if not os.path.exists(kref):
        return None
    env = dict(os.environ, OMP_NUM_THREADS=str(host_threads), OMP_PROC_BIND="close")
    reps = 100   # a few seconds of CPU work on 16 threads
    t0 = time.time()
    out = subprocess.run([kref, "time", mesh_path, str(host_threads), str(reps)], stdout=subprocess.PIPE, text=True,
                         env=env, timeout=900, check=True).stdout
    wall = time.time() - t0
    t = json.loads(out.strip().splitlines()[-1])
    est = (counts["flux"] * t["flux_s"] + counts["trsv"] * t["sptrsv_s"] + counts["fill"] * (t["fill_s"] + t["ilu_s"]) +
           counts["blas_passes"] * t["blas_pass_s"])
    return {"value": est, "unit": "s", "cores": host_threads, "kind": "reference",
            "sample": ("reference src/*.c + matrix/*.cpp (gcc -O3 -fopenmp, no FMA) on the same %dx%dx%d mesh, %d "
                       "threads, contiguous owner partition: per-call ComputeFlux %.4f s, ComputeSparseTriangularSolve "
                       "%.4f s, ComputeA %.4f s, ComputeNumericalILU %.4f s, one BLAS-1 pass %.5f s (%d reps, %.1f s "
                       "wall); value = those times x the GPU solve's call counts (%d residuals, %d solves, %d "
                       "fills, %d BLAS passes)" % (sizes + (host_threads, t["flux_s"], t["sptrsv_s"], t["fill_s"],
                                                           t["ilu_s"], t["blas_pass_s"], reps, wall, counts["flux"],
                                                           counts["trsv"], counts["fill"], counts["blas_passes"])))}


def blas_passes(hist, restart):
    """4N-vector passes of the driver's BLAS-1 per solve (MGS: a dot and an
    axpy per basis vector, plus norms, scalings, the FD combine and update)."""
    n = 0
    for h in hist:
        k = h["krylov"]
        cycles = max(1, (k + restart - 1) // restart)
        for j in range(k):
            jj = j % restart
            n += 2 * (jj + 1) + 2 + 4          # MGS dots + axpys, norm + scale, iJacVec norm/perturb/combine
        n += cycles * 3 + k + 4                # restart residual, x += V y, update, |r|, |q|
    return n


def mixed_precision(K, m64, mesh, hist64):
    """The same solve with the FP32 and the BF16-factor preconditioner
    (residual, Arnoldi, least squares and update stay FP64): wall-clock
    (one warm-up solve, one timed), Krylov count and force error against the
    plain FP64 path just timed."""
    import numpy as np
    f64 = np.array(hist64[-1]["forces"])
    k64 = sum(x["krylov"] for x in hist64)
    out = {}
    for prec in ("fp32", "bf16"):
        log("mixed leg: %s" % prec)
        m = K.Mesh(mesh, K.config_options(prec=prec))
        q0 = m.state()
        m.kernel()
        m.set_state(q0)
        m.synchronize()
        t0 = time.perf_counter()
        h = m.kernel()
        m.synchronize()
        t = time.perf_counter() - t0
        assert m.last_error() == 0, "device error word set (%s)" % prec
        f = np.array(h[-1]["forces"])
        out[prec] = {"seconds": t, "speedup_vs_fp64": None, "pseudo_steps": len(h),
                     "krylov_total": sum(x["krylov"] for x in h), "krylov_fp64": k64,
                     "residual_drop": h[0]["fnorm"] / h[-1]["fnorm"],
                     "force_rel_err_vs_fp64": (np.abs(f - f64) / np.abs(f64)).max()}
        m.close()
    return out


def log(msg):
    """a progress line on stderr (stdout carries the one JSON line)"""
    print("bench: %s" % msg, file=sys.stderr, flush=True)


def learned(K, mesh, hist64, rhs_per_step, train_iters, prec="fp64", verbose=0):
    """The learned correction on the same mesh: flexible GMRES with ILU(0)
    alone (the method without the network), then with the network trained
    online -- pairs from the first 10 pseudo-time steps (the Newton
    residual and rhs_per_step Gaussian right-hand sides, each solved to
    1e-10 or 300 iterations), Adam with learning rate 1e-3 for train_iters
    steps -- and used from step 11 on with judged steps and the ILU(0)
    fallback.  Wall-clock with and without the collection and training.
    prec: the preconditioner's precision (the combined form: fp32)."""
    import numpy as np
    f64 = np.array(hist64[-1]["forces"])
    m = K.Mesh(mesh, K.config_options(krylov="fgmres", prec=prec))
    log("learned leg: ILU(0)-FGMRES warm-up")
    q0 = m.state()
    m.kernel()
    m.set_state(q0)
    m.synchronize()
    t0 = time.perf_counter()
    hf = m.kernel()
    m.synchronize()
    t_ilu = time.perf_counter() - t0
    log("learned leg: ILU(0)-FGMRES %.3f s; collecting pairs, training, solving" % t_ilu)
    m.set_state(q0)
    if verbose:
        m.set_options(K.config_options(krylov="fgmres", prec=prec, verbose=1))
    net = K.Network(m, K.init_params(seed=0), nslot=11)
    m.attach(net, collect_steps=10, rhs_per_step=rhs_per_step, pair_rtol=1e-10, pair_maxit=300,
             train_iters=train_iters, lr=1e-3, judge_every=5, holdout=0.1, seed=3)
    m.synchronize()
    t0 = time.perf_counter()
    hl = m.kernel()
    m.synchronize()
    t_l = time.perf_counter() - t0
    st = m.learn_stats()
    log("learned leg: %.3f s" % t_l)
    assert m.last_error() == 0, "device error word set (learned)"
    m.attach(None)
    net.close()
    m.close()
    fl = np.array(hl[-1]["forces"])
    after = [x for x in hl if x["step"] >= 10]
    return {
        "precision": prec,
        "fgmres_ilu": {"seconds": t_ilu, "pseudo_steps": len(hf), "krylov_total": sum(x["krylov"] for x in hf)},
        "learned": {"seconds_total": t_l,
                    "seconds_without_collection_and_training": t_l - (st["pair_ms"] + st["train_ms"]) / 1e3,
                    "pseudo_steps": len(hl), "krylov_total": sum(x["krylov"] for x in hl),
                    "krylov_after_training": sum(x["krylov"] for x in after),
                    "krylov_ilu_same_steps": sum(x["krylov"] for x in hf if x["step"] >= 10),
                    "residual_drop": hl[0]["fnorm"] / hl[-1]["fnorm"],
                    "force_rel_err_vs_fp64": float((np.abs(fl - f64) / np.abs(f64)).max()),
                    "judged_steps": [[x["step"], x["krylov_ilu"], x["krylov_net"]] for x in hl
                                     if x["krylov_ilu"] >= 0 and x["krylov_net"] >= 0],
                    "steps_with_network": st["steps_learned"], "fallback": bool(st["fallback"]),
                    "pairs": st["npairs"], "rhs_per_step": rhs_per_step, "tight_solve_iterations": st["pair_iters"],
                    "collection_s": st["pair_ms"] / 1e3, "training_s": st["train_ms"] / 1e3,
                    "adam_steps": train_iters,
                    "heldout_residual_loss": {"ilu": st["loss_ilu"], "trained": st["loss_after"]}},
    }


COPY_BYTES = 1 << 28


def copy_probe(reps):
    """reps device-to-device copies of COPY_BYTES (the runtime's copy kernel):
    the reference point that scales FETCH_SIZE / WRITE_SIZE to bytes."""
    import ctypes as C
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        hip = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(COPY_BYTES)) == 0
    for _ in range(max(reps, 1)):
        assert hip.hipMemcpy(dst, src, C.c_size_t(COPY_BYTES), 3) == 0   # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"probe": "copy", "bytes_per_copy": COPY_BYTES, "copies": max(reps, 1)}))


def trsv_probe(K, m, reps, warmup):
    """The C2 Jacobian, its ILU(0) and reps triangular solves (x = U^-1 L^-1 b),
    for counter passes on the dominant kernels."""
    import numpy as np
    m.timestep(10.0)
    m.fill()
    m.ilu()
    b = K.DeviceVec(m.sz, np.random.default_rng(0).standard_normal(m.sz))
    x = K.DeviceVec(m.sz)
    for _ in range(warmup):
        m.sptrsv(b, x)
    m.profile(True)
    for _ in range(max(reps, 1)):
        m.sptrsv(b, x)
    m.synchronize()
    prof = {k: m.profile_get(k) for k in ("trsv_lower", "trsv_upper")}
    assert m.last_error() == 0
    print(json.dumps({"probe": "trsv", "solves": max(reps, 1),
                      "kernel_ms": {k: v[0] / v[1] for k, v in prof.items() if v[1]}}))
    m.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--prec", default="fp64", choices=["fp64", "fp32", "bf16"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the FP32/BF16-preconditioner comparison")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-correction comparison")
    ap.add_argument("--learn-rhs", type=int, default=8,
                    help="Gaussian right-hand sides per collecting step (the full recipe: 64)")
    ap.add_argument("--learn-iters", type=int, default=2000, help="Adam steps")
    ap.add_argument("--learn-prec", default="fp64", choices=["fp64", "fp32", "bf16"],
                    help="preconditioner precision of the learned leg (fp32: the combined form)")
    ap.add_argument("--progress", action="store_true",
                    help="also the solver's per-step lines in the learned leg (on stdout, before the JSON line)")
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS),
                    help="lattice: c2 144x48x48 seed 1 (the metric's), c3 216x72x72 seed 2, c5 288x96x96 seed 4")
    ap.add_argument("--kernel", default="solve", choices=["solve", "trsv", "copy"],
                    help="solve: the benchmark; trsv / copy: short runs for hardware-counter passes (the "
                         "triangular solves on the same mesh; a device-to-device copy of known size)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the path's multi-GPU sharding is a design, "
              "not an implementation); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.kernel == "copy":
        return copy_probe(args.steps)
    global LATTICE, SEED
    LATTICE, SEED = CONFIGS[args.config]
    import kfun3d_amd as K
    tmp = tempfile.mkdtemp(prefix="kfbench_")
    mesh = os.path.join(tmp, "%s.msh" % args.config)
    log("writing the %dx%dx%d lattice" % LATTICE)
    K.write_lattice(mesh, *LATTICE, SEED)
    opts = K.config_options(prec=args.prec)
    log("setup")
    m = K.Mesh(mesh, opts)
    q0 = m.state()
    h = m.host_arrays()
    N, E, S, Nf = m.N, m.E, len(h["fc_n0"]), len(h["f_nptr"])
    nnzb = N + 2 * E
    s_bytes = {"fp64": 8, "fp32": 4, "bf16": 2}[args.prec]

    if args.kernel == "trsv":
        return trsv_probe(K, m, args.steps, args.warmup)
    log("warm-up solves")
    for _ in range(args.warmup):
        m.set_state(q0)
        m.kernel()
    m.synchronize()
    log("timed solves")
    times, hist = [], None
    m.profile(True)
    for _ in range(args.steps):
        m.set_state(q0)
        m.synchronize()
        t0 = time.perf_counter()
        hist = m.kernel()
        m.synchronize()
        times.append(time.perf_counter() - t0)
    prof = {k: m.profile_get(k) for k in K.Mesh.PROFILED}
    m.profile(False)
    assert m.last_error() == 0, "device error word set"
    drop = hist[0]["fnorm"] / hist[-1]["fnorm"]
    value = max(times) if len(times) == 1 else sorted(times)[len(times) // 2]

    # dominant kernel and its roofline
    dom = max(prof, key=lambda k: prof[k][0])
    ms, n = prof[dom]
    kind = PROFILE_TO_KIND[dom]
    s_mat = 8 if kind not in ("fill", "ilu", "trsv_lower", "trsv_upper") else s_bytes
    ab = alg_bytes(kind, N, E, S, Nf, nnzb, s_mat)
    avg_s = ms / n / 1e3
    achieved = ab / avg_s / 1e9
    flux_s = sum(prof[k][0] / max(prof[k][1], 1) for k in ("grad", "edge_flux", "residual")) / 1e3
    flux_flops = 360 * E + 48 * S + 193 * Nf
    counts = {"flux": prof["edge_flux"][1] // max(args.steps, 1), "trsv": prof["trsv_lower"][1] // max(args.steps, 1),
              "fill": prof["fill"][1] // max(args.steps, 1), "blas_passes": blas_passes(hist, opts.gmres_restart)}
    host_threads = min(16, len(os.sched_getaffinity(0)))
    log("solve %.3f s" % max(times))
    cpu = None if args.no_cpu else cpu_baseline(mesh, LATTICE, counts, host_threads)
    mixed = None if (args.no_mixed or args.prec != "fp64") else mixed_precision(K, m, mesh, hist)
    learn = None if (args.no_learned or args.prec != "fp64") else learned(K, mesh, hist, args.learn_rhs,
                                                                           args.learn_iters, args.learn_prec,
                                                                           1 if args.progress else 0)

    out = {
        "metric": METRIC, "value": value, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": value * 1e3, "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f64" if args.prec == "fp64" else "f64 residual/Arnoldi, %s preconditioner" % args.prec,
        "data": "synthetic",
        "config": {"workload": "%dx%dx%d Kuhn lattice seed %d (%d vertices), second order, %s preconditioner, "
                               "GMRES(30) MGS, CFL ramp from 10, solve from free stream to a 1e-10 residual drop"
                               % (LATTICE + (SEED, N, args.prec)),
                   "lattice": list(LATTICE), "seed": SEED, "vertices": N, "edges": E},
        "pseudo_steps": len(hist), "krylov_total": sum(x["krylov"] for x in hist), "residual_drop": drop,
        "flux_gflops": flux_flops / flux_s / 1e9, "flux_fp64_frac": flux_flops / flux_s / 1e9 / FP64_PEAK_GFLOPS,
        "kernel_ms": {k: (v[0] / v[1] if v[1] else None) for k, v in prof.items()},
        "kernel_share": {k: v[0] / sum(x[0] for x in prof.values()) for k, v in prof.items()},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom, args.prec, args.config),
                     "alg_bytes_per_launch": ab, "avg_launch_ms": ms / n,
                     "traffic_source": "profiles/traffic_c2_fp64.md (FETCH_SIZE/WRITE_SIZE passes, copy-calibrated)"},
        "cpu_baseline": cpu,
        "mixed_precision": mixed,
        "learned_preconditioner": learn,
    }
    if mixed:
        for v in mixed.values():
            v["speedup_vs_fp64"] = value / v["seconds"]
    print(json.dumps(out))
    m.close()


if __name__ == "__main__":
    main()
