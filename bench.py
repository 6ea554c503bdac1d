"""Benchmark of the MI355X pseudo-transient Newton-Krylov path.

Workload (the headline configuration, one GPU): the 288x96x96 Kuhn lattice,
seed 4 (2,654,208 vertices, 18,322,367 edges), second-order residual,
GMRES(30) with modified Gram-Schmidt, CFL ramp from 10, stop at a 1e-10 drop
of the nonlinear residual.  One "step" is one complete solve from the free
stream; ``value`` is its wall-clock in seconds (mesh already resident in
HBM) for the fastest path whose final forces are within 1e-8 of the plain
FP64 path's: the candidates keep the residual, Arnoldi, least squares and
update in FP64 and change only the preconditioner -- its Jacobian fill,
ILU(0) factors and triangular solves in FP32 (or with the factors stored in
BF16), Jacobi sweeps instead of the exact triangular solves, or the
multilevel form (Galerkin coarse corrections over 2x2x2 vertex aggregates,
applied recursively, before the fine ILU(0) solve; path names
<prec>[_jacobi<K>][_coarse<L>[j<J>]], see parse_path).  The plain FP64 path,
the reference's algorithm, is timed beside it on the same mesh and context
(``paths``, ``speedup_vs_fp64``), with the force error of each path against
it.

roofline: the kernel class that takes the most time in one profiled solve of
the value path (HIP events on the context stream around every class,
kf_profile; that solve also gives kernel_ms and kernel_share); achieved =
its algorithmic bytes per launch (ALG_BYTES) / its average launch time over
the timed solves, which record events around that class only
(kf_profile_only); traffic = the HBM
bytes per launch from the committed FETCH_SIZE / WRITE_SIZE counter passes
(profiles/traffic_<config>.json), used only when the profile was taken on
the same kernel (its exact name) built from the same sources (hash of the
kernel source files), else null.

cpu_baseline: the reference's own Kernel() (oracle/_ref/kref, built from the
reference's sources) in its reference profile on the 48x16x16 lattice of
configuration 1, timed on all physical host cores and on one thread, next to
the GPU's reference-profile solve of the same mesh.

Also reported: the HBM footprint (hipMemGetInfo deltas), host setup time,
the BSR product's HBM rate (roofline_spmv), the 144x48x48 lattice's solve in
four modes (default, CGS2, flexible GMRES, the first-order BSR operator),
the cost of exact parity (exact_parity_cost: the first pseudo-time steps of
that lattice with index-order reductions against the default tree), and
learned_preconditioner -- configuration 4's recipe on the
144x48x48 lattice (seed 1): flexible GMRES with ILU(0) alone against ILU(0)
plus the learned correction trained online on pairs from the first 10
pseudo-time steps (the Newton residual and 64 seed-3 Gaussian right-hand
sides each, solved to 1e-10), 2000 Adam steps at learning rate 1e-3.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c5|c3|c2]
                    [--no-cpu] [--no-mixed] [--no-coarse] [--coarse-paths P,..] [--no-learned] [--no-seq]
                    [--jacobi K,..] [--jacobi-fp32] [--learn-rhs R] [--learn-iters T] [--learn-config c2|c5]
                    [--kernel solve|step|solve1|copy|traffic]
"""
import argparse
import ctypes as C
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "wall-clock to 1e-10 residual drop, mixed/learned vs plain FP64 GPU; flux GF/s"
CONFIGS = {"c1": ((48, 16, 16), 0), "c2": ((144, 48, 48), 1), "c3": ((216, 72, 72), 2), "c5": ((288, 96, 96), 4)}
HBM_PEAK_GBS = 8000.0
FP64_PEAK_GFLOPS = 78600.0
# the kernel sources a traffic profile is valid for (hashed into the profile)
KERNEL_SOURCES = ["kfun3d_amd/csrc/" + f for f in ("flux.hip", "fill.hip", "ilu.hip", "sptrsv.hip", "blas.hip",
                                                   "spmv.hip", "jacobi.hip", "handoff.h", "sync.h", "kf_dev.h")]


def alg_bytes(kind, N, E, S, Nf, nnzb, s):
    """Compulsory HBM bytes of one launch (each array moved once); s = bytes
    per stored matrix value."""
    return {
        # idx 8E, normal+area 32E, LSQ terms 48E; xyz 24N, q 32N, r 32N, grad written + read 2*96N
        "residual": 88 * E + 280 * N,
        "grad": 56 * E + 128 * N,                 # nbr 8E + terms 48E; q 32N, grad 96N
        # terms from the moments (k_grad<true>): nbr | side 8E; q 32N, xyz 32N, moments 56N, grad 96N
        "grad_mom": 8 * E + 216 * N,
        "edge_flux": 72 * E + 152 * N,            # ends 8E, normal 32E, flux out 32E; xyz 24N, q 32N, grad 96N
        "residual_gather": 40 * E + 32 * N,       # ent 8E, flux in 32E; r 32N
        # with iJacVec's combination (k_residual<false, true>): + base residual, state, base state 96N, cdt 8N
        "residual_fd": 40 * E + 136 * N,
        "timestep": 40 * E + 40 * N,
        "fill": 40 * E + 40 * N + 16 * s * nnzb,
        "ilu": 2 * 16 * s * nnzb + 4 * nnzb,
        "trsv_lower": E * (16 * s + 4) + 64 * N,  # lower blocks + cols; b in, x out
        "trsv_upper": (E + N) * (16 * s + 4) + 64 * N,
        "spmv": (N + 2 * E) * (16 * s + 4) + 64 * N,   # every block + col; x in, y out
        # one Jacobi sweep (per launch), the compulsory bytes: the half's blocks + cols, row bounds and
        # diagonal index, b (FP64) or y in, the new vector out, vectors in the sweep's precision (8 bytes
        # FP64, 4 otherwise); the old vector's gathers, the fused D^-1 y of the last forward sweep and
        # the FP64 store of the last backward one are not counted (a lower bound)
        "jacobi_lower": E * (16 * s + 4) + 8 * N + 4 * N * (8 + (8 if s == 8 else 4)),
        "jacobi_upper": (E + N) * (16 * s + 4) + 8 * N + 4 * N * 2 * (8 if s == 8 else 4),
    }[kind]


PROFILE_TO_KIND = {"grad": "grad_mom", "edge_flux": "edge_flux", "residual": "residual_fd", "timestep": "timestep",
                   "fill": "fill", "ilu": "ilu", "trsv_lower": "trsv_lower", "trsv_upper": "trsv_upper",
                   "spmv": "spmv", "jacobi_lower": "jacobi_lower", "jacobi_upper": "jacobi_upper"}
MATRIX_KINDS = ("fill", "ilu", "trsv_lower", "trsv_upper", "spmv", "jacobi_lower", "jacobi_upper")
TYPE = {"fp64": ("double", "double"), "fp32": ("float", "float"), "bf16": ("unsigned short", "float")}


def jacobi_names(dom, prec, sweeps, held, multilevel=False):
    """the rocprofv3 names of the launches of one Jacobi solve's forward
    (jacobi_lower) or backward (jacobi_upper) sweeps over the sliced factors,
    in order: the first forward sweep reads b, FP64 (or, as the fine solve of
    the multilevel preconditioner, in the sweeps' own type), the last writes
    D^-1 y too; the last backward sweep writes FP64"""
    tm, tv = TYPE[prec]
    tb = tv if multilevel else "double"
    if dom == "jacobi_lower":
        return ["k_jac_lower_sl<%d, %s, %s, %s, %s, %s>" % (held, tm, tv, tb, tb if i == 0 else tv,
                                                             "true" if i == sweeps - 1 else "false")
                for i in range(sweeps)]
    return ["k_jac_upper_sl<%d, %s, %s, %s>" % (held, tm, tv, "double" if i == sweeps - 1 else tv)
            for i in range(sweeps)]


def kernel_name(dom, prec, held, fold):
    """the rocprofv3 name of the kernel behind a kf_profile class (for the
    forward sweep, the instantiation GMRES's inner steps launch)"""
    tm, tv = TYPE[prec]
    if dom == "trsv_upper":
        return "k_trsv_upper<%d, %s, %s>" % (held, tm, tv)
    if dom == "trsv_lower":
        return "k_trsv_lower<%d, %s, %s, %s>" % (held, tm, tv, "true" if fold else "false")
    if dom == "ilu":
        return "k_ilu_rows<%s, %s>" % (("double", "true") if prec == "fp64" else ("float", "false"))
    if dom == "fill":
        return "k_fill<%s>" % ("double" if prec == "fp64" else "float")
    if dom == "spmv":
        return "k_bsr_spmv<%s>" % ("double" if prec == "fp64" else "float")
    if dom == "jacobi_lower":   # one sweep: reads b, writes D^-1 y too
        return "k_jac_lower_sl<%d, %s, %s, double, double, true>" % (held, tm, tv)
    if dom == "jacobi_upper":
        return "k_jac_upper_sl<%d, %s, %s, double>" % (held, tm, tv)
    return {"grad": "k_grad<true>", "edge_flux": "k_edge_flux<false>", "residual": "k_residual<false, true>",
            "timestep": "k_timestep"}[dom]


def source_hash():
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def measured_traffic(config, name):
    """HBM bytes per launch of one kernel from the committed counter passes
    (profiles/traffic_<config>.json), or None when the profile does not cover
    this exact kernel built from the current sources.  name may be a list of
    launches (a Jacobi solve's sweeps): their mean."""
    if isinstance(name, list):
        got = [measured_traffic(config, x) for x in name]
        if any(b is None for b, _ in got):
            return None, "; ".join(src for b, src in got if b is None)
        return sum(b for b, _ in got) / len(got), got[0][1]
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % config)
    if not os.path.exists(path):
        return None, "no profiles/traffic_%s.json" % config
    prof = json.load(open(path))
    if prof.get("source_hash") != source_hash():
        return None, "profiles/traffic_%s.json was taken on other kernel sources" % config
    rec = prof["kernels"].get(name)
    if rec is None:
        return None, "profiles/traffic_%s.json has no %s" % (config, name)
    return rec["bytes"], "profiles/traffic_%s.md (FETCH_SIZE/WRITE_SIZE passes, copy-calibrated)" % config


def held_entries(m):
    """entries per row the triangular solves hold in registers: 7 when no row
    has more than 7 blocks on either side of its diagonal, else 8"""
    import numpy as np
    rp = np.empty(m.N + 1, dtype=np.uint32)
    dg = np.empty(m.N, dtype=np.uint32)
    import kfun3d_amd as K
    K.lib().kf_pattern(m.ref, rp.ctypes.data, None, dg.ctypes.data)
    below = (dg - rp[:-1]).max()
    above = (rp[1:] - dg - 1).max()
    return 7 if max(below, above) <= 7 else 8


def log(msg):
    """a progress line on stderr (stdout carries the one JSON line)"""
    print("bench: %s" % msg, file=sys.stderr, flush=True)


def _hip():
    try:
        return C.CDLL("libamdhip64.so")
    except OSError:
        return C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))


def mem_used():
    """device memory in use (bytes), from hipMemGetInfo"""
    hip = _hip()
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return total.value - free.value, total.value


def lattice_mesh(K, tmp, config):
    dims, seed = CONFIGS[config]
    path = os.path.join(tmp, "%s.msh" % config)
    if not os.path.exists(path):
        log("writing the %dx%dx%d lattice (seed %d)" % (dims + (seed,)))
        K.write_lattice(path, *dims, seed)
    return path


def parse_path(path):
    """a path name: <prec>[_jacobi<K>][_coarse[L][j<J>]][_lag[R]] -- the
    preconditioner's precision (fp64, fp32, bf16), K Jacobi sweeps each way
    instead of the fine level's exact triangular solves, the multilevel form
    with L coarse levels (1 when L is omitted: two-level) whose ILU(0)
    solves take J Jacobi sweeps each way (exact when omitted), and _lag[R]:
    the preconditioner kept across pseudo-steps while the CFL stays below R
    times the CFL it was factored at (kf_options.lag_factors = R, 1 when
    omitted: only while unchanged); returns (prec, K, L, J) -- path_lag()
    reads R"""
    path = path.rpartition("_lag")[0] if path_lag(path) else path
    base, _, lv = path.partition("_coarse")
    lv, _, cj = lv.partition("j")
    coarse = (int(lv) if lv else 1) if "_coarse" in path else 0
    prec, _, jac = base.partition("_jacobi")
    return prec, int(jac) if jac else 0, coarse, int(cj) if cj else 0


def path_lag(path):
    """R of a _lag[R] suffix, 0 without one"""
    head, sep, tail = path.rpartition("_lag")
    if not sep or (tail and not tail.isdigit()):
        return 0
    return int(tail) if tail else 1


def path_options(K, path, **kw):
    """configuration-profile options of a path name (parse_path)"""
    prec, jac, coarse, cj = parse_path(path)
    return K.config_options(prec=prec, jacobi_sweeps=jac, coarse=coarse, coarse_sweeps=cj,
                            lag_factors=path_lag(path), **kw)


def path_precond(path):
    prec, jac, coarse, cj = parse_path(path)
    s = "%s block ILU(0)" % prec
    if jac:
        s += " with %d Jacobi sweeps each way" % jac
    if coarse:
        s += (", %d-level: a Galerkin coarse correction over 2x2x2 vertex aggregates before the fine solve, "
              "%s, %s, coarse arithmetic in %s"
              % (coarse + 1, "the coarse system by block ILU(0)" if coarse == 1 else
                 "each coarse system by the same scheme, the coarsest by block ILU(0)",
                 "coarse ILU(0) solves by %d Jacobi sweeps each way" % cj if cj else "exact coarse ILU(0) solves",
                 {"fp64": "fp64", "fp32": "fp32", "bf16": "fp32 on bf16 factors"}[prec]))
    if path_lag(path) == 1:
        s += ", kept across pseudo-steps once the CFL ramp reaches its cap"
    elif path_lag(path):
        s += ", kept across pseudo-steps while the CFL stays below %d times the CFL it was formed at" % path_lag(path)
    return s


def path_dtype(path):
    prec = parse_path(path)[0]
    return {"fp64": "f64", "fp32": "f64 residual/Arnoldi/update, f32 preconditioner",
            "bf16": "f64 residual/Arnoldi/update, f32 preconditioner arithmetic on bf16 factors"}[prec]


def timed_solve(m, q0):
    m.set_state(q0)
    m.synchronize()
    t0 = time.perf_counter()
    h = m.kernel()
    m.synchronize()
    return time.perf_counter() - t0, h


def summary(h, t, f_ref=None):
    import numpy as np
    out = {"seconds": t, "pseudo_steps": len(h), "krylov_total": sum(x["krylov"] for x in h),
           "residual_drop": h[0]["fnorm"] / h[-1]["fnorm"], "forces": h[-1]["forces"]}
    if f_ref is not None:
        f, f0 = np.array(h[-1]["forces"]), np.array(f_ref)
        out["force_rel_err_vs_fp64"] = float((np.abs(f - f0) / np.abs(f0)).max())
    return out


def roofline(prof, path, sizes, config, held, fold):
    """the dominant kernel class of a profile and its roofline object.  A
    Jacobi class times all sweeps of one solve as one event pair: its launch
    average divides by the sweep count, and its kernel is the list of the
    sweeps' instantiations.  The multilevel preconditioner's coarse chain
    (class "coarse": many small latency-bound launches) is not a kernel and
    is not a candidate."""
    prec, sweeps = parse_path(path)[:2]
    N, E, S, Nf = sizes
    nnzb = N + 2 * E
    dom = max((k for k in prof if k in PROFILE_TO_KIND), key=lambda k: prof[k][0])
    ms, n = prof[dom]
    kind = PROFILE_TO_KIND[dom]
    per = sweeps if kind in ("jacobi_lower", "jacobi_upper") else 1
    s = {"fp64": 8, "fp32": 4, "bf16": 2}[prec] if kind in MATRIX_KINDS else 8
    ab = alg_bytes(kind, N, E, S, Nf, nnzb, s)
    achieved = ab / (ms / (n * per) / 1e3) / 1e9
    name = jacobi_names(dom, prec, sweeps, held, parse_path(path)[2] > 0) if per > 1 else \
        kernel_name(dom, prec, held, fold)
    traffic, src = measured_traffic(config, name)
    return {"kernel": name, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "alg_bytes_per_launch": ab,
            "avg_launch_ms": ms / (n * per), "launches": n * per, "traffic_source": src}


def spmv_rate(K, m, q0, sizes, config, reps=20):
    """the BSR product y = A x (the first-order operator's kernel) with the
    FP64 Jacobian of the free stream at CFL 10 on the same mesh: reps
    profiled launches after three unprofiled ones"""
    import numpy as np
    m.set_options(K.config_options(prec="fp64", bsr_operator=1))
    m.set_state(q0)
    m.timestep(10.0)
    m.fill()
    xs = K.DeviceVec(m.sz, np.random.default_rng(0).standard_normal(m.sz))
    ys = K.DeviceVec(m.sz)
    for _ in range(3):
        m.spmv(xs, ys)
    m.profile(True)
    for _ in range(reps):
        m.spmv(xs, ys)
    prof = {"spmv": m.profile_get("spmv")}
    m.profile(False)
    assert m.last_error() == 0
    xs.free()
    ys.free()
    m.set_options(K.config_options(prec="fp64"))
    out = roofline(prof, "fp64", sizes, config, 7, False)
    log("BSR SpMV %.3f ms, %.0f GB/s" % (out["avg_launch_ms"], out["achieved"]))
    return out


def physical_cores():
    """(physical cores this process may run on, the host's physical cores by
    lscpu, CPUs in the affinity mask): the first counts distinct (package,
    core) pairs of the CPUs in the affinity mask (sysfs topology)"""
    cpus = sorted(os.sched_getaffinity(0))
    host = None
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True, timeout=30).stdout
        kv = {l.split(":", 1)[0].strip(): l.split(":", 1)[1].strip() for l in out.splitlines() if ":" in l}
        host = int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
    except Exception:
        pass
    try:
        cores = set()
        for c in cpus:
            base = "/sys/devices/system/cpu/cpu%d/topology/" % c
            cores.add((open(base + "physical_package_id").read().strip(), open(base + "core_id").read().strip()))
        return len(cores), host, len(cpus)
    except OSError:
        return len(cpus), host, len(cpus)


def cpu_baseline(K, tmp):
    """The reference's own Kernel() (reference profile: alpha 3, CFL 50 to
    1e5, at most 15 pseudo-time steps, stop at a 1e7 drop) on the 48x16x16
    lattice, on all physical cores and on one thread, beside the GPU's
    solve of the same mesh in the same profile."""
    kref = os.path.join(ROOT, "oracle", "_ref", "kref")
    if not os.path.exists(kref):
        return None
    mesh = lattice_mesh(K, tmp, "c1")
    cores, phys, aff = physical_cores()
    res = {}
    for T in (cores, 1):
        out = os.path.join(tmp, "kref_%d" % T)
        env = dict(os.environ, OMP_NUM_THREADS=str(T), OMP_PROC_BIND="close")
        log("cpu baseline: reference Kernel() on the 48x16x16 lattice, %d thread(s)" % T)
        subprocess.run([kref, "solve", mesh, out, str(T)], stdout=subprocess.DEVNULL, env=env, timeout=900,
                       check=True)
        lines = open(os.path.join(out, "history.txt")).read().split("\n")
        sec = [float(l.split()[1]) for l in lines if l.startswith("seconds")][0]
        forces = [[float(x) for x in l.split()[1:]] for l in lines if l.startswith("forces")]
        res[T] = (sec, len(forces), forces[-1])
    m = K.Mesh(mesh, K.options())
    q0 = m.state()
    m.kernel()
    t_gpu, h = timed_solve(m, q0)
    m.close()
    sec, steps, f = res[cores]
    return {"value": sec, "unit": "s", "cores": cores, "kind": "reference",
            "sample": ("the reference's Kernel() (src/kernel.c:47-232; src/*.c and matrix/*.cpp built with gcc -O3 "
                       "-fopenmp, no FMA, contiguous owner partition) in its reference profile (alpha 3, CFL 50 to "
                       "1e5, 15 pseudo-time steps at most, stop at a 1e7 drop) on the 48x16x16 lattice of "
                       "configuration 1 (12,288 vertices, seed 0), second order: one whole solve, %d pseudo-time "
                       "steps, on %d threads (physical cores %s, affinity %d CPUs)"
                       % (steps, cores, phys, aff)),
            "one_thread_s": res[1][0], "gpu_same_solve_s": t_gpu, "gpu_pseudo_steps": len(h),
            "gpu_krylov_total": sum(x["krylov"] for x in h), "cpu_forces": f, "gpu_forces": h[-1]["forces"]}


C2_VARIANTS = [("default", {}), ("cgs2", {"ortho": "cgs2"}), ("fgmres", {"krylov": "fgmres"}),
               ("bsr_operator", {"bsr_operator": 1})]


def c2_variants(K, mesh_c2):
    """The configuration-profile solve of the 144x48x48 lattice (FP64) in the
    default mode and with one option changed each: batched classical
    Gram-Schmidt with reorthogonalisation, flexible GMRES (the learned
    correction's frame) and the assembled first-order Jacobian as the
    operator.  One warm-up and one timed solve each, on one context."""
    import numpy as np
    m = K.Mesh(mesh_c2, K.config_options())
    q0 = m.state()
    out, f_ref = {}, None
    for name, kw in C2_VARIANTS:
        m.set_options(K.config_options(**kw))
        timed_solve(m, q0)
        t, h = timed_solve(m, q0)
        assert m.last_error() == 0, name
        f = np.array(h[-1]["forces"])
        if f_ref is None:
            f_ref = f
        out[name] = {"seconds": t, "krylov_total": sum(x["krylov"] for x in h), "pseudo_steps": len(h),
                     "residual_drop": h[0]["fnorm"] / h[-1]["fnorm"],
                     "force_rel_err_vs_default": float((np.abs(f - f_ref) / np.abs(f_ref)).max())}
        log("144x48x48 %s: %.3f s, %d Krylov, %d steps" % (name, t, out[name]["krylov_total"], len(h)))
    m.close()
    out["workload"] = "144x48x48 lattice seed 1, FP64 preconditioner, configuration profile, one solve to a 1e-10 drop"
    return out


def exact_parity_cost(K, mesh_c2, ts=2):
    """The cost of bit-exact parity: the first `ts` pseudo-time steps of the
    144x48x48 lattice's configuration-profile solve with the default
    fixed-order tree reductions, and with index-order reductions (reduce_seq:
    every dot and norm a sequential sum over the 1.3M entries in one wave, as
    the reference's one-thread loops add them; the mode whose whole history
    equals the reference's bit for bit).  A whole solve in the index-order
    mode takes about a minute on this lattice, so the comparison runs on its
    first steps.  One warm-up and one timed run each."""
    m = K.Mesh(mesh_c2, K.config_options(ts_max=ts))
    q0 = m.state()
    out = {}
    for name, kw in (("tree", {}), ("index_order", {"reduce_seq": 1})):
        m.set_options(K.config_options(ts_max=ts, **kw))
        timed_solve(m, q0)
        t, h = timed_solve(m, q0)
        assert m.last_error() == 0, name
        out[name] = {"seconds": t, "krylov_total": sum(x["krylov"] for x in h), "pseudo_steps": len(h),
                     "fnorm": [x["fnorm"] for x in h]}
        log("exact-parity cost, first %d pseudo-steps of 144x48x48, %s reductions: %.3f s, %d Krylov"
            % (ts, name, t, out[name]["krylov_total"]))
    m.close()
    out["ratio"] = out["index_order"]["seconds"] / out["tree"]["seconds"]
    out["workload"] = ("144x48x48 lattice seed 1, FP64, configuration profile, the first %d pseudo-time steps "
                       "(same Krylov counts in both modes)" % ts)
    return out


class stdout_to_stderr:
    """file descriptor 1 pointed at descriptor 2 for a block (the library's
    printf lines land on stderr)"""

    def __enter__(self):
        sys.stdout.flush()
        C.CDLL(None).fflush(None)
        self.saved = os.dup(1)
        os.dup2(2, 1)

    def __exit__(self, *exc):
        sys.stdout.flush()
        C.CDLL(None).fflush(None)
        os.dup2(self.saved, 1)
        os.close(self.saved)
        return False


def learned(K, mesh, rhs_per_step, train_iters, prec="fp64"):
    """The learned correction on one mesh: flexible GMRES with the path's
    preconditioner alone (prec, a path name; fp64 = ILU(0), the method
    without the network), then with the network trained online
    -- pairs from the first 10 pseudo-time steps (the Newton residual and
    rhs_per_step Gaussian right-hand sides, each solved to 1e-10 or 300
    iterations), Adam with learning rate 1e-3 for train_iters steps -- and
    used from step 11 on with judged steps and the ILU(0) fallback.
    Wall-clock with and without the collection and training."""
    import numpy as np
    m = K.Mesh(mesh, path_options(K, prec, krylov="fgmres"))
    log("learned leg: ILU(0)-FGMRES warm-up")
    q0 = m.state()
    m.kernel()
    t_ilu, hf = timed_solve(m, q0)
    log("learned leg: ILU(0)-FGMRES %.3f s; collecting pairs (%d per step), training (%d Adam steps), solving"
        % (t_ilu, rhs_per_step + 1, train_iters))
    m.set_state(q0)
    # the solver's per-step lines (and its collection and training lines) are
    # the leg's progress: printed, but on stderr (stdout carries the JSON line)
    m.set_options(path_options(K, prec, krylov="fgmres", verbose=1))
    net = K.Network(m, K.init_params(seed=0), nslot=11)
    m.attach(net, collect_steps=10, rhs_per_step=rhs_per_step, pair_rtol=1e-10, pair_maxit=300,
             train_iters=train_iters, lr=1e-3, judge_every=5, holdout=0.1, seed=3)
    m.synchronize()
    with stdout_to_stderr():
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
    ff, fl = np.array(hf[-1]["forces"]), np.array(hl[-1]["forces"])
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
                    "force_rel_err_vs_ilu_fgmres": float((np.abs(fl - ff) / np.abs(ff)).max()),
                    "judged_steps": [[x["step"], x["krylov_ilu"], x["krylov_net"]] for x in hl
                                     if x["krylov_ilu"] >= 0 and x["krylov_net"] >= 0],
                    "steps_with_network": st["steps_learned"], "fallback": bool(st["fallback"]),
                    "pairs": st["npairs"], "rhs_per_step": rhs_per_step, "tight_solve_iterations": st["pair_iters"],
                    "collection_s": st["pair_ms"] / 1e3, "training_s": st["train_ms"] / 1e3,
                    "adam_steps": train_iters,
                    "heldout_residual_loss": {"ilu": st["loss_ilu"], "before": st["loss_before"],
                                              "trained": st["loss_after"]}},
    }


COPY_BYTES = 1 << 28


def copy_probe(reps):
    """reps device-to-device copies of COPY_BYTES (the runtime's copy kernel):
    the reference point that scales FETCH_SIZE / WRITE_SIZE to bytes."""
    hip = _hip()
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


def _kernel_key(name):
    """rocprofv3's kernel name without the return type and the argument list"""
    name = name.strip()
    if name.startswith("void "):
        name = name[5:]
    depth = 0
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return name[:i]
    return name


def _pmc_means(root, counter, finest=False):
    """per kernel: mean value of one counter per dispatch over every
    *counter_collection.csv under root, and the dispatch count.  finest: only
    the dispatches with the kernel's largest grid (a multilevel path launches
    the same kernels on every coarse level; the fine level's grid is the
    largest)"""
    import csv
    import glob
    acc, grid = {}, {}
    files = glob.glob(os.path.join(root, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit("no counter_collection.csv under %s" % root)
    for f in files:
        for row in csv.DictReader(open(f)):
            if row.get("Counter_Name") != counter:
                continue
            k = _kernel_key(row["Kernel_Name"])
            a = acc.setdefault(k, {})
            d = row.get("Dispatch_Id") or row.get("Correlation_Id")
            a[d] = a.get(d, 0.0) + float(row["Counter_Value"])   # summed over the counter's instances
            grid.setdefault(k, {})[d] = int(row.get("Grid_Size") or row.get("Grid_Size_X") or 0)
    out = {}
    for k, v in acc.items():
        keep = list(v)
        if finest:
            top = max(grid[k].values())
            keep = [d for d in v if grid[k][d] == top]
        out[k] = (sum(v[d] for d in keep) / len(keep), len(keep))
    return out


def traffic_profile(K, config, root, precs):
    """profiles/traffic_<config>.json from rocprofv3 --pmc passes: under root,
    fetch_<prec>/ and write_<prec>/ hold FETCH_SIZE and WRITE_SIZE runs of
    `--kernel step --prec <prec>`, fetch_copy/ and write_copy/ those of
    `--kernel copy`.  Each counter is scaled by its reading for the copy
    kernel of COPY_BYTES read and written.  Prints the summary table."""
    import numpy as np
    cf, cw = _pmc_means(os.path.join(root, "fetch_copy"), "FETCH_SIZE"), _pmc_means(os.path.join(root, "write_copy"),
                                                                                   "WRITE_SIZE")
    copy_key = [k for k in cf if "copyBuffer" in k or "copy" in k.lower()]
    if not copy_key:
        raise SystemExit("no copy kernel in the calibration passes: %s" % sorted(cf))
    ck = copy_key[0]
    fscale = COPY_BYTES / (cf[ck][0] * 1024.0)
    wscale = COPY_BYTES / (cw[ck][0] * 1024.0)
    tmp = tempfile.mkdtemp(prefix="kftraffic_")
    m = K.Mesh(lattice_mesh(K, tmp, config), device=False)
    h = m.host_arrays()
    N, E, S, Nf = m.N, m.E, len(h["fc_n0"]), len(h["f_nptr"])
    m.close()
    nnzb = N + 2 * E
    kinds = {"k_grad": "grad", "k_edge_flux": "edge_flux", "k_residual": "residual_gather", "k_fill": "fill",
             "k_ilu_rows": "ilu", "k_trsv_lower": "trsv_lower", "k_trsv_upper": "trsv_upper",
             "k_bsr_spmv": "spmv", "k_timestep": "timestep", "k_jac_lower_sl": "jacobi_lower",
             "k_jac_upper_sl": "jacobi_upper"}
    out = {"workload": "%dx%dx%d Kuhn lattice seed %d: the first pseudo-time step of the configuration-profile "
                       "solve (bench.py --kernel step), per preconditioner precision" % (CONFIGS[config][0] +
                                                                                         (CONFIGS[config][1],)),
           "fetch_scale": fscale, "write_scale": wscale,
           "calibration": "%s of %d bytes: FETCH_SIZE %.1f KiB, WRITE_SIZE %.1f KiB per dispatch"
                          % (ck, COPY_BYTES, cf[ck][0], cw[ck][0]),
           "source_hash": source_hash(), "kernels": {}}
    rows = []
    for prec in precs:
        multi = parse_path(prec)[2] > 0
        f = _pmc_means(os.path.join(root, "fetch_%s" % prec), "FETCH_SIZE", finest=multi)
        w = _pmc_means(os.path.join(root, "write_%s" % prec), "WRITE_SIZE", finest=multi)
        s = {"fp64": 8, "fp32": 4, "bf16": 2}[parse_path(prec)[0]]
        for k in sorted(set(f) & set(w)):
            base = k.split("<")[0]
            if base not in kinds:
                continue
            kind = kinds[base]
            if k == "k_grad<true>":
                kind = "grad_mom"
            if k == "k_residual<false, true>":
                kind = "residual_fd"
            if multi and kind == "ilu":
                continue   # the fine and the coarse factorisations share the persistent grid: not separable
            if multi and kind in ("jacobi_lower", "jacobi_upper"):
                held = int(k.split("<")[1].split(",")[0])
                pp, sw = parse_path(prec)[:2]
                if k not in jacobi_names(kind, pp, sw, held, True):
                    continue   # a coarse level's sweeps (other instantiations of the same kernels)
            ss = s if kind in MATRIX_KINDS else 8
            if base in ("k_fill", "k_ilu_rows") and s != 8:
                ss = 4   # the fill and the factorisation run in FP32 for both single-precision modes
            ab = alg_bytes(kind, N, E, S, Nf, nnzb, ss)
            b = f[k][0] * 1024.0 * fscale + w[k][0] * 1024.0 * wscale
            if k in out["kernels"]:
                continue
            out["kernels"][k] = {"fetch_kib": f[k][0], "write_kib": w[k][0], "bytes": b, "alg_bytes": ab,
                                 "dispatches": min(f[k][1], w[k][1])}
            rows.append("| `%s` | %.1f | %.1f | %.0f | %d | %.2f |" % (k, f[k][0], w[k][0], b, ab, b / ab))
    json.dump(out, open(os.path.join(ROOT, "profiles", "traffic_%s.json" % config), "w"), indent=1)
    print("Calibration: %s; fetched bytes = FETCH_SIZE x %.4f x 1024, written bytes = WRITE_SIZE x %.4f x 1024."
          % (out["calibration"], fscale, wscale))
    print()
    print("| kernel | FETCH_SIZE (KiB) | WRITE_SIZE (KiB) | HBM bytes per launch | algorithmic bytes | ratio |")
    print("|---|---:|---:|---:|---:|---:|")
    for r in rows:
        print(r)


def step_probe(K, mesh, prec, reps, warmup, ts_max):
    """reps solves from the free stream (configuration profile) of ts_max
    pseudo-time steps (1: residual, time step, fill, ILU(0), one GMRES solve,
    update, forces -- every kernel of the path, for hardware-counter passes;
    500: whole solves, for kernel traces), nothing else timed or run"""
    m = K.Mesh(mesh, path_options(K, prec, ts_max=ts_max))
    q0 = m.state()
    for _ in range(warmup):
        timed_solve(m, q0)
    ts, h = [], None
    for _ in range(max(reps, 1)):
        t, h = timed_solve(m, q0)
        ts.append(t)
    assert m.last_error() == 0
    print(json.dumps({"probe": "step" if ts_max == 1 else "solve", "prec": prec, "solves": len(ts), "seconds": ts,
                      "pseudo_steps": len(h), "krylov_total": sum(x["krylov"] for x in h),
                      "source_hash": source_hash()}))
    m.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c5", choices=["c2", "c3", "c5"],
                    help="lattice: c5 288x96x96 seed 4 (the headline), c3 216x72x72 seed 2, c2 144x48x48 seed 1")
    ap.add_argument("--prec", default="auto",
                    help="the path `value` times: fp64, fp32, bf16, or one of those with _jacobiK (K Jacobi "
                         "sweeps each way instead of the exact triangular solves) and/or _coarse (the two-level "
                         "preconditioner); auto = the fastest other path of the warm-up solves whose forces are "
                         "within 1e-8 of the plain FP64 path's")
    ap.add_argument("--jacobi", default="3",
                    help="Jacobi sweep counts of the extra BF16-factor paths (comma list; empty: none)")
    ap.add_argument("--jacobi-fp32", action="store_true", help="also FP32-factor paths with the Jacobi sweeps")
    ap.add_argument("--no-coarse", action="store_true", help="skip the multilevel (coarse-corrected) paths")
    ap.add_argument("--coarse-paths", default="fp64_coarse8,bf16_coarse8,fp64_jacobi2_coarse8j3,bf16_jacobi2_coarse8j3,"
                                              "bf16_jacobi2_coarse8j3_lag4",
                    help="the multilevel paths timed beside the others (comma list of path names)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="time the FP64 path only")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-correction comparison")
    ap.add_argument("--no-seq", action="store_true",
                    help="skip the 144x48x48 legs (CGS2, FGMRES, BSR operator; the exact-parity cost)")
    ap.add_argument("--learn-rhs", type=int, default=64, help="Gaussian right-hand sides per collecting step")
    ap.add_argument("--learn-iters", type=int, default=2000, help="Adam steps")
    ap.add_argument("--learn-config", default="c2", choices=["c2", "c5"],
                    help="mesh of the learned leg: c2 (configuration 4) or c5 (the combined form)")
    ap.add_argument("--learn-prec", default="fp64",
                    help="the learned leg's preconditioner path (fp64: ILU(0), the recipe; fp32: the combined form; "
                         "a multilevel path name puts the network on top of the multilevel preconditioner)")
    ap.add_argument("--traffic-dir", default="gpurun_out/pmc",
                    help="--kernel traffic: the directory holding the counter passes")
    ap.add_argument("--traffic-paths", default="fp64,bf16,bf16_jacobi2_coarse8j3_lag4",
                    help="--kernel traffic: the paths whose passes it reads (fetch_<path>/, write_<path>/)")
    ap.add_argument("--kernel", default="solve", choices=["solve", "step", "solve1", "copy", "traffic"],
                    help="solve: the benchmark; step / solve1 / copy: short runs for profiles (first pseudo-time "
                         "steps, or whole solves, of one precision on the same mesh; a device-to-device copy of "
                         "known size)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the path's multi-GPU sharding is a design, "
              "not an implementation); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.kernel == "copy":
        return copy_probe(args.steps)
    if args.kernel == "traffic":
        import kfun3d_amd as K
        return traffic_profile(K, args.config, os.path.join(ROOT, args.traffic_dir),
                               [x for x in args.traffic_paths.split(",") if x.strip()])
    import numpy as np
    import kfun3d_amd as K
    tmp = tempfile.mkdtemp(prefix="kfbench_")
    mesh = lattice_mesh(K, tmp, args.config)
    if args.kernel in ("step", "solve1"):
        return step_probe(K, mesh, "fp64" if args.prec == "auto" else args.prec, args.steps, args.warmup,
                          1 if args.kernel == "step" else 500)
    dims, seed = CONFIGS[args.config]

    used0, total = mem_used()
    log("setup")
    t0 = time.perf_counter()
    m = K.Mesh(mesh, K.config_options(prec="fp64"))
    m.synchronize()
    setup_s = time.perf_counter() - t0
    q0 = m.state()
    h = m.host_arrays()
    N, E, S, Nf = m.N, m.E, len(h["fc_n0"]), len(h["f_nptr"])
    del h
    held = held_entries(m)
    log("setup %.2f s; FP64 warm-up solve" % setup_s)
    t64w, h64 = timed_solve(m, q0)
    used64, _ = mem_used()
    f64 = h64[-1]["forces"]
    log("FP64 %.3f s" % t64w)

    precs = []
    if not args.no_mixed:
        precs = ["fp32", "bf16"]
        for k in [int(x) for x in args.jacobi.split(",") if x.strip()]:
            precs += (["fp32_jacobi%d" % k] if args.jacobi_fp32 else []) + ["bf16_jacobi%d" % k]
        if not args.no_coarse:
            precs += [x for x in args.coarse_paths.split(",") if x.strip()]
    warm, warm_err = {"fp64": t64w}, {"fp64": 0.0}
    for p in precs:
        m.set_options(path_options(K, p))
        warm[p], hw = timed_solve(m, q0)
        f = np.array(hw[-1]["forces"])
        warm_err[p] = float((np.abs(f - np.array(f64)) / np.abs(np.array(f64))).max())
        log("%s warm-up %.3f s, %d Krylov, forces within %.1e of FP64" % (p, warm[p], sum(x["krylov"] for x in hw),
                                                                          warm_err[p]))
    used_all, _ = mem_used()
    ok = [p for p in precs if warm_err[p] <= 1e-8]
    vprec = args.prec if args.prec != "auto" else (min(ok, key=lambda p: warm[p]) if ok else "fp64")

    # the value path: W more warm-up solves, one solve with every kernel class
    # profiled (kernel_ms, kernel_share and the dominant class), then K timed
    # with HIP events on the dominant class only (its launch time for the
    # roofline, measured over the timed solves)
    m.set_options(path_options(K, vprec))
    for _ in range(args.warmup):
        timed_solve(m, q0)
    m.profile(True)
    timed_solve(m, q0)
    prof = {k: m.profile_get(k) for k in K.Mesh.PROFILED}
    m.profile(False)
    dom = max((k for k in prof if k in PROFILE_TO_KIND), key=lambda k: prof[k][0])
    log("timed solves (%s), events on %s" % (vprec, dom))
    times, hist = [], None
    m.profile_only(dom)
    for i in range(args.steps):
        t, hist = timed_solve(m, q0)
        times.append(t)
        log("solve %d/%d %.3f s" % (i + 1, args.steps, t))
    prof_timed = {dom: m.profile_get(dom)}
    m.profile(False)
    assert m.last_error() == 0, "device error word set"
    value = sum(times) / len(times)
    fold = {vprec: K.lib().kf_trsv_fold(m.ref)}

    # the other paths beside it (FP64 profiled too)
    paths = {vprec: summary(hist, value, f64 if vprec != "fp64" else None)}
    paths[vprec]["timed_solves"] = len(times)
    prof64 = None
    for p in ["fp64"] + [x for x in precs if x != vprec]:
        if p == vprec:
            continue
        m.set_options(path_options(K, p))
        reps = min(args.steps, 3) if p == "fp64" else 1
        if p == "fp64":
            m.profile(True)
            fold["fp64"] = K.lib().kf_trsv_fold(m.ref)
        ts = []
        for _ in range(reps):
            t, hp = timed_solve(m, q0)
            ts.append(t)
        if p == "fp64":
            prof64 = {k: m.profile_get(k) for k in K.Mesh.PROFILED}
            m.profile(False)
        paths[p] = summary(hp, sum(ts) / len(ts), f64 if p != "fp64" else None)
        paths[p]["timed_solves"] = reps
        log("%s %.3f s" % (p, paths[p]["seconds"]))
    assert m.last_error() == 0, "device error word set"
    spmv = spmv_rate(K, m, q0, sizes=(N, E, S, Nf), config=args.config)
    m.close()

    sizes = (N, E, S, Nf)
    flux_s = sum(prof[k][0] / max(prof[k][1], 1) for k in ("grad", "edge_flux", "residual")) / 1e3
    flux_flops = 360 * E + 48 * S + 193 * Nf
    out = {
        "metric": METRIC, "value": value, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": value * 1e3, "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": path_dtype(vprec),
        "data": "synthetic",
        "config": {"workload": "%dx%dx%d Kuhn lattice seed %d (%d vertices, %d edges), second order, "
                               "preconditioner %s, GMRES(30) MGS, CFL ramp from 10, one solve from the "
                               "free stream to a 1e-10 residual drop"
                               % (dims + (seed, N, E, path_precond(vprec))),
                   "lattice": list(dims), "seed": seed, "vertices": N, "edges": E, "path": vprec},
        "pseudo_steps": len(hist), "krylov_total": sum(x["krylov"] for x in hist),
        "residual_drop": hist[0]["fnorm"] / hist[-1]["fnorm"],
        "speedup_vs_fp64": paths["fp64"]["seconds"] / value if "fp64" in paths else None,
        "paths": paths, "warmup_seconds": warm, "setup_s": setup_s,
        "hbm_footprint": {"fp64_path_bytes": used64 - used0, "all_paths_bytes": used_all - used0,
                          "device_total_bytes": total,
                          "note": "hipMemGetInfo deltas: after the first FP64 solve (state, mesh, incidences, "
                                  "BSR FP64 Jacobian and the exact ILU's side arrays, plans, Krylov basis), and "
                                  "after the FP32/BF16 solves (their factors and vectors added)"},
        "flux_gflops": flux_flops / flux_s / 1e9, "flux_fp64_frac": flux_flops / flux_s / 1e9 / FP64_PEAK_GFLOPS,
        "kernel_ms": {k: (v[0] / v[1] if v[1] else None) for k, v in prof.items()},
        "kernel_share": {k: v[0] / sum(x[0] for x in prof.values()) for k, v in prof.items()},
        "roofline": roofline(prof_timed, vprec, sizes, args.config, held, fold[vprec]),
        "roofline_fp64": roofline(prof64, "fp64", sizes, args.config, held, fold["fp64"]) if prof64 else None,
        "roofline_spmv": spmv,
    }
    if not args.no_seq or not args.no_learned:
        mesh_c2 = lattice_mesh(K, tmp, "c2")
    out["variants_144x48x48"] = None if args.no_seq else c2_variants(K, mesh_c2)
    out["exact_parity_cost"] = None if args.no_seq else exact_parity_cost(K, mesh_c2)
    out["cpu_baseline"] = None if args.no_cpu else cpu_baseline(K, tmp)
    if args.no_learned:
        out["learned_preconditioner"] = None
    else:
        lmesh = mesh_c2 if args.learn_config == "c2" else mesh
        lr = learned(K, lmesh, args.learn_rhs, args.learn_iters, args.learn_prec)
        ldims, lseed = CONFIGS[args.learn_config]
        lr["mesh"] = "%dx%dx%d lattice seed %d" % (ldims + (lseed,))
        if args.learn_config == "c2" and out["variants_144x48x48"]:
            # configuration 4 reports against plain (left-preconditioned) ILU(0)-GMRES on the same mesh
            lr["plain_ilu_gmres"] = out["variants_144x48x48"]["default"]
        out["learned_preconditioner"] = lr
    print(json.dumps(out))


if __name__ == "__main__":
    main()
