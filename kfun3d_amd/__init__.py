"""kfun3d_amd -- MI355X (gfx950) pseudo-transient Newton-Krylov path of KFUN3D.

The product is the C-ABI shared library ``libkfun3d.so`` (include/kfun3d.h):
the reference's entry points (ComputeFlux, ComputeTimeStep, ComputeA,
ComputeNumericalILU, ComputeSparseTriangularSolve, ComputeForces, iJacVec,
Kernel, Compute2ndNorm, ComputeAXPY, ComputeNewAXPY, Normalize, Setup) backed
by hand-written HIP kernels.  This module is the ctypes binding used by the
tests and the benchmark; it computes nothing itself.  Loading fails loudly
when the library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libkfun3d.so")

PREC = {"fp64": 0, "fp32": 1, "bf16": 2}
ORTHO = {"mgs": 0, "cgs2": 1}
KRYLOV = {"gmres": 0, "fgmres": 1}


class Xyzn(C.Structure):
    _fields_ = [("x0", C.POINTER(C.c_double)), ("x1", C.POINTER(C.c_double)),
                ("x2", C.POINTER(C.c_double)), ("x3", C.POINTER(C.c_double))]


class Xyz(C.Structure):
    _fields_ = [("x0", C.POINTER(C.c_double)), ("x1", C.POINTER(C.c_double)), ("x2", C.POINTER(C.c_double))]


class Edge(C.Structure):
    _fields_ = [("n0", C.POINTER(C.c_uint)), ("n1", C.POINTER(C.c_uint))]


class Face(C.Structure):
    _fields_ = [("n0", C.POINTER(C.c_uint)), ("n1", C.POINTER(C.c_uint)), ("n2", C.POINTER(C.c_uint))]


class Weights(C.Structure):
    _fields_ = [("w0", C.POINTER(Xyz)), ("w1", C.POINTER(Xyz))]


class Etbl(C.Structure):
    _fields_ = [("sz", C.c_ulong), ("eptr", C.POINTER(Edge)), ("xyzn", C.POINTER(Xyzn)), ("w", C.POINTER(Weights))]


class Ntbl(C.Structure):
    _fields_ = [("sz", C.c_ulong), ("xyz", C.POINTER(Xyz)), ("part", C.POINTER(C.c_uint)),
                ("cdt", C.POINTER(C.c_double))]


class Stbl(C.Structure):
    _fields_ = [("sz", C.c_ulong), ("i", C.POINTER(C.c_uint))]


class Bntbl(C.Structure):
    _fields_ = [("sz", C.c_ulong), ("nptr", C.POINTER(C.c_uint)), ("xyz", C.POINTER(Xyz))]


class Bftbl(C.Structure):
    _fields_ = [("sz", C.c_ulong), ("fptr", C.POINTER(Face))]


class Btbl(C.Structure):
    _fields_ = [("s", C.POINTER(Bntbl)), ("f", C.POINTER(Bntbl)), ("fc", C.POINTER(Bftbl))]


class Ctbl(C.Structure):
    _fields_ = [("b", C.c_uint), ("b2", C.c_uint), ("sz", C.c_ulong)]


class Qtbl(C.Structure):
    _fields_ = [("q", C.POINTER(C.c_double))]


class Geometry(C.Structure):
    _fields_ = [("e", C.POINTER(Etbl)), ("n", C.POINTER(Ntbl)), ("b", C.POINTER(Btbl)), ("c", C.POINTER(Ctbl)),
                ("s", C.POINTER(Stbl)), ("t", C.POINTER(Stbl)), ("q", C.POINTER(Qtbl)), ("matrix", C.c_void_p)]


class Options(C.Structure):
    _fields_ = [("alpha_deg", C.c_double), ("cfl_init", C.c_double), ("cfl_max", C.c_double),
                ("ts_max", C.c_int), ("fnorm_ratio", C.c_double), ("gmres_restart", C.c_int),
                ("gmres_rtol", C.c_double), ("gmres_max_it", C.c_int), ("first_order", C.c_int),
                ("prec", C.c_int), ("ortho", C.c_int), ("verbose", C.c_int), ("ilu_exact", C.c_int),
                ("reduce_seq", C.c_int), ("krylov", C.c_int), ("bsr_operator", C.c_int),
                ("jacobi_sweeps", C.c_int), ("coarse", C.c_int),
                ("coarse_sweeps", C.c_int), ("lag_factors", C.c_int)]


class StepRecord(C.Structure):
    _fields_ = [("step", C.c_int), ("cfl", C.c_double), ("fnorm", C.c_double), ("krylov", C.c_int),
                ("forces", C.c_double * 3), ("learned", C.c_int), ("krylov_ilu", C.c_int),
                ("krylov_net", C.c_int)]


class LearnOptions(C.Structure):
    _fields_ = [("collect_steps", C.c_int), ("rhs_per_step", C.c_int), ("pair_rtol", C.c_double),
                ("pair_maxit", C.c_int), ("train_iters", C.c_int), ("lr", C.c_double), ("gclip", C.c_double),
                ("guard", C.c_double), ("judge_every", C.c_int), ("holdout", C.c_double), ("seed", C.c_uint64),
                ("ablate", C.c_int)]


class LearnStats(C.Structure):
    _fields_ = [("npairs", C.c_int), ("nheld", C.c_int), ("pair_iters", C.c_long), ("pair_ms", C.c_double),
                ("train_ms", C.c_double), ("loss_before", C.c_double), ("loss_after", C.c_double),
                ("loss_ilu", C.c_double), ("steps_learned", C.c_int), ("steps_judged", C.c_int),
                ("judged_won", C.c_int), ("fallback", C.c_int)]


GP = C.POINTER(Geometry)
DP = C.c_void_p   # device pointers travel as plain addresses


def _load():
    if not os.path.exists(LIBPATH):
        raise RuntimeError("kfun3d_amd: %s is not built (run python -m kfun3d_amd.build)" % LIBPATH)
    lib = C.CDLL(LIBPATH)
    sig = {
        "Setup": (None, [C.c_char_p, GP]),
        "kf_setup_ex": (None, [C.c_char_p, GP, C.POINTER(Options)]),
        "kf_setup_host": (None, [C.c_char_p, GP]),
        "kf_geometry_free": (None, [GP]),
        "kf_default_options": (Options, []),
        "kf_set_options": (None, [GP, C.POINTER(Options)]),
        "kf_get_options": (Options, [GP]),
        "kf_kernel": (C.c_int, [GP, C.POINTER(StepRecord), C.c_int]),
        "kf_set_state": (None, [GP, C.c_void_p]),
        "kf_get_state": (None, [GP, C.c_void_p]),
        "kf_memcpy_d2h": (None, [C.c_void_p, DP, C.c_size_t]),
        "kf_memcpy_h2d": (None, [DP, C.c_void_p, C.c_size_t]),
        "kf_device_alloc": (DP, [C.c_size_t]),
        "kf_device_free": (None, [DP]),
        "kf_synchronize": (None, [GP]),
        "kf_stream": (C.c_void_p, [GP]),
        "kf_nnzb": (C.c_size_t, [GP]),
        "kf_trsv_fold": (C.c_int, [GP]),
        "kf_grad_moments": (C.c_int, [GP]),
        "kf_bsr_spmv": (None, [GP, DP, DP]),
        "kf_pattern": (None, [GP, C.c_void_p, C.c_void_p, C.c_void_p]),
        "kf_get_blocks": (None, [GP, C.c_void_p]),
        "kf_write_lattice_msh": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_uint64]),
        "kf_last_error": (C.c_int, [GP]),
        "kf_bind": (None, [GP]),
        "kf_profile": (None, [GP, C.c_int]),
        "kf_profile_only": (C.c_int, [GP, C.c_char_p]),
        "kf_profile_get": (C.c_int, [GP, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_long)]),
        "kf_trsv_plan_host": (C.c_size_t, [GP, C.c_int, C.c_uint, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.c_size_t]),
        "kf_ilu_plan_host": (C.c_size_t, [GP, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
        "kf_coarse_plan_host": (C.c_size_t, [GP, C.c_void_p, C.POINTER(C.c_size_t), C.c_void_p, C.c_void_p]),
        "kf_coarse_matrix": (C.c_size_t, [GP, C.POINTER(C.c_size_t), C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p]),
        "ComputeFlux": (None, [GP, DP, C.POINTER(Xyz), DP]),
        "ComputeTimeStep": (None, [C.c_double, GP]),
        "ComputeA": (None, [GP]),
        "ComputeForces": (None, [GP, C.POINTER(C.c_double)]),
        "ComputeNumericalILU": (None, [GP]),
        "ComputeSparseTriangularSolve": (None, [GP, DP, DP]),
        "kf_sptrsv_jacobi_noalloc": (C.c_int, [GP, DP, DP]),
        "iJacVec": (None, [GP, DP, DP, C.c_double, C.POINTER(Xyz), DP, DP]),
        "Kernel": (None, [GP]),
        "kfs_flux_spectral": (None, [GP, DP, C.POINTER(Xyz), DP]),
        "kfs_timestep_spectral": (None, [GP, C.c_double]),
        "Compute2ndNorm": (C.c_double, [C.c_size_t, DP]),
        "ComputeAXPY": (None, [C.c_size_t, C.c_double, DP, DP]),
        "ComputeNewAXPY": (None, [C.c_size_t, C.c_double, DP, DP, DP]),
        "Normalize": (C.c_double, [C.c_size_t, DP]),
        "kf_mpnn_nparam": (C.c_size_t, []),
        "kf_mpnn_nhead": (C.c_size_t, []),
        "kf_mpnn_init_params": (None, [C.c_void_p, C.c_uint64]),
        "kf_mpnn_prepare": (None, [C.c_void_p, C.c_int]),
        "kf_mpnn_create": (C.c_void_p, [GP, C.c_void_p, C.c_int]),
        "kf_mpnn_destroy": (None, [C.c_void_p]),
        "kf_mpnn_set_params": (None, [C.c_void_p, C.c_void_p]),
        "kf_mpnn_get_params": (None, [C.c_void_p, C.c_void_p]),
        "kf_mpnn_take_diag": (None, [C.c_void_p, C.c_int]),
        "kf_mpnn_set_diag": (None, [C.c_void_p, C.c_int, C.c_void_p]),
        "kf_mpnn_forward": (None, [C.c_void_p, C.c_int, DP, DP, C.c_int]),
        "kf_mpnn_take_matrix": (None, [C.c_void_p, C.c_int]),
        "kf_mpnn_set_loss": (None, [C.c_void_p, C.c_int]),
        "kf_mpnn_loss_grad": (C.c_double, [C.c_void_p, C.c_int, DP, DP, DP, C.c_void_p]),
        "kf_mpnn_adam": (None, [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float]),
        "kf_mpnn_train_step": (None, [C.c_void_p, C.c_int, DP, DP, DP, C.c_float, C.c_float]),
        "kf_mpnn_losses": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
        "kf_mpnn_reset_losses": (None, [C.c_void_p]),
        "kf_mpnn_activation": (None, [C.c_void_p, C.c_int, C.c_void_p]),
        "kf_default_learn_options": (LearnOptions, []),
        "kf_attach_mpnn": (None, [GP, C.c_void_p, C.POINTER(LearnOptions)]),
        "kf_learn_get_stats": (None, [GP, C.POINTER(LearnStats)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


# exported C symbols the header declares (checked by the CPU tests)
ABI_SYMBOLS = ["ComputeA", "ComputeTimeStep", "ComputeFlux", "ComputeForces", "ComputeNumericalILU",
               "ComputeSparseTriangularSolve", "iJacVec", "Kernel", "Compute2ndNorm", "ComputeAXPY",
               "ComputeNewAXPY", "Normalize", "Setup", "kf_default_options", "kf_set_options", "kf_get_options",
               "kf_kernel", "kf_set_state", "kf_get_state", "kf_memcpy_d2h", "kf_memcpy_h2d", "kf_device_alloc",
               "kf_device_free", "kf_synchronize", "kf_stream", "kf_nnzb", "kf_trsv_fold", "kf_grad_moments", "kf_bsr_spmv", "kf_pattern", "kf_get_blocks",
               "kf_write_lattice_msh", "kf_sptrsv_jacobi_noalloc", "kf_geometry_free", "kf_last_error", "kf_setup_ex", "kf_setup_host",
               "kf_bind", "kf_profile", "kf_profile_only", "kf_profile_get", "kf_trsv_plan_host",
               "kf_ilu_plan_host", "kf_coarse_plan_host", "kf_coarse_matrix", "kf_mpnn_nparam", "kf_mpnn_nhead", "kf_mpnn_init_params", "kf_mpnn_prepare", "kf_mpnn_create", "kf_mpnn_destroy", "kf_mpnn_set_params",
               "kf_mpnn_get_params", "kf_mpnn_take_diag", "kf_mpnn_set_diag", "kf_mpnn_forward",
               "kf_mpnn_loss_grad", "kf_mpnn_adam", "kf_mpnn_train_step", "kf_mpnn_losses", "kf_mpnn_reset_losses",
               "kf_mpnn_activation", "kf_mpnn_take_matrix", "kf_mpnn_set_loss",
               "kf_default_learn_options", "kf_attach_mpnn", "kf_learn_get_stats"]


def options(**kw):
    o = lib().kf_default_options()
    for k, v in kw.items():
        if k == "prec":
            v = PREC[v] if isinstance(v, str) else v
        if k == "ortho":
            v = ORTHO[v] if isinstance(v, str) else v
        if k == "krylov":
            v = KRYLOV[v] if isinstance(v, str) else v
        setattr(o, k, v)
    return o


def config_options(**kw):
    """Options of the synthetic configurations: alpha 3.06 deg, CFL ramp from
    10, stop at a 1e-10 residual drop."""
    base = dict(alpha_deg=3.06, cfl_init=10.0, ts_max=500, fnorm_ratio=1e10)
    base.update(kw)
    return options(**base)


def write_lattice(path, nx, ny, nz, seed):
    rc = lib().kf_write_lattice_msh(path.encode(), nx, ny, nz, seed)
    if rc != 0:
        raise RuntimeError("kf_write_lattice_msh failed (%d)" % rc)
    return path


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    ct = {np.float64: C.c_double, np.uint32: C.c_uint}[dtype]
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n,)).copy()


class DeviceVec:
    """A device buffer of float64 (4N for a state vector)."""

    def __init__(self, n, host=None):
        self.n = int(n)
        self.ptr = lib().kf_device_alloc(8 * max(self.n, 1))
        if host is not None:
            self.upload(host)

    def upload(self, host):
        h = np.ascontiguousarray(host, dtype=np.float64)
        assert h.size == self.n
        lib().kf_memcpy_h2d(self.ptr, h.ctypes.data, 8 * self.n)

    def numpy(self):
        out = np.empty(self.n, dtype=np.float64)
        lib().kf_memcpy_d2h(out.ctypes.data, self.ptr, 8 * self.n)
        return out

    def free(self):
        if self.ptr:
            lib().kf_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Mesh:
    """A GEOMETRY built by Setup (device=True) or by its host part only."""

    def __init__(self, path, opts=None, device=True):
        self.path = path
        self.g = Geometry()
        self.device = device
        if device:
            o = opts if opts is not None else lib().kf_default_options()
            lib().kf_setup_ex(path.encode(), C.byref(self.g), C.byref(o))
        else:
            lib().kf_setup_host(path.encode(), C.byref(self.g))
        self.N = int(self.g.n.contents.sz)
        self.E = int(self.g.e.contents.sz)
        self.sz = 4 * self.N

    @property
    def ref(self):
        return C.byref(self.g)

    def close(self):
        if self.g.e:
            lib().kf_geometry_free(C.byref(self.g))

    # ---- host arrays (as the reference's Setup leaves them) ----
    def host_arrays(self):
        g = self.g
        e, n, b = g.e.contents, g.n.contents, g.b.contents
        E, N = self.E, self.N
        s, f, fc = b.s.contents, b.f.contents, b.fc.contents
        t = g.t.contents
        out = {
            "e_n0": _arr(e.eptr.contents.n0, E, np.uint32), "e_n1": _arr(e.eptr.contents.n1, E, np.uint32),
            "e_x0": _arr(e.xyzn.contents.x0, E, np.float64), "e_x1": _arr(e.xyzn.contents.x1, E, np.float64),
            "e_x2": _arr(e.xyzn.contents.x2, E, np.float64), "e_x3": _arr(e.xyzn.contents.x3, E, np.float64),
            "w0_x0": _arr(e.w.contents.w0.contents.x0, E, np.float64),
            "w0_x1": _arr(e.w.contents.w0.contents.x1, E, np.float64),
            "w0_x2": _arr(e.w.contents.w0.contents.x2, E, np.float64),
            "w1_x0": _arr(e.w.contents.w1.contents.x0, E, np.float64),
            "w1_x1": _arr(e.w.contents.w1.contents.x1, E, np.float64),
            "w1_x2": _arr(e.w.contents.w1.contents.x2, E, np.float64),
            "n_x0": _arr(n.xyz.contents.x0, N, np.float64), "n_x1": _arr(n.xyz.contents.x1, N, np.float64),
            "n_x2": _arr(n.xyz.contents.x2, N, np.float64),
            "fc_n0": _arr(fc.fptr.contents.n0, fc.sz, np.uint32), "fc_n1": _arr(fc.fptr.contents.n1, fc.sz, np.uint32),
            "fc_n2": _arr(fc.fptr.contents.n2, fc.sz, np.uint32),
            "t_i": _arr(t.i, t.sz + 1, np.uint32),
            "s_nptr": _arr(s.nptr, s.sz, np.uint32), "s_x0": _arr(s.xyz.contents.x0, s.sz, np.float64),
            "s_x1": _arr(s.xyz.contents.x1, s.sz, np.float64), "s_x2": _arr(s.xyz.contents.x2, s.sz, np.float64),
            "f_nptr": _arr(f.nptr, f.sz, np.uint32), "f_x0": _arr(f.xyz.contents.x0, f.sz, np.float64),
            "f_x1": _arr(f.xyz.contents.x1, f.sz, np.float64), "f_x2": _arr(f.xyz.contents.x2, f.sz, np.float64),
        }
        return out

    # ---- device entry points ----
    def set_options(self, opts):
        lib().kf_set_options(self.ref, C.byref(opts))

    def state(self):
        q = np.empty(self.sz)
        lib().kf_get_state(self.ref, q.ctypes.data)
        return q

    def set_state(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        lib().kf_set_state(self.ref, q.ctypes.data)

    def flux(self, q_dev, grad_devs, r_dev):
        grad = Xyz(C.cast(grad_devs[0].ptr, C.POINTER(C.c_double)), C.cast(grad_devs[1].ptr, C.POINTER(C.c_double)),
                   C.cast(grad_devs[2].ptr, C.POINTER(C.c_double)))
        lib().ComputeFlux(self.ref, q_dev.ptr, C.byref(grad), r_dev.ptr)

    def flux_spectral(self, q_dev, grad_devs, r_dev):
        """Residual of q that also leaves the time-step sums (the driver's
        first residual of a pseudo-time step)."""
        grad = Xyz(C.cast(grad_devs[0].ptr, C.POINTER(C.c_double)), C.cast(grad_devs[1].ptr, C.POINTER(C.c_double)),
                   C.cast(grad_devs[2].ptr, C.POINTER(C.c_double)))
        lib().kfs_flux_spectral(self.ref, q_dev.ptr, C.byref(grad), r_dev.ptr)

    def timestep_spectral(self, cfl):
        """cdt from the sums of the last flux_spectral (same state)."""
        lib().kfs_timestep_spectral(self.ref, cfl)

    def timestep(self, cfl):
        lib().ComputeTimeStep(cfl, self.ref)

    def cdt(self):
        out = np.empty(self.N)
        lib().kf_synchronize(self.ref)
        lib().kf_memcpy_d2h(out.ctypes.data, C.cast(self.g.n.contents.cdt, C.c_void_p), 8 * self.N)
        return out

    def fill(self):
        lib().ComputeA(self.ref)

    def ilu(self):
        lib().ComputeNumericalILU(self.ref)

    def sptrsv(self, b_dev, x_dev):
        lib().ComputeSparseTriangularSolve(self.ref, b_dev.ptr, x_dev.ptr)

    def sptrsv_noalloc(self, b_dev, x_dev):
        """ComputeSparseTriangularSolve with Jacobi sweeps, work vectors in the
        library's fixed device array (no allocation); returns 0 or
        KF_NOALLOC_ESIZE (1) / KF_NOALLOC_EMODE (2)"""
        return int(lib().kf_sptrsv_jacobi_noalloc(self.ref, b_dev.ptr, x_dev.ptr))

    def spmv(self, x_dev, y_dev):
        """y = A x with the Jacobian the last fill kept (options bsr_operator=1)"""
        lib().kf_bsr_spmv(self.ref, x_dev.ptr, y_dev.ptr)

    def forces(self):
        f = (C.c_double * 3)()
        lib().ComputeForces(self.ref, f)
        return np.array(list(f))

    def ijacvec(self, x_dev, r_dev, qnorm, grad_devs, w_dev, y_dev):
        grad = Xyz(C.cast(grad_devs[0].ptr, C.POINTER(C.c_double)), C.cast(grad_devs[1].ptr, C.POINTER(C.c_double)),
                   C.cast(grad_devs[2].ptr, C.POINTER(C.c_double)))
        lib().iJacVec(self.ref, x_dev.ptr, r_dev.ptr, qnorm, C.byref(grad), w_dev.ptr, y_dev.ptr)

    def blocks(self):
        out = np.empty(int(lib().kf_nnzb(self.ref)) * 16)
        lib().kf_get_blocks(self.ref, out.ctypes.data)
        return out.reshape(-1, 4, 4)

    def trsv_plan(self, upper, rows, cluster=True):
        """Host-side schedule of the triangular solves: (task table of shape
        (ntask, 16 // rows) in ticket order, each task's local depth, each
        row's local level in its task, per-task row table of shape
        (ntask, 16, 3), per-block-position slot of the column's row in the
        same task or 255)."""
        nt = int(lib().kf_trsv_plan_host(self.ref, int(upper), rows, int(cluster), None, None, None, None, None, 0))
        task = np.empty(nt * (16 // rows), dtype=np.uint32)
        tdep = np.empty(nt, dtype=np.uint8)
        lvl = np.empty(self.N, dtype=np.uint8)
        rowtab = np.empty(nt * 48, dtype=np.uint32)
        eslot = np.empty(self.N + 2 * self.E, dtype=np.uint8)   # one diagonal block per vertex, two per edge
        lib().kf_trsv_plan_host(self.ref, int(upper), rows, int(cluster), task.ctypes.data, tdep.ctypes.data,
                                lvl.ctypes.data, rowtab.ctypes.data, eslot.ctypes.data, nt)
        return task.reshape(nt, 16 // rows), tdep, lvl, rowtab.reshape(nt, 16, 3), eslot

    def ilu_plan(self, nnzb):
        """Host-side schedule of the ILU(0) factorisation: (records (N, 8),
        ldiag, moff (nnzb each), pair blocks, pair positions)."""
        nm = int(lib().kf_ilu_plan_host(self.ref, None, None, None, None, None, 0))
        rec = np.empty(8 * self.N, dtype=np.uint32)
        ldiag = np.empty(nnzb, dtype=np.uint32)
        moff = np.empty(nnzb, dtype=np.uint32)
        mjm = np.empty(max(nm, 1), dtype=np.uint32)
        mlo = np.empty(max(nm, 1), dtype=np.uint16)
        lib().kf_ilu_plan_host(self.ref, rec.ctypes.data, ldiag.ctypes.data, moff.ctypes.data, mjm.ctypes.data,
                               mlo.ctypes.data, nm)
        return rec.reshape(self.N, 8), ldiag, moff, mjm[:nm], mlo[:nm]

    def pattern(self):
        nnzb = int(lib().kf_nnzb(self.ref))
        rp = np.empty(self.N + 1, dtype=np.uint32)
        col = np.empty(nnzb, dtype=np.uint32)
        dg = np.empty(self.N, dtype=np.uint32)
        lib().kf_pattern(self.ref, rp.ctypes.data, col.ctypes.data, dg.ctypes.data)
        return rp, col, dg

    def coarse_plan(self):
        """The two-level preconditioner's coarse level as the host builds it:
        (aggregate of each vertex, coarse row pointers, coarse columns)."""
        nz = C.c_size_t(0)
        nc = int(lib().kf_coarse_plan_host(self.ref, None, C.byref(nz), None, None))
        agg = np.empty(self.N, dtype=np.uint32)
        rp = np.empty(nc + 1, dtype=np.uint32)
        col = np.empty(max(int(nz.value), 1), dtype=np.uint32)
        lib().kf_coarse_plan_host(self.ref, agg.ctypes.data, C.byref(nz), rp.ctypes.data, col.ctypes.data)
        return agg, rp, col[:int(nz.value)]

    def coarse_matrix(self):
        """The device's coarse level after ComputeNumericalILU with coarse=1:
        (aggregates, coarse row pointers, columns, Galerkin blocks (cnnzb, 4, 4)
        of the last ComputeA); None before."""
        nz = C.c_size_t(0)
        nc = int(lib().kf_coarse_matrix(self.ref, C.byref(nz), None, None, None, None))
        if nc == 0:
            return None
        n = int(nz.value)
        agg = np.empty(self.N, dtype=np.uint32)
        rp = np.empty(nc + 1, dtype=np.uint32)
        col = np.empty(max(n, 1), dtype=np.uint32)
        blocks = np.empty(max(16 * n, 1))
        lib().kf_coarse_matrix(self.ref, C.byref(nz), agg.ctypes.data, rp.ctypes.data, col.ctypes.data,
                               blocks.ctypes.data)
        return agg, rp, col[:n], blocks[:16 * n].reshape(-1, 4, 4)

    def kernel(self, max_rec=1000):
        rec = (StepRecord * max_rec)()
        n = lib().kf_kernel(self.ref, rec, max_rec)
        return [dict(step=r.step, cfl=r.cfl, fnorm=r.fnorm, krylov=r.krylov, forces=list(r.forces),
                     learned=r.learned, krylov_ilu=r.krylov_ilu, krylov_net=r.krylov_net)
                for r in rec[:min(n, max_rec)]]

    def attach(self, net, **kw):
        """Attach a Network to Kernel (None detaches); keywords override
        kf_default_learn_options()."""
        if net is None:
            lib().kf_attach_mpnn(self.ref, None, None)
            return
        lo = lib().kf_default_learn_options()
        for k, v in kw.items():
            setattr(lo, k, v)
        lib().kf_attach_mpnn(self.ref, net.h, C.byref(lo))

    def learn_stats(self):
        st = LearnStats()
        lib().kf_learn_get_stats(self.ref, C.byref(st))
        return {name: getattr(st, name) for name, _ in LearnStats._fields_}

    def last_error(self):
        return lib().kf_last_error(self.ref)

    def profile(self, on):
        lib().kf_profile(self.ref, 1 if on else 0)

    def profile_only(self, name):
        """profile one kernel class only (the others run without events)"""
        if lib().kf_profile_only(self.ref, name.encode()) != 0:
            raise KeyError(name)

    def profile_get(self, name):
        """(total milliseconds, launches) of one kernel class since profile(True)"""
        ms, n = C.c_double(), C.c_long()
        if lib().kf_profile_get(self.ref, name.encode(), C.byref(ms), C.byref(n)) != 0:
            raise KeyError(name)
        return ms.value, n.value

    PROFILED = ["grad", "edge_flux", "residual", "timestep", "fill", "ilu", "trsv_lower", "trsv_upper", "spmv",
                "jacobi_lower", "jacobi_upper", "coarse_factor", "coarse"]

    def synchronize(self):
        lib().kf_synchronize(self.ref)


def init_params(seed=0):
    """kf_mpnn_init_params: the network's initial parameters (float32)"""
    th = np.empty(int(lib().kf_mpnn_nparam()), dtype=np.float32)
    lib().kf_mpnn_init_params(th.ctypes.data, seed)
    return th


class Network:
    """The learned preconditioner's network bound to a device Mesh
    (kf_mpnn_*): parameters are kf_mpnn_nparam() float32 values, vectors are
    DeviceVec of 4N, diagonal blocks live in slots 0 .. nslot-1."""

    def __init__(self, mesh, theta, nslot=1):
        self.mesh = mesh
        self.nparam = int(lib().kf_mpnn_nparam())
        th = np.ascontiguousarray(theta, dtype=np.float32)
        assert th.size == self.nparam
        self.h = lib().kf_mpnn_create(mesh.ref, th.ctypes.data, int(nslot))

    def close(self):
        if self.h:
            lib().kf_mpnn_destroy(self.h)
            self.h = None

    def params(self):
        out = np.empty(self.nparam, dtype=np.float32)
        lib().kf_mpnn_get_params(self.h, out.ctypes.data)
        return out

    def set_params(self, theta):
        th = np.ascontiguousarray(theta, dtype=np.float32)
        assert th.size == self.nparam
        lib().kf_mpnn_set_params(self.h, th.ctypes.data)

    def take_diag(self, slot=0):
        lib().kf_mpnn_take_diag(self.h, slot)

    def set_diag(self, D, slot=0):
        d = np.ascontiguousarray(D, dtype=np.float64)
        assert d.size == 16 * self.mesh.N
        lib().kf_mpnn_set_diag(self.h, slot, d.ctypes.data)

    def prepare(self, slot=0):
        lib().kf_mpnn_prepare(self.h, slot)

    def forward(self, z_dev, out_dev, slot=0, add=False):
        """out = G z (or out += G z) with the operator of `slot`"""
        lib().kf_mpnn_forward(self.h, slot, z_dev.ptr, out_dev.ptr, 1 if add else 0)

    def take_matrix(self, slot):
        lib().kf_mpnn_take_matrix(self.h, slot)

    def set_loss(self, metric):
        """"residual" (|A (base + G base) - v|^2 / |v|^2, default), "a"
        (|A e|^2 / |v|^2) or "euclid" (|e|^2 / |target|^2), e = base + G base
        - target"""
        lib().kf_mpnn_set_loss(self.h, {"a": 1, "euclid": 0, "residual": 2}[metric])

    def loss_grad(self, v_dev, base_dev, target_dev, slot=0):
        g = np.empty(self.nparam, dtype=np.float32)
        loss = lib().kf_mpnn_loss_grad(self.h, slot, v_dev.ptr, base_dev.ptr, target_dev.ptr, g.ctypes.data)
        return loss, g

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, gclip=0.0):
        lib().kf_mpnn_adam(self.h, lr, beta1, beta2, eps, gclip)

    def train_step(self, v_dev, base_dev, target_dev, lr, slot=0, gclip=0.0):
        lib().kf_mpnn_train_step(self.h, slot, v_dev.ptr, base_dev.ptr, target_dev.ptr, lr, gclip)

    def losses(self):
        n = lib().kf_mpnn_losses(self.h, None, 0)
        out = np.empty(max(min(n, 1 << 16), 1), dtype=np.float32)
        lib().kf_mpnn_losses(self.h, out.ctypes.data, out.size)
        return out[:min(n, 1 << 16)]

    def reset_losses(self):
        lib().kf_mpnn_reset_losses(self.h)

    def activation(self, which):
        n = self.mesh.N * (16 if which == 0 else 32)
        out = np.empty(n, dtype=np.float32)
        lib().kf_mpnn_activation(self.h, which, out.ctypes.data)
        return out.reshape(self.mesh.N, -1)
