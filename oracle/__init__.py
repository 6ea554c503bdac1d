"""CPU oracle of the KFUN3D hot path -- TEST INFRASTRUCTURE ONLY.

A one-thread C restatement of the reference's algorithm (oracle.c) that
rounds exactly as the reference's one-thread run.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg use it, as the
checker; the product never does.  The GEOMETRY it reads is the host part of
kfun3d_amd's Setup (the reference's inc/geometry.h layout).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboracle.so")
KREF = os.path.join(HERE, "_ref", "kref")


class OrOpts(C.Structure):
    _fields_ = [("alpha_deg", C.c_double), ("cfl_init", C.c_double), ("cfl_max", C.c_double),
                ("fnorm_ratio", C.c_double), ("gmres_rtol", C.c_double), ("ts_max", C.c_int),
                ("gmres_restart", C.c_int), ("gmres_max_it", C.c_int), ("first_order", C.c_int),
                ("ilu_textbook", C.c_int), ("reduce_order", C.c_int)]


class OrStep(C.Structure):
    _fields_ = [("step", C.c_int), ("cfl", C.c_double), ("fnorm", C.c_double), ("krylov", C.c_int),
                ("forces", C.c_double * 3)]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("oracle: %s is not built (make -C oracle)" % LIBPATH)
        L = C.CDLL(LIBPATH)
        vp, dp = C.c_void_p, C.c_void_p
        L.or_create.restype = vp
        L.or_create.argtypes = [vp, C.POINTER(OrOpts)]
        L.or_destroy.argtypes = [vp]
        L.or_state.restype = C.POINTER(C.c_double)
        L.or_state.argtypes = [vp]
        L.or_cdt.restype = C.POINTER(C.c_double)
        L.or_cdt.argtypes = [vp]
        L.or_nnzb.restype = C.c_size_t
        L.or_nnzb.argtypes = [vp]
        L.or_blocks.restype = C.POINTER(C.c_double)
        L.or_blocks.argtypes = [vp]
        L.or_pattern.argtypes = [vp, dp, dp]
        L.or_flux.argtypes = [vp, dp, dp, dp, dp, dp]
        L.or_timestep.argtypes = [vp, C.c_double]
        L.or_fill.argtypes = [vp]
        L.or_ilu.argtypes = [vp]
        L.or_sptrsv.argtypes = [vp, dp, dp]
        L.or_sptrsv_jacobi.argtypes = [vp, dp, dp, C.c_int]
        L.or_spmv.argtypes = [vp, dp, dp]
        L.or_forces.argtypes = [vp, dp]
        L.or_ijacvec.argtypes = [vp, dp, dp, C.c_double, dp, dp, dp, dp, dp]
        L.or_norm.restype = C.c_double
        L.or_norm.argtypes = [C.c_size_t, dp]
        L.or_dot_tree.restype = C.c_double
        L.or_dot_tree.argtypes = [C.c_size_t, dp, dp]
        L.or_kernel.restype = C.c_int
        L.or_kernel.argtypes = [vp, C.POINTER(OrStep), C.c_int]
        L.or_inv4.argtypes = [dp, dp]
        _LIB = L
    return _LIB


def reference_options(**kw):
    """The reference's compile-time constants (inc/phy.h, src/kernel.c)."""
    o = dict(alpha_deg=3.0, cfl_init=50.0, cfl_max=1e5, fnorm_ratio=1e7, gmres_rtol=0.01, ts_max=15,
             gmres_restart=30, gmres_max_it=10000, first_order=0, ilu_textbook=0, reduce_order=0)
    o.update(kw)
    return o


def config_options(**kw):
    o = reference_options(alpha_deg=3.06, cfl_init=10.0, ts_max=500, fnorm_ratio=1e10)
    o.update(kw)
    return o


def _p(a):
    return a.ctypes.data


class Oracle:
    def __init__(self, mesh, **opts):
        o = reference_options()
        o.update(opts)
        self.opts = OrOpts(**o)
        self.mesh = mesh
        self.N = mesh.N
        self.sz = 4 * mesh.N
        self.c = lib().or_create(C.cast(C.byref(mesh.g), C.c_void_p), C.byref(self.opts))

    def close(self):
        if self.c:
            lib().or_destroy(self.c)
            self.c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def state(self):
        return np.ctypeslib.as_array(lib().or_state(self.c), shape=(self.sz,)).copy()

    def set_state(self, q):
        dst = np.ctypeslib.as_array(lib().or_state(self.c), shape=(self.sz,))
        dst[:] = q

    def cdt(self):
        return np.ctypeslib.as_array(lib().or_cdt(self.c), shape=(self.N,)).copy()

    def set_cdt(self, cdt):
        dst = np.ctypeslib.as_array(lib().or_cdt(self.c), shape=(self.N,))
        dst[:] = cdt

    def flux(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        r = np.empty(self.sz)
        gx, gy, gz = np.empty(self.sz), np.empty(self.sz), np.empty(self.sz)
        lib().or_flux(self.c, _p(q), _p(gx), _p(gy), _p(gz), _p(r))
        return r, gx, gy, gz

    def timestep(self, cfl):
        lib().or_timestep(self.c, cfl)
        return self.cdt()

    def fill(self):
        lib().or_fill(self.c)

    def ilu(self):
        lib().or_ilu(self.c)

    def blocks(self):
        n = int(lib().or_nnzb(self.c))
        return np.ctypeslib.as_array(lib().or_blocks(self.c), shape=(n * 16,)).copy().reshape(-1, 4, 4)

    def pattern(self):
        n = int(lib().or_nnzb(self.c))
        rp = np.empty(self.N + 1, dtype=np.uint32)
        col = np.empty(n, dtype=np.uint32)
        lib().or_pattern(self.c, _p(rp), _p(col))
        return rp, col

    def sptrsv(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty(self.sz)
        lib().or_sptrsv(self.c, _p(b), _p(x))
        return x

    def sptrsv_jacobi(self, b, sweeps):
        """x ~ U^-1 L^-1 b by `sweeps` Jacobi sweeps each way over the factors"""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty(self.sz)
        lib().or_sptrsv_jacobi(self.c, _p(b), _p(x), int(sweeps))
        return x

    def spmv(self, x):
        """y = A x with the current blocks (after fill, before ilu)"""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.sz)
        lib().or_spmv(self.c, _p(x), _p(y))
        return y

    def forces(self):
        f = np.empty(3)
        lib().or_forces(self.c, _p(f))
        return f

    def ijacvec(self, x, r, qnorm):
        x = np.ascontiguousarray(x, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        gx, gy, gz, w, y = (np.empty(self.sz) for _ in range(5))
        lib().or_ijacvec(self.c, _p(x), _p(r), qnorm, _p(gx), _p(gy), _p(gz), _p(w), _p(y))
        return w, y

    def kernel(self, max_rec=1000):
        rec = (OrStep * max_rec)()
        n = lib().or_kernel(self.c, rec, max_rec)
        return [dict(step=s.step, cfl=s.cfl, fnorm=s.fnorm, krylov=s.krylov, forces=list(s.forces))
                for s in rec[:min(n, max_rec)]]


def dot_tree(x, y):
    """x . y in the HIP path's fixed reduction tree"""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return lib().or_dot_tree(x.size, _p(x), _p(y))


def norm(v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().or_norm(v.size, _p(v))


def inv4(block):
    src = np.ascontiguousarray(block, dtype=np.float64).reshape(16)
    dst = np.empty(16)
    lib().or_inv4(_p(src), _p(dst))
    return dst.reshape(4, 4)
