"""Writes the golden fixtures of tests/golden/ from the reference itself.

Runs oracle/_ref/kref -- the reference's own src/*.c and matrix/*.cpp built
by oracle/Makefile -- on mesh A (mesh/uns3dA.msh of the reference, shipped
here as uns3dA.msh.gz) and on the 48x16x16 lattice of configuration 1, and
stores for each:
  setup    SHA-256 of every host array the reference's Setup produces
  kernels  SHA-256 of each output of one call of ComputeFlux, ComputeTimeStep,
           ComputeA, ComputeNumericalILU, ComputeSparseTriangularSolve,
           iJacVec, ComputeForces on a seeded state, a second
           ComputeA + ILU + solve on a second state, and ComputeFlux with zero
           LSQ terms (the first-order residual); plus norms for diagnostics
  solve    the Kernel() history: per step CFL, |R|, forces (all %.17g) and
           Krylov iterations
The seeded states are reproduced in numpy by the tests (kernel_inputs()).

Run in the development container (needs /root/reference):
    python tests/golden/make_golden.py
"""
import gzip
import hashlib
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
KREF = os.path.join(ROOT, "oracle", "_ref", "kref")

SETUP_ARRAYS = {"e_n0": np.uint32, "e_n1": np.uint32, "e_x0": np.float64, "e_x1": np.float64, "e_x2": np.float64,
                "e_x3": np.float64, "w0_x0": np.float64, "w0_x1": np.float64, "w0_x2": np.float64,
                "w1_x0": np.float64, "w1_x1": np.float64, "w1_x2": np.float64, "n_x0": np.float64,
                "fc_n0": np.uint32, "fc_n1": np.uint32, "fc_n2": np.uint32, "t_i": np.uint32,
                "s_nptr": np.uint32, "s_x0": np.float64, "f_nptr": np.uint32, "f_x2": np.float64}
KERNEL_ARRAYS = ["r", "grad_x0", "grad_x1", "grad_x2", "cdt", "a_fill", "a_ilu", "trsv_x", "jv_w", "jv_y",
                 "forces", "cdt2", "a_fill2", "a_ilu2", "trsv_x2"]
KERNEL_SEED = 7


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).view(np.uint8).tobytes()).hexdigest()


def run(args):
    subprocess.run([KREF] + args, check=True, stdout=subprocess.DEVNULL)


def golden_for(mesh, name):
    out = {"name": name}
    with tempfile.TemporaryDirectory() as td:
        run(["setup", mesh, td + "/s"])
        out["setup"] = {k: sha(np.fromfile(td + "/s/%s.bin" % k, dtype=dt)) for k, dt in SETUP_ARRAYS.items()}
        out["phy_uv"] = [float(x) for x in np.fromfile(td + "/s/phy_uv.bin")]
        run(["kernels", mesh, td + "/k", str(KERNEL_SEED)])
        ker = {}
        for k in KERNEL_ARRAYS:
            a = np.fromfile(td + "/k/%s.bin" % k)
            ker[k] = {"sha256": sha(a), "norm": float(np.linalg.norm(a)), "head": [float(x) for x in a[:4]]}
        ker["jv_qnorm"] = float(np.fromfile(td + "/k/jv_qnorm.bin")[0])
        ker["forces_values"] = [float(x) for x in np.fromfile(td + "/k/forces.bin")]
        run(["kernels", mesh, td + "/z", str(KERNEL_SEED), "zerow"])
        a = np.fromfile(td + "/z/r.bin")
        ker["r_first_order"] = {"sha256": sha(a), "norm": float(np.linalg.norm(a))}
        out["kernels"] = ker
        os.makedirs(td + "/h")
        with open(td + "/h/stdout.txt", "w") as f:
            subprocess.run([KREF, "solve", mesh, td + "/h", "1"], check=True, stdout=f)
        hist = open(td + "/h/history.txt").read().split("\n")
        krylov = [int(m) for m in re.findall(r"TOTAL ITERATIONS (\d+)", open(td + "/h/stdout.txt").read())]
        fn = [float(l.split()[1]) for l in hist if l.startswith("fnorm")]
        cf = [float(l.split()[1]) for l in hist if l.startswith("cfl")]
        fo = [[float(x) for x in l.split()[1:]] for l in hist if l.startswith("forces")]
        out["solve"] = [{"step": i, "cfl": cf[i], "fnorm": fn[i], "forces": fo[i], "krylov": krylov[i]}
                        for i in range(len(fo))]
    return out


def main():
    if not os.path.exists(KREF):
        sys.exit("build oracle/_ref/kref first: make -C oracle ref")
    import kfun3d_amd as K
    with tempfile.TemporaryDirectory() as td:
        mesh_a = os.path.join(td, "uns3dA.msh")
        with gzip.open(os.path.join(HERE, "uns3dA.msh.gz"), "rb") as s, open(mesh_a, "wb") as d:
            shutil.copyfileobj(s, d)
        c1 = K.write_lattice(os.path.join(td, "c1.msh"), 48, 16, 16, 0)
        c1_sha = hashlib.sha256(open(c1, "rb").read()).hexdigest()
        for path, name in [(mesh_a, "mesh_a"), (c1, "c1")]:
            g = golden_for(path, name)
            if name == "c1":
                g["msh_sha256"] = c1_sha
            with open(os.path.join(HERE, "ref_%s.json" % name), "w") as f:
                json.dump(g, f, indent=1)
            print("wrote", name, len(g["solve"]), "steps")


if __name__ == "__main__":
    main()
