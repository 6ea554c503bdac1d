"""The C-ABI library loads without a GPU and exports every function that
include/kfun3d.h declares; the host-only setup runs on the CPU."""
import ctypes
import os
import re
import subprocess

import kfun3d_amd as K
from conftest import ROOT


def declared_functions():
    text = open(os.path.join(ROOT, "include", "kfun3d.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set()
    for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M):
        head = m.group(0)
        if head.startswith(("typedef", "struct", "#")):
            continue
        names.add(m.group(1))
    return sorted(names)


def test_header_declares_reference_entry_points():
    names = declared_functions()
    for ref in ["ComputeA", "ComputeTimeStep", "ComputeFlux", "ComputeForces", "ComputeNumericalILU",
                "ComputeSparseTriangularSolve", "iJacVec", "Kernel", "Compute2ndNorm", "ComputeAXPY",
                "ComputeNewAXPY", "Normalize", "Setup"]:
        assert ref in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(K.LIBPATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", K.LIBPATH], stdout=subprocess.PIPE, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(declared_functions()) <= exported
    assert set(K.ABI_SYMBOLS) <= exported


def test_host_setup_runs_without_gpu(mesh_c1):
    m = K.Mesh(mesh_c1, device=False)
    assert (m.N, m.E) == (12288, 79007)
    h = m.host_arrays()
    assert len(h["fc_n0"]) == 1410 and len(h["s_nptr"]) == 768 and len(h["f_nptr"]) == 2628
    # edges come out sorted by n0 with n0 < n1 (the reference's index sort is the identity)
    assert (h["e_n0"] < h["e_n1"]).all() and (h["e_n0"][1:] >= h["e_n0"][:-1]).all()
    m.close()
