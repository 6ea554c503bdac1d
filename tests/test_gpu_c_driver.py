"""The library driven from C as the reference's own driver drives its path
(tests/c/kf_main.c: Setup, Kernel, ComputeForces, then Compute2ndNorm,
ComputeAXPY, ComputeNewAXPY and Normalize on device vectors, built only
against include/kfun3d.h), on the reference's mesh uns3dA.

  * seq (index-order sums): the lines Kernel prints equal what the
    reference prints for its own run (src/kernel.c:198-212, "%g" of the
    golden history), and the final forces equal the golden's bit for bit;
  * tree (the default reductions): the same Krylov counts, final forces
    within 1e-8;
  * the BLAS-1 results equal their exact restatements: a x + y rounded as
    the product, then the sum (src/utils.c:67-103); the norm equals the
    oracle's sum in the same order (sequential, or the device's tree); the
    normalised vector is x * (1.0 / |x|) (src/utils.c:106-123).
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "tests", "c", "kf_main")


def _run(mesh, mode, tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/c/kf_main is not built (python -c 'import __graft_entry__ as g; g.build()')")
    out = str(tmp_path / ("%s.bin" % mode))
    r = subprocess.run([DRIVER, mesh, mode, out], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=300)
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    lines = r.stdout.splitlines()
    vals = {ln.split()[0]: [float(x) for x in ln.split()[1:]] for ln in lines
            if ln.split() and ln.split()[0] in ("FORCES", "NORM2", "NORMALIZE", "ERROR")}
    raw = np.fromfile(out, dtype=np.float64)
    q, y0, y1, w1 = raw.reshape(4, -1)
    return lines, vals, (q, y0, y1, w1)


def _golden():
    with open(os.path.join(GOLDEN, "ref_mesh_a.json")) as f:
        return json.load(f)["solve"]


def _reference_lines(ref):
    """what the reference prints per step (src/kernel.c:198-212)"""
    out = []
    for s in ref:
        out.append(">> GMRES: TOTAL ITERATIONS %d" % s["krylov"])
        out.append("STEP: %d CFL: %g NORM: %g" % (s["step"], s["cfl"], s["fnorm"]))
        out.append("LIFT: %g DRAG: %g MOMN: %g" % tuple(s["forces"]))
    return out


def _check_blas(vals, arrays, seq):
    q, y0, y1, w1 = arrays
    y_exp = 0.37 * q + y0            # numpy: the product, then the sum (no contraction)
    assert np.array_equal(y1.view(np.uint64), y_exp.view(np.uint64))
    w_exp = -1.25 * q + y_exp
    nrm = O.norm(w_exp) if seq else float(np.sqrt(O.dot_tree(w_exp, w_exp)))
    assert vals["NORM2"][0] == nrm, (vals["NORM2"][0], nrm)
    assert vals["NORMALIZE"][0] == nrm
    w_norm = w_exp * (1.0 / nrm)
    assert np.array_equal(w1.view(np.uint64), w_norm.view(np.uint64))


def test_c_driver_reference_order(mesh_a, tmp_path):
    ref = _golden()
    lines, vals, arrays = _run(mesh_a, "seq", tmp_path)
    assert vals["ERROR"] == [0.0]
    printed = [ln for ln in lines if ln.startswith((">> GMRES", "STEP:", "LIFT:"))]
    assert printed == _reference_lines(ref)
    assert vals["FORCES"] == ref[-1]["forces"]
    _check_blas(vals, arrays, seq=True)


def test_c_driver_default_order(mesh_a, tmp_path):
    ref = _golden()
    lines, vals, arrays = _run(mesh_a, "tree", tmp_path)
    assert vals["ERROR"] == [0.0]
    kry = [int(ln.split()[-1]) for ln in lines if ln.startswith(">> GMRES: TOTAL ITERATIONS")]
    assert kry == [s["krylov"] for s in ref]
    f, f0 = np.array(vals["FORCES"]), np.array(ref[-1]["forces"])
    assert np.all(np.abs(f - f0) <= 1e-8 * np.abs(f0)), (f, f0)
    _check_blas(vals, arrays, seq=False)
