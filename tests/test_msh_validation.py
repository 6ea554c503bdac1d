"""The .msh reader rejects malformed files with a message instead of letting
an out-of-range index reach host arrays or device kernels.

The layout is the reference's (src/mesh.c:22-109, src/edges.c:22-87,
src/nodes.c:19-56, src/boundaries.c:105-204): a 13-word big-endian header
[cells, N, E, solid groups, 0, free groups, solid faces, 0, free faces,
solid nodes, 0, free nodes, 0], then n0[E] n1[E] (1-based), four double
blocks of E, four double blocks of N, then per boundary section a group
table of 2 x groups words, four face blocks, the node ids and three normal
blocks.  Each case below overwrites one index of a valid lattice file and
runs the host setup in a child process (the library aborts, as the
reference's asserts do).
"""
import os
import struct
import subprocess
import sys

import pytest

import kfun3d_amd as K
from conftest import ROOT, lattice


def _layout(path):
    with open(path, "rb") as f:
        h = struct.unpack(">13I", f.read(52))
    N, E, gs, gf, S, F, Ns, Nf = h[1], h[2], h[3], h[5], h[6], h[8], h[9], h[11]
    solid = 52 + 8 * E + 32 * E + 32 * N
    free = solid + 8 * gs + 16 * S + 4 * Ns + 24 * Ns
    return dict(N=N, E=E, S=S, F=F, Ns=Ns, Nf=Nf, edges=52, solid_faces=solid + 8 * gs,
                solid_nodes=solid + 8 * gs + 16 * S, free_faces=free + 8 * gf, free_nodes=free + 8 * gf + 16 * F)


def _setup_in_child(path):
    code = ("import sys; sys.path.insert(0, %r); import kfun3d_amd as K; m = K.Mesh(%r, device=False); "
            "print('N', m.N); m.close()" % (ROOT, path))
    return subprocess.run([sys.executable, "-c", code], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=120)


def _corrupt(src, dst, offset, value):
    data = bytearray(open(src, "rb").read())
    data[offset:offset + 4] = struct.pack(">I", value)
    open(dst, "wb").write(bytes(data))


@pytest.fixture(scope="module")
def small(mesh_dir):
    return lattice(mesh_dir, 7, 6, 5, 3)


def test_valid_file_reads(small):
    r = _setup_in_child(small)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["N", str(7 * 6 * 5)]


@pytest.mark.parametrize("case", ["edge_end_high", "edge_end_zero", "solid_corner", "solid_node_high",
                                  "free_corner_zero", "free_node_high"])
def test_out_of_range_index_is_rejected(case, small, tmp_path):
    L = _layout(small)
    off, val, msg = {
        "edge_end_high": (L["edges"] + 4 * (L["E"] + 3), L["N"] + 1, "edge 3 end 1 is vertex"),
        "edge_end_zero": (L["edges"] + 4 * 5, 0, "edge 5 end 0 is vertex 0"),
        "solid_corner": (L["solid_faces"] + 4 * (L["S"] + 2), L["Ns"] + 1, "solid face 2 corner 1"),
        "solid_node_high": (L["solid_nodes"] + 4 * 4, L["N"] + 7, "solid node 4 is vertex"),
        "free_corner_zero": (L["free_faces"] + 4 * (2 * L["F"] + 1), 0, "free-stream face 1 corner 2"),
        "free_node_high": (L["free_nodes"] + 4 * (L["Nf"] - 1), 0xFFFFFFFF, "free-stream node %d" % (L["Nf"] - 1)),
    }[case]
    bad = str(tmp_path / ("%s.msh" % case))
    _corrupt(small, bad, off, val)
    r = _setup_in_child(bad)
    assert r.returncode != 0
    assert msg in r.stderr, r.stderr


def test_truncated_file_is_rejected(small, tmp_path):
    bad = str(tmp_path / "truncated.msh")
    open(bad, "wb").write(open(small, "rb").read()[:-8])
    r = _setup_in_child(bad)
    assert r.returncode != 0
    assert "truncated mesh file" in r.stderr, r.stderr


def test_trailing_bytes_are_rejected(small, tmp_path):
    bad = str(tmp_path / "long.msh")
    open(bad, "wb").write(open(small, "rb").read() + b"\0" * 8)
    r = _setup_in_child(bad)
    assert r.returncode != 0
    assert "bytes parsed but the file holds" in r.stderr, r.stderr
