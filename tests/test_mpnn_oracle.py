"""The CPU oracle of the learned preconditioner's network (oracle/mpnn.py):
its hand-written backward pass against central finite differences, its
scale equivariance, and the message sum against a per-edge loop, on the
reference's mesh A.  These pin the oracle the GPU kernels are tested with."""
import numpy as np
import pytest

import kfun3d_amd as K
from oracle import mpnn as P


@pytest.fixture(scope="module")
def graph(mesh_a):
    m = K.Mesh(mesh_a, device=False)
    h = m.host_arrays()
    g = P.Graph(h["e_n0"], h["e_n1"], h["e_x0"], h["e_x1"], h["e_x2"], h["e_x3"], m.N)
    m.close()
    return g


def _data(g, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(4 * g.N)
    D = rng.standard_normal(16 * g.N) * 3.0
    return v, D


def test_gradient_matches_finite_differences(graph):
    """Every parameter block, along a random direction, and all of them
    together: central differences with the loss O(1), so that rounding
    stays far below the difference."""
    g = graph
    theta = P.init_params(seed=1, scale=1.0).astype(np.float64)
    theta[P.NPARAM - 132:] *= 100.0    # a decoder large enough that delta shapes the loss
    v, D = _data(g, 2)
    rng = np.random.default_rng(3)
    target = rng.standard_normal(4 * g.N)
    base = np.zeros(4 * g.N)
    loss, grad = P.loss_and_grad(theta, g, v, D, base, target)
    assert 0.5 < loss < 50.0, loss
    blocks, o = [], 0
    for name, shape in P.PARAM_SHAPES:
        n = int(np.prod(shape))
        d = np.zeros(P.NPARAM)
        d[o:o + n] = rng.standard_normal(n)
        blocks.append((name, d))
        o += n
    blocks.append(("all", rng.standard_normal(P.NPARAM)))
    for name, d in blocks:
        d = d / np.linalg.norm(d)
        eps = 1e-5
        lp, _ = P.loss_and_grad(theta + eps * d, g, v, D, base, target)
        lm, _ = P.loss_and_grad(theta - eps * d, g, v, D, base, target)
        fd = (lp - lm) / (2 * eps)
        an = float(grad @ d)
        assert abs(fd - an) <= 1e-4 * abs(an) + 1e-9, (name, fd, an)


def test_scale_equivariance(graph):
    g = graph
    theta = P.init_params(seed=4).astype(np.float64)
    v, D = _data(g, 5)
    d1 = P.forward(theta, g, v, D)
    d2 = P.forward(theta, g, 7.0 * v, D)
    # v and 7 v give inputs equal up to the rounding of the scale
    assert np.max(np.abs(d2 - 7.0 * d1)) <= 1e-12 * np.max(np.abs(7.0 * d1))


def test_message_sum_matches_edge_loop(graph):
    """m_v summed per vertex equals a loop over the edges scattering to both
    ends (the incidence lists hold each edge twice, signs opposite)."""
    g = graph
    assert len(g.owner) % 2 == 0
    E = len(g.owner) // 2
    # each edge appears once with + and once with - sign, at its two end points
    plus = g.sign > 0
    assert plus.sum() == E
    np.testing.assert_array_equal(np.sort(g.eid[plus]), np.arange(E))
    np.testing.assert_array_equal(np.sort(g.eid[~plus]), np.arange(E))
    # ascending edge order within each vertex
    same = g.owner[1:] == g.owner[:-1]
    assert np.all(g.eid[1:][same] > g.eid[:-1][same])
