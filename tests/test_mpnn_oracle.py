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


def test_bf16_rounding_matches_torch():
    """The oracle's bfloat16 rounding (nearest, ties to even) is torch's."""
    import torch
    x = np.random.default_rng(0).standard_normal(100000).astype(np.float32) * np.float32(3.7)
    x[:8] = np.array([1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8), 2 ** -130, 3e38, 0.0, 1.00390625, 7.0],
                     dtype=np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(P.bf16(x), ref)


def test_adam_matches_torch():
    """The oracle's float32 Adam step against torch.optim.Adam (an
    independent arrangement of the same update), three steps."""
    import torch
    rng = np.random.default_rng(1)
    th0 = rng.standard_normal(1000).astype(np.float32)
    grads = [rng.standard_normal(1000).astype(np.float32) * s for s in (1.0, 0.1, 3.0)]
    p = torch.nn.Parameter(torch.from_numpy(th0.copy()))
    opt = torch.optim.Adam([p], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    th, m, v = th0.copy(), np.zeros(1000, np.float32), np.zeros(1000, np.float32)
    for t, gr in enumerate(grads, start=1):
        p.grad = torch.from_numpy(gr.copy())
        opt.step()
        th, m, v = P.adam(th, m, v, gr, t, 1e-3)
        np.testing.assert_allclose(th, p.detach().numpy(), rtol=0, atol=2e-7 * np.abs(th0).max())


def test_bf16_gradient_close_to_exact(graph):
    """Rounding the dense operands to bfloat16 moves the gradient by well
    under 5 percent per parameter block."""
    g = graph
    theta = P.init_params(seed=6).astype(np.float64)
    theta[P.NPARAM - 132:] *= 30.0
    v, D = _data(g, 7)
    rng = np.random.default_rng(8)
    target, base = rng.standard_normal(4 * g.N), rng.standard_normal(4 * g.N) * 0.5
    l1, g1 = P.loss_and_grad(theta, g, v, D, base, target)
    l2, g2 = P.loss_and_grad(theta, g, v, D, base, target, round_bf16=True)
    assert abs(l1 - l2) <= 1e-2 * l1
    o = 0
    for name, shape in P.PARAM_SHAPES:
        n = int(np.prod(shape))
        a, b = g1[o:o + n], g2[o:o + n]
        assert np.linalg.norm(a - b) <= 5e-2 * np.linalg.norm(a), name
        o += n
