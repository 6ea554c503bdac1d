"""The CPU oracle of the learned preconditioner's network (oracle/mpnn.py):
its hand-written backward pass against central finite differences (all
three losses), linearity of the correction in z and its invariance to the scale
of the diagonal blocks, the vectorised correction against a per-vertex loop,
the incidence structure, bfloat16 rounding against torch and the Adam step
against torch.optim.Adam, on the reference's mesh A.  These pin the oracle
the GPU kernels are tested with."""
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


def _random_bsr(g, seed):
    """a random 4 x 4-block matrix on the mesh's vertex-edge pattern,
    diagonally weighted"""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    rows = np.concatenate([g.owner, np.arange(g.N)])
    cols = np.concatenate([g.nbr, np.arange(g.N)])
    r4 = np.repeat(4 * rows, 16) + np.tile(np.repeat(np.arange(4), 4), rows.size)
    c4 = np.repeat(4 * cols, 16) + np.tile(np.tile(np.arange(4), 4), cols.size)
    vals = rng.standard_normal(r4.size) + 8.0 * (r4 == c4)
    return sp.csr_matrix((vals, (r4, c4)), shape=(4 * g.N, 4 * g.N))


@pytest.mark.parametrize("metric", ["euclid", "a", "residual"])
def test_gradient_matches_finite_differences(graph, metric):
    """Every parameter block, along a random direction, and all of them
    together: central differences with the loss O(1), so that rounding
    stays far below the difference.  All three losses: |e|^2 / |target|^2,
    |A e|^2 / |v|^2 and |A (base + delta) - v|^2 / |v|^2 with a random block
    matrix on the mesh pattern."""
    g = graph
    A = _random_bsr(g, 12) * 0.05 if metric != "euclid" else None
    theta = P.init_params(seed=1, head_scale=3.0).astype(np.float64)
    v, D = _data(g, 2)
    rng = np.random.default_rng(3)
    target = rng.standard_normal(4 * g.N) if metric != "residual" else None
    base = rng.standard_normal(4 * g.N) * 0.5
    loss, grad = P.loss_and_grad(theta, g, v, D, base, target, A=A)
    assert 0.05 < loss < 50.0, loss
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
        eps = 2e-6   # small enough that few ReLU kinks fall inside the step
        lp, _ = P.loss_and_grad(theta + eps * d, g, v, D, base, target, A=A)
        lm, _ = P.loss_and_grad(theta - eps * d, g, v, D, base, target, A=A)
        fd = (lp - lm) / (2 * eps)
        an = float(grad @ d)
        assert abs(fd - an) <= 1e-4 * abs(an) + 1e-9, (name, fd, an)


def test_correction_is_linear_in_z(graph):
    """M v = z + G z with G independent of v: delta(z1 + 2 z2) =
    delta(z1) + 2 delta(z2); scaling the diagonal blocks changes nothing
    (they enter over their RMS)."""
    g = graph
    theta = P.init_params(seed=4, head_scale=3.0).astype(np.float64)
    _, D = _data(g, 5)
    rng = np.random.default_rng(9)
    z1, z2 = rng.standard_normal(4 * g.N), rng.standard_normal(4 * g.N)
    d1, d2 = P.forward(theta, g, z1, D), P.forward(theta, g, z2, D)
    d12 = P.forward(theta, g, z1 + 2.0 * z2, D)
    scale = np.abs(d12).max()
    assert np.abs(d12 - (d1 + 2.0 * d2)).max() <= 1e-12 * scale
    d5 = P.forward(theta, g, z1, 5.0 * D)
    assert np.abs(d5 - d1).max() <= 1e-12 * np.abs(d1).max()
    # zero heads: G = 0 (training starts from ILU(0))
    th0 = P.init_params(seed=4).astype(np.float64)
    assert np.all(P.forward(th0, g, z1, D) == 0.0)


def test_apply_matches_vertex_loop(graph):
    """The vectorised correction against a loop over vertices and their
    incidences for the first 200 vertices."""
    g = graph
    theta = P.init_params(seed=6, head_scale=3.0).astype(np.float64)
    _, D = _data(g, 7)
    z = np.random.default_rng(10).standard_normal((g.N, 4))
    op1, opq, _ = P.context(theta, g, D)
    delta = P.apply(theta, g, op1, opq, z)
    We = P.unpack(theta)["We"]
    start = np.searchsorted(g.owner, np.arange(g.N))
    stop = np.searchsorted(g.owner, np.arange(g.N), side="right")
    for v in range(200):
        acc = op1[v, :16].reshape(4, 4) @ z[v]
        s = np.zeros(4)
        for k in range(start[v], stop[v]):
            u = g.nbr[k]
            T = (op1[v, 16:] + opq[u] + We @ g.e[k]).reshape(4, 4)
            s += T @ z[u]
        acc = acc + s / max(stop[v] - start[v], 1)
        np.testing.assert_allclose(delta[v], acc, rtol=1e-12, atol=1e-12 * np.abs(acc).max())


def test_message_sum_matches_edge_loop(graph):
    """The incidence lists hold each edge twice, signs opposite, in
    ascending edge order per vertex."""
    g = graph
    assert len(g.owner) % 2 == 0
    E = len(g.owner) // 2
    plus = g.sign > 0
    assert plus.sum() == E
    np.testing.assert_array_equal(np.sort(g.eid[plus]), np.arange(E))
    np.testing.assert_array_equal(np.sort(g.eid[~plus]), np.arange(E))
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
    theta = P.init_params(seed=6, head_scale=3.0).astype(np.float64)
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


def test_library_initial_parameters():
    """kf_mpnn_init_params: the oracle's layout, normals at sqrt(2 / fan-in)
    (U and V halved), zero biases and zero operator heads (G = 0)."""
    th = K.init_params(seed=5).astype(np.float64)
    assert th.size == P.NPARAM
    assert np.all(th[-P.HEADS:] == 0.0)
    p = P.unpack(th)
    for name, shape in P.PARAM_SHAPES[:-4]:
        if len(shape) == 1:
            assert np.all(p[name] == 0.0), name
            continue
        want = np.sqrt(2.0 / shape[1]) * (0.5 if name[0] in "UV" else 1.0)
        assert abs(p[name].std() / want - 1.0) < 0.35, (name, p[name].std(), want)
    assert not np.array_equal(th, K.init_params(seed=6).astype(np.float64))
