"""The learned preconditioner's network on the GPU (kfun3d_amd/csrc/mpnn.hip)
against the CPU oracle (oracle/mpnn.py) on the reference's mesh A: the
context network's activations, the operator coefficients, the correction
G z, its linearity, the losses and parameter gradients, Adam and training.

The oracle's bf16 mode rounds both operands of every dense-layer product to
bfloat16 as the matrix cores do and keeps the sums exact; the GPU sums in
FP32.  What remains is FP32 summation order and, rarely, a value whose FP32
and FP64 versions round to neighbouring bfloat16 numbers.  Bars:
  * every activation and the output: 2-norm error <= 5e-4 relative, max
    error <= 5e-3 of the largest entry.  Measured: about 1e-4 and 2e-3 by
    the third round, the flips accumulating layer by layer; a transposed
    operand, a wrong lane or row index or a dropped term give errors of
    order 1e-1 or more;
  * loss (|e|^2 / |target|^2, the A-norm |A e|^2 / |v|^2 and the residual
    |A (base + G base) - v|^2 / |v|^2 with the assembled Jacobian) to 1e-5
    relative; gradient per parameter block to
    5e-3 relative in 2-norm, against the oracle's gradient of the same
    bf16-rounded network (the forward's rounding flips, and ReLU masks they
    move, carry into the backward); against the exact (unrounded) network's
    gradient, whose blocks bfloat16 rounding moves by up to about 10 percent
    with the A-norm loss, a cosine of at least 0.99 per block;
  * Adam: the oracle's float32 update applied to the GPU gradient, to 2 ulp
    (the clipping norm is a tree sum on the GPU);
  * repeated calls: bit-identical (every reduction has a fixed order).
"""
import numpy as np
import pytest

import kfun3d_amd as K
from oracle import mpnn as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(mesh_a):
    m = K.Mesh(mesh_a, K.options())
    h = m.host_arrays()
    g = P.Graph(h["e_n0"], h["e_n1"], h["e_x0"], h["e_x1"], h["e_x2"], h["e_x3"], m.N)
    theta = P.init_params(seed=11, head_scale=3.0)     # non-zero operator heads
    net = K.Network(m, theta, nslot=2)
    yield m, g, net, theta
    net.close()
    m.close()


def _data(N, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(4 * N) * 0.3
    D = rng.standard_normal(16 * N) * 2.0
    base = rng.standard_normal(4 * N) * 0.5
    target = rng.standard_normal(4 * N)
    return v, D, base, target


def _close(name, gpu, ref, rel=5e-4, mx=5e-3):
    gpu = np.asarray(gpu, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    scale = np.abs(ref).max()
    e2 = np.linalg.norm(gpu - ref) / max(np.linalg.norm(ref), 1e-300)
    emax = np.abs(gpu - ref).max() / max(scale, 1e-300)
    assert e2 <= rel and emax <= mx, "%s: 2-norm error %.3e, max error %.3e of max" % (name, e2, emax)


def test_forward_matches_oracle(setup):
    m, g, net, theta = setup
    _, D, z, _ = _data(m.N, 1)
    net.set_diag(D, slot=1)
    zd, od = K.DeviceVec(m.sz, z), K.DeviceVec(m.sz)
    net.forward(zd, od, slot=1)
    th = theta.astype(np.float64)
    op1, opq, cache = P.context(th, g, D, round_bf16=True)
    delta = P.apply(th, g, op1, opq, z)
    _close("features", net.activation(0), cache["x"], rel=1e-6, mx=1e-6)
    _close("z0", net.activation(1), cache["z0"])
    for l in range(P.NROUND):
        rc = cache["rounds"][l]
        for k, key in enumerate(["h", "P", "Q", "m", "zu"]):
            _close("round %d %s" % (l, key), net.activation(2 + 5 * l + k), rc[key])
    _close("h final", net.activation(17), cache["h"])
    _close("[S | P]", net.activation(18), op1)
    q = net.activation(19)
    _close("Q", q[:, :16], opq)
    assert np.all(q[:, 16:] == 0.0)
    _close("delta", od.numpy(), delta.ravel())
    # add mode accumulates into the output
    od2 = K.DeviceVec(m.sz, np.ones(m.sz))
    net.forward(zd, od2, slot=1, add=True)
    np.testing.assert_array_equal(od2.numpy(), 1.0 + od.numpy())


def test_correction_is_linear(setup):
    """G (z1 + 2 z2) = G z1 + 2 G z2 to FP32 rounding: the preconditioner
    stays a linear operator."""
    m, g, net, theta = setup
    _, D, _, _ = _data(m.N, 1)
    net.set_diag(D, slot=1)
    rng = np.random.default_rng(4)
    z1, z2 = rng.standard_normal(m.sz), rng.standard_normal(m.sz)
    outs = []
    for z in (z1, z2, z1 + 2.0 * z2):
        zd, od = K.DeviceVec(m.sz, z), K.DeviceVec(m.sz)
        net.forward(zd, od, slot=1)
        outs.append(od.numpy())
    d = outs[2] - (outs[0] + 2.0 * outs[1])
    assert np.linalg.norm(d) <= 1e-5 * np.linalg.norm(outs[2])


def test_forward_with_ilu_diagonal(setup):
    """kf_mpnn_take_diag reads the factor's diagonal blocks after ILU(0)."""
    m, g, net, theta = setup
    from conftest import perturbed_state
    m.set_state(perturbed_state(m.N, 3.0, 5))
    m.timestep(50.0)
    m.fill()
    m.ilu()
    net.take_diag(slot=0)
    rp, col, dg = m.pattern()
    D = m.blocks()[dg].reshape(-1)
    _, _, z, _ = _data(m.N, 2)
    zd, od = K.DeviceVec(m.sz, z), K.DeviceVec(m.sz)
    net.forward(zd, od, slot=0)
    delta = P.forward(theta.astype(np.float64), g, z, D, round_bf16=True)
    _close("delta (ILU diagonal)", od.numpy(), delta.ravel())


def _jacobian(m):
    """the assembled Jacobian of a perturbed state (before ILU) as a scipy
    BSR matrix, and the same blocks taken into network slot 1"""
    import scipy.sparse as sp
    from conftest import perturbed_state
    m.set_state(perturbed_state(m.N, 3.0, 7))
    m.timestep(20.0)
    m.fill()
    rp, col, _ = m.pattern()
    blocks = m.blocks().copy()
    # the network holds A in FP32: compare against the same rounding
    blocks = blocks.astype(np.float32).astype(np.float64)
    return sp.bsr_matrix((blocks, col.astype(np.int64), rp.astype(np.int64)), shape=(m.sz, m.sz)).tocsr()


@pytest.mark.parametrize("metric", ["euclid", "a", "residual"])
def test_gradient_matches_oracle(setup, metric):
    m, g, net, theta = setup
    v, D, base, target = _data(m.N, 3)
    net.set_diag(D, slot=1)
    A = None
    if metric != "euclid":
        A = _jacobian(m)
        net.take_matrix(1)
    net.set_loss(metric)
    vd, bd, td = K.DeviceVec(m.sz, v), K.DeviceVec(m.sz, base), K.DeviceVec(m.sz, target)
    loss, grad = net.loss_grad(vd, bd, td, slot=1)
    th = theta.astype(np.float64)
    tg = None if metric == "residual" else target
    lref, gref = P.loss_and_grad(th, g, v, D, base, tg, round_bf16=True, A=A)
    _, gexact = P.loss_and_grad(th, g, v, D, base, tg, A=A)
    assert abs(loss - lref) <= 1e-5 * abs(lref), (loss, lref)
    o = 0
    for name, shape in P.PARAM_SHAPES:
        n = int(np.prod(shape))
        a, b, c = grad[o:o + n].astype(np.float64), gref[o:o + n], gexact[o:o + n]
        nb = max(np.linalg.norm(b), 1e-30)
        assert np.linalg.norm(a - b) <= 5e-3 * nb, (name, np.linalg.norm(a - b) / nb)
        cos = float(a @ c) / max(np.linalg.norm(a) * np.linalg.norm(c), 1e-300)
        assert cos >= 0.99, (name, cos)
        o += n
    # deterministic: the same call gives the same bits
    loss2, grad2 = net.loss_grad(vd, bd, td, slot=1)
    net.set_loss("euclid")
    assert loss2 == loss
    np.testing.assert_array_equal(grad2.view(np.uint32), grad.view(np.uint32))


@pytest.mark.parametrize("gclip", [0.0, 1e-3])
def test_adam_matches_oracle(setup, gclip):
    m, g, net, theta = setup
    net.set_params(theta)
    net.set_loss("euclid")
    v, D, base, target = _data(m.N, 4)
    net.set_diag(D, slot=1)
    vd, bd, td = K.DeviceVec(m.sz, v), K.DeviceVec(m.sz, base), K.DeviceVec(m.sz, target)
    th = theta.copy()
    mo = np.zeros_like(th)
    vo = np.zeros_like(th)
    for t in (1, 2, 3):
        _, grad = net.loss_grad(vd, bd, td, slot=1)
        net.adam(1e-3, gclip=gclip)
        th, mo, vo = P.adam(th, mo, vo, grad, t, 1e-3, gclip=gclip)
        got = net.params()
        ulp = np.abs(got.view(np.int32).astype(np.int64) - th.view(np.int32).astype(np.int64))
        assert ulp.max() <= 2, (t, int(ulp.max()))
    net.set_params(theta)


def test_train_steps_reduce_loss(setup):
    """A few queued training steps on one pair lower its loss, and the loss
    log holds the loss before each step."""
    m, g, net, theta = setup
    net.set_params(theta)
    net.set_loss("euclid")
    net.reset_losses()
    v, D, base, target = _data(m.N, 5)
    net.set_diag(D, slot=1)
    vd, bd, td = K.DeviceVec(m.sz, v), K.DeviceVec(m.sz, base), K.DeviceVec(m.sz, target)
    l0, _ = net.loss_grad(vd, bd, td, slot=1)
    for _ in range(30):
        net.train_step(vd, bd, td, 1e-3, slot=1)
    log = net.losses()
    assert log.size == 30
    assert abs(log[0] - l0) <= 1e-6 * l0
    assert log[-1] < 0.98 * log[0], (log[0], log[-1])
    net.set_params(theta)
