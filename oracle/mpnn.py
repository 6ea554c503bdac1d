"""CPU oracle of the learned preconditioner's network -- TEST INFRASTRUCTURE ONLY.

A float64 numpy restatement of the edge message-passing network whose HIP
forward, backward and Adam kernels live in kfun3d_amd/csrc/mpnn.hip; the GPU
tests compare against it (activations, corrections, losses and parameter
gradients).  Nothing in the product imports it.

The network produces a sparse linear correction operator G from the ILU(0)
factor's diagonal blocks and the mesh geometry; the preconditioner is
M v = z + G z with z = ILU(0)^-1 v, so M is linear in v.

Context (per vertex v; u runs over the neighbours of v in ascending edge
order, e_vu = +-(nx, ny, nz) * area of the edge, + when v is the edge's
first vertex):
    x_v   = D_v / sd                                  16 features
    h0_v  = relu(W0 x_v + b0)                          32
    round l = 1..3:
      m_v = mean_u relu(A_l h_v + B_l h_u + C_l e_vu + c_l)
      h_v = h_v + relu(U_l h_v + V_l m_v + u_l)
Operator heads (16 values = one 4 x 4 block, row-major):
    [S_v; P_v] = WSA h_v + bSA           S_v the vertex block, P_v (with its
                                         bias) the owner part of the edge blocks
    Q_v        = Wb h_v                  the neighbour part
    T_vu       = P_v + Q_u + We e_vu
Correction:
    delta_v = S_v z_v + (1 / deg v) sum_u T_vu z_u
D_v is the diagonal block as stored (the inverse of U(v,v)), sd the RMS of
all the D blocks.  Parameters are one flat float32 vector in the order of
PARAM_SHAPES.

round_bf16=True rounds both operands of every dense-layer product (W0, A, B,
U, V, WSA, Wb and their weight and data gradients) to bfloat16 (round to
nearest even), as the GPU's matrix cores consume them; the sums themselves
stay exact, so the GPU differs from this mode only by its FP32 summation
order.
"""
import numpy as np

H = 32
NIN = 16
NROUND = 3

PARAM_SHAPES = [("W0", (H, NIN)), ("b0", (H,))]
for _l in range(NROUND):
    PARAM_SHAPES += [("A%d" % _l, (H, H)), ("B%d" % _l, (H, H)), ("C%d" % _l, (H, 3)), ("c%d" % _l, (H,)),
                     ("U%d" % _l, (H, H)), ("V%d" % _l, (H, H)), ("u%d" % _l, (H,))]
PARAM_SHAPES += [("WSA", (32, H)), ("bSA", (32,)), ("Wb", (16, H)), ("We", (16, 3))]
NPARAM = sum(int(np.prod(s)) for _, s in PARAM_SHAPES)
HEADS = 32 * H + 32 + 16 * H + 16 * 3      # trailing parameters: the operator heads


def bf16(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _mm(x, w, round_bf16):
    """x @ w.T with both operands rounded to bfloat16 when asked."""
    if round_bf16:
        return bf16(x) @ bf16(w).T
    return x @ w.T


def _wg(dy, x, round_bf16):
    """dy.T @ x (a weight gradient), operands rounded to bfloat16 when asked."""
    if round_bf16:
        return bf16(dy).T @ bf16(x)
    return dy.T @ x


def _dg(dy, w, round_bf16):
    """dy @ w (a data gradient through a dense layer), rounded likewise."""
    if round_bf16:
        return bf16(dy) @ bf16(w)
    return dy @ w


def unpack(theta):
    out, o = {}, 0
    for name, shape in PARAM_SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(theta[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], dtype=np.float64).ravel() for name, _ in PARAM_SHAPES])


def init_params(seed=0, scale=1.0, head_scale=0.0):
    """Deterministic initial parameters: He-style scaled normals and zero
    biases for the context network; the operator heads' normals scaled by
    0.1 head_scale (0: G = 0, so training starts from ILU(0) itself)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in PARAM_SHAPES:
        if len(shape) == 1:
            p[name] = np.zeros(shape)
        else:
            fan_in = shape[1]
            p[name] = rng.standard_normal(shape) * np.sqrt(2.0 / fan_in) * scale
    for l in range(NROUND):
        p["U%d" % l] *= 0.5
        p["V%d" % l] *= 0.5
    for name in ("WSA", "Wb", "We"):
        p[name] *= head_scale * 0.1
    return pack(p).astype(np.float32)


class Graph:
    """Incidence lists of a mesh in ascending edge order: for vertex v,
    (neighbour, sign) pairs; e[k] = sign * normal * area of that edge."""

    def __init__(self, n0, n1, nx, ny, nz, area, N):
        n0 = np.asarray(n0, dtype=np.int64)
        n1 = np.asarray(n1, dtype=np.int64)
        E = len(n0)
        owner = np.concatenate([n0, n1])
        nbr = np.concatenate([n1, n0])
        sign = np.concatenate([np.ones(E), -np.ones(E)])
        eid = np.concatenate([np.arange(E), np.arange(E)])
        order = np.lexsort((eid, owner))       # by vertex, then ascending edge id
        self.owner, self.nbr, self.sign, self.eid = owner[order], nbr[order], sign[order], eid[order]
        nrm = np.stack([np.asarray(nx), np.asarray(ny), np.asarray(nz)], axis=1) * np.asarray(area)[:, None]
        self.e = nrm[self.eid] * self.sign[:, None]
        self.N = N
        self.inv_deg = 1.0 / np.maximum(np.bincount(self.owner, minlength=N), 1)


def _rms(a):
    r = float(np.sqrt(np.mean(a * a)))
    return r if r > 0 and np.isfinite(r) else 1.0


def features(g, D):
    """x (N, 16): the diagonal blocks over their RMS."""
    D = np.asarray(D, dtype=np.float64).reshape(g.N, 16)
    return D / _rms(D)


def context(theta, g, D, round_bf16=False):
    """The operator coefficients (OP1 = [S | P + bias] (N, 32), Q (N, 16))
    and the activations backward needs."""
    p = unpack(theta)
    x = features(g, D)
    z0 = _mm(x, p["W0"], round_bf16) + p["b0"]
    h = np.maximum(z0, 0.0)
    cache = {"x": x, "z0": z0, "rounds": []}
    for l in range(NROUND):
        P = _mm(h, p["A%d" % l], round_bf16)
        Q = _mm(h, p["B%d" % l], round_bf16)
        Ce = g.e @ p["C%d" % l].T
        zm = P[g.owner] + Q[g.nbr] + Ce + p["c%d" % l]
        m = np.zeros((g.N, H))
        np.add.at(m, g.owner, np.maximum(zm, 0.0))
        m *= g.inv_deg[:, None]
        zu = _mm(h, p["U%d" % l], round_bf16) + _mm(m, p["V%d" % l], round_bf16) + p["u%d" % l]
        cache["rounds"].append({"h": h, "P": P, "Q": Q, "zm": zm, "m": m, "zu": zu})
        h = h + np.maximum(zu, 0.0)
    cache["h"] = h
    op1 = _mm(h, p["WSA"], round_bf16) + p["bSA"]
    opq = _mm(h, p["Wb"], round_bf16)
    return op1, opq, cache


def apply(theta, g, op1, opq, z):
    """delta (N, 4) = S_v z_v + (1/deg v) sum_u (P_v + Q_u + We e_vu) z_u"""
    p = unpack(theta)
    z = np.asarray(z, dtype=np.float64).reshape(g.N, 4)
    S = op1[:, :16].reshape(g.N, 4, 4)
    T = op1[g.owner, 16:] + opq[g.nbr] + g.e @ p["We"].T
    acc = np.zeros((g.N, 4))
    np.add.at(acc, g.owner, np.einsum("kij,kj->ki", T.reshape(-1, 4, 4), z[g.nbr]))
    return np.einsum("vij,vj->vi", S, z) + acc * g.inv_deg[:, None]


def forward(theta, g, z, D, round_bf16=False):
    """delta (N, 4) of z = ILU(0)^-1 v."""
    op1, opq, _ = context(theta, g, D, round_bf16)
    return apply(theta, g, op1, opq, z)


def backward(theta, g, cache, z, ddelta, round_bf16=False):
    """Gradient of sum(ddelta * delta) with respect to theta (flat, float64),
    for delta = apply(context(D), z)."""
    p = unpack(theta)
    rb = round_bf16
    gr = {name: np.zeros(shape) for name, shape in PARAM_SHAPES}
    z = np.asarray(z, dtype=np.float64).reshape(g.N, 4)
    gv = np.asarray(ddelta, dtype=np.float64).reshape(g.N, 4)
    idg = g.inv_deg[:, None]
    zs = np.zeros((g.N, 4))
    np.add.at(zs, g.owner, z[g.nbr])
    dS = np.einsum("vi,vj->vij", gv, z).reshape(g.N, 16)
    dP = np.einsum("vi,vj->vij", gv * idg, zs).reshape(g.N, 16)
    gbar = np.zeros((g.N, 4))
    np.add.at(gbar, g.nbr, (gv * idg)[g.owner])            # sum over v with u in N(v)
    dQ = np.einsum("vi,vj->vij", gbar, z).reshape(g.N, 16)
    # dWe[4i+j][k] = sum_v (g_v[i] / deg v) Ez_v[j][k], Ez_v = sum_u z_u e_vu^T
    Ez = np.zeros((g.N, 4, 3))
    np.add.at(Ez, g.owner, np.einsum("kj,kl->kjl", z[g.nbr], g.e))
    gr["We"] = _wg(gv * idg, Ez.reshape(g.N, 12), rb).reshape(16, 3)
    dop1 = np.concatenate([dS, dP], axis=1)
    h = cache["h"]
    gr["WSA"] = _wg(dop1, h, rb)
    gr["bSA"] = dop1.sum(axis=0)
    gr["Wb"] = _wg(dQ, h, rb)
    dh = _dg(dop1, p["WSA"], rb) + _dg(dQ, p["Wb"], rb)
    for l in reversed(range(NROUND)):
        rc = cache["rounds"][l]
        h, zm, m, zu = rc["h"], rc["zm"], rc["m"], rc["zu"]
        dzu = dh * (zu > 0)
        gr["U%d" % l] = _wg(dzu, h, rb)
        gr["V%d" % l] = _wg(dzu, m, rb)
        gr["u%d" % l] = dzu.sum(axis=0)
        dhn = dh + _dg(dzu, p["U%d" % l], rb)
        dm = _dg(dzu, p["V%d" % l], rb) * idg
        dzm = dm[g.owner] * (zm > 0)
        gr["C%d" % l] = dzm.T @ g.e
        gr["c%d" % l] = dzm.sum(axis=0)
        dPm = np.zeros((g.N, H))
        np.add.at(dPm, g.owner, dzm)
        dQm = np.zeros((g.N, H))
        np.add.at(dQm, g.nbr, dzm)
        gr["A%d" % l] = _wg(dPm, h, rb)
        gr["B%d" % l] = _wg(dQm, h, rb)
        dh = dhn + _dg(dPm, p["A%d" % l], rb) + _dg(dQm, p["B%d" % l], rb)
    dz0 = dh * (cache["z0"] > 0)
    gr["W0"] = _wg(dz0, cache["x"], rb)
    gr["b0"] = dz0.sum(axis=0)
    return pack(gr)


def loss_and_grad(theta, g, v, D, base, target, round_bf16=False, A=None):
    """Training loss of one pair and its gradient, with e = base + delta -
    target, base = ILU(0)^-1 v (the z the correction acts on):
    |e|^2 / |target|^2, or with A (a 4N x 4N matrix, the assembled Jacobian)
    |A e|^2 / |v|^2; target None with A: |A (base + delta) - v|^2 / |v|^2."""
    op1, opq, cache = context(theta, g, D, round_bf16)
    delta = apply(theta, g, op1, opq, base)
    tg = np.zeros(4 * g.N) if target is None else np.asarray(target)
    e = (np.asarray(base).reshape(g.N, 4) + delta - tg.reshape(g.N, 4)).ravel()
    if target is None:
        tn = float(np.sum(np.asarray(v) ** 2))
        y = A @ e - np.asarray(v)
        loss = float(y @ y) / tn
        dd = 2.0 * (A.T @ y) / tn
    elif A is None:
        tn = float(np.sum(np.asarray(target) ** 2))
        loss = float(e @ e) / tn
        dd = 2.0 * e / tn
    else:
        tn = float(np.sum(np.asarray(v) ** 2))
        y = A @ e
        loss = float(y @ y) / tn
        dd = 2.0 * (A.T @ y) / tn
    return loss, backward(theta, g, cache, base, dd.reshape(g.N, 4), round_bf16=round_bf16)


def adam(theta, m, v, grad, t, lr, b1=0.9, b2=0.999, eps=1e-8, gclip=0.0):
    """One Adam step in float32 (t counts from 1), with optional clipping of
    the gradient's global 2-norm to gclip; returns (theta, m, v)."""
    f = np.float32
    b1, b2, lr, eps, gclip = f(b1), f(b2), f(lr), f(eps), f(gclip)
    g = np.asarray(grad, dtype=f)
    gn = float(np.sqrt(np.sum(np.asarray(grad, dtype=np.float64) ** 2)))
    if gclip > 0 and gn > float(gclip):
        g = g * f(float(gclip) / gn)
    m = b1 * m + (f(1) - b1) * g
    v = b2 * v + (f(1) - b2) * (g * g)
    c1 = f(1.0 / (1.0 - float(b1) ** t))
    c2 = f(1.0 / (1.0 - float(b2) ** t))
    theta = theta - lr * (m * c1) / (np.sqrt(v * c2) + eps)
    return theta.astype(f), m.astype(f), v.astype(f)
