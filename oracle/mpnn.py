"""CPU oracle of the learned preconditioner's network -- TEST INFRASTRUCTURE ONLY.

A float64 numpy restatement of the edge message-passing network whose HIP
forward, backward and Adam kernels live in kfun3d_amd/csrc/mpnn.hip; the GPU
tests compare against it (outputs and parameter gradients).  Nothing in the
product imports it.

Network (per vertex v; u runs over the neighbours of v in ascending edge
order, e_vu = +-(nx, ny, nz) * area of the edge, + when v is the edge's
first vertex):
    x_v   = [v_v / s, D_v / sd]                       20 features
    h0_v  = relu(W0 x_v + b0)                          32
    round l = 1..3:
      m_v = mean_u relu(A_l h_v + B_l h_u + C_l e_vu + c_l)
      h_v = h_v + relu(U_l h_v + V_l m_v + u_l)
    delta_v = s * (Wo h_v + bo)                        4
where v is the vector being preconditioned (4 per vertex), D_v the ILU
factor's diagonal block (16 values as stored: the inverse of U(v,v)),
s = RMS of v over all entries and sd = RMS of all the D blocks.

Parameters are one flat float32 vector in the order of PARAM_SHAPES.

bf16=True rounds both operands of every dense layer (W0, A, B, U, V; the
small C, Wo products stay exact) to bfloat16 (round to nearest even), as the
GPU's matrix cores consume them; the sums themselves stay exact, so the GPU
differs from this mode only by its FP32 summation order.
"""
import numpy as np

H = 32
NIN = 20
NROUND = 3

PARAM_SHAPES = [("W0", (H, NIN)), ("b0", (H,))]
for _l in range(NROUND):
    PARAM_SHAPES += [("A%d" % _l, (H, H)), ("B%d" % _l, (H, H)), ("C%d" % _l, (H, 3)), ("c%d" % _l, (H,)),
                     ("U%d" % _l, (H, H)), ("V%d" % _l, (H, H)), ("u%d" % _l, (H,))]
PARAM_SHAPES += [("Wo", (4, H)), ("bo", (4,))]
NPARAM = sum(int(np.prod(s)) for _, s in PARAM_SHAPES)


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


def unpack(theta):
    out, o = {}, 0
    for name, shape in PARAM_SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(theta[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], dtype=np.float64).ravel() for name, _ in PARAM_SHAPES])


def init_params(seed=0, scale=1.0):
    """Deterministic initial parameters (He-style scaled normals, zero
    biases; the decoder starts small so the untrained correction is tiny)."""
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
    p["Wo"] *= 0.01
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


def features(g, v, D):
    """x (N, 20) and the two scales."""
    v = np.asarray(v, dtype=np.float64).reshape(g.N, 4)
    D = np.asarray(D, dtype=np.float64).reshape(g.N, 16)
    s = np.sqrt(np.mean(v * v))
    sd = np.sqrt(np.mean(D * D))
    s = s if s > 0 else 1.0
    sd = sd if sd > 0 else 1.0
    return np.concatenate([v / s, D / sd], axis=1), s, sd


def forward(theta, g, v, D, keep=False, round_bf16=False):
    """delta (N, 4); with keep, also the activations backward needs."""
    p = unpack(theta)
    x, s, sd = features(g, v, D)
    z0 = _mm(x, p["W0"], round_bf16) + p["b0"]
    h = np.maximum(z0, 0.0)
    cache = {"x": x, "s": s, "z0": z0, "rounds": []}
    for l in range(NROUND):
        P = _mm(h, p["A%d" % l], round_bf16)
        Q = _mm(h, p["B%d" % l], round_bf16)
        Ce = g.e @ p["C%d" % l].T
        zm = P[g.owner] + Q[g.nbr] + Ce + p["c%d" % l]
        mm = np.maximum(zm, 0.0)
        m = np.zeros((g.N, H))
        np.add.at(m, g.owner, mm)
        m *= g.inv_deg[:, None]
        zu = _mm(h, p["U%d" % l], round_bf16) + _mm(m, p["V%d" % l], round_bf16) + p["u%d" % l]
        cache["rounds"].append({"h": h, "P": P, "Q": Q, "zm": zm, "m": m, "zu": zu})
        h = h + np.maximum(zu, 0.0)
    cache["h"] = h
    delta = s * (h @ p["Wo"].T + p["bo"])
    return (delta, cache) if keep else delta


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


def backward(theta, g, cache, ddelta, round_bf16=False):
    """Gradient of sum(ddelta * delta) with respect to theta (flat, float64).
    round_bf16 rounds the operands of the dense-layer products (the weight
    gradients and the data gradients through W0, A, B, U, V) to bfloat16;
    the decoder's data gradient, the message backward and the bias sums stay
    exact."""
    p = unpack(theta)
    rb = round_bf16
    gr = {name: np.zeros(shape) for name, shape in PARAM_SHAPES}
    dout = np.asarray(ddelta, dtype=np.float64).reshape(g.N, 4) * cache["s"]
    gr["Wo"] = _wg(dout, cache["h"], rb)
    gr["bo"] = dout.sum(axis=0)
    dh = dout @ p["Wo"]
    for l in reversed(range(NROUND)):
        rc = cache["rounds"][l]
        h, zm, m, zu = rc["h"], rc["zm"], rc["m"], rc["zu"]
        dzu = dh * (zu > 0)
        gr["U%d" % l] = _wg(dzu, h, rb)
        gr["V%d" % l] = _wg(dzu, m, rb)
        gr["u%d" % l] = dzu.sum(axis=0)
        dhn = dh + _dg(dzu, p["U%d" % l], rb)
        dm = _dg(dzu, p["V%d" % l], rb) * g.inv_deg[:, None]
        dzm = dm[g.owner] * (zm > 0)
        gr["C%d" % l] = dzm.T @ g.e
        gr["c%d" % l] = dzm.sum(axis=0)
        dP = np.zeros((g.N, H))
        np.add.at(dP, g.owner, dzm)
        dQ = np.zeros((g.N, H))
        np.add.at(dQ, g.nbr, dzm)
        gr["A%d" % l] = _wg(dP, h, rb)
        gr["B%d" % l] = _wg(dQ, h, rb)
        dh = dhn + _dg(dP, p["A%d" % l], rb) + _dg(dQ, p["B%d" % l], rb)
    dz0 = dh * (cache["z0"] > 0)
    gr["W0"] = _wg(dz0, cache["x"], rb)
    gr["b0"] = dz0.sum(axis=0)
    return pack(gr)


def loss_and_grad(theta, g, v, D, base, target, round_bf16=False):
    """Training loss of one pair: |base + delta - target|^2 / |target|^2,
    where base = ILU(0)^-1 v; and its gradient."""
    delta, cache = forward(theta, g, v, D, keep=True, round_bf16=round_bf16)
    res = np.asarray(base).reshape(g.N, 4) + delta - np.asarray(target).reshape(g.N, 4)
    tn = float(np.sum(np.asarray(target) ** 2))
    loss = float(np.sum(res * res)) / tn
    return loss, backward(theta, g, cache, 2.0 * res / tn, round_bf16=round_bf16)


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
