"""CPU FP64 restatement of the learned closure's network and its gradient.

Test infrastructure only: the checker of the device kernels in
gabe_amd/csrc/gabe_closure.hip.  The network has no counterpart in the
reference; its definition is the one the C ABI documents:

    a1 = SiLU(conv3(x;  W1) + b1),  a2 = SiLU(conv3(a1; W2) + b2),
    u  = conv3(a2; W3) + b3,        y = lap lap u (lap u = 6 neighbours - 6 u),
    conv3(u; W)[s, co] = sum_t sum_ci W[co, ci, t] u[s + off(t), ci],
    off(t) = (t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) on a periodic N^3 grid,
    loss   = sum_{s,c} (lap u - target)^2 / (2 N^3),
the target being that of lap u (the force target carried through lap^-1).

With bf16=True the rounding points of the device path are reproduced (inputs,
weights, stored pre-activations, activations and back-propagated gradients
rounded to bfloat16, round to nearest even); arithmetic stays FP64 otherwise.
The gradient is pinned in the tests by central finite differences of the FP64
loss.
"""
import numpy as np

NIN, NH, NOUT, TAPS = 4, 16, 2, 27
SHAPES = [("W1", (NH, NIN, TAPS)), ("b1", (NH,)), ("W2", (NH, NH, TAPS)), ("b2", (NH,)),
          ("W3", (NOUT, NH, TAPS)), ("b3", (NOUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)
OFFSETS = [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(TAPS)]


def unpack(p):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def bf16_round(a):
    """float64 -> nearest bfloat16 (ties to even) -> float64, via float32."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def shifted(u, off):
    """u[s + off] for every site s (periodic)."""
    return np.roll(u, shift=(-off[0], -off[1], -off[2]), axis=(0, 1, 2))


def conv3(u, W):
    """u [N,N,N,C], W [CO,C,27] -> [N,N,N,CO]."""
    out = np.zeros(u.shape[:3] + (W.shape[0],))
    for t, off in enumerate(OFFSETS):
        out += shifted(u, off) @ W[:, :, t].T
    return out


def conv3_adjoint(g, W):
    """sum_t sum_co W[co, ci, t] g[s - off(t), co]: [N,N,N,CO] -> [N,N,N,C]."""
    out = np.zeros(g.shape[:3] + (W.shape[1],))
    for t, off in enumerate(OFFSETS):
        out += shifted(g, (-off[0], -off[1], -off[2])) @ W[:, :, t]
    return out


def weight_grad(g, u):
    """dW[co, ci, t] = sum_s g[s, co] u[s + off(t), ci]."""
    C = g.shape[-1]
    dW = np.zeros((C, u.shape[-1], TAPS))
    gf = g.reshape(-1, C)
    for t, off in enumerate(OFFSETS):
        dW[:, :, t] = gf.T @ shifted(u, off).reshape(-1, u.shape[-1])
    return dW


def lap(u):
    """sum of the 6 periodic neighbours minus 6 u, per channel (self-adjoint)"""
    out = -6.0 * u
    for ax in range(3):
        out = out + np.roll(u, 1, axis=ax) + np.roll(u, -1, axis=ax)
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward_u(p, x, bf16=False):
    """x [N,N,N,4] -> the network output u [N,N,N,2] and the cache of the
    backward pass."""
    P = unpack(p)
    rd = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    W1, W2, W3 = rd(P["W1"]), rd(P["W2"]), rd(P["W3"])
    x = rd(x)
    z1 = conv3(x, W1) + P["b1"]
    a1 = rd(silu(z1))
    z1 = rd(z1)
    z2 = conv3(a1, W2) + P["b2"]
    a2 = rd(silu(z2))
    z2 = rd(z2)
    u = conv3(a2, W3) + P["b3"]
    return u, dict(x=x, z1=z1, a1=a1, z2=z2, a2=a2, W1=W1, W2=W2, W3=W3)


def forward(p, x, bf16=False):
    """x [N,N,N,4] -> y = lap lap u [N,N,N,2] (the correction applied to the
    velocities) and the cache of the backward pass."""
    u, c = forward_u(p, x, bf16)
    return lap(lap(u)), c


def loss_and_grad(p, x, target, bf16=False):
    """Loss sum (lap u - t)^2 / (2 N^3) and its gradient; target [2][N,N,N]
    (channel-major, as the device stores it) is the target of lap u."""
    u, c = forward_u(p, x, bf16)
    w = lap(u)
    rd = bf16_round if bf16 else (lambda a: a)
    n3 = w.shape[0] * w.shape[1] * w.shape[2]
    t = np.moveaxis(np.asarray(target, dtype=np.float64), 0, -1)
    d = w - t
    loss = float((d * d).sum() / (2.0 * n3))
    gy = rd(lap(d / n3))      # dL/du: lap is self-adjoint
    g = {}
    g["W3"] = weight_grad(gy, c["a2"])
    g["b3"] = gy.reshape(-1, NOUT).sum(0)
    gz2 = rd(conv3_adjoint(gy, c["W3"]) * dsilu(c["z2"]))
    g["W2"] = weight_grad(gz2, c["a1"])
    g["b2"] = gz2.reshape(-1, NH).sum(0)
    gz1 = rd(conv3_adjoint(gz2, c["W2"]) * dsilu(c["z1"]))
    g["W1"] = weight_grad(gz1, c["x"])
    g["b1"] = gz1.reshape(-1, NH).sum(0)
    return loss, pack(g)


def loss_only(p, x, target):
    u, _ = forward_u(p, x)
    w = lap(u)
    t = np.moveaxis(np.asarray(target, dtype=np.float64), 0, -1)
    n3 = w.shape[0] * w.shape[1] * w.shape[2]
    return float(((w - t) ** 2).sum() / (2.0 * n3))
