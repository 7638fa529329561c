"""Pins of the closure network's CPU FP64 restatement (oracle/closure_ref.py).

* conv3 of a delta input returns the kernel at the mirrored offsets: fixes the
  tap numbering t = 9 (dz+1) + 3 (dy+1) + (dx+1) and the direction of off(t);
* conv3_adjoint is the adjoint of conv3: <conv3(u), g> = <u, conv3_adjoint(g)>;
* the analytic gradient equals central finite differences of the FP64 loss
  (every bias and a random subset of every weight tensor); b3 drops out of
  lap u, so its gradient is zero;
* bf16_round rounds to nearest with ties to even.
"""
import numpy as np

from oracle import closure_ref as cr


def random_params(rng, scale=0.3):
    return rng.standard_normal(cr.NPARAMS) * scale


def test_delta_input_returns_the_mirrored_kernel():
    rng = np.random.default_rng(0)
    N = 5
    W = rng.standard_normal((3, 2, cr.TAPS))
    u = np.zeros((N, N, N, 2))
    u[0, 0, 0, 1] = 1.0
    out = cr.conv3(u, W)
    for t, (dz, dy, dx) in enumerate(cr.OFFSETS):
        assert t == 9 * (dz + 1) + 3 * (dy + 1) + (dx + 1)
        assert np.allclose(out[-dz % N, -dy % N, -dx % N], W[:, 1, t])
    assert np.count_nonzero(np.abs(out).sum(-1)) == cr.TAPS


def test_adjoint_identity():
    rng = np.random.default_rng(1)
    N = 6
    W = rng.standard_normal((4, 3, cr.TAPS))
    u = rng.standard_normal((N, N, N, 3))
    g = rng.standard_normal((N, N, N, 4))
    lhs = (cr.conv3(u, W) * g).sum()
    rhs = (u * cr.conv3_adjoint(g, W)).sum()
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    N = 4
    p = random_params(rng)
    x = rng.standard_normal((N, N, N, 4))
    t = rng.standard_normal((2, N, N, N))
    loss, grad = cr.loss_and_grad(p, x, t)
    assert abs(loss - cr.loss_only(p, x, t)) < 1e-14 * loss
    sizes = [int(np.prod(s)) for _, s in cr.SHAPES]
    starts = np.cumsum([0] + sizes[:-1])
    pick = []
    for (name, _), st, n in zip(cr.SHAPES, starts, sizes):
        if name == "b3":
            assert np.abs(grad[st:st + n]).max() < 1e-12 * max(loss, 1.0)
            continue
        pick += list(range(st, st + n)) if name.startswith("b") else list(st + rng.choice(n, 40, replace=False))
    h = 1e-5
    for e in pick:
        pp, pm = p.copy(), p.copy()
        pp[e] += h
        pm[e] -= h
        fd = (cr.loss_only(pp, x, t) - cr.loss_only(pm, x, t)) / (2 * h)
        assert abs(fd - grad[e]) <= 1e-6 * max(abs(fd), 1e-3), (e, fd, grad[e])


def test_bf16_rounding_ties_to_even():
    e7, e8 = 2.0 ** -7, 2.0 ** -8
    vals = np.array([1.0, 1 + e8, 1 + 3 * e8, 1 + e7, -(1 + e8), 3.0e-30])
    got = cr.bf16_round(vals)
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1 + 2 * e7 and got[3] == 1 + e7 and got[4] == -1.0
    assert abs(got[5] - 3.0e-30) <= 3.0e-30 * 2 ** -8
