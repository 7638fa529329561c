"""Learned closure on the device against the CPU FP64 network, the spectral
restriction against its closed form, the plain step as the closure's
fallback, and the gain of a trained closure over the plain coarse step.

Tolerances: forward and gradient against the CPU network with the device's
BF16 rounding points reproduced: output within 1e-2 of its largest value
(rms 3e-3), each gradient tensor within 3e-2 (relative 2-norm), loss within
1e-4; against the pure FP64 network (no BF16 anywhere): 5e-2 and 1e-1.
"""
import math

import numpy as np
import pytest

from gabe_amd import closure as cm
from gabe_amd import configs, host
from gabe_amd.lattice import GABE_FP32, Lattice
from oracle import closure_ref as cr

pytestmark = pytest.mark.gpu


def tensors(v):
    out, o = {}, 0
    for name, shp in cr.SHAPES:
        n = int(np.prod(shp))
        out[name] = v[o:o + n]
        o += n
    return out


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module", params=[32, 128])
def net(request):
    """A random network with nonzero output biases and random inputs and
    targets, at 32^3 and at the closure workload's 128^3."""
    N = request.param
    cl = cm.Closure(N, 0.01, seed=5)
    p = cl.params().astype(np.float64)
    p[cr.NPARAMS - 2:] = [0.03, -0.02]   # nonzero output biases
    cl.set_params(p)
    rng = np.random.default_rng(11)
    x = rng.standard_normal((N, N, N, 4)).astype(np.float32)
    t = rng.standard_normal((2, N, N, N)).astype(np.float32) * 0.3
    yield cl, cl.params().astype(np.float64), x, t
    cl.close()


def test_forward_matches_cpu_network(net):
    cl, p, x, _ = net
    y = cl.forward(x.reshape(-1, 4)).reshape(x.shape[:3] + (2,))
    ref, _ = cr.forward(p, x, bf16=True)
    assert np.abs(y - ref).max() <= 1e-2 * np.abs(ref).max()
    assert np.sqrt(((y - ref) ** 2).mean()) <= 3e-3 * np.sqrt((ref ** 2).mean())
    ref64, _ = cr.forward(p, x, bf16=False)
    assert rel(y, ref64) < 5e-2


def test_gradient_matches_cpu_network(net):
    cl, p, x, t = net
    loss, g = cl.gradient(x.reshape(-1, 4), t.reshape(2, -1))
    lref, gref = cr.loss_and_grad(p, x, t, bf16=True)
    assert abs(loss - lref) <= 1e-4 * lref
    G, R = tensors(g.astype(np.float64)), tensors(gref)
    # b3 drops out of lap u: its gradient is the sum of N^3 BF16-rounded
    # terms of zero sum, i.e. rounding noise
    assert np.abs(G["b3"]).max() <= 2e-2 * np.abs(G["b2"]).max()
    for name in G:
        if name != "b3":
            assert rel(G[name], R[name]) < 3e-2, name
    l64, g64 = cr.loss_and_grad(p, x, t, bf16=False)
    assert abs(loss - l64) <= 1e-2 * l64
    G64 = tensors(g64)
    for name in G:
        if name != "b3":
            assert rel(G[name], G64[name]) < 1e-1, name


def plane_waves(N, waves, L):
    x = np.arange(N) * (L / N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    out = np.zeros((N, N, N))
    for amp, n in waves:
        out += amp * np.cos(2 * np.pi / L * (n[0] * X + n[1] * Y + n[2] * Z) + 0.3)
    return out


def test_restriction_keeps_the_coarse_band():
    L = 20.0
    Nf, Nc = 64, 32
    fine = Lattice(N=Nf, nflds=2, L=L, precision=GABE_FP32, m2_phi=1.0)
    coarse = Lattice(N=Nc, nflds=2, L=L, precision=GABE_FP32, m2_phi=1.0)
    keep = [(1e-3, (1, 2, 3)), (5e-4, (15, -7, 2)), (2e-4, (-3, 0, 12))]
    drop = [(7e-4, (20, 1, 0)), (3e-4, (0, 16, 1))]   # beyond or on the coarse Nyquist planes
    f = np.stack([0.3 + plane_waves(Nf, keep + drop, L), plane_waves(Nf, keep[:1] + drop[:1], L)])
    d = np.stack([-0.2 + plane_waves(Nf, keep[1:], L), plane_waves(Nf, drop, L)])
    fine.upload_state(f, d)
    fine.set_scalars((1.25, 1.25), (0.4, 0.4), t=0.75)
    cm.restrict(fine, coarse)
    fc, dc = coarse.download_state()
    want_f = np.stack([0.3 + plane_waves(Nc, keep, L), plane_waves(Nc, keep[:1], L)])
    want_d = np.stack([-0.2 + plane_waves(Nc, keep[1:], L), np.zeros((Nc, Nc, Nc))])
    assert np.abs(fc - want_f).max() < 2e-9 and np.abs(dc - want_d).max() < 2e-9
    sc = coarse.get_scalars()
    assert sc["a"][0] == 1.25 and sc["t"] == 0.75


def small_setup(N=32):
    return cm.ClosureSetup.make(configs.lambda_phi4_config, N, 4)


def test_zero_closure_is_the_plain_step_and_fallback_switches_it_off():
    setup = small_setup()
    cfg = setup.coarse_cfg
    plain = host.initialized_lattice(cfg, precision=GABE_FP32)
    zero = host.initialized_lattice(cfg, precision=GABE_FP32)
    cl = cm.Closure(cfg.N, cfg.dt, seed=3)
    cl.set_params(np.zeros(cm.NPARAMS, dtype=np.float32))
    cl.set_scales([1e-6] * 4, [1e-3, 1e-3])
    cm.attach(zero, cl, 1.0)
    plain.step(10)
    zero.step(10)
    for a, b in zip(plain.download_state(), zero.download_state()):
        assert np.array_equal(a, b)
    assert cm.status(zero)[:2] == (1, -1)
    # a random closure with a threshold nothing meets: applied after steps 1
    # and 2, switched off by the check after step 3 (the first with three
    # history rows); equals attaching for two steps and detaching
    rnd = cm.Closure(cfg.N, cfg.dt, seed=4)
    rnd.set_scales([1e-6] * 4, [1e-3, 1e-3])
    A = host.initialized_lattice(cfg, precision=GABE_FP32)
    B = host.initialized_lattice(cfg, precision=GABE_FP32)
    cm.attach(A, rnd, 1e-300)
    cm.attach(B, rnd, 1e300)
    A.step(10)
    B.step(2)
    cm.detach(B)
    B.step(8)
    active, fb, _ = cm.status(A)
    assert active == 0 and fb == 3
    for a, b in zip(A.download_state(), B.download_state()):
        assert np.array_equal(a, b)
    assert not np.array_equal(A.download_state()[1], plain.download_state()[1])   # it was applied


def stored_arrays(lat):
    """The FP32 lattice's field and velocity arrays as stored on the device
    (fluctuations about the FP64 means): the tail of its checkpoint blob when
    the lattice carries no tensor modes and no closure (an attached closure's
    state follows the arrays)."""
    n = 2 * lat.N ** 3
    tail = np.frombuffer(lat.checkpoint()[-8 * n:].tobytes(), dtype=np.float32)
    shp = (2, lat.N, lat.N, lat.N)
    return tail[:n].reshape(shp), tail[n:].reshape(shp)


@pytest.mark.parametrize("N", [32, 128])
def test_step_correction_is_the_network_kick(N):
    """After one step with the closure attached the fields equal the plain
    step's, and the stored velocities differ by dt s_out[c] y_c / a^2: y the
    network output (the forward call, itself pinned to the CPU network above)
    for the post-step stored arrays (fluctuations about the FP64 means, as in
    the training samples) scaled by 1 / s_in and rounded to BF16; the step
    path reads the FP32 arrays inside its first convolution."""
    setup = small_setup(N)
    cfg = setup.coarse_cfg
    assert cfg.calc_gws == 0 and cfg.N == N
    A = host.initialized_lattice(cfg, precision=GABE_FP32)
    B = host.initialized_lattice(cfg, precision=GABE_FP32)
    f0, v0 = stored_arrays(B)
    s_in = [float(np.sqrt((f0[0].astype(np.float64) ** 2).mean())) or 1.0,
            float(np.sqrt((f0[1].astype(np.float64) ** 2).mean())) or 1.0,
            float(np.sqrt((v0[0].astype(np.float64) ** 2).mean())) or 1.0,
            float(np.sqrt((v0[1].astype(np.float64) ** 2).mean())) or 1.0]
    s_out = [2e-2, 3e-2]
    cl = cm.Closure(N, cfg.dt, seed=9)
    cl.set_scales(s_in, s_out)
    cm.attach(A, cl, 1e300)
    A.step(1)
    B.step(1)
    for a, b in zip(A.download_state()[0], B.download_state()[0]):
        assert np.array_equal(a, b)
    fB, vB = stored_arrays(B)
    cm.detach(A)
    _, vA = stored_arrays(A)
    inv = [np.float32(1.0 / s) for s in s_in]
    x = np.stack([fB[0] * inv[0], fB[1] * inv[1], vB[0] * inv[2], vB[1] * inv[3]], axis=-1)
    y = cl.forward(x).reshape(N, N, N, 2).astype(np.float64)
    a = A.get_scalars()["a"][0]
    ia2 = float(np.float32(1.0 / (a * a)))
    eps = float(np.finfo(np.float32).eps)
    for c in range(2):
        kick = float(np.float32(cfg.dt * s_out[c])) * ia2 * y[..., c]
        got = vA[c].astype(np.float64) - vB[c]
        assert np.abs(kick).max() > 1e3 * eps * np.abs(vB[c]).max()    # the kick is well resolved
        assert np.all(np.abs(got - kick) <= 2 * eps * (np.abs(vB[c]) + np.abs(kick)) + 1e-30)
    cl.close()


def test_training_reduces_loss_and_the_closure_beats_the_plain_step():
    setup = small_setup()
    cl = cm.Closure(setup.coarse_cfg.N, setup.coarse_cfg.dt, seed=7)
    cm.generate(cl, setup, seeds=[101, 202], samples_per_seed=24, stride=4)
    assert cl.num_samples() == 48
    l0 = np.mean([cl.loss(s) for s in range(cl.num_samples())])
    losses = cl.train(600, lr=2e-3, seed=1)
    l1 = np.mean([cl.loss(s) for s in range(cl.num_samples())])
    assert np.all(np.isfinite(losses)) and l1 < 0.5 * l0, (l0, l1)
    res = cm.evaluate(cl, setup, seed=303, coarse_steps=96, gws=False)
    assert res["closure_status"][0] == 1
    assert res["closure_field_err_mean"] < 0.7 * res["plain_field_err_mean"], res


def test_committed_closure_meets_the_128_vs_512_target_on_a_held_out_seed():
    """Config 5 judged by the plain fine method: the committed network
    (gabe_amd/closure_128.npz) on a 128^3 lattice against the plain 512^3 run
    of held-out seed 3 (training used seeds 1, 2, 6, 7 and chose among
    networks on seed 8), with tensor modes, 1000 coarse / 4000 fine steps:
      * every field-spectrum bin (1 .. 63, both fields) within 5 % of 512^3;
      * the closure run's worst and mean bins better than the plain 128^3
        run's;
      * the fallback never engaged (the Friedmann residual stayed under 10x
        the plain run's);
      * at most 12.8x the plain 128^3 run's wall time, and at least 20x
        faster than the 512^3 run."""
    import os
    setup = cm.ClosureSetup.make(configs.lambda_phi4_config, 128, 4)
    path = os.path.join(os.path.dirname(cm.__file__), "closure_128.npz")
    cl = cm.Closure.from_file(path)
    r = cm.evaluate(cl, setup, 3, 1000, log=lambda m: print(m, flush=True))
    cl.close()
    print(f"field worst bin {r['closure_field_err_max']:.4f} (plain {r['plain_field_err_max']:.4f}), "
          f"mean {r['closure_field_err_mean']:.4f} (plain {r['plain_field_err_mean']:.4f}); "
          f"GW mean {r['closure_gw_err_mean']:.3f} (plain {r['plain_gw_err_mean']:.3f}); "
          f"cost {r['closure_cost_vs_plain']:.2f}x plain, {r['speedup_vs_fine']:.0f}x faster than 512^3",
          flush=True)
    assert r["closure_field_err_max"] <= 0.05
    assert r["closure_field_err_max"] < r["plain_field_err_max"]
    assert r["closure_field_err_mean"] < r["plain_field_err_mean"]
    assert r["closure_status"][0] == 1, r["closure_status"]   # (active, fallback step, max residual)
    assert r["closure_cost_vs_plain"] <= 12.8
    assert r["speedup_vs_fine"] >= 20.0


def test_closure_kernel_times_hook():
    """gabe_closure_kernel_times: four positive per-kernel times of an
    attached closure, the lattice advanced by the timed steps, and a state
    error on a lattice without a closure."""
    from gabe_amd.lattice import GABE_ERR_STATE, GabeError
    cfg = configs.lambda_phi4_config(N=32, calc_gws=1)
    lat = host.initialized_lattice(cfg, precision=GABE_FP32)
    with pytest.raises(GabeError) as e:
        cm.kernel_times(lat, 2)
    assert e.value.code == GABE_ERR_STATE
    cl = cm.Closure(32, cfg.dt, seed=3)
    cl.set_scales([1e6] * 4, [1e-12] * 2)   # fluctuation-sized inputs, a negligible kick
    cm.attach(lat, cl, 1e300)
    n0 = lat.get_scalars()["nstep"]
    kt = cm.kernel_times(lat, 3)
    assert lat.get_scalars()["nstep"] == n0 + 3
    assert set(kt) == set(cm.CLOSURE_KERNELS) and all(v > 0 for v in kt.values()), kt
    assert cm.status(lat)[0] == 1
    print({k: f"{v * 1e3:.1f} us" for k, v in kt.items()}, flush=True)
    lat.close()
    cl.close()
