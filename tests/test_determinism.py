"""The same work twice from fresh contexts in one process gives the same bits:
the pipelined step (sweep A in chunks with the j-transform on a stream of its
own, the tensor modes on the side stream, sweep B on LDS-staged tiles), the
batched ensemble, and the restriction of a fine lattice followed by steps
with a learned closure attached.  A missing wait between streams shows up
here as results that change from run to run.  A fresh process repeats the
same bits as well."""
import hashlib
import os
import subprocess
import sys

import numpy as np
import pytest

from gabe_amd import closure as cm
from gabe_amd import configs, host
from gabe_amd.lattice import GABE_FP32, Ensemble, Lattice

pytestmark = pytest.mark.gpu


def _twice(run):
    a, b = run(), run()
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert np.all(np.isfinite(x))
        assert np.array_equal(x, y)


def test_pipelined_step_with_chunked_sweep_a_repeats_bitwise():
    cfg = configs.lambda_phi4_config(N=128, calc_gws=1)

    def run():
        lat = host.initialized_lattice(cfg, precision=GABE_FP32, planes_per_block=16, transform_tile=16,
                                       mode_tile=1, stress_chunk=32)
        assert lat.query().stress_chunk == 32
        lat.step(30)
        f, v = lat.download_state()
        h, l = lat.download_gw()
        a = np.array([lat.get_scalars()["a"][0]])
        lat.close()
        return [f, v, h, l, a]

    _twice(run)


def test_ensemble_repeats_bitwise():
    cfgs = configs.ensemble_configs(sizes=(32, 48, 64), per_size=2)

    def run():
        ens = Ensemble([dict(precision=GABE_FP32, **c.lattice_kwargs()) for c in cfgs])
        for c, m in zip(cfgs, ens.members):
            host.initialize_member(m, c)
        ens.step(20)
        ens.sync()
        out = [np.concatenate([x.ravel() for x in m.download_state()]) for m in ens.members]
        ens.close()
        return out

    _twice(run)


def test_restriction_and_closure_steps_repeat_bitwise():
    setup = cm.ClosureSetup.make(configs.lambda_phi4_config, 32, 4)
    cfg = setup.coarse_cfg
    cl = cm.Closure(cfg.N, cfg.dt, seed=5)
    # an untrained network with a small output scale: a kick that changes the
    # velocities without driving the run unstable over its 20 steps
    cl.set_scales([1e-6] * 4, [1e-9, 1e-9])

    def run():
        fine = host.initialized_lattice(setup.fine_cfg, precision=GABE_FP32)
        fine.step(8)
        coarse = Lattice(precision=GABE_FP32, **cfg.lattice_kwargs())
        cm.restrict(fine, coarse)
        start = list(coarse.download_state())
        cm.attach(coarse, cl, 1e300)
        before = coarse.download_state()[1]
        coarse.step(20)
        out = start + list(coarse.download_state())
        assert cm.status(coarse)[0] == 1 and not np.array_equal(before, out[-1])
        fine.close()
        coarse.close()
        return out

    _twice(run)
    cl.close()


def _digest(*arrs):
    m = hashlib.sha256()
    for a in arrs:
        m.update(np.ascontiguousarray(a).tobytes())
    return m.hexdigest()


def _hashes():
    """Digests of a 128^3 step run with tensor modes and of a 128^3 run with
    the trained closure attached, 20 steps each."""
    out = []
    cfg = configs.lambda_phi4_config(N=128, calc_gws=1)
    lat = host.initialized_lattice(cfg, precision=GABE_FP32, planes_per_block=32, transform_tile=16, mode_tile=1,
                                   stress_chunk=64)
    lat.step(20)
    out.append(_digest(*lat.download_state(), *lat.download_gw(), np.array([lat.get_scalars()["a"][0]])))
    lat.close()
    cl = cm.Closure.from_file(os.path.join(os.path.dirname(cm.__file__), "closure_128.npz"))
    lat = host.initialized_lattice(cfg, precision=GABE_FP32, planes_per_block=32, transform_tile=16, mode_tile=1,
                                   stress_chunk=-1)
    cm.attach(lat, cl, 1e-3)
    lat.step(20)
    assert cm.status(lat)[0] == 1
    out.append(_digest(*lat.download_state()))
    lat.close()
    cl.close()
    return out


def test_fresh_process_repeats_bitwise():
    here = os.path.dirname(os.path.abspath(__file__))
    code = ("import sys; sys.path[:0] = [%r, %r]; import test_determinism as t; print(' '.join(t._hashes()))"
            % (os.path.dirname(here), here))
    child = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240)
    assert child.returncode == 0, child.stderr[-2000:]
    assert child.stdout.split() == _hashes()
