"""Checkpoint / resume: a context restored from a checkpoint continues bit for
bit (fields, velocities, tensor modes, scalars and history), a blob from a
context with other parameters or a truncated one is refused, and the host
program stopped after a checkpoint and resumed in a new process writes the
same files as an uninterrupted run."""
import filecmp
import os
import subprocess

import numpy as np
import pytest

from gabe_amd.lattice import GABE_FP32, GABE_FP64, GabeError, Lattice
from tests.helpers import configs, lattice_from_oracle, seeded_oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def assert_same(a, b, gw):
    fa, da = a.download_state()
    fb, db = b.download_state()
    assert np.array_equal(fa, fb) and np.array_equal(da, db)
    if gw:
        ha, la = a.download_gw()
        hb, lb = b.download_gw()
        assert np.array_equal(ha, hb) and np.array_equal(la, lb)
    sa, sb = a.get_scalars(), b.get_scalars()
    for k in ("a", "adot", "edkin", "edpot", "edgrad", "edrho", "t", "nstep"):
        assert np.array_equal(np.asarray(sa[k]), np.asarray(sb[k])), k
    assert np.array_equal(a.history(), b.history())


@pytest.mark.parametrize("precision,N,gw", [(GABE_FP32, 32, 1), (GABE_FP64, 16, 1), (GABE_FP32, 24, 0)])
def test_resume_is_bitwise(precision, N, gw):
    cfg = configs.lambda_phi4_config(N=N, calc_gws=gw)
    o = seeded_oracle(cfg, amp=2e-3)
    whole = lattice_from_oracle(o, cfg, precision)
    whole.step(11)
    first = lattice_from_oracle(o, cfg, precision)
    first.step(5)
    blob = first.checkpoint()
    first.close()
    second = Lattice(precision=precision, **cfg.lattice_kwargs())
    second.restore(bytes(blob))
    second.step(6)
    assert_same(whole, second, gw)


def test_foreign_and_truncated_blobs_are_refused():
    cfg = configs.lambda_phi4_config(N=16, calc_gws=1)
    o = seeded_oracle(cfg, amp=2e-3)
    lat = lattice_from_oracle(o, cfg, GABE_FP32)
    lat.step(2)
    blob = lat.checkpoint()
    other = Lattice(precision=GABE_FP64, **cfg.lattice_kwargs())
    with pytest.raises(GabeError):
        other.restore(blob)
    kw = cfg.lattice_kwargs()
    kw["dt"] = cfg.dt / 2
    with pytest.raises(GabeError):
        Lattice(precision=GABE_FP32, **kw).restore(blob)
    with pytest.raises(GabeError):
        Lattice(precision=GABE_FP32, **cfg.lattice_kwargs()).restore(blob[:-8])
    bad = blob.copy()
    bad[0] ^= 1
    with pytest.raises(GabeError):
        Lattice(precision=GABE_FP32, **cfg.lattice_kwargs()).restore(bad)


def test_host_program_resume_writes_the_same_files(tmp_path):
    exe = os.path.join(ROOT, "gabe_amd", "gabe_test16")
    whole, split = tmp_path / "whole", tmp_path / "split"
    whole.mkdir()
    split.mkdir()
    r = subprocess.run([exe], cwd=whole, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe, "--checkpoint-at", "6"], cwd=split, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "Checkpoint written after step 6" in r.stdout, r.stderr
    r = subprocess.run([exe, "--resume", "checkpoint.bin"], cwd=split, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "Resumed from checkpoint.bin at step 6" in r.stdout, r.stderr
    names = sorted(os.listdir(whole / "slices"))
    assert names == sorted(os.listdir(split / "slices")) and len(names) > 4
    match, mismatch, errors = filecmp.cmpfiles(whole / "slices", split / "slices", names, shallow=False)
    assert not mismatch and not errors, mismatch


def closure_lattices(n, zero=False):
    """n FP32 32^3 lattices of the closure's coarse workload from one seeded
    start, and a random closure (zero weights with zero=True, so its kick
    is zero) with scales for them."""
    from gabe_amd import closure as cm
    from gabe_amd import host
    cfg = cm.ClosureSetup.make(configs.lambda_phi4_config, 32, 4).coarse_cfg
    lats = [host.initialized_lattice(cfg, precision=GABE_FP32) for _ in range(n)]
    cl = cm.Closure(cfg.N, cfg.dt, seed=4)
    if zero:
        cl.set_params(np.zeros(cm.NPARAMS, dtype=np.float32))
    cl.set_scales([1e-6] * 4, [1e-3, 1e-3])
    return cm, cfg, lats, cl


def test_closure_state_travels_with_the_checkpoint():
    """A checkpoint of a lattice with a learned closure attached carries the
    attachment: after a fallback the restored lattice stays on the plain
    step (same gate, fallback step and largest residual) even if a closure
    was attached to it before the load; an active closure is re-attached on
    a lattice that had none; a blob without a closure detaches one.  Each
    restored lattice then continues bit for bit."""
    cm, cfg, (A, B, C, D, E, F), cl = closure_lattices(6)
    cm.attach(A, cl, 1e-300)      # switched off by the check after step 3
    A.step(5)
    assert cm.status(A)[:2] == (0, 3)
    cm.attach(B, cl, 1e300)
    B.restore(A.checkpoint())
    assert cm.status(B) == cm.status(A)
    A.step(4)
    B.step(4)
    assert_same(A, B, 0)
    assert cm.status(B) == cm.status(A)

    cm.attach(C, cl, 1e300)
    C.step(2)
    D.restore(C.checkpoint())
    assert cm.status(D) == cm.status(C) and cm.status(D)[0] == 1
    C.step(3)
    D.step(3)
    assert_same(C, D, 0)

    E.step(1)
    cm.attach(F, cl, 1e300)
    F.restore(E.checkpoint())
    assert cm.status(F) == (0, -1, 0.0)
    E.step(2)
    F.step(2)
    assert_same(E, F, 0)
    cl.close()


def test_set_scalars_restarts_the_history():
    """gabe_set_scalars starts a new trajectory: the history holds only rows
    recorded after it, checkpoints carry only those, and an attached
    closure's residual check waits for three rows of the new trajectory
    instead of reading rows that were never written (a zero-weight closure:
    only a non-finite residual, such as 0/0 from an unwritten row, could
    switch it off)."""
    cm, cfg, (A, B), cl = closure_lattices(2, zero=True)
    cm.attach(A, cl, 1e300)
    A.step(4)
    assert A.history().shape[0] == 4
    sc = A.get_scalars()
    A.set_scalars(sc["a"], sc["adot"], sc["edkin"], sc["edpot"], sc["edgrad"], sc["edrho"], t=sc["t"], nstep=1000)
    assert A.history().shape[0] == 0
    A.step(5)
    h = A.history()
    assert h.shape[0] == 5 and list(h[:, 7]) == [1000, 1001, 1002, 1003, 1004]
    active, fb, rmax = cm.status(A)
    assert active == 1 and fb == -1 and np.isfinite(rmax)
    B.restore(A.checkpoint())
    assert np.array_equal(B.history(), h)
    cl.close()
