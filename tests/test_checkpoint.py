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
