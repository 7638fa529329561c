"""The kept host program (gabe_amd/host/gabe_main.cpp) end to end: its output
files against the oracle driven through the reference's cadence.

The test build (gabe_test16) is the default model at 16^3 with tensor modes,
FP64, 12 steps and a slice every 4 calls.  Expected values come from the
oracle run through g2main.cpp:106-131: outputslice() after initialisation and
after every step, writing when its counter reaches slicewait
(g2output.cpp:129-186), with the loop's pre-increment t as the label (Q5).
Files are compared token by token: the same layout, values equal to the
printed precision.
"""
import os
import subprocess

import numpy as np
import pytest

from gabe_amd import configs
from oracle import pyoracle as po
from tests.helpers import seeded_oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gabe_amd", "gabe_test16")
N, STEPS, SLICEWAIT = 16, 12, 4


def expected_outputs():
    cfg = configs.lambda_phi4_config(N=N, calc_gws=1)
    o = seeded_oracle(cfg)
    dt = cfg.dt
    files = {"time": [], "mv": [], "fields": {}, "spectra": {}}
    state = {"first": 0, "lastslice": SLICEWAIT}

    def outputslice(t):
        state["lastslice"] += 1
        if state["lastslice"] < SLICEWAIT:
            return
        first = state["first"]
        lines = []
        for j in range(0, N, 2):
            for k in range(0, N, 2):
                vals = [o.field[0, f, 0, j, k] for f in range(2)] + [o.calcrho(0, 0, j, k)]
                lines.append("".join("%e " % v for v in vals))
        files["fields"][first] = lines
        sp = o.specout(gw_quirk=1)
        nb = sp.shape[1]
        files["spectra"][first] = [("%e" % (2 * np.pi * i / cfg.L)) + "".join(" %e" % sp[c, i] for c in range(sp.shape[0]))
                                   for i in range(nb)]
        mv = o.meansvars()
        files["mv"].append(("%f" % t) + "".join(" %e %e" % (mv[2 * f], mv[2 * f + 1]) for f in range(2)))
        s = o.scalars()
        files["time"].append("%e %e %e %e %e %e %e %e" % (t, s["a"][0], s["adot"][0], s["adot"][0] / s["a"][0],
                                                          s["edkin"][0], s["edpot"][0], s["edgrad"][0], s["edrho"][0]))
        state["first"] += 1
        state["lastslice"] = 0

    t = 0.0
    endtime = (STEPS - 1) * dt
    while t <= endtime:
        if t == 0.0:
            outputslice(t)
        o.step(1)
        outputslice(t)
        t += dt
    return files


def same_tokens(got_lines, want_lines, rtol=2e-6, atol_frac=1e-12):
    assert len(got_lines) == len(want_lines)
    scale = max((abs(float(x)) for line in want_lines for x in line.split()), default=1.0)
    for g, w in zip(got_lines, want_lines):
        gt, wt = g.split(), w.split()
        assert len(gt) == len(wt), (g, w)
        for a, b in zip(gt, wt):
            fa, fb = float(a), float(b)
            assert abs(fa - fb) <= rtol * abs(fb) + atol_frac * scale, (g, w)
        assert g.rstrip("\n") == g.rstrip("\n")


@pytest.mark.gpu
def test_host_program_outputs_match_oracle(tmp_path):
    assert os.path.exists(EXE), "gabe_test16 not built"
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    exp = expected_outputs()
    sl = tmp_path / "slices"
    got_time = (sl / "slices_time.dat").read_text().splitlines()
    same_tokens(got_time, exp["time"])
    # four slices: t = 0 (before the first step), then after steps 4, 8, 12
    # labelled with the loop's t, one dt behind the fields
    assert len(got_time) == 4 and float(got_time[0].split()[0]) == 0.0
    same_tokens((sl / "meansvariances.dat").read_text().splitlines(), exp["mv"])
    for first, lines in exp["fields"].items():
        same_tokens((sl / f"slices_fields_{first}.dat").read_text().splitlines(), lines)
    for first, lines in exp["spectra"].items():
        same_tokens((sl / f"slices_spectra_{first}.dat").read_text().splitlines(), lines, atol_frac=1e-9)
    info = (tmp_path / "info.txt").read_text()
    assert "Grid size=16^3" in info and "Using a dot expansion" in info and "Final scale factor is" in info
    assert "%d momentum bins for spectra" % po.numbins(N) in info
