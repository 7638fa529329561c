"""The kept host program (gabe_amd/host/gabe_main.cpp) end to end: its output
files against the oracle driven through the reference's cadence.

The test builds run 16^3 with tensor modes, FP64, 12 steps and a slice every
4 calls.  gabe_ref16 is GABE's shipped parameter file (same names and
spelling, #define switches included) with a user's edits for that short run,
under GABE's own model (m^2 phi^2/2 + g^2 phi^2 chi^2/2): the drop-in of a
GABE parameter file.  The others use the benchmark model (lambda phi^4 +
g^2 phi^2 chi^2): gabe_test16 with the reference's default switches (random
modes, plane slices), gabe_test16_sym3d with symmetric modes
(field_full_rand 0), the sqrt(3)/8 window and whole-lattice slices
(field_outdim 3), gabe_test16_cut1d with windowed random modes and line
slices (field_outdim 1).  Expected values come from the
oracle run through g2main.cpp:106-131: outputslice() after initialisation and
after every step, writing when its counter reaches slicewait
(g2output.cpp:129-186), with the loop's pre-increment t as the label (Q5).
Files are compared token by token: the same layout, values equal to the
printed precision.  One deliberate difference: meansvariances.dat carries the
exact variance (the device reduces shifted sums), where the reference's
sum of (f^2 - <f>^2) (g2output.cpp:90-98) loses everything to cancellation
once the variance falls below about eps <f>^2 sqrt(N^3): for phi at physical
amplitudes it prints roundoff, negative values included.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from gabe_amd import configs
from oracle import pyoracle as po
from tests.helpers import seeded_oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
N, STEPS, SLICEWAIT = 16, 12, 4
CUT = math.sqrt(3.0) / 8.0
BUILDS = {   # executable: (field_full_rand, spec_cut_off, field_outdim, model)
    "gabe_test16": (1, 0.0, 2, "lambda_phi4"),
    "gabe_test16_sym3d": (0, CUT, 3, "lambda_phi4"),
    "gabe_test16_cut1d": (1, CUT, 1, "lambda_phi4"),
    "gabe_ref16": (1, 0.0, 2, "reference_default"),
}


def expected_outputs(full_rand, cut, outdim, model):
    if model == "lambda_phi4":
        cfg = configs.lambda_phi4_config(N=N, calc_gws=1)
    else:   # GABE's shipped model and parameters (g2parameters.h:43-62), N = 16
        cfg = configs.reference_default_config(N=N, calc_gws=1)
    cfg.field_full_rand = full_rand
    cfg.spec_cut_off = cut
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
        for i in range(0, N if outdim > 2 else 1, 2):
            for j in range(0, N if outdim > 1 else 1, 2):
                for k in range(0, N, 2):
                    vals = [o.field[0, f, i, j, k] for f in range(2)] + [o.calcrho(0, i, j, k)]
                    lines.append("".join("%e " % v for v in vals))
        files["fields"][first] = lines
        sp = o.specout(gw_quirk=1)
        nb = sp.shape[1]
        files["spectra"][first] = [("%e" % (2 * np.pi * i / cfg.L)) + "".join(" %e" % sp[c, i] for c in range(sp.shape[0]))
                                   for i in range(nb)]
        # the device reduces with shifted sums: the exact mean and variance of
        # the oracle's state are what it must print.  The reference's own
        # sum of (f^2 - <f>^2) is within its cancellation bound of them.
        mv = o.meansvars()
        exact = []
        for f in range(2):
            x = o.field[0, f]
            m = x.mean()
            v = ((x - m) ** 2).mean()
            assert abs(mv[2 * f + 1] - v) <= 2 * m * m * np.finfo(float).eps * x.size + 1e-9 * v
            exact += [m, v]
        files["mv"].append(("%f" % t) + "".join(" %e %e" % (exact[2 * f], exact[2 * f + 1]) for f in range(2)))
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


def same_meansvars(got_lines, want_lines, rtol=2e-6):
    """t, then (mean, variance) per field: variances to the printed precision,
    means to it or to 1e-9 of the field's rms fluctuation (a zero-mean field's
    mean is summation roundoff)."""
    assert len(got_lines) == len(want_lines)
    for g, w in zip(got_lines, want_lines):
        gt, wt = [float(x) for x in g.split()], [float(x) for x in w.split()]
        assert len(gt) == len(wt) and gt[0] == wt[0], (g, w)
        for f in range((len(wt) - 1) // 2):
            gm, gv, wm, wv = gt[1 + 2 * f], gt[2 + 2 * f], wt[1 + 2 * f], wt[2 + 2 * f]
            assert wv > 0 and abs(gv - wv) <= rtol * wv, (g, w)
            assert abs(gm - wm) <= rtol * abs(wm) + 1e-9 * math.sqrt(wv), (g, w)


@pytest.mark.gpu
@pytest.mark.parametrize("build", sorted(BUILDS))
def test_host_program_outputs_match_oracle(tmp_path, build):
    exe = os.path.join(ROOT, "gabe_amd", build)
    assert os.path.exists(exe), f"{build} not built"
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    exp = expected_outputs(*BUILDS[build])
    sl = tmp_path / "slices"
    got_time = (sl / "slices_time.dat").read_text().splitlines()
    same_tokens(got_time, exp["time"])
    # four slices: t = 0 (before the first step), then after steps 4, 8, 12
    # labelled with the loop's t, one dt behind the fields
    assert len(got_time) == 4 and float(got_time[0].split()[0]) == 0.0
    same_meansvars((sl / "meansvariances.dat").read_text().splitlines(), exp["mv"])
    for first, lines in exp["fields"].items():
        same_tokens((sl / f"slices_fields_{first}.dat").read_text().splitlines(), lines)
    for first, lines in exp["spectra"].items():
        same_tokens((sl / f"slices_spectra_{first}.dat").read_text().splitlines(), lines, atol_frac=1e-9)
    info = (tmp_path / "info.txt").read_text()
    assert "Grid size=16^3" in info and "Using a dot expansion" in info and "Final scale factor is" in info
    assert "%d momentum bins for spectra" % po.numbins(N) in info
