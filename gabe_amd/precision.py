"""FP32 form against the FP64 form over a long run with periodic spectra.

The lattice of the benchmark model (lambda phi^4 + g^2 phi^2 chi^2, L = 20,
dt = dx/20, tensor modes on) is started from the same seeded initial state in
both storage forms and stepped in lockstep.  Every `every` steps both write
their spectra (the reference's specOut columns); the per-bin relative error
of the FP32 form is taken over the bins above 1e-9 of each column's maximum
(bin 0 excluded).  At the end the field and velocity L2 errors relative to the
fluctuation, the energy density, a and both runs' Friedmann residuals are
reported.

    python -m gabe_amd.precision --N 512 --steps 4000 --every 50

prints one JSON line; progress goes to stderr after each output.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import configs, host
from .lattice import GABE_FP32, GABE_FP64

FIELD_COLS = (0, 2)   # power spectrum of each field
GW_COL = 4            # GW power spectrum


def bin_errors(got, ref):
    """max over live bins of |got/ref - 1|, per column group."""
    out = {}
    for name, cols in (("field", FIELD_COLS), ("gw", (GW_COL,))):
        worst = 0.0
        for c in cols:
            if c >= ref.shape[0]:
                continue
            r, g = ref[c], got[c]
            live = np.abs(r) > 1e-9 * np.abs(r).max()
            live[0] = False
            if live.any():
                worst = max(worst, float(np.max(np.abs(g[live] / r[live] - 1.0))))
        out[name] = worst
    return out


def run(N, steps, every, log=None):
    cfg = configs.lambda_phi4_config(N=N, calc_gws=1)
    t0 = time.perf_counter()
    lat32 = host.initialized_lattice(cfg, precision=GABE_FP32)
    lat64 = host.initialized_lattice(cfg, precision=GABE_FP64)
    series = []
    t32 = t64 = 0.0
    done = 0
    while done < steps:
        n = min(every, steps - done)
        a = time.perf_counter()
        lat32.step(n)
        lat32.sync()
        b = time.perf_counter()
        lat64.step(n)
        lat64.sync()
        c = time.perf_counter()
        t32 += b - a
        t64 += c - b
        done += n
        e = bin_errors(lat32.specOut(), lat64.specOut())
        series.append({"step": done, "field_max_rel": e["field"], "gw_max_rel": e["gw"]})
        if log:
            log(f"step {done}: field spectra {e['field']:.3e}, GW spectrum {e['gw']:.3e}")
    f32, v32 = lat32.download_state()
    f64, v64 = lat64.download_state()
    res = {"N": N, "steps": steps, "every": every}
    for name, a, b in (("field", f32, f64), ("velocity", v32, v64)):
        res[f"{name}_l2_rel"] = [float(np.linalg.norm(a[f] - b[f]) / np.linalg.norm(b[f] - b[f].mean()))
                                 for f in range(b.shape[0])]
    s32, s64 = lat32.get_scalars(), lat64.get_scalars()
    res["edrho_rel"] = abs(s32["edrho"][0] - s64["edrho"][0]) / abs(s64["edrho"][0])
    res["a_rel"] = abs(s32["a"][0] - s64["a"][0]) / s64["a"][0]
    res["friedmann_max_fp32"] = float(host.friedmann_residual(lat32.history())[0])
    res["friedmann_max_fp64"] = float(host.friedmann_residual(lat64.history())[0])
    res["field_spectra_max_rel"] = max(s["field_max_rel"] for s in series)
    res["gw_spectrum_max_rel"] = max(s["gw_max_rel"] for s in series)
    res["series"] = series
    res["fp32_ms_per_step"] = 1e3 * t32 / steps
    res["fp64_ms_per_step"] = 1e3 * t64 / steps
    res["wall_seconds"] = time.perf_counter() - t0
    lat32.close()
    lat64.close()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m gabe_amd.precision")
    ap.add_argument("--N", type=int, default=512)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--every", type=int, default=50, help="steps between spectra")
    args = ap.parse_args(argv)

    def log(msg):
        print(f"[precision] {msg}", file=sys.stderr, flush=True)

    print(json.dumps(run(args.N, args.steps, args.every, log)))


if __name__ == "__main__":
    main()
