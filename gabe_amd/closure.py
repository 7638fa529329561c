"""Learned closure of a coarse lattice (no counterpart in the reference):
binding of the gabe_closure_* entry points of libgabe_amd.so, training data
from the plain method, training and evaluation against fine runs.

The network, its kernels and the per-step application live in
csrc/gabe_closure.hip; this module drives them.  Training pairs come from a
fine lattice (e.g. 512^3) stepped by the plain method and spectrally
restricted onto the coarse lattice (e.g. 128^3): the input is the restricted
state, the target what the coarse lattice's Laplacian misses of the fine
one, R(lap_fine f) - lap_coarse(R f).  Applied as a velocity kick after each
plain step, a correction of this form moves each mode's frequency from the
coarse lattice's to the fine lattice's at leading order in dt.  The plain coarse
step is the fallback, and the fine run is the yardstick: spectra per bin,
energies and the Friedmann residual.
"""
import ctypes as C
import dataclasses
import time

import numpy as np

from . import host
from .lattice import GABE_FP32, Lattice, _check, load

NPARAMS = 9538
_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)


def _f32(a):
    return a.ctypes.data_as(_fp)


def _f64(a):
    return a.ctypes.data_as(_dp)


class Closure:
    """The network for lattices of size N stepped with dt."""

    def __init__(self, N, dt, seed=1):
        self.N, self.dt = N, dt
        h = C.c_void_p()
        _check(load().gabe_closure_create(N, dt, seed, C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            load().gabe_closure_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- parameters ----------------------------------------------------------
    def params(self):
        p = np.empty(NPARAMS, dtype=np.float32)
        _check(load().gabe_closure_params(self._h, _f32(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == NPARAMS
        _check(load().gabe_closure_set_params(self._h, _f32(p)))

    def scales(self):
        si, so = np.empty(4), np.empty(2)
        _check(load().gabe_closure_scales(self._h, _f64(si), _f64(so)))
        return si, so

    def set_scales(self, s_in, s_out):
        si = np.ascontiguousarray(s_in, dtype=np.float64)
        so = np.ascontiguousarray(s_out, dtype=np.float64)
        _check(load().gabe_closure_set_scales(self._h, _f64(si), _f64(so)))

    def save(self, path):
        si, so = self.scales()
        np.savez(path, params=self.params(), s_in=si, s_out=so, N=self.N, dt=self.dt)

    @classmethod
    def from_file(cls, path):
        z = np.load(path, allow_pickle=False)
        cl = cls(int(z["N"]), float(z["dt"]))
        cl.set_params(z["params"])
        cl.set_scales(z["s_in"], z["s_out"])
        return cl

    # -- data and training ---------------------------------------------------
    def sample(self, fine, coarse, fine_steps, restrict_first=True):
        return _check(load().gabe_closure_sample(self._h, fine._h, coarse._h, fine_steps, int(restrict_first)))

    def num_samples(self):
        return _check(load().gabe_closure_num_samples(self._h))

    def train(self, iters, lr=1e-3, seed=0, record=True):
        out = np.zeros(iters) if record else None
        _check(load().gabe_closure_train(self._h, iters, lr, seed, _f64(out) if record else None))
        return out

    def loss(self, sample):
        v = C.c_double()
        _check(load().gabe_closure_loss(self._h, sample, C.byref(v)))
        return v.value

    def forward(self, x):
        """x [N^3][4] in scaled units -> y [N^3][2]."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 4)
        y = np.empty((x.shape[0], 2), dtype=np.float32)
        _check(load().gabe_closure_forward(self._h, _f32(x), _f32(y)))
        return y

    def gradient(self, x, t):
        """loss and d loss / d params for x [N^3][4] and targets t [2][N^3]."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 4)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(2, -1)
        g = np.empty(NPARAMS, dtype=np.float32)
        v = C.c_double()
        _check(load().gabe_closure_gradient(self._h, _f32(x), _f32(t), _f32(g), C.byref(v)))
        return v.value, g


def restrict(fine, coarse):
    _check(load().gabe_restrict(fine._h, coarse._h))


def attach(lat, closure, residual_threshold):
    _check(load().gabe_closure_attach(lat._h, closure._h, residual_threshold))


def detach(lat):
    _check(load().gabe_closure_detach(lat._h))


def status(lat):
    """(active, fallback_step, max_residual)"""
    a, f, r = C.c_int(), C.c_longlong(), C.c_double()
    _check(load().gabe_closure_status(lat._h, C.byref(a), C.byref(f), C.byref(r)))
    return a.value, f.value, r.value


# ----------------------------------------------------------------- pipeline
@dataclasses.dataclass
class ClosureSetup:
    """A coarse/fine pair of the same model and box: the fine lattice is
    `ratio` times finer in space and in time (dt = dx / 20 on both)."""
    coarse_cfg: object
    fine_cfg: object
    ratio: int

    @classmethod
    def make(cls, cfg_fn, N_coarse, ratio, **kw):
        c = cfg_fn(N=N_coarse, calc_gws=0, **kw)
        f = cfg_fn(N=N_coarse * ratio, calc_gws=0, **kw)
        return cls(c, f, ratio)


def _with_seed(cfg, seed, gws=None):
    c = dataclasses.replace(cfg, randseed=seed)
    if gws is not None:
        c.calc_gws = gws
    return c


def generate(cl, setup, seeds, samples_per_seed, stride=1, log=None):
    """Training pairs from fine runs seeded with `seeds`: one pair every
    `stride` coarse steps of each run."""
    for seed in seeds:
        fine = host.initialized_lattice(_with_seed(setup.fine_cfg, seed), precision=GABE_FP32)
        coarse = Lattice(precision=GABE_FP32, **setup.coarse_cfg.lattice_kwargs())
        for n in range(samples_per_seed):
            cl.sample(fine, coarse, setup.ratio * stride, restrict_first=(n == 0))
        if log:
            log(f"seed {seed}: {samples_per_seed} samples, fine t = {fine.get_scalars()['t']:.4f}")
        fine.close()
        coarse.close()


def field_spectra(lat):
    """P_b of each field: columns 0 and 2 of the spectra output."""
    sp = lat.specOut()
    return sp[[0, 2]]


def spectral_error(sp, ref, nbins):
    """per-bin relative error |P - P_ref| / P_ref over bins 1 .. nbins-1, per field"""
    b = slice(1, nbins)
    return np.abs(sp[:, b] - ref[:, b]) / np.abs(ref[:, b])


def evaluate(cl, setup, seed, coarse_steps, threshold_factor=10.0, gws=True, log=None):
    """Held-out comparison: the fine run over coarse_steps * ratio steps, the
    plain coarse run and the closure run over coarse_steps steps, all from the
    same seeded initial state (the coarse ones from its restriction).  The
    closure's fallback threshold is threshold_factor times the largest
    Friedmann residual of the plain coarse run.  Returns per-bin spectral
    errors (field and, with gws, tensor spectra), energies, Friedmann
    residuals and wall times."""
    out = {}
    fine_cfg = _with_seed(setup.fine_cfg, seed, gws=int(gws))
    coarse_cfg = _with_seed(setup.coarse_cfg, seed, gws=int(gws))
    fine = host.initialized_lattice(fine_cfg, precision=GABE_FP32)
    plain = Lattice(precision=GABE_FP32, **coarse_cfg.lattice_kwargs())
    withc = Lattice(precision=GABE_FP32, **coarse_cfg.lattice_kwargs())
    restrict(fine, plain)
    restrict(fine, withc)
    runs = {"fine": (fine, coarse_steps * setup.ratio), "plain": (plain, coarse_steps),
            "closure": (withc, coarse_steps)}
    for name, (lat, n) in runs.items():
        if name == "closure":
            out["residual_threshold"] = threshold_factor * out["plain_friedmann_max"]
            attach(withc, cl, out["residual_threshold"])
        lat.sync()
        t0 = time.perf_counter()
        lat.step(n)
        lat.sync()
        out[f"{name}_seconds"] = time.perf_counter() - t0
        out[f"{name}_friedmann_max"] = float(host.friedmann_residual(lat.history())[0])
        out[f"{name}_edrho"] = lat.get_scalars()["edrho"][0]
        if log:
            log(f"{name}: {n} steps in {out[f'{name}_seconds']:.3f} s")
    nb = setup.coarse_cfg.N // 2
    spf = fine.specOut()
    for name in ("plain", "closure"):
        sp = runs[name][0].specOut()
        err = spectral_error(sp[[0, 2]], spf[[0, 2]], nb)
        out[f"{name}_field_err_max"] = float(err.max())
        out[f"{name}_field_err_mean"] = float(err.mean())
        out[f"{name}_field_err_bins"] = err.tolist()
        out[f"{name}_bins_within_5pct"] = float((err <= 0.05).mean())
        if gws:
            g = spectral_error(sp[[4]], spf[[4]], nb)
            out[f"{name}_gw_err_max"] = float(g.max())
            out[f"{name}_gw_err_mean"] = float(g.mean())
    out["closure_status"] = status(withc)
    out["speedup_vs_fine"] = out["fine_seconds"] / out["closure_seconds"]
    out["closure_cost_vs_plain"] = out["closure_seconds"] / out["plain_seconds"]
    for lat in (fine, plain, withc):
        lat.close()
    return out


def main(argv=None):
    """Train a closure on fine runs and evaluate it against a held-out fine run.

    Defaults: 128^3 coarse lattice, 512^3 fine lattice (ratio 4, dt = dx/20
    on both), lambda phi^4 + g^2 phi^2 chi^2 model of the benchmark, training
    seeds 1 and 2, held-out seed 3, 1000 coarse steps with tensor modes in the
    evaluated runs.  Prints one JSON line and writes the trained network."""
    import argparse
    import json
    import sys

    from . import configs
    ap = argparse.ArgumentParser(prog="python -m gabe_amd.closure")
    ap.add_argument("--N", type=int, default=128)
    ap.add_argument("--ratio", type=int, default=4)
    ap.add_argument("--train-seeds", type=int, nargs="+", default=[1, 2])
    ap.add_argument("--samples", type=int, default=100, help="training pairs per seed")
    ap.add_argument("--stride", type=int, default=10, help="coarse steps between pairs")
    ap.add_argument("--iters", type=int, default=2000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval-seed", type=int, default=3)
    ap.add_argument("--steps", type=int, default=1000, help="coarse steps of the evaluated runs")
    ap.add_argument("--threshold-factor", type=float, default=10.0,
                    help="fallback once the Friedmann residual exceeds this times the plain run's")
    ap.add_argument("--no-gw", action="store_true")
    ap.add_argument("--out", default=None, help="write the trained network here (.npz)")
    ap.add_argument("--load", default=None, help="skip training: use this network")
    args = ap.parse_args(argv)

    def log(msg):
        print(f"[closure] {msg}", file=sys.stderr, flush=True)

    setup = ClosureSetup.make(configs.lambda_phi4_config, args.N, args.ratio)
    res = {"N_coarse": args.N, "N_fine": args.N * args.ratio, "coarse_steps": args.steps,
           "fine_steps": args.steps * args.ratio, "gw": not args.no_gw}
    if args.load:
        cl = Closure.from_file(args.load)
    else:
        cl = Closure(args.N, setup.coarse_cfg.dt, seed=7)
        t0 = time.perf_counter()
        generate(cl, setup, args.train_seeds, args.samples, stride=args.stride, log=log)
        res["data_seconds"] = time.perf_counter() - t0
        res["samples"] = cl.num_samples()
        l0 = float(np.mean([cl.loss(s) for s in range(cl.num_samples())]))
        log(f"{cl.num_samples()} samples, initial loss {l0:.4f}")
        t0 = time.perf_counter()
        losses = cl.train(args.iters, lr=args.lr, seed=1)
        res["train_seconds"] = time.perf_counter() - t0
        l1 = float(np.mean([cl.loss(s) for s in range(cl.num_samples())]))
        log(f"trained {args.iters} iterations in {res['train_seconds']:.1f} s: loss {l0:.4f} -> {l1:.4f}")
        res.update(loss_initial=l0, loss_final=l1, loss_curve=[float(v) for v in losses[:: max(1, args.iters // 20)]])
        if args.out:
            cl.save(args.out)
    res.update(evaluate(cl, setup, args.eval_seed, args.steps, args.threshold_factor, gws=not args.no_gw, log=log))
    print(json.dumps(res))


if __name__ == "__main__":
    main()
