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


CLOSURE_KERNELS = ("first", "hidden", "output", "kick")


def kernel_times(lat, nsteps):
    """mean ms per step of the attached closure's kernels (first, hidden and
    output convolutions, the bi-Laplacian kick), timed with HIP events over
    nsteps steps launched one kernel after another"""
    out = (C.c_double * len(CLOSURE_KERNELS))()
    _check(load().gabe_closure_kernel_times(lat._h, nsteps, out))
    return dict(zip(CLOSURE_KERNELS, list(out)))


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


def validation_case(setup, seed, coarse_steps):
    """A held-out run for choosing among trained networks: the restricted
    coarse start state (a checkpoint blob) and the fine run's spectra after
    coarse_steps * ratio steps, fields only."""
    fine = host.initialized_lattice(_with_seed(setup.fine_cfg, seed, gws=0), precision=GABE_FP32)
    coarse = Lattice(precision=GABE_FP32, **_with_seed(setup.coarse_cfg, seed, gws=0).lattice_kwargs())
    restrict(fine, coarse)
    start = coarse.checkpoint()
    coarse.close()
    fine.step(coarse_steps * setup.ratio)
    spf = fine.specOut()
    fine.close()
    return start, spf


def validation_error(cl, setup, case, coarse_steps):
    """Worst and mean per-bin field-spectrum error of a closure run from the
    case's start state (no fallback), against the case's fine spectra."""
    start, spf = case
    lat = Lattice(precision=GABE_FP32, **setup.coarse_cfg.lattice_kwargs())
    lat.restore(start)
    attach(lat, cl, 1e300)
    lat.step(coarse_steps)
    err = spectral_error(lat.specOut()[[0, 2]], spf[[0, 2]], setup.coarse_cfg.N // 2)
    lat.close()
    return float(err.max()), float(err.mean())


def train_validated(cl, setup, schedule, case, coarse_steps, check_every, log=None):
    """Adam phases (lr, iterations); every check_every iterations the network
    is run on the validation case, and the parameters with the smallest worst
    bin are kept (the training loss weights the short wavelengths, where the
    coarse Laplacian's error is largest, and is not the rollout's error).
    Returns the loss per iteration and the validation history."""
    losses, hist = [], []
    best = (float("inf"), None)
    for n, (lr, iters) in enumerate(schedule):
        done = 0
        while done < iters:
            k = min(check_every, iters - done)
            losses += list(cl.train(k, lr=lr, seed=1000 * n + done))
            done += k
            worst, mean = validation_error(cl, setup, case, coarse_steps)
            hist.append({"iteration": len(losses), "lr": lr, "worst": worst, "mean": mean})
            if log:
                log(f"iteration {len(losses)} (lr {lr:g}): validation worst bin {worst:.4f}, mean {mean:.4f}")
            if worst < best[0]:
                best = (worst, cl.params().copy())
    if best[1] is not None:
        si, so = cl.scales()
        cl.set_params(best[1])
        cl.set_scales(si, so)
    return losses, hist


def summarize(per_seed, gws=True):
    """Worst and mean per-bin errors over all held-out seeds, both fields."""
    out = {}
    for name in ("plain", "closure"):
        errs = np.array([r[f"{name}_field_err_bins"] for r in per_seed.values()])   # [seed][field][bin]
        out[f"{name}_field_err_max"] = float(errs.max())
        out[f"{name}_field_err_mean"] = float(errs.mean())
        out[f"{name}_bins_within_5pct"] = float((errs <= 0.05).mean())
        worst = np.unravel_index(np.argmax(errs), errs.shape)
        out[f"{name}_worst_at"] = {"seed": list(per_seed)[worst[0]], "field": int(worst[1]), "bin": int(worst[2]) + 1}
        if gws:
            out[f"{name}_gw_err_mean"] = float(np.mean([r[f"{name}_gw_err_mean"] for r in per_seed.values()]))
            out[f"{name}_gw_err_max"] = float(np.max([r[f"{name}_gw_err_max"] for r in per_seed.values()]))
    out["closure_cost_vs_plain"] = float(np.mean([r["closure_cost_vs_plain"] for r in per_seed.values()]))
    out["speedup_vs_fine"] = float(np.mean([r["speedup_vs_fine"] for r in per_seed.values()]))
    return out


def main(argv=None):
    """Train a closure on fine runs and evaluate it against a held-out fine run.

    Defaults: 128^3 coarse lattice, 512^3 fine lattice (ratio 4, dt = dx/20
    on both), lambda phi^4 + g^2 phi^2 chi^2 model of the benchmark, training
    seeds 1 and 2, held-out seeds 3, 4 and 5, 1000 coarse steps with tensor
    modes in the evaluated runs.  Prints one JSON line (per seed and a summary
    over the seeds) and writes the trained network."""
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
    ap.add_argument("--schedule", default="1e-3:2000",
                    help="Adam phases as lr:iterations, comma separated (the moments carry over)")
    ap.add_argument("--val-seed", type=int, default=0,
                    help="validation seed: keep the parameters whose closure run matches it best (0: last)")
    ap.add_argument("--check-every", type=int, default=250, help="iterations between validation runs")
    ap.add_argument("--eval-seeds", type=int, nargs="+", default=[3, 4, 5], help="held-out seeds")
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
        schedule = [(float(p.split(":")[0]), int(p.split(":")[1])) for p in args.schedule.split(",")]
        if args.val_seed:
            case = validation_case(setup, args.val_seed, args.steps)
            losses, vhist = train_validated(cl, setup, schedule, case, args.steps, args.check_every, log=log)
            res["validation"] = {"seed": args.val_seed, "history": vhist,
                                 "chosen_worst": min(h["worst"] for h in vhist)}
        else:
            losses = []
            for n, (lr, iters) in enumerate(schedule):
                losses += list(cl.train(iters, lr=lr, seed=1 + n))
        iters = len(losses)
        res["train_seconds"] = time.perf_counter() - t0
        l1 = float(np.mean([cl.loss(s) for s in range(cl.num_samples())]))
        log(f"trained {iters} iterations in {res['train_seconds']:.1f} s: loss {l0:.4f} -> {l1:.4f}")
        res.update(schedule=args.schedule, loss_initial=l0, loss_final=l1,
                   loss_curve=[float(v) for v in losses[:: max(1, iters // 20)]])
        if args.out:
            cl.save(args.out)
    per_seed = {}
    for seed in args.eval_seeds:
        per_seed[str(seed)] = evaluate(cl, setup, seed, args.steps, args.threshold_factor, gws=not args.no_gw, log=log)
        r = per_seed[str(seed)]
        log(f"seed {seed}: field worst bin {r['closure_field_err_max']:.4f} (plain {r['plain_field_err_max']:.4f}), "
            f"mean {r['closure_field_err_mean']:.4f} (plain {r['plain_field_err_mean']:.4f}), "
            f"within 5% {r['closure_bins_within_5pct']:.3f}")
    res["eval"] = per_seed
    res["summary"] = summarize(per_seed, gws=not args.no_gw)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
