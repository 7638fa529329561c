"""Throughput of the lattice step on one MI355X.

Workload: a 256^3 periodic lattice with two scalar fields (V = lambda phi^4/4
+ g^2 phi^2 chi^2/2, g^2/lambda = 100, L = 20, dt = dx/20) and the six
tensor-perturbation components, midpoint RK2 with self-consistent expansion.
Initial state: homogeneous phi0 = 0.342 with slow-roll velocity plus the seeded
vacuum fluctuations, produced by the kept host's start-up sequence.

value  = site-updates/s of the FP32 form (FP32 fluctuations, FP64 background,
         a(t) and reductions) = N^3 x steps / seconds, inputs resident in HBM.
Also reported: the same for the FP64 form, per-kernel mean durations (HIP
events on the lattice's stream), the roofline of the dominant kernel, and the
CPU oracle timed on the host's cores on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "lattice site-updates/s (RK2, 2 fields+GW) at 256³ FP32 vs FP64; % HBM peak"


def alg_bytes_per_site(s, N, rows_in_sweep=True):
    """Compulsory HBM bytes per site-update of each kernel class of this
    pipeline (s = bytes per real, r = (N/2+1)/N modes per site).

    The stress is stored as 10 components (xx - zz, xy, xz, yy - zz, yz of
    both RK2 stages): the TT projection removes any multiple of delta_ij.

    sweep_a    reads the 4 state arrays; with the k-rows transformed in the
               sweep writes the 10 half-spectrum rows as N/2 complex each
               (kz = 0 and N/2, both real, packed in one slot), otherwise the
               10 real stress arrays
    stress_fft j columns: reads the packed rows and writes them back with
               kz = N/2 split off into a plane of its own (the 10 half-spectra);
               without rows in the sweep, the library k-row R2C and j columns
               in place
    gw_modes   i columns fused with the mode update: reads the 10 half-spectra
               and h, l (two TT polarisations, complex), writes h, l
    sweep_b    reads and writes the 4 state arrays
    """
    r = (N // 2 + 1) / N
    return {
        "sweep_a": 14 * s,   # 10 packed rows carry as many bytes as 10 real arrays
        "finalize": 0.0,
        "stress_fft": (10 * s + 20 * s * r) if rows_in_sweep else (10 * s + 20 * s * r + 40 * s * r),
        "gw_modes": 20 * s * r + 16 * s * r,
        "sweep_b": 8 * s,
    }


def b_alg(s, N):
    """Bytes per site-update of the minimal dataflow used for the step
    roofline: sweeps 12s, stress 24s, one read + write per transform 24s r,
    six-component tensor modes 72s r (36s + 96s r)."""
    r = (N // 2 + 1) / N
    return 36 * s + 96 * s * r


def read_traffic(kernel, N):
    """HBM bytes per launch from the committed rocprofv3 counter summary, if any."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{N}.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def run_device(cfg, precision, steps, warmup, profile_steps, ppb=0):
    """Time `steps` steps after `warmup`; then `profile_steps` more with
    per-kernel events.  Returns the time, kernel means, the tiling in use and
    the final state, scalars, spectra and Friedmann residual for the precision
    comparison."""
    from gabe_amd import host
    lat = host.initialized_lattice(cfg, precision=precision, planes_per_block=ppb)
    q = lat.query()
    tiling = {"planes_per_block": q.planes_per_block, "transform_tile": q.transform_tile, "mode_tile": q.mode_tile,
              "stress_chunk": q.stress_chunk}
    lat.step(warmup)
    lat.sync()
    t0 = time.perf_counter()
    lat.step(steps)
    lat.sync()
    t1 = time.perf_counter()
    ktimes = lat.kernel_times(profile_steps) if profile_steps else {}
    final = {"scalars": lat.get_scalars(), "state": lat.download_state(), "spectra": lat.specOut(),
             "friedmann_max": host.friedmann_residual(lat.history())[0], "tiling": tiling}
    lat.close()
    return t1 - t0, ktimes, final


def precision_report(f32, f64):
    """FP32 form against the FP64 form after the same steps from the same
    initial state: field L2 error relative to the fluctuation, per-bin
    spectra (bins above 1e-9 of the column maximum), energy and the Friedmann
    residual of each."""
    import numpy as np
    out = {}
    for name, k in (("field", 0), ("velocity", 1)):
        a, b = f32["state"][k], f64["state"][k]
        errs = []
        for f in range(b.shape[0]):
            fl = b[f] - b[f].mean()
            errs.append(float(np.linalg.norm(a[f] - b[f]) / np.linalg.norm(fl)))
        out[f"{name}_l2_rel"] = errs
    sp32, sp64 = f32["spectra"], f64["spectra"]
    cols = {"field_spectra": [0, 2], "gw_spectrum": [4]}
    for name, cs in cols.items():
        worst = 0.0
        for c in cs:
            ref, got = sp64[c], sp32[c]
            if not (np.all(np.isfinite(ref)) and np.all(np.isfinite(got))):
                out[f"{name}_nonfinite_column"] = c
                worst = float("nan")
                continue
            live = np.abs(ref) > 1e-9 * np.abs(ref).max()
            live[0] = False
            if live.any():
                worst = max(worst, float(np.max(np.abs(got[live] / ref[live] - 1.0))))
        out[f"{name}_max_rel"] = worst
    r32, r64 = f32["scalars"]["edrho"][0], f64["scalars"]["edrho"][0]
    out["edrho_rel"] = abs(r32 - r64) / abs(r64)
    out["a_rel"] = abs(f32["scalars"]["a"][0] - f64["scalars"]["a"][0]) / f64["scalars"]["a"][0]
    out["friedmann_max_fp32"] = float(f32["friedmann_max"])
    out["friedmann_max_fp64"] = float(f64["friedmann_max"])
    return out


def run_ensemble(steps, warmup):
    """Aggregate site-updates/s of the batched ensemble against the same 64
    lattices run one after another, each to completion, on their own contexts;
    per-phase HIP-event times of the batched launches and the roofline of the
    longest (fields only: sweep A reads the 4 state arrays, sweep B reads and
    writes them)."""
    from gabe_amd import configs, host
    from gabe_amd.lattice import GABE_FP32, Ensemble, Lattice
    cfgs = configs.ensemble_configs()
    sites = sum(c.N ** 3 for c in cfgs)
    ens = Ensemble([dict(precision=GABE_FP32, **c.lattice_kwargs()) for c in cfgs])
    for c, m in zip(cfgs, ens.members):
        host.initialize_member(m, c)
    ens.step(warmup)
    ens.sync()
    t0 = time.perf_counter()
    ens.step(steps)
    ens.sync()
    t_ens = time.perf_counter() - t0
    ktimes = ens.kernel_times(20)
    ens.sync()
    ppb = [m.query().planes_per_block for m in ens.members]
    tiling = {str(c.N): p for c, p in zip(cfgs, ppb)}
    ens.close()

    singles = []
    for c in cfgs:
        s = Lattice(precision=GABE_FP32, **c.lattice_kwargs())
        host.initialize(s, c)
        s.step(warmup)
        s.sync()
        singles.append(s)
    t0 = time.perf_counter()
    for s in singles:
        s.step(steps)
        s.sync()
    t_seq = time.perf_counter() - t0
    for s in singles:
        s.close()
    rate = sites * steps / t_ens
    alg = {"sweep_a": 16.0, "sweep_b": 32.0}   # bytes per site of the FP32 state arrays
    dom = max(alg, key=lambda k: ktimes[k])
    achieved = alg[dom] * sites / (ktimes[dom] * 1e-3) / 1e9
    return {
        "value": rate,
        "unit": "site-updates/s",
        "sequential_value": sites * steps / t_seq,
        "speedup_vs_sequential": t_seq / t_ens,
        "ms_per_step": 1e3 * t_ens / steps,
        "sequential_ms_per_step": 1e3 * t_seq / steps,
        "kernels_ms": ktimes,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(f"ensemble_{dom}", "64lat"),
                     "alg_bytes_per_launch": alg[dom] * sites},
        "fields_frac": 48.0 * rate / (HBM_PEAK_GBS * 1e9),
        "planes_per_block": tiling,
        "lattices": len(cfgs),
        "sites": sites,
        "steps": steps,
        "warmup": warmup,
        "workload": "64 lattices (16 each of 32^3, 64^3, 96^3, 128^3), L=20, dt=dx/20, 2 fields, no GW, "
                    "distinct seeds, g^2/lambda log-uniform in [1,200], FP32 fields / FP64 per-lattice a(t)",
    }


def run_closure(steps, warmup):
    """128^3 with tensor modes: the plain step against the step with the
    committed trained closure (gabe_amd/closure_128.npz) attached."""
    from gabe_amd import closure as cm
    from gabe_amd import configs, host
    from gabe_amd.lattice import GABE_FP32
    path = os.path.join(ROOT, "gabe_amd", "closure_128.npz")
    cfg = configs.lambda_phi4_config(N=128, calc_gws=1)
    cl = cm.Closure.from_file(path)
    out = {}
    for name in ("plain", "closure"):
        lat = host.initialized_lattice(cfg, precision=GABE_FP32)
        if name == "closure":
            cm.attach(lat, cl, 1e-3)
        lat.step(warmup)
        lat.sync()
        t0 = time.perf_counter()
        lat.step(steps)
        lat.sync()
        out[f"{name}_ms_per_step"] = 1e3 * (time.perf_counter() - t0) / steps
        if name == "closure":
            out["closure_active"] = cm.status(lat)[0]
            out["kernels_ms"] = cm.kernel_times(lat, 20)
        lat.close()
    cl.close()
    out["cost_vs_plain"] = out["closure_ms_per_step"] / out["plain_ms_per_step"]
    # useful multiply-adds on the matrix cores, 2 x c_in x c_out x 27 flop per site and convolution
    n3 = cfg.N ** 3
    conv_flops = {"first": 2.0 * 4 * 16 * 27 * n3, "hidden": 2.0 * 16 * 16 * 27 * n3, "output": 2.0 * 16 * 2 * 27 * n3}
    # algorithmic HBM bytes per site: the first layer reads the four FP32 lattice arrays and writes 16 BF16
    # channels, the hidden one reads and writes 16 BF16 channels, the output one reads 16 and writes 2 FP32
    conv_bytes = {"first": 48.0 * n3, "hidden": 64.0 * n3, "output": 40.0 * n3}
    kt = out["kernels_ms"]
    dom = max(conv_flops, key=lambda k: kt[k])
    sec = kt[dom] * 1e-3
    tf = conv_flops[dom] / sec / 1e12
    gbs = conv_bytes[dom] / sec / 1e9
    # the bound is the roof that allows the longer time: 216 flop/B for the hidden convolution, below the
    # BF16 ridge of 2.5 PFLOP/s over 8 TB/s = 312 flop/B, so HBM
    hbm_bound = conv_bytes[dom] / 8e12 >= conv_flops[dom] / 2.5e15
    out["roofline"] = {"bound": "hbm" if hbm_bound else "mfma", "kernel": f"{dom} convolution",
                       "dtype": "bf16 (fp32 accumulate)", "alg_bytes_per_launch": conv_bytes[dom],
                       "flop_per_launch": conv_flops[dom],
                       "achieved": gbs if hbm_bound else tf, "peak": 8000.0 if hbm_bound else 2500.0,
                       "unit": "GB/s" if hbm_bound else "TFLOP/s",
                       "frac": gbs / 8000.0 if hbm_bound else tf / 2500.0,
                       "mfma": {"achieved": tf, "peak": 2500.0, "unit": "TFLOP/s", "frac": tf / 2500.0},
                       "hbm": {"achieved": gbs, "peak": 8000.0, "unit": "GB/s", "frac": gbs / 8000.0},
                       "note": "the longest convolution over its mean HIP-event duration in steps launched one "
                               "kernel after another; bound = the roof allowing the longer time"}
    # the whole network per step, from the step times (includes the kick and the side stream's share of the GPU)
    flops = sum(conv_flops.values())
    t = (out["closure_ms_per_step"] - out["plain_ms_per_step"]) * 1e-3
    out["per_step"] = {"flop_per_eval": flops, "achieved": flops / t / 1e12 if t > 0 else None, "unit": "TFLOP/s",
                       "frac": flops / t / 2.5e15 if t > 0 else None,
                       "note": "time = closure step minus plain step (includes the input packing and the "
                               "bi-Laplacian kick)"}
    out["workload"] = "128^3, 2 fields + GW, FP32 form, trained closure (3x 3^3 conv, 16 ch, BF16) after every step"
    out["steps"] = steps
    return out


def _oracle_for(cfg, threads):
    from oracle import pyoracle as po
    return po.Oracle(cfg.N, nflds=cfg.nflds, L=cfg.L, dt=cfg.dt, calc_gws=cfg.calc_gws,
                     expansion_type=cfg.expansion_type, m2_phi=cfg.m2_phi, lam=cfg.lam, g2=cfg.g2,
                     m2_chi=cfg.m2_chi, rescale_B=cfg.rescale_B, nthreads=threads)


def _time_oracle(cfg, state, threads, nsteps):
    """nsteps oracle steps from `state` (fields and velocities, FP64) with
    `threads` OpenMP threads; returns site-updates/s and seconds."""
    o = _oracle_for(cfg, threads)
    f, d = state
    o.field[0] = f
    o.dfield[0] = d
    o.set_scalar("a", 0, 1.0)
    o.calcEnergy(0)
    o.set_scalar("adot", 0, o.adf(0))
    t0 = time.perf_counter()
    o.step(nsteps)
    dt = time.perf_counter() - t0
    del o
    return cfg.N ** 3 * nsteps / dt, dt


def cpu_baseline(cfg, state, nsteps=2):
    """The CPU FP64 oracle (the reference's loops restated in C, OpenMP over
    the outer index) on this host's cores: the benchmark lattice (2 steps) on
    the host's CPU share and on 4 threads (the reference's default thread
    count, tot_num_thrds in g2parameters.h:54), and the plain 64^3 workload
    without tensor modes over its full 200 steps on both thread counts, with
    the same 200 steps on the GPU in FP64 beside them."""
    from gabe_amd import configs, host
    from gabe_amd.lattice import GABE_FP64
    nproc = os.cpu_count() or 1
    share = min(16, nproc)   # one GPU's share of the box's host cores
    main_rate, main_s = _time_oracle(cfg, state, share, nsteps)
    rate4, s4 = _time_oracle(cfg, state, 4, nsteps)
    c1 = configs.lambda_phi4_config(N=64, calc_gws=0)
    lat = host.initialized_lattice(c1, precision=GABE_FP64)
    f1, d1 = lat.download_state()
    lat.sync()
    t0 = time.perf_counter()
    lat.step(200)
    lat.sync()
    gpu_s = time.perf_counter() - t0
    lat.close()
    c1_share, c1_share_s = _time_oracle(c1, (f1, d1), share, 200)
    c1_four, c1_four_s = _time_oracle(c1, (f1, d1), 4, 200)
    del f1, d1
    return {
        "value": main_rate,
        "unit": "site-updates/s",
        "cores": share,
        "kind": "port",
        "host_cpus": nproc,
        "sample": f"{nsteps} RK2 steps of the same {cfg.N}^3 2-field+GW lattice in FP64 (oracle restatement, "
                  f"OpenMP over i, {share} threads = one GPU's share of the {nproc} host CPUs), {main_s:.1f} s",
        "threads_4": {"value": rate4, "cores": 4, "seconds": s4,
                      "sample": f"the same {nsteps} steps on 4 threads (tot_num_thrds, g2parameters.h:54)"},
        "plain_64_200_steps": {
            "workload": "64^3, 2 fields, no GW, FP64, 200 RK2 steps from the seeded vacuum state",
            "cpu": {f"threads_{share}": {"value": c1_share, "seconds": c1_share_s},
                    "threads_4": {"value": c1_four, "seconds": c1_four_s}},
            "gpu_fp64": {"value": 64 ** 3 * 200 / gpu_s, "seconds": gpu_s},
            "unit": "site-updates/s",
        },
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=1000, help="steps of the 64-lattice ensemble (0: skip)")
    ap.add_argument("--closure-steps", type=int, default=200, help="steps of the 128^3 closure timing (0: skip)")
    ap.add_argument("--ppb", type=int, default=0, help="lattice planes per sweep workgroup (0: autotuned)")
    ap.add_argument("--no-autotune", action="store_true", help="use the built-in default tiles")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs every lattice on a single MI355X "
              "(multi-GPU sharding is a design note only); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    from gabe_amd import configs
    from gabe_amd.lattice import GABE_FP32, GABE_FP64

    N = args.N
    cfg = configs.lambda_phi4_config(N=N, calc_gws=1)
    sites = N ** 3

    autotune = None
    if not args.no_autotune:
        from gabe_amd.lattice import autotune_sweep
        autotune = {}
        for name, prec in (("fp32", GABE_FP32), ("fp64", GABE_FP64)):
            if name == "fp64" and args.no_fp64:
                continue
            print(f"[bench] autotuning the {name} {N}^3 tiles", file=sys.stderr, flush=True)
            t0 = time.perf_counter()
            ppb, ms = autotune_sweep(precision=prec, **cfg.lattice_kwargs())
            autotune[name] = {"best_ms_per_step": ms, "seconds": time.perf_counter() - t0}
    print(f"[bench] FP32 {N}^3: warmup {args.warmup}, timed {args.steps}", file=sys.stderr, flush=True)
    dt32, k32, fin32 = run_device(cfg, GABE_FP32, args.steps, args.warmup, args.profile_steps, args.ppb)
    sc32, state, tiling32 = fin32["scalars"], fin32["state"], fin32["tiling"]
    rate32 = sites * args.steps / dt32
    fp64 = None
    if not args.no_fp64:
        print(f"[bench] FP64 {N}^3", file=sys.stderr, flush=True)
        dt64, k64, fin64 = run_device(cfg, GABE_FP64, args.steps, args.warmup, args.profile_steps, args.ppb)
        fp64 = {"value": sites * args.steps / dt64, "ms_per_step": 1e3 * dt64 / args.steps,
                "tiling": fin64["tiling"],
                "kernels_ms": k64,
                "step_hbm_frac": b_alg(8, N) * sites * args.steps / dt64 / (HBM_PEAK_GBS * 1e9),
                "fp32_vs_fp64": dict(steps=args.warmup + args.steps + args.profile_steps,
                                     **precision_report(fin32, fin64))}
        del fin64

    ab = alg_bytes_per_site(4, N, rows_in_sweep=16 <= N <= 512 and (N & (N - 1)) == 0)
    roofline = None
    if k32:
        dom = max((k for k in k32 if k != "finalize"), key=lambda k: k32[k])
        per_launch = ab[dom] * sites
        achieved = per_launch / (k32[dom] * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(dom, N), "kernel": dom,
                    "alg_bytes_per_launch": per_launch}
    kernels = {k: {"ms": v, "alg_GBs": (ab[k] * sites / (v * 1e-3) / 1e9) if v > 0 and ab[k] else None}
               for k, v in k32.items()}

    ensemble = None
    if args.ensemble_steps:
        print(f"[bench] ensemble of 64 lattices, {args.ensemble_steps} steps", file=sys.stderr, flush=True)
        ensemble = run_ensemble(args.ensemble_steps, min(args.warmup, 20))

    del fin32
    closure = None
    if args.closure_steps:
        print(f"[bench] closure at 128^3, {args.closure_steps} steps", file=sys.stderr, flush=True)
        closure = run_closure(args.closure_steps, min(args.warmup, 20))

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] CPU oracle sample", file=sys.stderr, flush=True)
        cpu = cpu_baseline(cfg, state)

    out = {
        "metric": METRIC,
        "value": rate32,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt32 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: seeded vacuum fluctuations (reference IC recipe, randseed 44463132) on the homogeneous "
                "slow-roll background",
        "config": {"workload": f"{N}^3 lattice, 2 scalar fields + 6 GW tensor components, midpoint RK2, "
                               "self-consistent expansion, FP32 fields / FP64 background, a(t) and sums",
                   "N": N, "L": cfg.L, "dt": cfg.dt, "model": "lambda phi^4/4 + g^2 phi^2 chi^2/2, g^2/lambda=100"},
        "roofline": roofline,
        "step_hbm_frac": b_alg(4, N) * rate32 / (HBM_PEAK_GBS * 1e9),
        "step_bytes_per_site": {"b_alg": b_alg(4, N), "this_pipeline": sum(ab.values())},
        "step_frac_own_bytes": sum(ab.values()) * rate32 / (HBM_PEAK_GBS * 1e9),
        "tiling": tiling32,
        "autotune": autotune,
        "kernels": kernels,
        "fp64": fp64,
        "ensemble": ensemble,
        "closure": closure,
        "cpu_baseline": cpu,
        "final_a": sc32["a"][0],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
