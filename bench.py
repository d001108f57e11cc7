"""Benchmark: sheared Heun particle-steps/s on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload shear_large]

A step is one Heun sub-step (predictor force pass + corrector force pass with
Lees-Edwards shear and cell re-binning) over the whole packing; inputs are
resident in HBM before the timed region.  Prints one JSON line.

Workloads (seeded synthetic packings, spherocyls_amd/configs.py):
  shear_large  N=262144, R uniform [0.5, 0.7], aspect 1.5, phi=0.845,
               strain rate 1e-6, dt 0.01  (default; largest single-GPU case)
  compress     N=16384 bidisperse aspect 2, phi=0.80, strain rate 1e-5
  shear_small  N=512 bidisperse aspect 2, phi=0.85, strain rate 1e-5

Beside the headline fields the line carries, unless switched off: the
reference CPU path on the same packing (cpu_baseline), packings/s minimised
over the ragged 256-packing ensemble (minimization), the quasi-static
compression curve (quasistatic), and the large shear in mixed precision with
minimisation checkpoints every 1000 steps (checkpoints).
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from spherocyls_amd import configs, native  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per particle and launch (FP64 SoA, compulsory arrays once):
#   predictor: read x,y,phi,R,A (40) + write x~,y~,phi~ (24) + write v0 (24)      =  88
#   corrector: read x~ (24) + R,A (16) + v0 (24) + x (24), write x (24) + moved (32) = 144
BYTES_PRED = 88
BYTES_CORR = 144
# HBM bytes per corrector launch from rocprofv3 FETCH_SIZE / WRITE_SIZE runs
# (list path, 2 lanes per particle, 256-thread workgroups, median of the 100
# event-timed launches after 1000 steps), scaled by a plain copy's readings
# (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0; profiles/k_list_heun_shear_N262144.md):
# (workload, path, precision) -> bytes
MEASURED_TRAFFIC = {("shear_large", "list", "fp64", 0.1): (22705.47 * 2.0 + 18303.91) * 1024}

WORKLOADS = {
    "shear_large": (lambda: configs.shear_large(1), 1e-6, 0.01),
    "compress": (lambda: configs.compress_start(1), 1e-5, 0.01),
    "shear_small": (lambda: configs.shear_small(1), 1e-5, 0.01),
}


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(pk, gamma, strain_rate, dt, steps=200, runs=3):
    """Reference CPU path (oracle/_ref: the reference's own src/cpp compiled
    here; else the oracle restatement), 1 thread, on the same relaxed state the
    GPU line is timed from: `runs` runs of `steps` Heun steps each, each from
    that state; the median run's rate."""
    import oracle

    kind = "reference" if oracle.ref_available() else "port"
    cls = oracle.RefBox if kind == "reference" else oracle.OracleBox
    rates = []
    for _ in range(runs):
        b = cls.from_packing(pk, 0.1)
        b.set_params(gamma, 0.0, strain_rate, 0.0, dt)
        b.find_neighbors()
        b.calc_forces()
        t0 = time.perf_counter()
        b.run_steps(steps, 0)
        rates.append(pk.n * steps / (time.perf_counter() - t0))
        print(f"cpu baseline run {len(rates)}/{runs}: {rates[-1]:.4g} particle-steps/s",
              file=sys.stderr, flush=True)
        del b
    return {"value": float(np.median(rates)), "unit": "particle-steps/s", "cores": 1,
            "kind": kind, "runs": [float(r) for r in rates],
            "sample": f"median of {runs} runs of {steps} Heun steps of the same relaxed "
                      f"N={pk.n} packing, 1 thread ({cpu_model()}, nproc={os.cpu_count()})"}


# FIRE, per particle-iteration, non-fused force + update passes: 112 B for a
# fused pass (read x, y, phi, R, A and v; write x, y, phi and v) plus 48 B for
# writing the forces in the force pass and reading them back in the update pass
BYTES_FIRE = 160


def minimization(n_packings, reps, cpu_budget_s):
    """packings/s minimised to F_rms < 1e-10 over a ragged ensemble: N log-uniform
    in [64, 4096], phi U[0.80, 0.86], aspect {1.2, 2, 4}; each packing first
    gets 2000 sheared Heun steps (strain rate 1e-4, dt 0.01), then FIRE with the
    two-point contact model.  GPU paths from the same states: all packings
    batched in FP64 (the headline), batched in mixed precision, batched mixed
    precision from the learned warm start (when trained weights exist), and
    one packing at a time in FP64 (the plain path).  Each rate counts only
    packings that reached the tolerance."""
    from spherocyls_amd import ensemble

    pks = configs.ensemble(n_packings)
    t0 = time.perf_counter()
    # cell size from the cell-size autotuner (200 timed Heun steps per
    # candidate padding), launch shape from the shape autotuner
    work = ensemble.prepare(pks, heun_steps=2000, strain_rate=1e-4, dt=0.01,
                            tune_pad=(0.05, 0.1, 0.2, 0.3), tune_steps=200)
    prep_s = time.perf_counter() - t0
    params = native.fire_params(max_iter=20000)
    out = {"metric": "packings/s minimized to F_rms<1e-10", "unit": "packings/s",
           "packings": n_packings, "particles": work.particles, "contact_model": "two_point",
           "prepare_s": prep_s,
           "cell_pad": {"chosen": work.pad, "trial_ms_200_heun_steps": work.pad_trials}}
    runs = {}
    for label, prec in (("batched_mixed", "mixed"), ("batched_fp64", "fp64")):
        best = None
        for _ in range(reps):
            sec, it, fr, st = ensemble.minimize_batched(work, params, precision=prec)
            if best is None or sec < best[0]:
                best = (sec, it, fr, st)
        sec, it, fr, st = best
        conv = int(np.sum(fr < 1e-10))
        pit = float(np.sum(np.array(work.sizes) * it))
        # only packings that reached F_rms < 1e-10 count
        runs[label] = {"value": conv / sec, "seconds": sec, "converged": conv,
                       "max_iterations": int(it.max()), "mean_iterations": float(it.mean()),
                       "particle_iterations_per_s": pit / sec,
                       "hbm_frac_fire_bytes": pit * BYTES_FIRE / sec / (HBM_PEAK_GBS * 1e9)}
    from spherocyls_amd import warmstart
    if os.path.exists(warmstart.DEFAULT_WEIGHTS):
        net = warmstart.load()
        try:
            sec, it, fr, fb = warmstart.minimize_with_warm_start(work, params, net,
                                                                 precision="mixed")
        finally:
            net.close()
        conv = int(np.sum(fr < 1e-10))
        runs["batched_mixed_warm_start"] = {"value": conv / sec, "seconds": sec,
                                            "converged": conv,
                                            "mean_iterations": float(it.mean()),
                                            "fallbacks": len(fb)}
    # what a warm start could save at best: exact displacements to the minimum
    # applied with quality q (q = 0: the plain run)
    out["warm_start_ceiling"] = warmstart.ceiling(work, params)
    sec, it, fr = ensemble.minimize_one_by_one(work, params, precision="fp64")
    conv = int(np.sum(fr < 1e-10))
    runs["one_by_one_fp64"] = {"value": conv / sec, "seconds": sec, "converged": conv,
                               "mean_iterations": float(it.mean())}
    out.update(runs)
    # headline: every packing in one batched FP64 system (the fastest path)
    plain = runs["one_by_one_fp64"]["value"]
    out["value"] = runs["batched_fp64"]["value"]
    out["speedup_vs_one_by_one_fp64"] = runs["batched_fp64"]["value"] / plain
    if "batched_mixed_warm_start" in runs:
        # batched mixed precision with the learned warm start against the plain
        # FP64 GPU path (one packing at a time, no warm start): target >= 3
        r = runs["batched_mixed_warm_start"]["value"] / plain
        out["mixed_warm_start_vs_plain_fp64"] = {"ratio": r, "target": 3.0, "met": r >= 3.0}
        # the same path against batched FP64 without the warm start: what the
        # learned part and mixed precision add on top of batching
        out["mixed_warm_start_vs_batched_fp64"] = (runs["batched_mixed_warm_start"]["value"] /
                                                   runs["batched_fp64"]["value"])
    if cpu_budget_s > 0:
        out["cpu_baseline"] = cpu_minimization(work, params, cpu_budget_s)
    return out


def quasistatic_compression(eps=1e-4, phi0=0.80, phi1=0.86, n=16384, samples=13):
    """Quasi-static compression: N=16384 bidisperse (R 0.5 / 0.7) aspect-2
    packing, seeded random start at phi 0.80, steps x, y, L *= 1 - 1e-4 (the
    reference's relative length step) until phi >= 0.86, FIRE (two-point
    contacts) to F_rms < 1e-10 after each step; FP64 and mixed precision.
    Reports the pressure-vs-packing-fraction curve (sampled), the FIRE
    iterations and the mixed-mode pressure error against FP64 along the curve."""
    from spherocyls_amd import quasistatic

    pk = configs.compress_start(1, n)
    params = native.fire_params(max_iter=200000)
    out = {"workload": f"quasi-static compression N={n} phi {phi0}->{phi1} eps={eps} aspect 2",
           "contact_model": "two_point"}
    curves = {}
    for prec in ("fp64", "mixed"):
        s = native.System(pk, pads=0.1)
        try:
            s.set_contact_model("two_point")
            s.set_precision(prec)
            s.autotune(reps=3)
            def progress(step, _s):
                if step % 100 == 0:
                    print(f"quasi-static {prec}: step {step}", file=sys.stderr, flush=True)
            c = quasistatic.compress(s, eps, phi1, params, on_step=progress, tune_fire=True)
        finally:
            s.close()
        curves[prec] = c
        it = c.iterations[:, 0]
        out[prec] = {"seconds": c.seconds, "steps": int(len(it) - 1),
                     "converged_minimisations": int(np.sum(c.frms[:, 0] < 1e-10)),
                     "fire_iterations_first": int(it[0]),
                     "fire_iterations_steps": int(it[1:].sum()),
                     "mean_iterations_per_step": float(it[1:].mean()),
                     "steps_per_s": (len(it) - 1) / c.seconds,
                     "particle_iterations_per_s": n * float(it.sum()) / c.seconds,
                     "launch_shape": c.tuned}
    f64, mix = curves["fp64"], curves["mixed"]
    k = min(len(f64.packing), len(mix.packing))
    idx = np.unique(np.linspace(0, len(f64.packing) - 1, samples).astype(int))
    out["curve_fp64"] = [[float(f64.packing[i, 0]), float(f64.pressure[i, 0]),
                          int(f64.iterations[i, 0])] for i in idx]
    p64, pmx = f64.pressure[:k, 0], mix.pressure[:k, 0]
    jam = (p64 > 1e-8) & (pmx > 1e-8)  # finite pressure; unjammed states read ~1e-11
    out["mixed_pressure_rel_error"] = {
        "max": float(np.max(np.abs(pmx[jam] - p64[jam]) / p64[jam])) if np.any(jam) else None,
        "median": float(np.median(np.abs(pmx[jam] - p64[jam]) / p64[jam])) if np.any(jam) else None,
        "jammed_steps": int(np.sum(jam))}
    return out


def jammed_compression(n=1024, phi0=0.86, phi1=0.90, eps=1e-3, seed=3):
    """The compression continued past jamming (bidisperse aspect 2, seeded
    random start at phi0, x, y, L *= 1 - eps per step to phi1, FIRE with the
    two-point model to F_rms < 1e-10 after each step), in FP64 and in mixed
    precision.  Reports, in the jammed steps (finite pressure): the mixed run's
    reported pressure against the oracle's FP64 evaluation of the same state
    (the precision error of the report), and against the FP64 run's pressure
    at the same step (the two runs reach different minima once FP32 pair
    forces set FIRE on another path: the spread of the jammed states)."""
    import oracle
    from spherocyls_amd import quasistatic

    R, A = configs.bidisperse(n)
    pk = configs.make_packing(R, A, phi0, seed, f"jammed compression N={n}")
    params = native.fire_params(max_iter=400000)
    curves, final = {}, {}
    for prec in ("fp64", "mixed"):
        s = native.System(pk, pads=0.1)
        try:
            s.set_contact_model("two_point")
            s.set_precision(prec)
            curves[prec] = quasistatic.compress(s, eps, phi1, params)
            final[prec] = (s.positions(), s.box())
        finally:
            s.close()
    f64, mix = curves["fp64"], curves["mixed"]
    k = min(len(f64.packing), len(mix.packing))
    p64, pmx = f64.pressure[:k, 0], mix.pressure[:k, 0]
    jam = (p64 > 1e-8) & (pmx > 1e-8)
    (x, y, phi), bx = final["mixed"]
    ob = oracle.OracleBox(bx["Lx"], x, y, phi, pk.A, pk.R, 0.1)
    ob.set_contact_model(1)
    ob.find_neighbors()
    ob.calc_se()
    o = ob.scalars()
    p_orc = -0.5 * (o["cPxx"] + o["cPyy"])
    p_rep = -0.5 * (bx["cPxx"] + bx["cPyy"])
    rel = np.abs(pmx[jam] - p64[jam]) / p64[jam]
    return {"workload": f"N={n} bidisperse aspect 2, phi {phi0} -> {phi1}, eps {eps}, "
                        "two-point FIRE to 1e-10",
            "steps": int(k - 1), "jammed_steps": int(np.sum(jam)),
            "final_pressure_fp64": float(p64[-1]), "final_pressure_mixed": float(pmx[-1]),
            "mixed_report_vs_oracle_rel_error": abs(p_rep - p_orc) / abs(p_orc),
            "mixed_final_frms_oracle": float(ob.frms()),
            "mixed_vs_fp64_run_rel_diff": {"max": float(rel.max()) if len(rel) else None,
                                           "median": float(np.median(rel)) if len(rel) else None},
            "converged": {p: int(np.sum(c.frms[:, 0] < 1e-10)) for p, c in curves.items()},
            "mean_iterations": {p: float(c.iterations[1:, 0].mean()) for p, c in curves.items()},
            "seconds": {p: c.seconds for p, c in curves.items()}}


def _cpu_fire_one(args):
    """One packing minimised by the oracle restatement's FIRE (a worker of
    cpu_minimization; runs in its own process)."""
    import oracle

    L, x, y, phi, A, R, gamma, params = args
    b = oracle.OracleBox(L, x, y, phi, A, R, 0.1)
    b.set_contact_model(1)
    b.set_params(gamma, 0.0, 0.0, 0.0, 0.05)
    it, fr = b.fire(params)
    return it, fr


def cpu_minimization(work, params, budget_s, workers=16):
    """The oracle restatement's FIRE (two-point model, FP64) on the same
    pre-minimisation states, one packing per process on `workers` host cores
    (the ensemble is embarrassingly parallel), until every packing is done or
    the wall-clock budget is spent."""
    import concurrent.futures as cf
    import multiprocessing as mp

    workers = max(1, min(workers, os.cpu_count() or 1))
    jobs = [(pk.L, pk.x, pk.y, pk.phi, pk.A, pk.R, g, params)
            for pk, g in zip(work.packings, work.gammas)]
    done, iters = 0, 0
    with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        # a first task per worker starts the interpreters before the clock does
        list(ex.map(_cpu_fire_one, [jobs[0]] * workers))
        t0 = time.perf_counter()
        futs = [ex.submit(_cpu_fire_one, j) for j in jobs]
        for f in cf.as_completed(futs, timeout=None):
            it, fr = f.result()
            done += 1
            iters += it
            elapsed = time.perf_counter() - t0
            if elapsed > budget_s:
                break
        elapsed = time.perf_counter() - t0
        for f in futs:
            f.cancel()
    return {"value": done / elapsed, "unit": "packings/s", "cores": workers, "kind": "port",
            "sample": f"{done} of the {len(jobs)} packings of the same ensemble in "
                      f"{elapsed:.1f} s, FIRE two-point FP64, one process per core on "
                      f"{workers} cores ({cpu_model()}), {iters} iterations"}


def copy_calibration(mib):
    """Plain device copy of a known size: its FETCH_SIZE / WRITE_SIZE counter
    readings give the factor that turns those counters into HBM bytes."""
    import torch

    n = mib * (1 << 20) // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 5}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", default="shear_large", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--path", default="auto", choices=["auto", "list", "tile"])
    ap.add_argument("--tile", type=int, default=12, help="tile path: columns per tile")
    ap.add_argument("--lanes", type=int, default=1, help="list path: lanes per particle")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--pad", type=float, default=0.1,
                    help="cell padding (list skin) of the Heun leg; 0: the cell-size autotuner "
                         "picks among 0.05, 0.1 and 0.2")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=200,
                    help="CPU baseline: Heun steps per run (from the relaxed state)")
    ap.add_argument("--cpu-runs", type=int, default=3, help="CPU baseline: runs (median)")
    ap.add_argument("--relax-steps", type=int, default=2000,
                    help="untimed Heun steps in setup before every timed leg")
    ap.add_argument("--no-minimization", action="store_true",
                    help="skip the packings/s minimisation measurement")
    ap.add_argument("--packings", type=int, default=256, help="minimisation ensemble size")
    ap.add_argument("--n", type=int, default=0,
                    help="shear_large with this many particles (HBM-resident sizes)")
    ap.add_argument("--no-quasistatic", action="store_true",
                    help="skip the quasi-static compression measurement")
    ap.add_argument("--checkpoints", type=int, default=5, metavar="K",
                    help="large near-jamming shear in mixed precision with K minimisation "
                         "checkpoints, one every 1000 steps (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported: ensembles "
              "would split by packing and a single large box into y-slabs with halo "
              "exchange, neither of which is built", file=sys.stderr)
        sys.exit(2)

    make, strain_rate, dt = WORKLOADS[args.workload]
    pk0 = configs.shear_large(1, n=args.n) if args.n and args.workload == "shear_large" else make()
    s = native.System(pk0, pads=args.pad or 0.1)
    s.set_dynamics(strain_rate=strain_rate, step=dt)
    if args.precision == "mixed":
        s.set_precision("mixed")
    # The seeded placement only avoids crossings: its overlaps relax over the
    # first steps, which move particles fast and rebuild the neighbour rows
    # often.  Setup runs those steps untimed; every timed leg (GPU, host to
    # host, CPU baseline) starts from the relaxed state.
    t0 = time.perf_counter()
    if args.relax_steps:
        s.strain_steps(args.relax_steps)
    relax_s = time.perf_counter() - t0
    xr, yr, pr = s.positions()
    b0 = s.box()
    gamma_r = b0["gamma"]
    pk = configs.Packing(pk0.L, xr, yr, pr, pk0.A, pk0.R, pk0.packing_fraction, pk0.name)
    if args.path == "auto":
        tuned = s.autotune(reps=10)  # the project's autotuner: path and launch shape
    else:
        s.set_path(args.path)
        if args.path == "tile":
            s.set_tiling(args.tile, args.block or 128)
        else:
            s.set_list_block(args.block or 256, args.lanes)
        tuned = None
    pad, pad_trials = args.pad, None
    if not pad:  # the cell-size autotuner, from the relaxed state
        pad, pad_trials = s.autotune_pad((0.05, 0.1, 0.2), steps=200)
    path = s.path()
    shape, block = s.tiling()

    s.strain_steps(args.warmup)
    s.synchronize()
    t0 = time.perf_counter()
    s.strain_steps(args.steps)  # returns after the stream is synchronised
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = pk.n * args.steps / elapsed

    # kernel-level figures: same steps again with HIP events on the system's stream
    s.timing(True)
    s.timing_reset()
    s.strain_steps(args.steps)
    pred_ms, pred_n = s.kernel_time(native.SC_KERNEL_PREDICT)
    corr_ms, corr_n = s.kernel_time(native.SC_KERNEL_CORRECT)
    s.timing(False)
    s.close()

    # the same steps from the caller's host arrays to host results: system
    # creation (upload, cell sort, lists) + steps + download of the positions
    t0 = time.perf_counter()
    s2 = native.System(pk, pads=pad)
    try:
        s2.set_gamma(gamma_r, 0.0)
        s2.set_dynamics(strain_rate=strain_rate, step=dt)
        if args.precision == "mixed":
            s2.set_precision("mixed")
        s2.set_path(path)
        if path == "tile":
            s2.set_tiling(shape, block)
        else:
            s2.set_list_block(block, shape)
        s2.strain_steps(args.steps)
        s2.positions()
    finally:
        s2.close()
    host_s = time.perf_counter() - t0
    corr_avg = corr_ms / max(corr_n, 1) * 1e-3
    pred_avg = pred_ms / max(pred_n, 1) * 1e-3
    achieved = BYTES_CORR * pk.n / corr_avg / 1e9

    out = {
        "metric": "particle-steps/s (sheared Heun)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-pair/f64-sum",
        "data": f"synthetic (seeded non-crossing random packing, relaxed by {args.relax_steps} "
                "untimed Heun steps in setup)",
        "config": {"workload": pk.name, "n_particles": pk.n, "strain_rate": strain_rate, "dt": dt,
                   "cell_pad": pad, "cell_pad_trials_ms_200_steps": pad_trials,
                   "force_path": path, "relax_steps": args.relax_steps,
                   "relax_s": relax_s,
                   ("tile_cols" if path == "tile" else "lanes_per_particle"): shape,
                   "block_threads": block, "autotuned": tuned is not None},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None if args.n else MEASURED_TRAFFIC.get((args.workload, path,
                                                                             args.precision, pad)),
                     "traffic_source": "steady-state constant: rocprofv3 FETCH_SIZE / WRITE_SIZE "
                                       "per corrector launch after the relaxation, scaled by a "
                                       "plain copy's readings (profiles/), not this run",
                     "kernel": f"k_{path}<corrector>", "avg_us": corr_avg * 1e6,
                     "predictor_avg_us": pred_avg * 1e6,
                     "bytes_per_particle": BYTES_CORR},
        "host_to_host": {"value": pk.n * args.steps / host_s, "unit": "particle-steps/s",
                         "seconds": host_s,
                         "includes": "system creation from host arrays (upload, cell sort, "
                                     "lists), the same steps, download of the positions"},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(pk, gamma_r, strain_rate, dt, steps=args.cpu_steps,
                                           runs=args.cpu_runs)
    if not args.no_minimization:
        out["minimization"] = minimization(args.packings, 1,
                                           0.0 if args.no_cpu_baseline else 15.0)
    if not args.no_quasistatic:
        out["quasistatic"] = quasistatic_compression()
        out["quasistatic"]["jammed"] = jammed_compression()
    if args.checkpoints > 0:
        from spherocyls_amd import checkpoints

        chk = checkpoints.run(checkpoints=args.checkpoints, every=1000, precision="mixed")
        chk.pop("records")
        out["checkpoints"] = chk
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
