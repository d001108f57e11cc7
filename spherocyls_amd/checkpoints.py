"""Long shear of one large packing with periodic minimisation checkpoints.

The large near-jamming workload: N = 262144, R uniform in [0.5, 0.7],
A = 0.5 R, phi = 0.845, sheared Heun steps at strain rate 1e-6 (dt 0.01) under
Lees-Edwards boundaries with FP32 pair forces and FP64 sums.  Every `every`
steps the current state is copied out and minimised (learned warm start, then
FIRE to F_rms < 1e-10 with the two-point contact model, the plain FIRE run
from the unmodified state as fallback), while the shear continues from the
unminimised state.  The same copy is also minimised by plain FIRE, without the
warm start, for comparison.

Stress and energy records of a mixed-precision run are FP64 evaluations of the
state it reached, so the error that FP32 pair forces bring into the reported
stress is that of the trajectory: an FP64 run from the same start steps beside
the mixed one, and at each checkpoint the relative difference of their energy
and stress is recorded (for the shear stress also relative to the pressure,
since Pxy is small beside Pxx and Pyy).  Near jamming the sheared dynamics is
chaotic, so two trajectories that start a rounding error apart separate
anyway: a second FP64 run, from the start with every x displaced by a seeded
1e-12, gives that intrinsic spread at each checkpoint, beside which the
mixed-precision difference is read.  The final minimised state of each
checkpoint is evaluated in FP64 by both paths alike.

    python -m spherocyls_amd.checkpoints [--checkpoints 10] [--every 1000]
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

from . import configs, ensemble, native

SE_KEYS = ("E", "Pxx", "Pyy", "Pxy")


def run(pk=None, checkpoints=10, every=1000, strain_rate=1e-6, dt=0.01, precision="mixed",
        warm_start=True, params=None, log=None):
    """Shear pk for checkpoints * every steps with a minimisation every `every`
    steps.  Returns a dict: Heun particle-steps/s (the shear alone), per
    checkpoint the precision-vs-FP64 stress error and the minimisation's
    seconds, iterations and F_rms."""
    from . import warmstart

    pk = configs.shear_large(1) if pk is None else pk
    params = native.fire_params(max_iter=200000) if params is None else params
    net = warmstart.load() if warm_start and os.path.exists(warmstart.DEFAULT_WEIGHTS) else None
    s = native.System(pk, pads=0.1)
    # the FP64 trajectory from the same start, stepped beside the run, and an
    # FP64 trajectory from a start displaced by 1e-12 (the intrinsic spread)
    ref = native.System(pk, pads=0.1) if precision != "fp64" else None
    kick = np.random.default_rng(12).standard_normal(pk.n) * 1e-12
    pk_kicked = configs.Packing(pk.L, pk.x + kick, pk.y, pk.phi, pk.A, pk.R,
                                pk.packing_fraction, pk.name)
    ref2 = native.System(pk_kicked, pads=0.1) if precision != "fp64" else None
    recs = []
    heun_s = 0.0
    try:
        s.set_dynamics(strain_rate=strain_rate, step=dt)
        s.set_precision(precision)
        s.autotune(reps=5)
        for r in (ref, ref2):
            if r is not None:
                r.set_dynamics(strain_rate=strain_rate, step=dt)
                r.autotune(reps=5)
        for c in range(checkpoints):
            s.synchronize()
            t0 = time.perf_counter()
            s.strain_steps(every)
            heun_s += time.perf_counter() - t0
            # the sheared state's stress and energy (an FP64 evaluation) against
            # the FP64 trajectory's at the same strain
            s.calc_se()
            b_run = s.box()
            if ref is not None:
                ref.strain_steps(every)
                ref.calc_se()
                b64 = ref.box()
                ref2.strain_steps(every)
                ref2.calc_se()
                b_kick = ref2.box()
            else:
                b64 = b_kick = b_run

            def rel(b):
                e = {k: abs(b[k] - b64[k]) / max(abs(b64[k]), 1e-300) for k in SE_KEYS}
                pressure = 0.5 * abs(b64["Pxx"] + b64["Pyy"])
                e["Pxy_vs_pressure"] = abs(b["Pxy"] - b64["Pxy"]) / max(pressure, 1e-300)
                return e

            err, spread = rel(b_run), rel(b_kick)
            b64 = b_run  # the state minimised below is the run's
            # minimisation checkpoint of a copy of the state
            x, y, phi = s.positions()
            work = ensemble.Workload([configs.Packing(b64["Lx"], x, y, phi, pk.A, pk.R,
                                                      pk.packing_fraction, pk.name)],
                                     [b64["gamma"]])
            psec, pit, pfr, _ = ensemble.minimize_batched(work, params, precision=precision)
            if net is not None:
                sec, it, fr, fb = warmstart.minimize_with_warm_start(work, params, net,
                                                                     precision=precision)
            else:
                sec, it, fr, fb = psec, pit, pfr, []
            rec = {"step": (c + 1) * every, "gamma_total": b64["gamma_total"],
                   "se_rel_error_vs_fp64_trajectory": err,
                   "se_rel_spread_fp64_from_kicked_start": spread, "E": b64["E"], "Pxx": b64["Pxx"],
                   "Pyy": b64["Pyy"], "Pxy": b64["Pxy"], "min_seconds": sec,
                   "min_iterations": int(it[0]), "min_frms": float(fr[0]),
                   "min_converged": bool(fr[0] < params[8]), "fallback": bool(fb),
                   "plain_seconds": psec, "plain_iterations": int(pit[0]),
                   "plain_converged": bool(pfr[0] < params[8])}
            recs.append(rec)
            if log:
                log(json.dumps(rec))
    finally:
        s.close()
        for r in (ref, ref2):
            if r is not None:
                r.close()
        if net is not None:
            net.close()
    steps = checkpoints * every
    worst = {k: max(r["se_rel_error_vs_fp64_trajectory"][k] for r in recs)
             for k in (*SE_KEYS, "Pxy_vs_pressure")}
    worst_spread = {k: max(r["se_rel_spread_fp64_from_kicked_start"][k] for r in recs)
                    for k in (*SE_KEYS, "Pxy_vs_pressure")}
    return {"workload": pk.name, "n_particles": pk.n, "precision": precision,
            "steps": steps, "checkpoints": checkpoints, "heun_seconds": heun_s,
            "particle_steps_per_s": pk.n * steps / heun_s,
            "se_rel_error_vs_fp64_trajectory_max": worst,
            "se_rel_spread_fp64_from_kicked_start_max": worst_spread,
            "minimisations_converged": sum(r["min_converged"] for r in recs),
            "min_seconds_mean": float(np.mean([r["min_seconds"] for r in recs])),
            "min_iterations_mean": float(np.mean([r["min_iterations"] for r in recs])),
            "plain_seconds_mean": float(np.mean([r["plain_seconds"] for r in recs])),
            "plain_iterations_mean": float(np.mean([r["plain_iterations"] for r in recs])),
            "warm_start": net is not None, "records": recs}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--checkpoints", type=int, default=10)
    ap.add_argument("--every", type=int, default=1000)
    ap.add_argument("--precision", choices=("mixed", "fp64"), default="mixed")
    ap.add_argument("--no-warm-start", action="store_true")
    args = ap.parse_args(argv)
    out = run(checkpoints=args.checkpoints, every=args.every, precision=args.precision,
              warm_start=not args.no_warm_start, log=lambda m: print(m, flush=True))
    out.pop("records")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
