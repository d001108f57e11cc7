"""Minimisation workloads: ensembles of ragged packings brought to their
pre-minimisation state on the GPU, then minimised to F_rms < 1e-10.

A workload is a list of packings plus the Lees-Edwards strain of each; the
pre-minimisation state of a packing is its seeded random configuration after a
number of sheared Heun steps (all packings advanced together in one batched
system).
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np

from . import configs, native


@dataclasses.dataclass
class Workload:
    packings: list
    gammas: list
    pad: float = 0.1  # cell padding (chosen by the cell-size autotuner in prepare)
    pad_trials: dict = dataclasses.field(default_factory=dict)  # padding -> ms

    @property
    def sizes(self):
        return [p.n for p in self.packings]

    @property
    def particles(self):
        return sum(self.sizes)


def prepare(packings, heun_steps=2000, strain_rate=1e-4, dt=0.01, contact_model="two_point",
            pad=0.1, tune_pad=None, tune_steps=200) -> Workload:
    """Advance every packing by heun_steps sheared Heun steps (one batched
    system) and return the resulting states.  tune_pad: candidate cell
    paddings for the cell-size autotuner (native.System.autotune_pad), run
    with the launch-shape autotuner before the steps; the chosen padding is
    the workload's, for its minimisation too."""
    s = native.System(packings, pads=pad)
    trials = {}
    try:
        s.set_contact_model(contact_model)
        s.set_dynamics(strain_rate=strain_rate, step=dt)
        if tune_pad:
            s.autotune(reps=5)
            pad, trials = s.autotune_pad(tune_pad, steps=tune_steps)
        if heun_steps:
            s.strain_steps(heun_steps)
        x, y, phi = s.positions()
        gammas = [s.box(k)["gamma"] for k in range(s.npack)]
    finally:
        s.close()
    out, o = [], 0
    for p in packings:
        sl = slice(o, o + p.n)
        out.append(configs.Packing(p.L, x[sl].copy(), y[sl].copy(), phi[sl].copy(), p.A, p.R,
                                   p.packing_fraction, p.name))
        o += p.n
    return Workload(out, gammas, pad, trials)


def system_for(work: Workload, indices=None, precision="fp64", contact_model="two_point",
               pad=None, autotune=True):
    idx = list(range(len(work.packings))) if indices is None else list(indices)
    s = native.System([work.packings[i] for i in idx], pads=work.pad if pad is None else pad)
    s.set_contact_model(contact_model)
    for k, i in enumerate(idx):
        s.set_gamma(work.gammas[i], 0.0, k=k)
    s.set_precision(precision)
    if autotune:
        s.autotune(reps=5)
    return s


def minimize_batched(work: Workload, params, precision="mixed", switch_frms=1e-6):
    """All packings in one system; returns (seconds, iterations[], frms[], status)."""
    s = system_for(work, precision=precision)
    try:
        s.synchronize()
        t0 = time.perf_counter()
        it, fr, st = s.fire(params, switch_frms=switch_frms if precision == "mixed" else 0.0,
                            check=False)
        t1 = time.perf_counter()
    finally:
        s.close()
    return t1 - t0, it, fr, st


def minimize_one_by_one(work: Workload, params, precision="fp64", limit=None):
    """One system per packing, minimised in turn (systems are built and tuned
    before the clock starts); returns (seconds, iterations[], frms[])."""
    n = len(work.packings) if limit is None else min(limit, len(work.packings))
    systems = [system_for(work, [i], precision=precision) for i in range(n)]
    its, frs = [], []
    try:
        for s in systems:
            s.synchronize()
        t0 = time.perf_counter()
        for s in systems:
            it, fr, _ = s.fire(params, check=False)
            its.append(int(it[0]))
            frs.append(float(fr[0]))
        t1 = time.perf_counter()
    finally:
        for s in systems:
            s.close()
    return t1 - t0, np.array(its), np.array(frs)


def main(argv=None):
    """Minimise the ragged benchmark ensemble once, all packings in one system:
    python -m spherocyls_amd.ensemble [--packings 256] [--precision fp64|mixed]"""
    import argparse
    import json

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--packings", type=int, default=256)
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--max-iter", type=int, default=20000)
    args = ap.parse_args(argv)
    work = prepare(configs.ensemble(args.packings), heun_steps=2000, strain_rate=1e-4, dt=0.01)
    sec, it, fr, st = minimize_batched(work, native.fire_params(max_iter=args.max_iter),
                                       precision=args.precision)
    conv = int(np.sum(fr < 1e-10))
    print(json.dumps({"packings": args.packings, "particles": work.particles,
                      "precision": args.precision, "seconds": sec, "converged": conv,
                      "packings_per_s": conv / sec, "max_iterations": int(it.max()),
                      "mean_iterations": float(it.mean())}), flush=True)


if __name__ == "__main__":
    main()
