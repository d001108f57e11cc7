"""Batched quasi-static compression: every packing of one system takes the
compress_qs2p step of the reference's GPU code (src/cuda/2point_dynamics.cu:
4300-4340, resize_step_2p :1187-1285) -- x, y and L multiplied by 1 - eps,
then one unsheared Heun step of dt 0.01 -- and is then minimised to
F_rms < ftol, all packings in one batched FIRE.  A packing that has reached
its final packing fraction takes scale 1 and, already minimised, costs its
FIRE nothing.  The per-packing loop of SpherocylBox::compress_qs2p is the same
protocol for one box, with its files.

Optional learned warm start: after each affine step the network's proposal
(damped so that no packing's energy rises, see warmstart.damped_warm_start)
moves the particles before FIRE; FIRE judges the result as usual.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np

from . import native


@dataclasses.dataclass
class Curve:
    """One row per step (row 0: the minimised start) and packing."""
    packing: np.ndarray     # [steps+1, npack]
    pressure: np.ndarray    # -(cPxx + cPyy) / 2, corrected virial: positive under compression
    energy: np.ndarray
    iterations: np.ndarray  # FIRE iterations of each minimisation
    frms: np.ndarray
    seconds: float
    tuned: dict = None      # launch shape kept by the FIRE-timed tuner (tune_fire)


def _record(s):
    b = [s.box(k) for k in range(s.npack)]
    # the CPU code's stress is the virial sum of f (x_A - x_B), negative for
    # repulsive contacts: the pressure is minus its mean diagonal
    return (np.array([x["packing"] for x in b]), np.array([-0.5 * (x["cPxx"] + x["cPyy"]) for x in b]),
            np.array([x["E"] for x in b]))


def compress(s, eps, phi_final, params, switch_frms=1e-6, max_steps=100000, warm=None,
             on_step=None, after_step=None, tune_fire=False):
    """Compress every packing of system s in steps of x, y, L *= 1 - eps until
    its packing fraction reaches phi_final (a scalar or one per packing),
    minimising after each step (mixed precision: FP32 pair math down to
    switch_frms).  warm: a native.Mpnn whose damped proposal precedes FIRE.
    on_step(step, s) runs after each affine and Heun step, before the warm
    start and FIRE, and after_step(step, s) after FIRE (used to record pre-
    and post-minimisation states).  tune_fire: at the first compression step,
    after its affine and Heun step, time FIRE iterations per launch shape from
    that state (System.autotune_fire, which restores it) and keep the fastest
    for the rest of the run: a quasi-static minimisation moves particles far
    less than the first one from a random start, and prefers other shapes.
    Leaves the system's dynamics at strain rate 0 and step 0.01.  Returns a
    Curve (its `tuned` field: the shape kept, or None)."""
    from . import warmstart

    fin = np.broadcast_to(np.asarray(phi_final, dtype=np.float64), (s.npack,))
    sw = switch_frms if s.precision == "mixed" else 0.0
    rows = []
    tuned = None
    s.set_dynamics(strain_rate=0.0, resize_rate=0.0, step=0.01)
    s.synchronize()
    t0 = time.perf_counter()
    it, fr, _ = s.fire(params, switch_frms=sw, check=False)
    rows.append(_record(s) + (it, fr))
    for step in range(1, max_steps + 1):
        packing = rows[-1][0]
        active = packing < fin
        if not np.any(active):
            break
        s.affine_resize(np.where(active, 1.0 - eps, 1.0))
        s.strain_steps(1)
        if on_step is not None:
            on_step(step, s)
        if tune_fire and step == 1:
            tuned = s.autotune_fire(params, iterations=60, switch_frms=sw)
        if warm is not None:
            warmstart.damped_warm_start(s, s.sizes, warm)
        it, fr, _ = s.fire(params, switch_frms=sw, check=False)
        if after_step is not None:
            after_step(step, s)
        it = np.where(active, it, 0)
        rows.append(_record(s) + (it, fr))
    s.synchronize()
    sec = time.perf_counter() - t0
    cols = list(zip(*rows))
    return Curve(*(np.array(c) for c in cols), seconds=sec, tuned=tuned)
