"""Learned warm starts for the minimiser.

A small message-passing network over the contact graph (native.Mpnn, kernels
in csrc/sc_mpnn.hip) predicts each particle's relaxation displacement
(dx, dy, dphi) from a pre-minimisation state.  Each packing takes the
proposal scaled by whichever of 1, 1/2, 1/4 (or none) lowers its energy most,
and the minimiser starts from there; the plain minimiser is the fallback and
the judge: a warm-started packing is accepted only when FIRE takes it to the
same F_rms < ftol, otherwise that packing is minimised again from its
original state.

Training data are (pre-minimisation, minimised) pairs the simulator makes
itself.  The committed weights come from quasi-static compression: the state
right after a compression step x, y, L *= 1 - 1e-4 as the pre state and the
state FIRE relaxes it to as the target, N = 4096 bidisperse aspect-2
packings compressed from phi 0.80 to 0.86, 576 pairs.  (The compression step
now also takes one unsheared Heun step, as the reference's protocols do;
weights trained on 11584 pairs made that way reached a higher held-out loss,
0.845 against 0.807, so the 576-pair weights stay.)  The `train` variant uses ragged packings
after sheared Heun steps instead.  Training runs on the GPU through the
network's own forward, backward and Adam kernels.

    python -m spherocyls_amd.warmstart train-qs --out spherocyls_amd/data/mpnn_warmstart.npz
"""
from __future__ import annotations

import argparse
import dataclasses
import math
import os
import time

import numpy as np

from . import configs, ensemble, native

REF_PI = 3.14159265358979
DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_WEIGHTS = os.path.join(DATA, "mpnn_warmstart.npz")


def workload(count, n, seed, phi=(0.80, 0.86), aspect=2.0):
    """Bidisperse (R 0.5 / 0.7) packings of n particles, phi uniform in the
    range, from seeds seed*100003 + k."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(count):
        R, A = configs.bidisperse(n, aspect=aspect)
        f = float(rng.uniform(*phi))
        out.append(configs.make_packing(R, A, f, seed * 100003 + k, f"warm[{seed}:{k}] phi={f:.4f}"))
    return out


def split_graph(g, sizes):
    """Split a concatenated graph (caller order) into per-packing graphs."""
    out, o = [], 0
    for n in sizes:
        e0, e1 = int(g.offsets[o]), int(g.offsets[o + n])
        out.append(native.Graph(g.node_feat[o:o + n], g.offsets[o:o + n + 1] - e0,
                                g.src[e0:e1] - o, g.edge_feat[e0:e1]))
        o += n
    return out


def displacement(pre, post_xyz, L):
    x, y, phi = post_xyz
    dx = x - pre.x
    dy = y - pre.y
    dx -= L * np.round(dx / L)
    dy -= L * np.round(dy / L)
    dp = phi - pre.phi
    dp -= 2 * REF_PI * np.round(dp / (2 * REF_PI))
    return np.column_stack([dx, dy, dp])


@dataclasses.dataclass
class Pairs:
    graphs: list      # native.Graph per packing (pre state)
    targets: list     # [n, 3] displacement to the minimised state
    iterations: np.ndarray


def make_pairs(packings, params, heun_steps=2000, mpnn=None):
    """Pre states (sheared Heun steps), their contact graphs, and the FIRE
    displacement of every particle (batched, FP64, two-point contacts)."""
    work = ensemble.prepare(packings, heun_steps=heun_steps)
    s = ensemble.system_for(work, precision="fp64", autotune=False)
    own = mpnn is None
    m = native.Mpnn() if own else mpnn
    try:
        g = m.graph_of(s)
        it, fr, st = s.fire(params, check=False)
        x, y, phi = s.positions()
    finally:
        s.close()
        if own:
            m.close()
    graphs = split_graph(g, work.sizes)
    targets, o = [], 0
    for pk in work.packings:
        sl = slice(o, o + pk.n)
        targets.append(displacement(pk, (x[sl], y[sl], phi[sl]), pk.L).astype(np.float32))
        o += pk.n
    ok = fr < params[8]
    keep = [i for i in range(len(graphs)) if ok[i]]
    return Pairs([graphs[i] for i in keep], [targets[i] for i in keep], it[ok]), work


def qs_packings(count, n, seed, phi0=0.80, aspect=2.0):
    """Bidisperse packings of n particles at phi0 from seeds seed*100003 + k
    (the quasi-static compression start of the N=16384 workload, smaller)."""
    R, A = configs.bidisperse(n, aspect=aspect)
    return [configs.make_packing(R, A, phi0, seed * 100003 + k, f"qs[{seed}:{k}] N={n}")
            for k in range(count)]


def qs_pairs(packings, params, eps=1e-4, phi_final=0.86, every=10, mpnn=None, log=None):
    """(pre-step, minimised) pairs from batched quasi-static compression: every
    `every`-th step, the state right after the compression step (x, y, L *=
    1 - eps and one unsheared Heun step; its contact graph) and the
    displacement FIRE then makes."""
    from . import quasistatic

    s = native.System(packings, pads=0.1)
    own = mpnn is None
    m = native.Mpnn() if own else mpnn
    graphs, targets, iters = [], [], []
    pre = {}
    sizes = s.sizes

    def on_step(step, sy):
        pre.clear()
        if step % every == 0:
            pre["g"] = m.graph_of(sy)
            pre["x"] = sy.positions()
            pre["L"] = [sy.box(k)["Lx"] for k in range(sy.npack)]

    def after_step(step, sy):
        if not pre:
            return
        post = sy.positions()
        it = [int(sy.box(k)["iters"]) for k in range(sy.npack)]
        o = 0
        for k, (g, nk) in enumerate(zip(split_graph(pre["g"], sizes), sizes)):
            sl = slice(o, o + nk)
            o += nk
            if sy.box(k)["frms"] >= params[8]:
                continue  # not minimised: no target
            graphs.append(g)
            targets.append(_displacement(pre["x"], post, sl, pre["L"][k]).astype(np.float32))
            iters.append(it[k])
        if log and step % (10 * every) == 0:
            log(f"step {step}: {len(graphs)} pairs")

    try:
        s.set_contact_model("two_point")
        s.set_precision("fp64")
        curve = quasistatic.compress(s, eps, phi_final, params, on_step=on_step,
                                     after_step=after_step)
    finally:
        s.close()
        if own:
            m.close()
    return Pairs(graphs, targets, np.array(iters)), curve


def _displacement(pre, post, sl, L):
    dx = post[0][sl] - pre[0][sl]
    dy = post[1][sl] - pre[1][sl]
    dx -= L * np.round(dx / L)
    dy -= L * np.round(dy / L)
    dp = post[2][sl] - pre[2][sl]
    dp -= 2 * REF_PI * np.round(dp / (2 * REF_PI))
    return np.column_stack([dx, dy, dp])


def train(pairs, steps=2000, batch=8, lr=1e-3, seed=1, log_every=200, mpnn=None):
    """Adam on the mean squared error of the scaled displacement; returns
    (network, output scale, loss history)."""
    tall = np.concatenate(pairs.targets)
    scale = tall.std(axis=0).astype(np.float32) + 1e-12
    m = native.Mpnn(seed=seed) if mpnn is None else mpnn
    m.set_output_scale(scale)
    rng = np.random.default_rng(seed)
    hist = []
    t0 = time.perf_counter()
    for step in range(1, steps + 1):
        idx = rng.choice(len(pairs.graphs), size=min(batch, len(pairs.graphs)), replace=False)
        g = native.Graph.concat([pairs.graphs[i] for i in idx])
        t = np.concatenate([pairs.targets[i] for i in idx]) / scale
        m.load_graph(g)
        m.forward()
        loss = m.backward(t)
        if not np.isfinite(loss):
            raise FloatingPointError(f"non-finite training loss at step {step}")
        # cosine decay of the learning rate
        m.adam(lr * 0.5 * (1 + math.cos(math.pi * (step - 1) / steps)), step)
        hist.append(loss)
        if log_every and step % log_every == 0:
            print(f"step {step} loss {np.mean(hist[-log_every:]):.4f} "
                  f"({time.perf_counter() - t0:.1f} s)", flush=True)
    return m, scale, np.array(hist)


def save(m, path):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=m.params(), scale=m.output_scale())


def load(path=DEFAULT_WEIGHTS, device=0):
    with np.load(path, allow_pickle=False) as f:
        m = native.Mpnn(device=device)
        m.set_params(f["params"])
        m.set_output_scale(f["scale"])
    return m


def _energies(s):
    s.calc_se()
    return np.array([s.box(k)["E"] for k in range(s.npack)])


def damped_warm_start(s, sizes, mpnn, lambdas=(1.0, 0.5, 0.25)):
    """Apply the network's proposal to every packing of system s, scaled by the
    factor in lambdas (or none) that gives that packing the lowest energy; the
    state never starts above the unrelaxed one.  The trials run on the device
    (sc_mpnn_warm_start_damped).  Returns (energy before, energy after,
    chosen factor per packing)."""
    return mpnn.warm_start_damped(s, lambdas)


def proposal_energy(work, mpnn, contact_model="two_point"):
    """Energy of every packing before the warm start and after the damped
    proposal, and the damping chosen (FP64 force passes)."""
    s = ensemble.system_for(work, precision="fp64", autotune=False, contact_model=contact_model)
    try:
        return damped_warm_start(s, work.sizes, mpnn)
    finally:
        s.close()


def ceiling(work, params, qualities=(0.0, 0.5, 0.9, 0.99), precision="fp64"):
    """What a warm start of a given quality could save: each packing's exact
    displacement d to its minimum (a FIRE run from the pre state) is applied
    as q d, and batched FIRE runs again from there.  q = 0 is the plain
    minimisation; q = 0.9 is a proposal with 10% error along the true
    displacement, far better than a learned one.  Returns one record per q:
    seconds, mean / max iterations and packings that reached ftol."""
    s = ensemble.system_for(work, precision="fp64")
    try:
        x0, y0, p0 = s.positions()
        s.fire(params, check=False)
        x1, y1, p1 = s.positions()
    finally:
        s.close()
    L = np.concatenate([np.full(pk.n, pk.L) for pk in work.packings])
    dx = x1 - x0
    dx -= L * np.round(dx / L)
    dy = y1 - y0
    dy -= L * np.round(dy / L)
    dp = p1 - p0
    dp -= 2 * REF_PI * np.round(dp / (2 * REF_PI))
    out = []
    for q in qualities:
        s = ensemble.system_for(work, precision=precision)
        try:
            if q:
                s.set_positions(x0 + q * dx, y0 + q * dy, p0 + q * dp)
            s.synchronize()
            t0 = time.perf_counter()
            it, fr, st = s.fire(params, check=False)
            sec = time.perf_counter() - t0
        finally:
            s.close()
        out.append({"quality": q, "seconds": sec, "mean_iterations": float(it.mean()),
                    "max_iterations": int(it.max()),
                    "converged": int(np.sum(fr < params[8]))})
    return out


def minimize_with_warm_start(work, params, mpnn, precision="mixed", switch_frms=1e-6):
    """Batched minimisation from warm starts with per-packing fallback.

    Returns (seconds, iterations[], frms[], fallbacks): iterations count the
    warm-started run plus, for a packing that did not reach ftol, its rerun
    from the original state."""
    s = ensemble.system_for(work, precision=precision)
    fallbacks = []
    try:
        s.synchronize()
        t0 = time.perf_counter()
        try:
            damped_warm_start(s, work.sizes, mpnn)
            it, fr, st = s.fire(params, switch_frms=switch_frms if precision == "mixed" else 0.0,
                                check=False)
        except native.SpherocylsError:
            # a proposal the system cannot hold (e.g. an over-full cell): every
            # packing falls back to its original state
            it = np.zeros(len(work.packings), dtype=np.int32)
            fr = np.full(len(work.packings), np.inf)
        bad = [k for k in range(len(work.packings)) if not fr[k] < params[8]]
        if bad:
            sub = ensemble.Workload([work.packings[k] for k in bad], [work.gammas[k] for k in bad],
                                    work.pad)
            s2 = ensemble.system_for(sub, precision=precision, autotune=False)
            try:
                it2, fr2, _ = s2.fire(params, switch_frms=switch_frms if precision == "mixed" else 0.0,
                                      check=False)
            finally:
                s2.close()
            for j, k in enumerate(bad):
                it[k] += it2[j]
                fr[k] = fr2[j]
            fallbacks = bad
        t1 = time.perf_counter()
    finally:
        s.close()
    return t1 - t0, it, fr, fallbacks


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    for name in ("train-qs", "train"):
        t = sub.add_parser(name)
        t.add_argument("--out", default=DEFAULT_WEIGHTS)
        t.add_argument("--packings", type=int, default=16 if name == "train-qs" else 512)
        t.add_argument("--n", type=int, default=4096)
        t.add_argument("--steps", type=int, default=3000)
        t.add_argument("--batch", type=int, default=8)
        t.add_argument("--lr", type=float, default=1e-3)
        t.add_argument("--seed", type=int, default=1)
        t.add_argument("--every", type=int, default=10,
                       help="train-qs: record every k-th compression step")
    args = ap.parse_args()
    t0 = time.perf_counter()
    if args.cmd == "train-qs":
        # (pre-step, minimised) pairs of quasi-static compression, phi 0.80 -> 0.86
        # in steps of 1e-4, every k-th step
        params = native.fire_params(max_iter=200000)
        pairs, _ = qs_pairs(qs_packings(args.packings, args.n, seed=10 + args.seed), params,
                            every=args.every, log=lambda m: print(m, flush=True))
    else:
        # (sheared, minimised) pairs of the ragged-ensemble physics
        params = native.fire_params(max_iter=20000)
        pairs = None
        chunk = 64
        for c0 in range(0, args.packings, chunk):
            pk = workload(min(chunk, args.packings - c0), args.n, seed=1000 + args.seed + c0)
            p, _ = make_pairs(pk, params)
            if pairs is None:
                pairs = p
            else:
                pairs.graphs += p.graphs
                pairs.targets += p.targets
                pairs.iterations = np.concatenate([pairs.iterations, p.iterations])
    print(f"pairs {len(pairs.graphs)} ({time.perf_counter() - t0:.1f} s), "
          f"FIRE iterations mean {pairs.iterations.mean():.0f}", flush=True)
    m, scale, hist = train(pairs, steps=args.steps, batch=args.batch, lr=args.lr, seed=args.seed)
    save(m, args.out)
    print(f"saved {args.out}: scale {scale}, final loss {hist[-100:].mean():.4f}", flush=True)


if __name__ == "__main__":
    main()
