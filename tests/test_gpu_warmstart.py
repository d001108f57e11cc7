"""Learned warm start: the trained network (spherocyls_amd/data/mpnn_warmstart.npz,
trained on the GPU by `python -m spherocyls_amd.warmstart train-qs` on
quasi-static compression steps) on held-out packings, against the plain
minimiser: in the quasi-static compression it was trained for, and on the
sheared ragged ensemble.

The plain minimiser is the judge: every packing must end at F_rms < 1e-10
(checked by the oracle on a sample), whether its warm start was accepted or it
fell back.  The test prints what the network buys -- the energy of its
damped proposal against the unrelaxed state, and FIRE iterations and wall
time with and without it -- and requires that the damped proposal never
raises a packing's energy.
"""
import os

import numpy as np
import pytest

import oracle
from spherocyls_amd import ensemble, native, quasistatic, warmstart

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net():
    if not os.path.exists(warmstart.DEFAULT_WEIGHTS):
        pytest.skip("trained weights not present")
    m = warmstart.load()
    yield m
    m.close()


@pytest.fixture(scope="module")
def held_out():
    pks = warmstart.workload(24, 1024, seed=4242)  # seeds disjoint from training
    return ensemble.prepare(pks, heun_steps=2000)


def test_damped_proposal_never_raises_energy(net, held_out):
    before, after, lam = warmstart.proposal_energy(held_out, net)
    print("energy before", before.mean(), "after damped proposal", after.mean(),
          "damping chosen", np.unique(lam, return_counts=True), flush=True)
    assert np.all(np.isfinite(after))
    assert np.all(after <= before)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_warm_start_reaches_the_same_residual(net, held_out, precision):
    params = native.fire_params(max_iter=20000)
    sec0, it0, fr0, _ = ensemble.minimize_batched(held_out, params, precision=precision)
    sec1, it1, fr1, fb = warmstart.minimize_with_warm_start(held_out, params, net,
                                                            precision=precision)
    print(f"{precision}: plain {sec0 * 1e3:.1f} ms, {it0.mean():.0f} iterations; "
          f"warm start {sec1 * 1e3:.1f} ms, {it1.mean():.0f} iterations, "
          f"{len(fb)} fallbacks", flush=True)
    assert np.all(fr0 < 1e-10) and np.all(fr1 < 1e-10)
    # the oracle judges a sample of the warm-started final states
    s = ensemble.system_for(held_out, precision=precision, autotune=False)
    try:
        warmstart.damped_warm_start(s, held_out.sizes, net)
        s.fire(params, switch_frms=1e-6 if precision == "mixed" else 0.0, check=False)
        x, y, phi = s.positions()
    finally:
        s.close()
    o = 0
    for pk, g in zip(held_out.packings[:3], held_out.gammas[:3]):
        sl = slice(o, o + pk.n)
        b = oracle.OracleBox(pk.L, x[sl], y[sl], phi[sl], pk.A, pk.R, 0.1)
        b.set_contact_model(1)
        b.set_params(g, 0.0, 0.0, 0.0, 0.05)
        assert b.frms() < 1e-10
        o += pk.n


def test_quasistatic_warm_start_gain(net):
    """Held-out seeds, 60 compression steps of x, y, L *= 1 - 1e-4 from phi 0.80:
    every minimisation converges with and without the warm start, and the test
    prints the FIRE iterations and wall time of both (the gain)."""
    pks = warmstart.qs_packings(4, 2048, seed=777)
    params = native.fire_params(max_iter=200000)
    phi_final = 0.80 / (1 - 1e-4) ** 120
    res = {}
    for label, warm in (("plain", None), ("warm start", net)):
        s = native.System(pks, pads=0.1)
        try:
            s.set_contact_model("two_point")
            c = quasistatic.compress(s, 1e-4, phi_final, params, warm=warm)
        finally:
            s.close()
        assert np.all(c.frms < 1e-10), label
        res[label] = c
        print(f"quasi-static {label}: {c.iterations[1:].mean():.1f} FIRE iterations per step, "
              f"{c.seconds:.2f} s for {c.iterations.shape[0] - 1} steps", flush=True)
    gain = res["plain"].iterations[1:].sum() / res["warm start"].iterations[1:].sum()
    print(f"iteration ratio plain / warm start: {gain:.3f}", flush=True)
    assert res["plain"].iterations.shape == res["warm start"].iterations.shape
    # the damped warm start never makes FIRE do materially more work
    assert gain > 0.9


def test_exact_proposal_ceiling(held_out):
    """What a warm start could buy at best on the held-out ensemble: each
    packing's exact displacement to the minimum plain FIRE reaches, applied
    at 90%, must cut the mean FIRE iterations by at least 10% (so the
    measurement above would show a gain where one exists); the printed table
    is the ceiling the learned proposal is read against."""
    params = native.fire_params(max_iter=20000)
    rows = warmstart.ceiling(held_out, params, qualities=(0.0, 0.5, 0.9))
    for r in rows:
        print(f"q = {r['quality']}: {r['mean_iterations']:.0f} mean / {r['max_iterations']} max "
              f"iterations, {r['seconds'] * 1e3:.1f} ms, {r['converged']} converged", flush=True)
    plain, q9 = rows[0], rows[2]
    n = len(held_out.packings)
    assert plain["converged"] == n and q9["converged"] == n
    assert q9["mean_iterations"] < 0.9 * plain["mean_iterations"]

