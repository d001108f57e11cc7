"""Cell padding after creation (sc_set_pad) and the cell-size autotuner
(sc_autotune_pad).  A padding set on a live system gives the same cell grid,
storage order and rows as a system built with that padding, so forces and
trajectories agree bit for bit; the autotuner puts the state back exactly and
keeps one of its candidates."""
import numpy as np
import pytest

from spherocyls_amd import configs, native

pytestmark = pytest.mark.gpu


def _packs():
    R, A = configs.bidisperse(3000, aspect=2.0)
    R2, A2 = configs.bidisperse(900, aspect=4.0)
    return [configs.make_packing(R, A, 0.84, 71, "cells N=3000"),
            configs.make_packing(R2, A2, 0.80, 72, "cells N=900")]


def _run(s, steps):
    s.set_dynamics(strain_rate=1e-4, step=0.01)
    s.set_gamma(0.1, 0.0)
    s.calc_se()
    f0 = s.forces()
    s.strain_steps(steps)
    return f0, s.positions(), [s.box(k)["gamma"] for k in range(s.npack)]


def test_set_pad_equals_a_system_built_with_it():
    packs = _packs()
    print("set_pad 0.1 -> 0.3 against a system built at 0.3", flush=True)
    a = native.System(packs, pads=0.1)
    b = native.System(packs, pads=0.3)
    try:
        a.set_pad(0.3)
        for k in range(2):
            assert a.box(k)["rows"] == b.box(k)["rows"] and a.box(k)["cols"] == b.box(k)["cols"]
        ra, rb = _run(a, 300), _run(b, 300)
        for u, v in zip(ra[0] + ra[1], rb[0] + rb[1]):
            np.testing.assert_array_equal(u, v)
        assert ra[2] == rb[2]
        for u, v in zip(a.contacts(0), b.contacts(0)):
            np.testing.assert_array_equal(u, v)
    finally:
        a.close()
        b.close()


def test_autotune_pad_restores_the_state_and_keeps_a_candidate():
    packs = _packs()
    s = native.System(packs, pads=0.1)
    try:
        s.set_dynamics(strain_rate=1e-4, step=0.01)
        s.set_gamma(0.1, 0.0)
        before = s.positions()
        cands = (0.05, 0.1, 0.2)
        print("autotune_pad", cands, flush=True)
        chosen, ms = s.autotune_pad(cands, steps=50)
        print("chosen", chosen, "ms", ms, flush=True)
        assert chosen in cands and set(ms) == set(cands) and all(v > 0 for v in ms.values())
        after = s.positions()
        for u, v in zip(before, after):
            np.testing.assert_array_equal(u, v)
        assert all(s.box(k)["gamma"] == 0.1 for k in range(2))
        # from here the tuned system steps exactly as one built with its padding
        ref = native.System([configs.Packing(p.L, *[x[o:o + p.n] for x in before], p.A, p.R,
                                             p.packing_fraction, p.name)
                             for p, o in zip(packs, (0, packs[0].n))], pads=chosen)
        try:
            r1, r2 = _run(s, 200), _run(ref, 200)
            for u, v in zip(r1[0] + r1[1], r2[0] + r2[1]):
                np.testing.assert_array_equal(u, v)
        finally:
            ref.close()
    finally:
        s.close()
