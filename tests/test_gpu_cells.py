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


def test_autotune_fire_restores_positions_and_keeps_a_shape():
    """The minimisation autotuner (FIRE iterations per candidate path and
    launch shape) puts the positions back exactly, and the system then
    minimises as one built from those positions with the shape it kept."""
    packs = _packs()
    params = native.fire_params(max_iter=20000)
    s = native.System(packs, pads=0.1)
    try:
        s.set_contact_model("two_point")
        before = s.positions()
        print("autotune_fire 30 iterations per candidate", flush=True)
        pick = s.autotune_fire(params, iterations=30)
        print("kept", pick, flush=True)
        assert pick["path"] in ("list", "tile") and pick["ms_per_trial"] > 0
        after = s.positions()
        for u, v in zip(before, after):
            np.testing.assert_array_equal(u, v)
        it1, fr1, _ = s.fire(params)
        p1 = s.positions()
    finally:
        s.close()
    ref = native.System(packs, pads=0.1)
    try:
        ref.set_contact_model("two_point")
        ref.set_path(pick["path"])
        if pick["path"] == "list":
            ref.set_list_block(pick["block"], pick["shape"])
        else:
            ref.set_tiling(pick["shape"], pick["block"])
        it2, fr2, _ = ref.fire(params)
        p2 = ref.positions()
    finally:
        ref.close()
    np.testing.assert_array_equal(it1, it2)
    for u, v in zip(p1, p2):
        np.testing.assert_array_equal(u, v)
    assert np.all(fr1 < 1e-10)


def test_compression_tunes_at_its_first_step():
    """quasistatic.compress(tune_fire=True) times the launch shapes at its
    first quasi-static step, keeps one, and every minimisation of the run
    still reaches F_rms < 1e-10."""
    from spherocyls_amd import quasistatic

    R, A = configs.bidisperse(1200)
    pk = configs.make_packing(R, A, 0.80, 73, "compress N=1200")
    s = native.System(pk, pads=0.1)
    try:
        s.set_contact_model("two_point")
        print("compress with tune_fire, 4 steps", flush=True)
        c = quasistatic.compress(s, 1e-3, 0.8032, native.fire_params(max_iter=100000),
                                 tune_fire=True)
    finally:
        s.close()
    print("tuned", c.tuned, "steps", len(c.packing) - 1, flush=True)
    assert c.tuned is not None and c.tuned["path"] in ("list", "tile")
    assert len(c.packing) >= 3 and np.all(c.frms[:, 0] < 1e-10)


def test_autotune_fire_reports_a_failing_trial_and_keeps_the_state():
    """A trial that fails (here FIRE parameters the minimiser rejects:
    max_move must stay below pad / 2) is reported with the minimiser's
    status, and the positions are those the tuner started from."""
    packs = _packs()
    s = native.System(packs, pads=0.1)
    try:
        before = s.positions()
        with pytest.raises(native.SpherocylsError) as e:
            s.autotune_fire(native.fire_params(max_move=0.06), iterations=10)
        assert e.value.status == native.SC_ERR_ARG
        for u, v in zip(before, s.positions()):
            np.testing.assert_array_equal(u, v)
    finally:
        s.close()
