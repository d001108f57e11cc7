"""FIRE of small packings in one launch (one workgroup per packing runs every
iteration) against the oracle, and against the two-launch path it replaces."""
import numpy as np
import pytest

import oracle
from spherocyls_amd import configs, native

pytestmark = pytest.mark.gpu

FIRE = dict(dt0=0.05, dt_max=0.5, n_min=5, f_inc=1.1, f_dec=0.5, alpha0=0.1, f_alpha=0.99,
            max_move=0.02, ftol=1e-10)
MODELS = {"closest": 0, "two_point": 1}


def wrap_diff(a, b, L):
    d = a - b
    return d - L * np.round(d / L)


def _oracle_box(pk, x, y, phi, model):
    ob = oracle.OracleBox(pk.L, x, y, phi, pk.A, pk.R, 0.1)
    ob.set_contact_model(MODELS[model])
    return ob


@pytest.mark.parametrize("n,phi,model", [(64, 0.80, "closest"), (200, 0.84, "closest"),
                                         (100, 0.84, "two_point"), (256, 0.80, "two_point")])
@pytest.mark.parametrize("small", [256, 0])
def test_first_iterations_match_oracle(n, phi, model, small):
    """40 FIRE iterations (max_iter = 40) of one small packing, in one launch
    (small = 256) and in the two-launch path (small = 0): positions and angles
    within 1e-9 of the oracle's 40 iterations."""
    R, A = configs.bidisperse(n)
    pk = configs.make_packing(R, A, phi, 61 + n)
    k = 40
    params = native.fire_params(**FIRE, max_iter=k)
    print("small FIRE first iterations", n, phi, model, small, flush=True)
    s = native.System(pk, pads=0.1)
    try:
        s.set_contact_model(model)
        s.set_small_fire(small)
        it, fr, st = s.fire(params, check=False)
        assert st == native.SC_ERR_NOT_CONVERGED and it[0] == k
        x, y, phi_ = s.positions()
    finally:
        s.close()
    ob = oracle.OracleBox.from_packing(pk, 0.1)
    ob.set_contact_model(MODELS[model])
    ob.fire(params)
    p = ob.particles()
    assert np.max(np.abs(wrap_diff(x, p["x"], pk.L))) < 1e-9
    assert np.max(np.abs(wrap_diff(y, p["y"], pk.L))) < 1e-9
    assert np.max(np.abs(wrap_diff(phi_, p["phi"], 2 * np.pi))) < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("model,n,aspect,phi,seed", [("closest", 160, 1.0, 0.86, 7),
                                                     ("two_point", 256, 2.0, 0.90, 3)])
def test_converges_to_an_oracle_minimum(precision, model, n, aspect, phi, seed):
    """A small jammed packing minimised in one launch: F_rms < 1e-10 as the
    oracle judges the final state, the reported energy and stress are the
    oracle's at that state (1e-8 relative), and the iteration count is of the
    order of the two-launch path's on the same packing.  Closest-point FIRE on
    rods can stall on the energy's kinks, whatever the launch, so that model
    runs on discs; the two-point packing is the jammed one of
    test_two_point_fire_jammed_stress, which converges from perturbed starts."""
    R, A = configs.bidisperse(n, aspect=aspect)
    pk = configs.make_packing(R, A, phi, seed, f"N={n} aspect {aspect} phi {phi}")
    res = {}
    for small in (256, 0):
        s = native.System(pk, pads=0.1)
        try:
            s.set_contact_model(model)
            if precision == "mixed":
                s.set_precision("mixed")
            s.set_small_fire(small)
            it, fr, st = s.fire(native.fire_params(**FIRE, max_iter=50000), switch_frms=1e-6,
                                check=False)
            x, y, ang = s.positions()
            bx = s.box()
        finally:
            s.close()
        print("small FIRE", model, precision, "small", small, "iterations", int(it[0]),
              "F_rms", float(fr[0]), flush=True)
        assert st == 0 and fr[0] < 1e-10
        assert bx["E"] > 1e-8  # jammed: stress well above the residual's scale
        ob = _oracle_box(pk, x, y, ang, model)
        assert ob.frms() < 1e-10
        ob.calc_se()
        o = ob.scalars()
        scale = max(abs(o["cPxx"]), abs(o["cPyy"]), 1e-300)
        for key in ("E", "cPxx", "cPyy", "Pxx", "Pyy"):
            assert bx[key] == pytest.approx(o[key], rel=1e-8, abs=1e-8 * scale), key
        assert bx["cPxy"] == pytest.approx(o["cPxy"], rel=1e-8, abs=1e-8 * scale)
        res[small] = int(it[0])
    assert 0.5 * res[0] < res[256] < 2.0 * res[0]


def test_tail_of_a_batch_and_tiny_packings():
    """A batch whose large packing (dilute, quickly relaxed) finishes first, so
    that the small packings still running finish in the one-launch kernel, with
    one- and two-particle packings in small boxes: every packing converges and
    the oracle judges each final state a minimum."""
    packs = []
    R, A = configs.bidisperse(1500)
    packs.append(configs.make_packing(R, A, 0.55, 11, "N=1500 phi 0.55"))
    for k, (n, phi) in enumerate([(90, 0.85), (150, 0.86), (256, 0.84)]):
        R, A = configs.bidisperse(n)
        packs.append(configs.make_packing(R, A, phi, 20 + k, f"N={n} phi {phi}"))
    # one rod alone, and two overlapping rods, in boxes of a single cell
    packs.append(configs.Packing(4.4, np.array([1.0]), np.array([2.0]), np.array([0.3]),
                                 np.array([0.5]), np.array([0.5]), 0.0, "one rod"))
    packs.append(configs.Packing(4.4, np.array([1.0, 1.6]), np.array([2.0, 2.3]),
                                 np.array([0.3, -0.4]), np.array([0.5, 0.5]),
                                 np.array([0.5, 0.5]), 0.0, "two rods"))
    s = native.System(packs, pads=0.1)
    try:
        s.set_contact_model("two_point")
        it, fr, st = s.fire(native.fire_params(**FIRE, max_iter=50000))
        pos = s.positions()
    finally:
        s.close()
    print("small FIRE batch iterations", it.tolist(), flush=True)
    assert st == 0 and np.all(fr < 1e-10)
    assert it[0] < max(it[1:4])  # the large packing finished before the slowest small one
    off = np.concatenate([[0], np.cumsum([p.n for p in packs])])
    for k, pk in enumerate(packs):
        x, y, phi = (a[off[k]:off[k + 1]] for a in pos)
        ob = _oracle_box(pk, x, y, phi, "two_point")
        assert ob.frms() < 1e-10, pk.name


@pytest.mark.parametrize("small", [256, 0])
def test_fire_in_a_non_square_box(small):
    """After rotate_by_gamma (Lx / Ly = 1 + gamma^2) and at a strain just below
    the box's remap strain Lx / (2 Ly) > 1/2, where the rows' x margin must
    exceed 1.5 skins: FIRE in one launch and in two launches reaches F_rms <
    1e-10 as the oracle, in the same box at the same strain, judges it."""
    R, A = configs.bidisperse(200)
    gamma = 0.4
    pk = configs.make_packing(R, A, 0.86, 31, "non-square", gamma=gamma)
    params = native.fire_params(**FIRE, max_iter=50000)
    s = native.System(pk, pads=0.1)
    try:
        s.set_contact_model("two_point")
        s.set_gamma(gamma, 0.0)
        s.rotate_by_gamma()
        b = s.box()
        g1 = 0.5 * b["Lx"] / b["Ly"] - 0.01
        assert g1 > 0.55
        s.set_gamma(g1, 0.0)
        s.set_small_fire(small)
        it, fr, st = s.fire(params)
        x, y, phi = s.positions()
        print("non-square FIRE, small", small, "iterations", int(it[0]), flush=True)
    finally:
        s.close()
    ob = oracle.OracleBox(pk.L, pk.x, pk.y, pk.phi, pk.A, pk.R, 0.1)
    ob.set_params(gamma, 0.0, 0.0, 0.0, 0.05)
    ob.rotate_by_gamma()
    ob.set_contact_model(1)
    ob.set_params(g1, 0.0, 0.0, 0.0, 0.05)
    ob.set_positions(x, y, phi)
    ob.find_neighbors()
    assert ob.scalars()["Lx"] == pytest.approx(b["Lx"], rel=1e-15)
    assert ob.frms() < 1e-10


def test_tile_path_keeps_two_launches():
    """The one-launch kernel runs on the list path only (its row capacity is
    the list kernels'): on the tile path a small packing minimises the same,
    bit for bit, whatever the small-packing setting."""
    R, A = configs.bidisperse(120)
    pk = configs.make_packing(R, A, 0.84, 9)
    out = []
    for small in (256, 0):
        s = native.System(pk, pads=0.1)
        try:
            s.set_contact_model("two_point")
            s.set_path("tile")
            s.set_small_fire(small)
            it, fr, _ = s.fire(native.fire_params(**FIRE, max_iter=50000))
            out.append((it.copy(), s.positions()))
        finally:
            s.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    for u, v in zip(out[0][1], out[1][1]):
        np.testing.assert_array_equal(u, v)


def test_launches_of_a_few_iterations_each():
    """With 7 iterations per launch the minimiser launches the one-launch
    kernel again and again (a few hundred times here); every launch resumes
    from the state, velocities and FIRE scalars the last one left, so the
    packing still reaches a minimum the oracle accepts, in about as many
    iterations as in one long launch."""
    R, A = configs.bidisperse(160, aspect=2.0)
    pk = configs.make_packing(R, A, 0.90, 3, "N=160 phi 0.90")
    out = {}
    for steps in (0, 7):
        s = native.System(pk, pads=0.1)
        try:
            s.set_contact_model("two_point")
            s.set_small_fire(256, steps)
            it, fr, st = s.fire(native.fire_params(**FIRE, max_iter=50000))
            x, y, phi = s.positions()
        finally:
            s.close()
        print("launches of", steps or "50000", "iterations:", int(it[0]), "iterations", flush=True)
        assert st == 0 and fr[0] < 1e-10
        assert _oracle_box(pk, x, y, phi, "two_point").frms() < 1e-10
        out[steps] = int(it[0])
    assert 0.5 * out[0] < out[7] < 2.0 * out[0]


def test_small_fire_argument_range():
    R, A = configs.bidisperse(64)
    pk = configs.make_packing(R, A, 0.80, 5)
    s = native.System(pk, pads=0.1)
    try:
        for bad in (-1, 257):
            with pytest.raises(native.SpherocylsError):
                s.set_small_fire(bad)
        with pytest.raises(native.SpherocylsError):
            s.set_small_fire(256, -1)
        s.set_small_fire(0)
        s.set_small_fire(256)
    finally:
        s.close()
