"""CPU tests that pin the oracle restatement (oracle/sc_oracle.cpp).

* against the golden vectors written by the compiled reference (tests/golden),
* against two-body closed forms and the isolated-particle shear rotation law,
* against invariants the physics fixes (Newton III, F = -N dE/dx, translation,
  periodic images, the Lees-Edwards remap),
* against an O(N^2) brute-force contact search.
FIRE has no reference counterpart; it is pinned by its defining properties.
"""
import math

import numpy as np
import pytest

import oracle
from conftest import golden, golden_names, wrap_diff
from spherocyls_amd import configs


def box_from_golden(g, pad=None):
    b = oracle.OracleBox(float(g["L"][0]), g["x"], g["y"], g["phi"], g["A"], g["R"],
                         float(g["pad"][0]) if pad is None else pad)
    return b


# ---------------------------------------------------------------------------
# golden vectors (compiled reference)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", golden_names("static_"))
def test_golden_static(name):
    g = golden(name)
    b = box_from_golden(g)
    b.set_params(float(g["gamma"][0]), 0.0, 0.0, 0.0, 0.05)
    b.find_neighbors()
    b.calc_se()
    p = b.particles()
    s = b.scalars()
    # same algorithm, same operation order: agreement to the last bit
    for k in ("fx", "fy", "tau"):
        np.testing.assert_array_equal(p[k], g[k])
    np.testing.assert_array_equal([s["E"], s["Pxx"], s["Pyy"], s["Pxy"]], g["se"])
    assert (s["rows"], s["cols"]) == (g["rows"][0], g["cols"][0])
    assert s["pad"] == g["eff_pad"][0]
    np.testing.assert_array_equal(p["x"], g["wx"])
    np.testing.assert_array_equal(p["phi"], g["wphi"])
    np_, nq = b.neighbors()
    np.testing.assert_array_equal(np_, g["nbr_p"])
    np.testing.assert_array_equal(nq, g["nbr_q"])


@pytest.mark.parametrize("name", golden_names("heun_"))
def test_golden_heun(name):
    g = golden(name)
    b = box_from_golden(g)
    b.set_params(float(g["gamma"][0]), 0.0, float(g["strain_rate"][0]), 0.0, float(g["dt"][0]))
    b.calc_forces()
    b.strain_step()
    p = b.particles()
    np.testing.assert_array_equal(p["x"], g["x1"])
    np.testing.assert_array_equal(p["phi"], g["phi1"])
    assert b.scalars()["gamma"] == g["gamma1"][0]
    b.calc_forces()
    se = b.run_steps(int(g["nsteps"][0]), int(g["se_every"][0]))
    np.testing.assert_array_equal(se, g["se"])
    p = b.particles()
    for k in ("x", "y", "phi", "fx", "fy", "tau"):
        np.testing.assert_array_equal(p[k], g[k + "N"])
    s = b.scalars()
    assert s["gamma"] == g["gammaN"][0]
    assert s["gamma_total"] == g["gamma_totalN"][0]


def test_golden_lees_edwards_crossed():
    g = golden("heun_lees_edwards")
    g0 = float(g["gamma"][0])
    steps = 1 + int(g["nsteps"][0])
    # the run crosses gamma = 0.5 once: gamma_N = gamma_0 + n dt rate - 1
    assert g["gammaN"][0] == pytest.approx(g0 + steps * g["dt"][0] * g["strain_rate"][0] - 1.0,
                                           abs=1e-12)


def test_golden_resize():
    g = golden("resize128")
    b = box_from_golden(g)
    b.set_params(0.0, 0.0, 0.0, float(g["rate"][0]), float(g["dt"][0]))
    b.calc_forces()
    for _ in range(int(g["nsteps"][0])):
        b.resize_step(shrink=True)
        b.calc_forces()
    p = b.particles()
    for k in ("x", "y", "phi", "fx", "fy", "tau"):
        np.testing.assert_array_equal(p[k], g[k + "N"])
    assert b.scalars()["Lx"] == g["LN"][0]


def test_golden_gd():
    g = golden("gd128")
    b = box_from_golden(g)
    b.set_params(0.0, 0.0, 0.0, 0.0, float(g["dt"][0]))
    es = []
    for _ in range(int(g["nsteps"][0])):
        b.calc_se()
        es.append(b.scalars()["E"])
        b.gd_step()
    b.calc_se()
    np.testing.assert_array_equal(es, g["energies"])
    p = b.particles()
    for k in ("x", "y", "phi", "fx", "fy", "tau"):
        np.testing.assert_array_equal(p[k], g[k + "N"])
    # gradient descent lowers the energy monotonically at this step size
    assert np.all(np.diff(es) < 0)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference library not built here")
@pytest.mark.parametrize("seed", [21, 22])
def test_reference_direct(seed):
    """Where the compiled reference is present, compare on fresh seeded inputs."""
    R, A = configs.bidisperse(300, aspect=3.0)
    pk = configs.make_packing(R, A, 0.84, seed, gamma=-0.3)
    r = oracle.RefBox.from_packing(pk, 0.2)
    o = oracle.OracleBox.from_packing(pk, 0.2)
    for b in (r, o):
        b.set_params(-0.3, 0.0, 1e-3, 0.0, 0.02)
        b.find_neighbors()
        b.calc_se()
    pr, po = r.particles(), o.particles()
    for k in ("fx", "fy", "tau"):
        np.testing.assert_array_equal(pr[k], po[k])
    sr, so = r.run_steps(50, 10), o.run_steps(50, 10)
    np.testing.assert_array_equal(sr, so)


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------
def two_body(L, p, q, gamma=0.0, pad=0.1):
    """p, q = (x, y, phi, A, R)."""
    arr = np.array([p, q], dtype=float)
    b = oracle.OracleBox(L, arr[:, 0], arr[:, 1], arr[:, 2], arr[:, 3], arr[:, 4], pad)
    b.set_params(gamma, 0.0, 0.0, 0.0, 0.05)
    b.find_neighbors()
    b.calc_se()
    return b.particles(), b.scalars()


def harmonic(h, sigma):
    return (1.0 - h / sigma) / sigma


def test_kat_discs():
    # A = 0 reduces to harmonic discs: |F| = (1 - d/sigma)/sigma along the centre line
    L, d = 20.0, 0.9
    p, s = two_body(L, (5.0, 5.0, 0.3, 0.0, 0.5), (5.0 + d * 0.6, 5.0 + d * 0.8, 1.1, 0.0, 0.6))
    f = harmonic(d, 1.1)
    np.testing.assert_allclose([p["fx"][0], p["fy"][0]], [-f * 0.6, -f * 0.8], rtol=1e-14)
    np.testing.assert_allclose([p["fx"][1], p["fy"][1]], [f * 0.6, f * 0.8], rtol=1e-14)
    assert p["tau"][0] == 0.0 and p["tau"][1] == 0.0
    assert s["E"] == pytest.approx((1 - d / 1.1) ** 2 / 4, rel=1e-14)


def test_kat_t_shape():
    # rod p along x, rod q along y above it: contact at p's centre and q's lower end
    L, A, Aq, R, h = 30.0, 1.0, 0.8, 0.5, 0.7
    p, s = two_body(L, (10.0, 10.0, 0.0, A, R), (10.0, 10.0 + Aq + h, math.pi / 2, Aq, R))
    f = harmonic(h, 2 * R)
    np.testing.assert_allclose(p["fy"], [-f, f], rtol=1e-13)
    np.testing.assert_allclose(p["fx"], [0.0, 0.0], atol=1e-15)
    np.testing.assert_allclose(p["tau"], [0.0, 0.0], atol=1e-15)
    # corrected virial: P_yy = -dy F_y / area with dy the centre separation
    assert s["cPyy"] == pytest.approx(-(-(Aq + h)) * (-f) / L ** 2, rel=1e-12)


def test_kat_end_to_end():
    L, Ap, Aq, R, h = 30.0, 1.0, 0.6, 0.5, 0.4
    p, _ = two_body(L, (10.0, 10.0, 0.0, Ap, R), (10.0 + Ap + Aq + h, 10.0, 0.0, Aq, R))
    f = harmonic(h, 2 * R)
    np.testing.assert_allclose(p["fx"], [-f, f], rtol=1e-13)
    np.testing.assert_allclose(p["tau"], [0.0, 0.0], atol=1e-15)


def test_kat_parallel_side_by_side():
    # exactly parallel rods: delta = ac - b^2 = 0 and (bd - ae)/delta is NaN; the
    # reference's fmax(NaN, -1) = -1 picks q's end t = -1 and the matching s.
    L, A, R, h = 30.0, 1.0, 0.5, 0.6
    p, _ = two_body(L, (10.0, 10.0, 0.0, A, R), (10.0, 10.0 + h, 0.0, A, R))
    f = harmonic(h, 2 * R)
    np.testing.assert_allclose(p["fy"], [-f, f], rtol=1e-13)
    np.testing.assert_allclose(p["fx"], [0.0, 0.0], atol=1e-15)
    # contact point at the ends s = t = -1: torques -A F_y and +A F_y on p, q
    np.testing.assert_allclose(p["tau"], [A * f, -A * f], rtol=1e-13)


def test_kat_lees_edwards_image():
    # across the y boundary the lab separation picks up gamma * L
    L, R, gamma = 20.0, 0.5, 0.3
    # q sits just below p's periodic image across y=L, shifted by gamma*L in x
    p, _ = two_body(L, (5.0, 0.2, 0.0, 0.0, R), (5.0 - gamma * L + 0.0, L - 0.5, 0.0, 0.0, R),
                    gamma=gamma)
    # separation p - q in sheared coords: dx = gamma L, dy = 0.7 - L -> +0.7 after min image
    # lab dx = gamma*L + gamma*0.7 - ... the minimum image is taken before the shear shift
    dx_s = gamma * L
    dx_s -= L * round(dx_s / L)
    dy = 0.2 - (L - 0.5) + L
    dx = dx_s + gamma * dy
    d = math.hypot(dx, dy)
    if d < 2 * R:
        f = harmonic(d, 2 * R)
        np.testing.assert_allclose(p["fx"][0], f * dx / d, rtol=1e-12)
        np.testing.assert_allclose(p["fy"][0], f * dy / d, rtol=1e-12)
    else:
        assert p["fx"][0] == 0.0


def test_kat_isolated_rotation_orbit():
    # No contacts: phi' = -gdot (1 - RC cos 2phi) / 2 has the closed-form orbit
    # tan(phi) = sqrt((1-RC)/(1+RC)) tan(-gdot/2 sqrt(1-RC^2) t + c0).
    R, A = 0.5, 0.5
    b = oracle.OracleBox(50.0, [25.0], [25.0], [0.4], [A], [R], 0.1)
    gdot, dt, n = 1.0, 0.01, 200
    b.set_params(0.0, 0.0, gdot, 0.0, dt)
    b.calc_forces()
    for _ in range(n):
        b.strain_step()
        b.calc_forces()
    p = b.particles()
    RC = p["RC"][0]
    k = math.sqrt((1 - RC) / (1 + RC))
    c0 = math.atan(math.tan(0.4) / k)
    t = n * dt
    phi_exact = math.atan(k * math.tan(-0.5 * gdot * math.sqrt(1 - RC ** 2) * t + c0))
    # Heun is second order: global error ~ dt^2
    assert p["phi"][0] == pytest.approx(phi_exact, abs=5e-5)
    assert p["x"][0] == 25.0 and p["y"][0] == 25.0
    # rotational constants from the spherocylinder's area and inertia
    al = A / R
    C = 3 * math.pi + 24 * al + 6 * math.pi * al ** 2 + 8 * al ** 3
    assert RC == pytest.approx((8 * al + 6 * math.pi * al ** 2 + 8 * al ** 3) / C, rel=1e-12)


# ---------------------------------------------------------------------------
# invariants
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def dense():
    R, A = configs.bidisperse(160, aspect=2.5)
    return configs.make_packing(R, A, 0.86, 31, gamma=0.17)


def forces_of(pk, gamma, pad=0.1, x=None):
    b = oracle.OracleBox(pk.L, pk.x if x is None else x, pk.y, pk.phi, pk.A, pk.R, pad)
    b.set_params(gamma, 0.0, 0.0, 0.0, 0.05)
    b.find_neighbors()
    b.calc_se()
    return b


def test_newton_third_law(dense):
    b = forces_of(dense, 0.17)
    p = b.particles()
    sx, sy, _ = b.force_scale()
    assert abs(p["fx"].sum()) < 1e-13 * sx.sum()
    assert abs(p["fy"].sum()) < 1e-13 * sy.sum()


def test_force_is_minus_n_grad_energy(dense):
    b = forces_of(dense, 0.17)
    p = b.particles()
    n = dense.n
    h = 1e-6
    rng = np.random.default_rng(0)
    for i in rng.choice(n, 6, replace=False):
        for comp, key in (("x", "fx"), ("y", "fy"), ("phi", "tau")):
            e = []
            for sgn in (1, -1):
                arrs = {k: p[k].copy() for k in ("x", "y", "phi")}
                arrs[comp][i] += sgn * h
                b.set_positions(arrs["x"], arrs["y"], arrs["phi"])
                b.calc_se()
                e.append(b.scalars()["E"])
            b.set_positions(p["x"], p["y"], p["phi"])
            grad = (e[0] - e[1]) / (2 * h)
            # positions are sheared coordinates: x_lab = x_s + gamma y, so
            # -N dE/dy_s = F_y + gamma F_x, while x and phi map directly
            expect = p[key][i] + (0.17 * p["fx"][i] if comp == "y" else 0.0)
            assert -n * grad == pytest.approx(expect, abs=1e-7)


def test_translation_and_image_invariance(dense):
    b0 = forces_of(dense, 0.17)
    p0 = b0.particles()
    sx, sy, st = b0.force_scale()
    shifted = np.mod(dense.x + 3.3, dense.L)
    p1 = forces_of(dense, 0.17, x=shifted).particles()
    for k, sc in (("fx", sx), ("fy", sy), ("tau", st)):
        assert np.all(np.abs(p1[k] - p0[k]) <= 1e-12 * (sc + 1e-300))
    b2 = oracle.OracleBox(dense.L, dense.x, dense.y, dense.phi, dense.A, dense.R, 0.1)
    b2.set_params(0.17, 0.0, 0.0, 0.0, 0.05)
    b2.find_neighbors()
    x = b2.particles()["x"].copy()
    x[::7] += dense.L  # periodic images of every 7th particle, same neighbour list
    b2.set_positions(x=x)
    b2.calc_se()
    p2 = b2.particles()
    for k, sc in (("fx", sx), ("fy", sy), ("tau", st)):
        assert np.all(np.abs(p2[k] - p0[k]) <= 1e-12 * (sc + 1e-300))


def test_lees_edwards_remap_invariance():
    R, A = configs.bidisperse(160, aspect=2.5)
    dense = configs.make_packing(R, A, 0.86, 33, gamma=0.5)
    b = oracle.OracleBox(dense.L, dense.x, dense.y, dense.phi, dense.A, dense.R, 0.1)
    b.set_params(0.5, 0.5, 0.0, 0.0, 0.05)
    b.find_neighbors()
    b.calc_se()
    p0 = b.particles()
    s0 = b.scalars()
    b.set_back_gamma()
    assert b.scalars()["gamma"] == -0.5
    b.calc_se()
    p1 = b.particles()
    sx, sy, st = b.force_scale()
    for k, sc in (("fx", sx), ("fy", sy), ("tau", st)):
        assert np.all(np.abs(p1[k] - p0[k]) <= 1e-11 * (sc + 1e-300))
    assert b.scalars()["E"] == pytest.approx(s0["E"], rel=1e-12)


def test_brute_force_contacts(dense):
    b = forces_of(dense, 0.17)
    c = set(zip(*b.contacts()))
    bf = set(zip(*b.brute_contacts()))
    assert c == bf and len(c) > dense.n


def test_stress_reference_slips_documented(dense):
    """The reference's Pxx/Pyy/Pxy carry two slips (src/cpp/spherocyl_box.cpp:486,
    493); the corrected tensor is the symmetric-in-sums virial -sum dx_a F_b."""
    b = forces_of(dense, 0.17)
    s = b.scalars()
    assert s["Pxx"] != s["cPxx"]
    # the corrected virial traces agree with -sum r.F / area from the forces
    assert s["cPxx"] + s["cPyy"] < 0  # repulsive contacts: negative in this sign convention


# ---------------------------------------------------------------------------
# FIRE (project specification)
# ---------------------------------------------------------------------------
FIRE = np.array([0.05, 0.5, 5, 1.1, 0.5, 0.1, 0.99, 0.02, 1e-10, 200000], dtype=float)


def test_fire_converges_small():
    # an overlap-free minimum without aligned overlapping pairs on the way:
    # F_rms reaches the tolerance with every particle still finite
    R, A = configs.bidisperse(256)
    pk = configs.make_packing(R, A, 0.70, 41)
    b = oracle.OracleBox.from_packing(pk, 0.1)
    it, fr = b.fire(FIRE)
    assert fr < 1e-10 and it < 5000
    assert b.frms() < 1e-10
    p = b.particles()
    assert np.all(np.isfinite(p["x"])) and np.all(np.isfinite(p["phi"]))


def test_kat_crossed_rods_are_pushed_apart():
    """Rods crossing in an X: the overlap energy is continued to negative
    distance -delta, delta the smallest distance of an endpoint from the other
    segment's line; the rod owning that endpoint is pushed back across the
    line with (1 + delta/sigma)/sigma (the reference normalises rounding noise
    there instead)."""
    # p along x through (10, 10); q vertical through (10.5, 10.2), reaching
    # y = 10.2 - 1 = 9.2 below p's line and 11.2 above: q's lower end is 0.8
    # past p's line, q's upper end 1.2; p's ends are 0.5 and 1.5 from q's line
    p, s = two_body(20.0, (10.0, 10.0, 0.0, 1.0, 0.5), (10.5, 10.2, math.pi / 2, 1.0, 0.5))
    delta, sig = 0.5, 1.0  # p's right end is 0.5 from q's line: move p left
    f = (1 + delta / sig) / sig
    np.testing.assert_allclose(p["fx"], [-f, f], rtol=1e-12)
    np.testing.assert_allclose(p["fy"], [0.0, 0.0], atol=1e-12)
    # the force acts at p's end (arm +1 along x: no torque for a force along x)
    # and at the point of q level with p, arm t n_q = (0, -0.2): torque on q
    # = arm x F_q = -(-0.2) * f
    np.testing.assert_allclose(p["tau"][0], 0.0, atol=1e-12)
    assert p["tau"][1] == pytest.approx(0.2 * f, rel=1e-12)
    assert s["E"] == pytest.approx((1 + delta / sig) ** 2 / 4, rel=1e-12)


def test_crossed_force_is_continuous_where_rods_uncross():
    """Moving p left until its right end leaves q's line: the force just before
    (crossed, d = -delta) and just after (ordinary contact, d = +delta) agree
    to O(delta)."""
    eps = 1e-7
    f = []
    for x in (10.5 - 1.0 + eps, 10.5 - 1.0 - eps):  # p's right end at 10.5 -+ eps
        p, _ = two_body(20.0, (x, 10.0, 0.0, 1.0, 0.5), (10.5, 10.2, math.pi / 2, 1.0, 0.5))
        f.append((p["fx"][0], p["fy"][0]))
    np.testing.assert_allclose(f[0], f[1], atol=1e-6)
    assert f[0][0] == pytest.approx(-1.0, rel=1e-6)


def test_crossed_force_is_minus_n_grad_energy():
    """Inside the crossed region the continued energy is smooth away from the
    switch between endpoints: F = -N dE/dx by central differences."""
    rng = np.random.default_rng(9)
    checked = 0
    for _ in range(40):
        a = (10.0, 10.0, rng.uniform(0, math.pi), rng.uniform(0.6, 1.2), 0.5)
        ph = a[2] + rng.uniform(0.6, 2.5)
        b = (10.0 + rng.uniform(-0.3, 0.3), 10.0 + rng.uniform(-0.3, 0.3), ph,
             rng.uniform(0.6, 1.2), 0.6)
        p, s = two_body(20.0, a, b)
        if s["E"] < 0.125:  # not crossed (E >= 1/4 per crossed pair of 2 particles)
            continue
        for comp, key in ((0, "fx"), (1, "fy"), (2, "tau")):
            h = 1e-6
            e = []
            for sg in (1, -1):
                aa = list(a)
                aa[comp] += sg * h
                e.append(two_body(20.0, tuple(aa), b)[1]["E"])
            g = -2 * (e[0] - e[1]) / (2 * h)
            if abs(g - p[key][0]) > 1e-5 * (1 + abs(g)):
                # straddles the endpoint switch: a finite difference there
                # is no derivative
                continue
            checked += 1
    assert checked > 30


def test_fire_stalls_at_parallel_kink():
    """The closest-point energy has |theta| kinks at aligned overlapping pairs
    (distance h - A|theta|): the torque flips sign across alignment, FIRE keeps
    resetting, and the residual does not vanish even where the two-point model
    relaxes the same packing to zero energy (tests/test_two_point_cpu.py); the
    minimiser reports non-convergence instead of a false pass."""
    R, A = configs.bidisperse(128)
    pk = configs.make_packing(R, A, 0.86, 47)
    b = oracle.OracleBox.from_packing(pk, 0.1)
    params = FIRE.copy()
    params[9] = 3000
    it, fr = b.fire(params)
    assert it == 3000 and fr > 1e-6


def test_fire_state_machine():
    """dt grows by f_inc only after more than n_min uphill-free steps, is capped
    at dt_max, and halves (with alpha reset) whenever P = F.v turns negative."""
    R, A = configs.bidisperse(64)
    pk = configs.make_packing(R, A, 0.84, 43)
    b = oracle.OracleBox.from_packing(pk, 0.1)
    n = pk.n
    v = [np.zeros(n) for _ in range(3)]
    st = np.array([FIRE[0], FIRE[5], 0.0, 0.0])
    b.calc_se()
    e0 = b.scalars()["E"]
    seen_reset = seen_growth = False
    for _ in range(400):
        dt, alpha, npos = st[:3].copy()
        b.fire_iter(FIRE, *v, st)
        if st[3] == 1:  # converged at entry: state untouched
            assert (st[0], st[1], st[2]) == (dt, alpha, npos)
            break
        if st[2] == 0:  # retreat branch
            seen_reset = True
            assert st[0] == pytest.approx(dt * FIRE[4]) and st[1] == FIRE[5]
            # velocities restart from zero: v = dt_new g with g the forces of the
            # iteration (fire_iter leaves them in the box)
            f = b.particles()
            np.testing.assert_array_equal(v[0], st[0] * f["fx"])
            np.testing.assert_array_equal(v[1], st[0] * f["fy"])
        elif st[2] > FIRE[2]:
            assert st[0] == pytest.approx(min(dt * FIRE[3], FIRE[1]))
            assert st[1] == pytest.approx(alpha * FIRE[6])
            seen_growth |= st[0] > dt
        else:
            assert st[0] == dt and st[1] == alpha
        assert st[0] <= FIRE[1]
    b.calc_se()
    assert b.scalars()["E"] < 0.1 * e0
    assert seen_reset and seen_growth


def test_fire_alpha_one_is_gradient_flow():
    """With alpha fixed at 1, f_alpha = 1 and no dt growth, the mixed velocity is
    |v| g_hat; starting from v = 0 the first step is plain gradient descent."""
    R, A = configs.bidisperse(32)
    pk = configs.make_packing(R, A, 0.8, 45)
    params = FIRE.copy()
    params[5], params[6], params[3] = 1.0, 1.0, 1.0
    params[7] = 10.0
    a = oracle.OracleBox.from_packing(pk, 0.1)
    a.calc_forces()
    p0 = a.particles()
    v = [np.zeros(pk.n) for _ in range(3)]
    st = np.array([params[0], 1.0, 0.0, 0.0])
    a.fire_iter(params, *v, st)
    p1 = a.particles()
    dt = params[0]
    np.testing.assert_allclose(wrap_diff(p1["x"], p0["x"], pk.L), dt * dt * p0["fx"], atol=1e-15)
    np.testing.assert_allclose(p1["phi"] - p0["phi"], dt * dt * p0["tau"] / p0["I"], atol=1e-15)


def _fire_start(n=32, phi=0.84, seed=47):
    R, A = configs.bidisperse(n)
    pk = configs.make_packing(R, A, phi, seed)
    b = oracle.OracleBox.from_packing(pk, 0.1)
    b.calc_forces()
    return pk, b, b.particles()


def test_fire_uphill_step_retreats_half_a_step():
    """P = g.v < 0: FIRE 2.0's retreat first moves every coordinate back by
    dt v / 2 (dt the step that produced v), then restarts from v = 0 with the
    halved step: v = dt' g, displacement dt' v = dt'^2 g.  With v = -0.3 g the
    new position is x0 + (0.15 dt + dt'^2) g, and psi = phi sqrt(I) likewise
    (tau / sqrt(I) is its generalised force)."""
    pk, b, p0 = _fire_start()
    sq = np.sqrt(p0["I"])
    g = [p0["fx"], p0["fy"], p0["tau"] / sq]
    v = [-0.3 * gi for gi in g]
    dt = 0.04
    st = np.array([dt, 0.07, 3.0, 0.0])
    b.fire_iter(FIRE, *[vi.copy() for vi in v], st)
    p1 = b.particles()
    dtn = dt * FIRE[4]
    assert st[0] == dtn and st[1] == FIRE[5] and st[2] == 0
    assert np.max(dtn * dtn * np.hypot(g[0], g[1])) < FIRE[7]  # no clamp in this step
    # (coordinates of ~10 carry ~2e-15 of rounding; the retreat term is ~1e-3)
    np.testing.assert_allclose(wrap_diff(p1["x"], p0["x"], pk.L), (0.15 * dt + dtn * dtn) * g[0],
                               rtol=1e-9, atol=1e-13)
    np.testing.assert_allclose(wrap_diff(p1["y"], p0["y"], pk.L), (0.15 * dt + dtn * dtn) * g[1],
                               rtol=1e-9, atol=1e-13)
    np.testing.assert_allclose((p1["phi"] - p0["phi"]) * sq, (0.15 * dt + dtn * dtn) * g[2],
                               rtol=1e-9, atol=1e-13)


def test_fire_displacement_is_clamped_to_max_move():
    """A step whose translation dt v is longer than max_move moves the particle
    by exactly max_move along dt v; the rotation is not clamped.  From v = 0
    with a large step dt, v = dt g and the translation is dt^2 g."""
    pk, b, p0 = _fire_start()
    sq = np.sqrt(p0["I"])
    g = [p0["fx"], p0["fy"], p0["tau"] / sq]
    v = [np.zeros(pk.n) for _ in range(3)]
    dt = 0.2
    st = np.array([dt, FIRE[5], 0.0, 0.0])
    b.fire_iter(FIRE, *v, st)
    p1 = b.particles()
    dxy = np.hypot(g[0], g[1]) * dt * dt
    big, small = dxy > FIRE[7], (dxy > 0) & (dxy < FIRE[7])
    assert big.sum() >= 4 and small.sum() >= 4
    mx = wrap_diff(p1["x"], p0["x"], pk.L)
    my = wrap_diff(p1["y"], p0["y"], pk.L)
    np.testing.assert_allclose(np.hypot(mx, my)[big], FIRE[7], rtol=1e-10)
    # direction along g
    np.testing.assert_allclose((mx * g[1] - my * g[0])[big], 0.0, atol=1e-13)
    assert np.all(mx[big] * g[0][big] + my[big] * g[1][big] > 0)
    np.testing.assert_allclose(mx[small], dt * dt * g[0][small], rtol=1e-9, atol=1e-13)
    np.testing.assert_allclose((p1["phi"] - p0["phi"]) * sq, dt * dt * g[2], rtol=1e-9, atol=1e-13)
