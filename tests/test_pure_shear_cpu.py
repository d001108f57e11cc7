"""Pure shear and the shear-direction flip in the oracle.

The reference's pure-shear step exists only in its GPU code
(src/cuda/2point_dynamics.cu:1663-2285), which cannot run (it reads an
unallocated buffer), and its flip only there too (find_neighbors.cu:552-579).
The restatement is pinned by what the mathematics fixes:
  * at zero strain rate a pure-shear step is the Heun step of the CPU code,
    which the golden vectors pin;
  * free particles move with the affine field: x / Lx and y / Ly stay fixed,
    and the box area shrinks by (1 - h^2) per step, h = dt rate / 2;
  * a free rod turns towards the extension axis as the exact solution of
    dphi/dt = -(rate / 2) RC sin(2 phi) does: tan(phi) decays as
    exp(-rate RC t);
  * the flip is a mirror image: energy and normal stresses unchanged, Fx, tau
    and the shear stress change sign, and two flips restore the state.
"""
import math

import numpy as np
import pytest

import oracle
from spherocyls_amd import configs


def _box(pk, rate, dt, gamma=0.0, model=1):
    b = oracle.OracleBox(pk.L, pk.x, pk.y, pk.phi, pk.A, pk.R, 0.1)
    b.set_contact_model(model)
    b.set_params(gamma, 0.0, rate, 0.0, dt)
    b.find_neighbors()
    return b


def _wrapdiff(a, b, L):
    d = a - b
    return np.max(np.abs(d - L * np.round(d / L)))


@pytest.mark.parametrize("model", [0, 1])
def test_zero_rate_is_the_heun_step(model):
    R, A = configs.bidisperse(256)
    pk = configs.make_packing(R, A, 0.85, 5, "pure")
    a = _box(pk, 0.0, 0.01, model=model)
    b = _box(pk, 0.0, 0.01, model=model)
    a.calc_forces()
    a.run_steps(20)  # strain_step, then the forces of the new state
    for _ in range(20):
        b.pure_step()
    pa, pb = a.particles(), b.particles()
    assert _wrapdiff(pa["x"], pb["x"], pk.L) < 1e-12
    assert _wrapdiff(pa["y"], pb["y"], pk.L) < 1e-12
    assert _wrapdiff(pa["phi"], pb["phi"], 2 * math.pi) < 1e-12
    assert b.scalars()["Lx"] == pk.L and b.scalars()["Ly"] == pk.L


def test_free_particles_follow_the_affine_field():
    # four particles far apart in a big box: no contacts
    L = 200.0
    x = np.array([20.0, 120.0, 70.0, 170.0])
    y = np.array([30.0, 40.0, 150.0, 110.0])
    phi = np.zeros(4)  # along the extension axis: no rotation
    A = np.full(4, 0.5)
    R = np.full(4, 0.5)
    b = oracle.OracleBox(L, x, y, phi, A, R, 0.1)
    rate, dt, n = 2e-3, 0.01, 300
    b.set_params(0.0, 0.0, rate, 0.0, dt)
    b.find_neighbors()
    for _ in range(n):
        b.pure_step()
    s, p = b.scalars(), b.particles()
    h = 0.5 * dt * rate
    np.testing.assert_allclose(p["x"] / s["Lx"], x / L, rtol=1e-12)
    np.testing.assert_allclose(p["y"] / s["Ly"], y / L, rtol=1e-12)
    assert s["Lx"] == pytest.approx(L * (1 + h) ** n, rel=1e-12)
    assert s["Ly"] == pytest.approx(L * (1 - h) ** n, rel=1e-12)
    assert s["Lx"] * s["Ly"] == pytest.approx(L * L * (1 - h * h) ** n, rel=1e-12)
    assert s["gamma_total"] == pytest.approx(n * dt * rate, rel=1e-12)
    np.testing.assert_array_equal(p["phi"], phi)


def test_free_rod_turns_to_the_extension_axis():
    L = 100.0
    phi0 = np.array([1.0, -0.4, 2.5])
    b = oracle.OracleBox(L, np.array([20.0, 50.0, 80.0]), np.array([20.0, 60.0, 30.0]), phi0,
                         np.full(3, 1.0), np.full(3, 0.5), 0.1)
    rate, dt, n = 0.05, 0.01, 2000
    b.set_params(0.0, 0.0, rate, 0.0, dt)
    b.find_neighbors()
    for _ in range(n):
        b.pure_step()
    p = b.particles()
    rc = p["RC"][0]
    t = n * dt
    want = np.tan(phi0) * math.exp(-rate * rc * t)
    np.testing.assert_allclose(np.tan(p["phi"]), want, rtol=2e-3)
    # and they did turn towards the x axis
    assert np.all(np.abs(np.tan(p["phi"])) < np.abs(np.tan(phi0)))


def test_flip_is_a_mirror_image():
    R, A = configs.bidisperse(256)
    pk = configs.make_packing(R, A, 0.86, 9, "flip", gamma=0.2)
    b = _box(pk, 0.0, 0.01, gamma=0.2)
    b.calc_se()
    p0, s0 = b.particles(), b.scalars()
    b.flip_shear()
    b.calc_se()
    p1, s1 = b.particles(), b.scalars()
    assert s1["gamma"] == -0.2
    sc = b.force_scale()
    np.testing.assert_allclose(p1["fx"], -p0["fx"], atol=1e-12 * sc[0].max())
    np.testing.assert_allclose(p1["fy"], p0["fy"], atol=1e-12 * sc[1].max())
    np.testing.assert_allclose(p1["tau"], -p0["tau"], atol=1e-12 * sc[2].max())
    assert s1["E"] == pytest.approx(s0["E"], rel=1e-12)
    for k in ("cPxx", "cPyy"):
        assert s1[k] == pytest.approx(s0[k], rel=1e-12)
    assert s1["cPxy"] == pytest.approx(-s0["cPxy"], rel=1e-10, abs=1e-12 * abs(s0["cPxx"]))
    b.flip_shear()
    p2 = b.particles()
    np.testing.assert_allclose(p2["x"], p0["x"], rtol=0, atol=1e-12)
    np.testing.assert_array_equal(p2["phi"], p0["phi"])
    assert b.scalars()["gamma"] == 0.2


# ---------------------------------------------------------------------------
# rotate_by_gamma (src/cuda/find_neighbors.cu:580-645) and the Lees-Edwards
# remap of a non-square box.  The rotation is a rigid turn of the lab frame, so
# the physics fixes the answer: the contact set, the energy, |F| and the
# torques are unchanged, forces turn with the frame and the stress tensor
# transforms as R s R^T.
# ---------------------------------------------------------------------------
def _pairset(pq):
    return set(zip(pq[0].tolist(), pq[1].tolist()))


def _sheared_box(gamma, seed=71, n=200, aspect=2.0, phi=0.84):
    R, A = configs.bidisperse(n, aspect=aspect)
    pk = configs.make_packing(R, A, phi, seed, gamma=gamma)
    b = oracle.OracleBox(pk.L, pk.x, pk.y, pk.phi, pk.A, pk.R, 0.1)
    b.set_params(gamma, 0.0, 0.0, 0.0, 0.05)
    b.find_neighbors()
    b.calc_se()
    return pk, b


@pytest.mark.parametrize("gamma", [0.3, -0.2, 0.0])
def test_rotate_by_gamma_is_a_rigid_rotation(gamma):
    pk, b = _sheared_box(gamma)
    p0, s0, c0 = b.particles(), b.scalars(), _pairset(b.contacts())
    b.rotate_by_gamma()
    b.find_neighbors()
    b.calc_se()
    p1, s1, c1 = b.particles(), b.scalars(), _pairset(b.contacts())
    root = math.sqrt(1 + gamma * gamma)
    assert s1["Lx"] == pytest.approx(pk.L * root, rel=1e-15)
    assert s1["Ly"] == pytest.approx(pk.L / root, rel=1e-15)
    assert s1["gamma"] == -gamma
    assert c1 == c0 and len(c0) > 100
    assert s1["E"] == pytest.approx(s0["E"], rel=1e-11)
    # the rotation that takes (gamma, 1) to the x axis (for gamma < 0 the frame
    # turns the other way: the new x axis is -(gamma, 1))
    sg = -1.0 if math.copysign(1.0, gamma) < 0 else 1.0
    Rm = sg * np.array([[gamma, 1.0], [-1.0, gamma]]) / root
    F0 = np.stack([p0["fx"], p0["fy"]])
    F1 = np.stack([p1["fx"], p1["fy"]])
    sx, sy, st = b.force_scale()
    scale = np.hypot(sx, sy) + 1e-300
    assert np.all(np.abs(F1 - Rm @ F0) <= 1e-10 * scale)
    assert np.all(np.abs(p1["tau"] - p0["tau"]) <= 1e-10 * (st + 1e-300))
    S0 = np.array([[s0["cPxx"], s0["cPxy"]], [s0["cPyx"], s0["cPyy"]]])
    S1 = np.array([[s1["cPxx"], s1["cPxy"]], [s1["cPyx"], s1["cPyy"]]])
    np.testing.assert_allclose(S1, Rm @ S0 @ Rm.T, rtol=1e-9, atol=1e-12 * np.abs(S0).max())


def test_lees_edwards_remap_of_a_non_square_box():
    """After the rotation Lx / Ly = 1 + gamma^2: the remap at gamma = Lx / (2 Ly),
    x += (Lx / Ly) y, gamma -= Lx / Ly is a lattice translation, so every force
    and the energy are unchanged by it."""
    pk, b = _sheared_box(0.3, seed=72, phi=0.86)
    b.rotate_by_gamma()
    s = b.scalars()
    r = s["Lx"] / s["Ly"]
    assert r == pytest.approx(1.09, rel=1e-12)
    b.set_params(0.5 * r, 0.0, 0.0, 0.0, 0.05)
    b.find_neighbors()
    b.calc_se()
    p0, s0 = b.particles(), b.scalars()
    b.set_back_gamma()
    assert b.scalars()["gamma"] == pytest.approx(-0.5 * r, rel=1e-15)
    b.calc_se()
    p1 = b.particles()
    sx, sy, st = b.force_scale()
    for k, sc in (("fx", sx), ("fy", sy), ("tau", st)):
        assert np.all(np.abs(p1[k] - p0[k]) <= 1e-11 * (sc + 1e-300)), k
    assert b.scalars()["E"] == pytest.approx(s0["E"], rel=1e-12)
    assert len(b.contacts()[0]) == s0["ncontacts"]


def test_non_square_cells_cover_contacts_at_the_remap_strain():
    """In the widened cell grid of a box with Lx / Ly = 1.25 at the largest
    strain it reaches, Lx / (2 Ly), the neighbour-list contacts equal the O(N^2)
    brute-force contacts."""
    pk, b = _sheared_box(0.5, seed=73, n=400, aspect=3.0, phi=0.86)
    b.rotate_by_gamma()
    s = b.scalars()
    r = s["Lx"] / s["Ly"]
    for g in (0.5 * r - 1e-9, -0.5 * r):
        b.set_params(g, 0.0, 0.0, 0.0, 0.05)
        b.find_neighbors()
        assert _pairset(b.contacts()) == _pairset(b.brute_contacts())
