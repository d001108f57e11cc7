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
