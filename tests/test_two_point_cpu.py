"""Two-point contact model in the oracle (src/cuda/2point_dynamics.cu:391-546).

The reference's GPU code cannot run (it reads an unallocated buffer), so no
golden vectors exist for this model.  The restatement is pinned by what the
mathematics fixes:
  * a hand-derived two-body answer (parallel rods overlapping over part of
    their length touch at both ends of the overlap);
  * reduction to the closest-point model when only one point touches;
  * Newton's third law and angular-momentum balance per pair (the torques
    act at the contact points);
  * F = -N grad E by finite differences;
  * continuity of the torque at alignment, where the closest-point model has a
    jump;
  * FIRE reaching F_rms < 1e-10 on a packing where the closest-point model
    stalls.
"""
import math

import numpy as np
import pytest

import oracle
from spherocyls_amd import configs

FIRE = np.array([0.05, 0.5, 5, 1.1, 0.5, 0.1, 0.99, 0.02, 1e-10, 200000], dtype=float)


def two_body(a, b, model=1, L=30.0, gamma=0.0):
    x, y, phi, A, R = (np.array(v, dtype=float) for v in zip(a, b))
    box = oracle.OracleBox(L, x, y, phi, A, R, 0.1)
    box.set_contact_model(model)
    box.set_params(gamma, 0.0, 0.0, 0.0, 0.05)
    box.find_neighbors()
    box.calc_se()
    return box.particles(), box.scalars()


def test_kat_parallel_partial_overlap():
    # p spans x in [9, 11] at y = 10; q spans x in [9.5, 11.5] at y = 10.8.  Their
    # overlap [9.5, 11] ends at q's left end (s = -1/2, t = -1) and p's right end
    # (s = 1, t = 1/2); both contacts are 0.8 apart, force (1 - 0.8)/1 each.
    p, s = two_body((10.0, 10.0, 0.0, 1.0, 0.5), (10.5, 10.8, 0.0, 1.0, 0.5))
    f = 0.2
    np.testing.assert_allclose(p["fy"], [-2 * f, 2 * f], rtol=1e-13)
    np.testing.assert_allclose(p["fx"], [0.0, 0.0], atol=1e-15)
    # tau_p = s (nx Fy - ny Fx): -1/2 * (-f) + 1 * (-f); tau_q = t (ny Fx - nx Fy) with
    # F the force on p: -1 * f + 1/2 * f
    np.testing.assert_allclose(p["tau"], [-0.5 * f, -0.5 * f], rtol=1e-13)
    assert s["E"] == pytest.approx(2 * f ** 2 / 2 / 2, rel=1e-13)
    assert s["ncontacts"] == 1


def test_one_contact_cases_match_closest_point_model():
    cases = [
        ((10.0, 10.0, 0.0, 1.0, 0.5), (10.0, 11.5, math.pi / 2, 0.8, 0.5)),  # T shape
        ((10.0, 10.0, 0.0, 1.0, 0.5), (12.0, 10.0, 0.0, 0.6, 0.5)),  # end to end
        ((10.0, 10.0, 0.3, 1.0, 0.5), (11.2, 10.9, 1.4, 0.7, 0.6)),  # skew
        ((10.0, 10.0, 0.0, 0.0, 0.5), (10.5, 10.6, 0.0, 0.0, 0.6)),  # discs
    ]
    for a, b in cases:
        p1, s1 = two_body(a, b, model=0)
        p2, s2 = two_body(a, b, model=1)
        for k in ("fx", "fy", "tau"):
            np.testing.assert_array_equal(p1[k], p2[k])
        assert s1["E"] == s2["E"]


def _random_pairs(n, rng, parallel_bias=True):
    for _ in range(n):
        A1, A2 = rng.uniform(0.3, 1.5, 2)
        R1, R2 = rng.uniform(0.4, 0.7, 2)
        ph1 = rng.uniform(0, math.pi)
        ph2 = ph1 + (rng.normal(0, 0.05) if parallel_bias else rng.uniform(0, math.pi))
        # centre offset: along the rod by up to its length, across by < sigma
        u = np.array([math.cos(ph1), math.sin(ph1)])
        v = np.array([-u[1], u[0]])
        off = u * rng.uniform(-1.2, 1.2) * A1 + v * rng.uniform(0.3, 0.95) * (R1 + R2)
        yield (10.0, 10.0, ph1, A1, R1), (10.0 + off[0], 10.0 + off[1], ph2, A2, R2)


def test_newton_and_angular_momentum_per_pair():
    rng = np.random.default_rng(3)
    for a, b in _random_pairs(200, rng):
        p, _ = two_body(a, b)
        scale = np.abs(p["fx"]).sum() + np.abs(p["fy"]).sum() + 1e-300
        assert abs(p["fx"].sum()) <= 1e-14 * scale and abs(p["fy"].sum()) <= 1e-14 * scale
        # sum of torques about the origin: tau_p + tau_q + r_p x F_p + r_q x F_q = 0
        Lz = p["tau"].sum() + np.sum(p["x"] * p["fy"] - p["y"] * p["fx"])
        assert abs(Lz) <= 1e-12 * (scale * 20 + np.abs(p["tau"]).sum())


def test_force_is_minus_n_grad_energy():
    rng = np.random.default_rng(5)
    h = 1e-6
    checked = 0
    for a, b in _random_pairs(60, rng):
        p, s = two_body(a, b)
        if s["E"] == 0:
            continue
        for comp, key in (("x", 0), ("y", 1), ("phi", 2)):
            e = []
            for sgn in (1, -1):
                aa = list(a)
                aa[key] += sgn * h
                e.append(two_body(tuple(aa), b)[1]["E"])
            grad = (e[0] - e[1]) / (2 * h)
            f = {0: p["fx"][0], 1: p["fy"][0], 2: p["tau"][0]}[key]
            # a finite difference across a switch of the second point's
            # definition is not a derivative: skip the rare straddling case
            if abs(-2 * grad - f) > 1e-6 * (abs(f) + 1e-3):
                e2 = []
                for sgn in (1, -1):
                    aa = list(a)
                    aa[key] += sgn * h / 10
                    e2.append(two_body(tuple(aa), b)[1]["E"])
                grad2 = (e2[0] - e2[1]) / (2 * h / 10)
                assert abs(grad2 - grad) > 1e-3 * abs(grad), (a, b, comp, -2 * grad, f)
                continue
            checked += 1
    assert checked > 100


def test_torque_is_continuous_at_alignment():
    # rotate q by +-eps about alignment with p (overlapping side by side)
    def tau_q(eps, model):
        p, _ = two_body((10.0, 10.0, 0.0, 1.0, 0.5), (10.3, 10.8, eps, 1.0, 0.5), model=model)
        return p["tau"][1]

    eps = 1e-7
    jump1 = abs(tau_q(eps, 0) - tau_q(-eps, 0))
    jump2 = abs(tau_q(eps, 1) - tau_q(-eps, 1))
    assert jump1 > 0.05  # closest point hops between the rod ends
    assert jump2 < 1e-5  # two-point torque varies continuously


@pytest.mark.parametrize("aspect,phi,seed", [(2.0, 0.83, 1), (4.0, 0.80, 2)])
def test_fire_converges_where_closest_point_stalls(aspect, phi, seed):
    R, A = configs.bidisperse(128, aspect=aspect)
    pk = configs.make_packing(R, A, phi, seed, "two-point fire")
    params = FIRE.copy()
    params[9] = 6000
    b1 = oracle.OracleBox.from_packing(pk, 0.1)
    it1, fr1 = b1.fire(params)
    b2 = oracle.OracleBox.from_packing(pk, 0.1)
    b2.set_contact_model(1)
    it2, fr2 = b2.fire(params)
    assert fr2 < 1e-10 and it2 < 6000
    assert fr1 > 1e-6
    # the relaxed two-point state is overlap-free here (below jamming)
    b2.calc_se()
    assert b2.scalars()["E"] < 1e-18
