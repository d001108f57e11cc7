"""Edge cases of the batched system: a packing of one particle and one of two
particles, each in a box of a single cell (every neighbour cell of the 3 x 3 block
is the same cell), batched with an ordinary packing.  The lone particle feels
no force and turns by the isolated rotation alone; the pair's contact is the
one the oracle's pair routine gives in an open box (counted once: the device
grid never visits a cell twice); the large packing matches the oracle as it
does on its own; FIRE brings all three to F_rms < 1e-10."""
import numpy as np
import pytest

import oracle
from spherocyls_amd import configs, native

pytestmark = pytest.mark.gpu
REF_PI = 3.14159265358979


def _packs():
    R, A = configs.bidisperse(800, aspect=2.0)
    big = configs.make_packing(R, A, 0.84, 81, "edge N=800")
    one = configs.Packing(3.0, np.array([1.0]), np.array([1.2]), np.array([0.3]),
                          np.array([0.5]), np.array([0.5]), 0.0, "edge N=1")
    # side 4.4: a single cell, and still more than twice the largest contact
    # reach, so the minimum image is the only image in contact
    two = configs.Packing(4.4, np.array([1.0, 1.6]), np.array([1.0, 1.3]), np.array([0.2, 1.4]),
                          np.array([0.5, 0.35]), np.array([0.5, 0.6]), 0.0, "edge N=2")
    return one, two, big


@pytest.mark.parametrize("model", ["closest", "two_point"])
def test_one_and_two_particle_packings_in_a_batch(model):
    one, two, big = _packs()
    code = 1 if model == "two_point" else 0
    print("edge batch", model, flush=True)
    s = native.System([one, two, big], pads=0.1)
    try:
        s.set_contact_model(model)
        for k in (0, 1):
            assert s.box(k)["rows"] == 1 and s.box(k)["cols"] == 1
        s.calc_se()
        fx, fy, tau = s.forces()
        boxes = [s.box(k) for k in range(3)]
    finally:
        s.close()
    # the lone particle
    assert fx[0] == 0.0 and fy[0] == 0.0 and tau[0] == 0.0 and boxes[0]["E"] == 0.0
    # the pair: the oracle's contacts of (p, q) in a box of side 3, p first
    p = (two.x[0], two.y[0], two.phi[0], two.A[0], two.R[0])
    q = (two.x[1], two.y[1], two.phi[1], two.A[1], two.R[1])
    cs = oracle.pair_contacts(p, q, model=code, L=4.4)
    assert len(cs) >= 1 and boxes[1]["ncontacts"] == 1  # contacts count pairs
    want_f = (sum(c["fx"] for c in cs), sum(c["fy"] for c in cs))
    want_t = (sum(c["tau_p"] for c in cs), sum(c["tau_q"] for c in cs))
    scale = max(abs(want_f[0]), abs(want_f[1]))
    np.testing.assert_allclose([fx[1], fy[1]], want_f, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose([fx[2], fy[2]], [-want_f[0], -want_f[1]], rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose([tau[1], tau[2]], want_t, rtol=0, atol=1e-12 * scale)
    # the large packing, as on its own
    b = oracle.OracleBox.from_packing(big, 0.1)
    b.set_contact_model(code)
    b.find_neighbors()
    b.calc_se()
    o = b.particles()
    sx, sy, st = b.force_scale()
    sl = slice(3, 3 + big.n)
    for got, ref, sc in ((fx[sl], o["fx"], sx), (fy[sl], o["fy"], sy), (tau[sl], o["tau"], st)):
        assert np.all(np.abs(got - ref) <= 1e-10 * sc + 1e-15)
    assert boxes[2]["E"] == pytest.approx(b.scalars()["E"], rel=1e-10)


def test_lone_particle_heun_and_batch_fire():
    one, two, big = _packs()
    s = native.System([one, two, big], pads=0.1)
    try:
        s.set_contact_model("two_point")
        s.set_dynamics(strain_rate=1e-3, step=0.01)
        s.strain_steps(50)
        x, y, phi = s.positions()
        # the lone particle: the oracle's Heun steps (isolated rotation only,
        # no affine drift in sheared coordinates)
        b = oracle.OracleBox(one.L, one.x, one.y, one.phi, one.A, one.R, 0.1)
        b.set_params(0.0, 0.0, 1e-3, 0.0, 0.01)
        b.calc_forces()
        b.run_steps(50, 0)
        o = b.particles()
        assert x[0] == pytest.approx(o["x"][0], abs=1e-14)
        assert y[0] == pytest.approx(o["y"][0], abs=1e-14)
        dphi = (phi[0] - o["phi"][0] + REF_PI) % (2 * REF_PI) - REF_PI
        assert abs(dphi) < 1e-12
        it, fr, st = s.fire(native.fire_params(max_iter=20000), check=False)
        print("FIRE iterations", it, "F_rms", fr, flush=True)
        assert np.all(fr < 1e-10)
    finally:
        s.close()
