"""Gamma-distributed aspect ratios (set_shapes with an aspect width,
src/cuda/spherocyl_box.cpp:1266-1300) on the GPU against the oracle.  The
cells are sized by the longest rod, so a cell holds many short ones: the
buckets grow past the 16 slots a monodisperse packing needs."""
import numpy as np
import pytest

import oracle
from conftest import force_tol, wrap_diff
from spherocyls_amd import configs, native

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pk():
    p = configs.gamma_aspect(2000, aspect_mean=2.0, rel_sigma=0.5, phi=0.84, seed=11)
    a = p.A / p.R
    assert a.max() > 3 * a.mean() / 2  # a long tail
    return p


def _oracle(pk, model=0, pad=0.1):
    b = oracle.OracleBox(pk.L, pk.x, pk.y, pk.phi, pk.A, pk.R, pad)
    b.set_contact_model(model)
    b.find_neighbors()
    return b


def test_crowded_cells_forces_and_stress(pk):
    print("gamma shapes: forces", pk.n, flush=True)
    s = native.System(pk, pads=0.1)
    try:
        s.calc_se()
        fx, fy, tau = s.forces()
        bx = s.box()
        nb = s.neighbors()
    finally:
        s.close()
    b = _oracle(pk)
    b.calc_se()
    p, o = b.particles(), b.scalars()
    # the occupancy the cell rule gives this packing exceeds 16 per cell
    assert pk.n / (bx["rows"] * bx["cols"]) > 8
    sc = b.force_scale()
    for a_, r_, c in zip((fx, fy, tau), (p["fx"], p["fy"], p["tau"]), sc):
        assert np.all(np.abs(a_ - r_) <= force_tol(c))
    for k in ("E", "Pxx", "Pyy", "Pxy", "cPxx", "cPyy"):
        assert bx[k] == pytest.approx(o[k], rel=1e-10, abs=1e-10 * abs(o["E"])), k
    on = b.neighbors()
    assert set(zip(*nb)) == set(zip(on[0].tolist(), on[1].tolist()))


def test_crowded_cells_heun_trajectory(pk):
    s = native.System(pk, pads=0.1)
    try:
        s.set_dynamics(strain_rate=1e-3, step=0.01)
        s.strain_steps(40)
        x, y, phi = s.positions()
    finally:
        s.close()
    b = _oracle(pk, pad=1.0)
    b.set_params(0.0, 0.0, 1e-3, 0.0, 0.01)
    b.calc_forces()
    b.run_steps(40, 0)
    p = b.particles()
    assert np.max(np.abs(wrap_diff(x, p["x"], pk.L))) < 1e-10
    assert np.max(np.abs(wrap_diff(phi, p["phi"], 2 * np.pi))) < 1e-10


def test_crowded_cells_fire_converges(pk):
    s = native.System(pk, pads=0.1)
    try:
        s.set_contact_model("two_point")
        it, fr, st = s.fire(native.fire_params(max_iter=40000), check=False)
        x, y, phi = s.positions()
    finally:
        s.close()
    print("gamma shapes: FIRE iterations", int(it[0]), "F_rms", float(fr[0]), flush=True)
    assert fr[0] < 1e-10
    b = oracle.OracleBox(pk.L, x, y, phi, pk.A, pk.R, 0.1)
    b.set_contact_model(1)
    b.find_neighbors()
    assert b.frms() < 1e-10
