"""Parity at the full sizes of the large-shear and compression workloads:
N = 262144 sheared Heun steps against the oracle's trajectory, and the
N = 16384 two-point quasi-static compression, each minimised state judged
by the oracle (residual, contact set, energy and stress)."""
import numpy as np
import pytest

import oracle
from spherocyls_amd import configs, native

pytestmark = pytest.mark.gpu


def wrap_diff(a, b, L):
    d = a - b
    return d - L * np.round(d / L)


def pairset(p, q):
    return set(zip(np.minimum(p, q).tolist(), np.maximum(p, q).tolist()))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_large_shear_trajectory_full_size(precision):
    """10 Heun steps (strain rate 1e-6, dt 0.01) of the N = 262144 packing
    (radii uniform in [0.5, 0.7], aspect 1.5, phi 0.845, from its random
    start) against the oracle's 10 steps: FP64 positions within 1e-10 and the
    stress/energy records within 1e-9 relative; mixed precision within 1e-6
    (positions) and 1e-5 (records), its error printed.  The oracle's steps
    start from forces it has evaluated, as run_strain does before its loop."""
    pk = configs.shear_large(1)
    print("full-size shear", precision, pk.n, flush=True)
    s = native.System(pk, pads=0.1)
    try:
        s.set_precision(precision)
        s.set_dynamics(strain_rate=1e-6, step=0.01)
        s.set_gamma(0.0, 0.0)
        se = s.strain_steps(10, 5)[:, 0, :4]
        x, y, phi = s.positions()
        gam = s.box()["gamma"]
    finally:
        s.close()
    ob = oracle.OracleBox.from_packing(pk, 0.1)
    ob.set_params(0.0, 0.0, 1e-6, 0.0, 0.01)
    ob.find_neighbors()
    ob.calc_forces()
    ose = ob.run_steps(10, 5)
    p = ob.particles()
    dx = np.max(np.abs(wrap_diff(x, p["x"], pk.L)))
    dy = np.max(np.abs(wrap_diff(y, p["y"], pk.L)))
    dp = np.max(np.abs(wrap_diff(phi, p["phi"], 2 * np.pi)))
    rel = np.max(np.abs(se - ose) / np.maximum(np.abs(ose), 1e-300))
    print(f"  max |dx| {dx:.3g} |dy| {dy:.3g} |dphi| {dp:.3g}; SE records max rel {rel:.3g}",
          flush=True)
    assert gam == pytest.approx(ob.scalars()["gamma"], abs=1e-15)
    tol_x, tol_se = (1e-10, 1e-9) if precision == "fp64" else (1e-6, 1e-5)
    assert dx < tol_x and dy < tol_x and dp < tol_x
    assert rel < tol_se


def test_compression_full_size_states_judged_by_the_oracle():
    """The N = 16384 aspect-2 compression start (phi 0.80): FIRE from the
    random start, then two quasi-static steps (x, y, L *= 1 - 1e-4, one
    unsheared Heun step, FIRE), two-point contacts.  After each minimisation
    the oracle, at the GPU's state, finds F_rms < 1e-10, the same contact set,
    and the energy and corrected stress the GPU reports.  The packing is
    unjammed here (pressure ~1e-12, asserted): energy and stress are sums of
    contact forces at the scale of the 1e-10 residual, which the order of
    summation alone moves by ~1e-7 relative, so they agree to 1e-6."""
    pk = configs.compress_start(1)
    params = native.fire_params(max_iter=200000)
    s = native.System(pk, pads=0.1)
    try:
        s.set_contact_model("two_point")
        s.set_dynamics(strain_rate=0.0, resize_rate=0.0, step=0.01)
        for step in range(3):
            if step > 0:
                s.affine_resize(np.array([1.0 - 1e-4]))
                s.strain_steps(1)
            it, fr, _ = s.fire(params)
            x, y, phi = s.positions()
            bx = s.box()
            cp, cq = s.contacts()
            print(f"compression step {step}: {int(it[0])} FIRE iterations, F_rms {fr[0]:.3g}, "
                  f"{len(cp)} contacts", flush=True)
            assert bx["Lx"] == pytest.approx(pk.L * (1 - 1e-4) ** step, rel=1e-14)
            ob = oracle.OracleBox(bx["Lx"], x, y, phi, pk.A, pk.R, 0.1)
            ob.set_contact_model(1)
            assert ob.frms() < 1e-10
            ob.calc_se()
            o = ob.scalars()
            op, oq = ob.contacts()
            assert pairset(cp, cq) == pairset(op, oq)
            scale = max(abs(o["cPxx"]), abs(o["cPyy"]))
            assert scale < 1e-9  # unjammed
            worst = max(abs(bx[k] - o[k]) / scale for k in ("cPxx", "cPyy", "cPxy"))
            print(f"  pressure scale {scale:.3g}, stress max diff / scale {worst:.3g}", flush=True)
            assert bx["E"] == pytest.approx(o["E"], rel=1e-6)
            assert worst < 1e-6
    finally:
        s.close()
