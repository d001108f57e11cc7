"""Mixed precision (FP32 pair forces) on the configurations that need care:
nearly parallel rods, whose closest-approach parameters come from differences
of nearly equal products, and the aspect-4 packings where FP32 parameters
made FIRE bounce without converging."""
import numpy as np
import pytest

import oracle
from spherocyls_amd import configs, ensemble, native

pytestmark = pytest.mark.gpu


def nematic_lattice(seed=5):
    """272 rods (R 0.5, A 1.5: aspect 4) along x in 34 staggered rows of 8;
    neighbouring rows overlap by 5% of the contact distance along half their
    length.  Angles are +-10^U(-9, -2): nearly parallel side-by-side pairs at
    every scale of misalignment."""
    rng = np.random.default_rng(seed)
    nx, ny = 8, 34
    L = 8 * 4.05
    h = L / ny
    R = np.full(nx * ny, 0.5)
    A = np.full(nx * ny, 1.5)
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    x = (i * 4.05 + (j % 2) * 2.025 + 1.0 + rng.uniform(-0.01, 0.01, i.shape)).ravel() % L
    y = (j * h + 0.3 + rng.uniform(-0.005, 0.005, j.shape)).ravel() % L
    phi = (10.0 ** rng.uniform(-9, -2, nx * ny)) * rng.choice([-1.0, 1.0], nx * ny)
    return configs.Packing(L, x, y, phi, A, R, 0.0, "nematic lattice aspect 4")


@pytest.mark.parametrize("path", ["list", "tile"])
@pytest.mark.parametrize("model", ["closest", "two_point"])
def test_mixed_forces_on_nearly_parallel_rods(model, path):
    """Every side-by-side pair is nearly parallel.  FP32 closest-approach
    parameters called 1.5% of such pairs crossed (a force of 1/sigma instead
    of ~0.05/sigma) and put others' contacts anywhere along the overlap; with
    FP64 parameters the mixed forces stay within 1e-4 of each particle's
    contact-force scale, and the FP64 path stays at the 1e-10 bar."""
    pk = nematic_lattice()
    b = oracle.OracleBox(pk.L, pk.x, pk.y, pk.phi, pk.A, pk.R, 0.1)
    b.set_contact_model(1 if model == "two_point" else 0)
    b.find_neighbors()
    b.calc_se()
    p = b.particles()
    sfx, sfy, stau = b.force_scale()
    # per particle: sum of contact-force magnitudes, and that times the arm A
    # for the torque
    fs = sfx + sfy
    sc = (fs, fs, stau + pk.A * fs)
    assert b.scalars()["ncontacts"] >= pk.n  # every rod touches its neighbouring rows
    for prec, tol in (("fp64", 1e-10), ("mixed", 1e-4)):
        s = native.System(pk, pads=0.1)
        try:
            s.set_path(path)
            s.set_contact_model(model)
            s.set_precision(prec)
            s.calc_forces()
            got = s.forces()
        finally:
            s.close()
        worst = 0.0
        for name, a, r, c in zip(("fx", "fy", "tau"), got, (p["fx"], p["fy"], p["tau"]), sc):
            err = np.abs(a - r) / (2 * c + 1e-300)
            worst = max(worst, float(err.max()))
            assert np.all(err <= tol), (prec, name, float(err.max()))
        print(f"nearly parallel rods {model} {path} {prec}: worst error / force scale {worst:.2e}",
              flush=True)


def test_mixed_fire_converges_on_aspect4_packings():
    """Ensemble packings (aspect 4, phi 0.85, 2000 sheared Heun steps) on which
    mixed-mode FIRE with FP32 closest-approach parameters bounced at F_rms
    1e-4..1e-2 until max_iter; FP64 takes a few hundred iterations.  Mixed mode
    now converges to the same F_rms < 1e-10 in a comparable count."""
    pks = configs.ensemble(256)
    sub = [pks[53], pks[41]]
    work = ensemble.prepare(sub, heun_steps=2000, strain_rate=1e-4, dt=0.01)
    params = native.fire_params(max_iter=20000)
    its = {}
    for prec in ("fp64", "mixed"):
        s = ensemble.system_for(work, precision=prec, autotune=False)
        try:
            it, fr, st = s.fire(params, switch_frms=1e-6 if prec == "mixed" else 0.0, check=False)
        finally:
            s.close()
        print(prec, [w.name for w in work.packings], "iterations", it.tolist(), "F_rms",
              fr.tolist(), flush=True)
        assert np.all(fr < 1e-10), prec
        its[prec] = it
    assert np.all(its["mixed"] <= 2 * its["fp64"] + 200)
