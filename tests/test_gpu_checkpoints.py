"""The long-shear protocol with minimisation checkpoints, on a small packing:
the stress and energy of the mixed-precision trajectory against the FP64
trajectory from the same start, and every checkpoint minimised to
F_rms < 1e-10 (with and without the learned warm start)."""
import pytest

from spherocyls_amd import checkpoints, configs

pytestmark = pytest.mark.gpu


def test_shear_with_checkpoints_small():
    pk = configs.shear_large(3, n=4096)
    out = checkpoints.run(pk, checkpoints=2, every=200, precision="mixed")
    assert out["steps"] == 400 and out["n_particles"] == 4096
    assert out["particle_steps_per_s"] > 0
    err = out["se_rel_error_vs_fp64_trajectory_max"]
    # FP32 pair forces in the dynamics: the trajectories part slowly at this
    # strain rate; the records themselves are FP64 evaluations
    spread = out["se_rel_spread_fp64_from_kicked_start_max"]
    print("mixed vs FP64 trajectory, relative:", err, flush=True)
    print("FP64 from a start displaced by 1e-12, relative:", spread, flush=True)
    for k in ("E", "Pxx", "Pyy", "Pxy_vs_pressure"):
        assert err[k] < 1e-4, (k, err[k])
        assert spread[k] < 1e-4, (k, spread[k])
    assert all(r["se_rel_spread_fp64_from_kicked_start"]["E"] > 0 for r in out["records"])
    assert out["minimisations_converged"] == 2
    for r in out["records"]:
        assert r["min_frms"] < 1e-10 and r["plain_converged"]
        assert r["step"] in (200, 400)
