"""Pure shear and the shear-direction flip on the GPU against the oracle
(tests/test_pure_shear_cpu.py pins the oracle's restatement)."""
import math

import numpy as np
import pytest

import oracle
from spherocyls_amd import configs, native

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _wrapdiff(a, b, L):
    d = a - b
    return np.max(np.abs(d - L * np.round(d / L)))


@pytest.mark.parametrize("model", ["closest", "two_point"])
def test_pure_shear_trajectory(model):
    R, A = configs.bidisperse(400)
    pk = configs.make_packing(R, A, 0.85, 17, "pure shear aspect 2 phi 0.85")
    rate, dt, nsteps = 2e-2, 0.01, 60
    print("pure shear", model, pk.n, "steps", nsteps, flush=True)
    s = native.System(pk, pads=0.1)
    try:
        s.set_path("list")
        s.set_contact_model(model)
        s.set_dynamics(strain_rate=rate, step=dt)
        se = s.pure_strain_steps(nsteps, se_every=20)
        x, y, phi = s.positions()
        fx, fy, tau = s.forces()
        bx = s.box()
    finally:
        s.close()
    b = oracle.OracleBox(pk.L, pk.x, pk.y, pk.phi, pk.A, pk.R, 0.1)
    b.set_contact_model(1 if model == "two_point" else 0)
    b.set_params(0.0, 0.0, rate, 0.0, dt)
    b.find_neighbors()
    for _ in range(nsteps):
        b.pure_step()
    b.calc_se()
    o, p = b.scalars(), b.particles()
    assert bx["Lx"] == pytest.approx(o["Lx"], rel=1e-14)
    assert bx["Ly"] == pytest.approx(o["Ly"], rel=1e-14)
    assert bx["Lx"] > pk.L * (1 + 0.4 * rate * dt * nsteps)  # the box did stretch
    assert bx["gamma_total"] == pytest.approx(o["gamma_total"], rel=1e-12)
    assert _wrapdiff(x, p["x"], o["Lx"]) < TOL
    assert _wrapdiff(y, p["y"], o["Ly"]) < TOL
    assert _wrapdiff(phi, p["phi"], 2 * math.pi) < TOL
    sc = b.force_scale()
    for a_, r_, c in zip((fx, fy, tau), (p["fx"], p["fy"], p["tau"]), sc):
        assert np.all(np.abs(a_ - r_) <= TOL * 2 * c + 1e-15)
    for k in ("E", "cPxx", "cPyy"):
        assert bx[k] == pytest.approx(o[k], rel=1e-9), k
    assert se.shape == (nsteps // 20, 1, 9)


def test_pure_shear_needs_the_list_path():
    R, A = configs.bidisperse(256)
    pk = configs.make_packing(R, A, 0.8, 3, "pure shear tile")
    s = native.System(pk, pads=0.1)
    try:
        s.set_path("tile")
        s.set_dynamics(strain_rate=1e-3, step=0.01)
        with pytest.raises(native.SpherocylsError):
            s.pure_strain_steps(1)
    finally:
        s.close()


def test_flip_shear_matches_the_oracle():
    R, A = configs.bidisperse(300)
    pk = configs.make_packing(R, A, 0.86, 23, "flip", gamma=0.15)
    s = native.System(pk, pads=0.1)
    try:
        s.set_gamma(0.15, 0.4)
        s.flip_shear()
        s.calc_se()
        x, y, phi = s.positions()
        fx, fy, tau = s.forces()
        bx = s.box()
    finally:
        s.close()
    b = oracle.OracleBox(pk.L, pk.x, pk.y, pk.phi, pk.A, pk.R, 0.1)
    b.set_params(0.15, 0.4, 0.0, 0.0, 0.05)
    b.flip_shear()
    b.calc_se()
    o, p = b.scalars(), b.particles()
    assert bx["gamma"] == -0.15 and o["gamma"] == -0.15
    assert bx["gamma_total"] == 0.4
    assert _wrapdiff(x, p["x"], pk.L) < 1e-12
    assert _wrapdiff(phi, p["phi"], 2 * math.pi) < 1e-12
    sc = b.force_scale()
    for a_, r_, c in zip((fx, fy, tau), (p["fx"], p["fy"], p["tau"]), sc):
        assert np.all(np.abs(a_ - r_) <= TOL * 2 * c + 1e-15)
    assert bx["cPxy"] == pytest.approx(o["cPxy"], rel=1e-9, abs=1e-12 * abs(o["cPxx"]))


def _sd(path):
    with open(path) as f:
        head = [float(v) for v in f.readline().split()]
    rows = np.loadtxt(path, skiprows=1, ndmin=2)
    return head, rows


def test_box_run_pure_strain_and_flip(tmp_path):
    """SpherocylBox::run_pure_strain and flip_shear_direction through a small
    C++ program (tests/cpp/box_pure_shear.cpp): the final and flipped states
    in its position files agree with the oracle's pure-shear steps and flip
    from the same start, and the SE file has the reference's cadence."""
    import os
    import subprocess
    from conftest import ROOT
    pkg = os.path.join(ROOT, "spherocyls_amd")
    exe = str(tmp_path / "box_pure_shear")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "box_pure_shear.cpp"), "-L", pkg,
                    "-lspherocyls", f"-Wl,-rpath,{pkg}", "-o", exe], check=True)
    rate, dt = 0.05, 0.01
    r = subprocess.run([exe, str(tmp_path), "128", "0.8", "5", str(rate), str(dt), "1.05"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    t = int(r.stdout.split()[0])
    assert t == 1  # Lx / Ly passes 1.05 within the first 100 steps
    lines = open(tmp_path / "pure_se.dat").read().splitlines()
    assert [int(l.split()[0]) for l in lines] == [0] * 10 + [1]
    (n, Lx0, Ly0, *_), start = _sd(tmp_path / "sd0000000999.dat")
    b = oracle.OracleBox(Lx0, start[:, 0], start[:, 1], start[:, 2], start[:, 4], start[:, 3], 0.1)
    b.set_contact_model(1)
    b.set_params(0.0, 0.0, rate, 0.0, dt)
    b.find_neighbors()
    b.calc_se()
    e0 = b.scalars()["E"]
    assert float(lines[0].split()[1]) == pytest.approx(e0, rel=1e-6)
    for _ in range(100):
        b.pure_step()
    (_, Lx1, Ly1, *_), fin = _sd(tmp_path / "sd0000000001.dat")
    o = b.scalars()
    assert Lx1 == pytest.approx(o["Lx"], rel=1e-15) and Ly1 == pytest.approx(o["Ly"], rel=1e-15)
    assert Lx1 / Ly1 >= 1.05
    p = b.particles()
    assert _wrapdiff(fin[:, 0], p["x"], Lx1) < TOL
    assert _wrapdiff(fin[:, 1], p["y"], Ly1) < TOL
    assert _wrapdiff(fin[:, 2], p["phi"], 2 * math.pi) < TOL
    b.calc_se()
    assert float(lines[-1].split()[1]) == pytest.approx(b.scalars()["E"], rel=1e-6)
    b.flip_shear()
    _, flipped = _sd(tmp_path / "sd0000001000.dat")
    p = b.particles()
    assert _wrapdiff(flipped[:, 0], p["x"], Lx1) < TOL
    assert _wrapdiff(flipped[:, 2], p["phi"], 2 * math.pi) < TOL


def test_rotate_by_gamma_then_shear_through_the_remap():
    """rotate_by_gamma leaves a box with Lx / Ly = 1 + gamma^2; Heun shear steps
    then cross the general remap strain Lx / (2 Ly).  Positions, forces and
    stress against the oracle after the rotation and after the steps."""
    R, A = configs.bidisperse(300)
    gamma = 0.3
    pk = configs.make_packing(R, A, 0.85, 29, "rotate", gamma=gamma)
    rate, dt, nsteps = 2.0, 0.01, 30
    print("rotate_by_gamma + shear", pk.n, flush=True)
    s = native.System(pk, pads=0.1)
    try:
        s.set_gamma(gamma, 0.0)
        s.rotate_by_gamma()
        s.calc_se()
        x0, y0, p0 = s.positions()
        f0 = s.forces()
        b0 = s.box()
        # shear from just below the remap strain, through it
        r = b0["Lx"] / b0["Ly"]
        g1 = 0.5 * r - 12 * rate * dt
        s.set_gamma(g1, 0.0)
        s.set_dynamics(strain_rate=rate, step=dt)
        s.strain_steps(nsteps)
        x1, y1, p1 = s.positions()
        f1 = s.forces()
        b1 = s.box()
    finally:
        s.close()
    b = oracle.OracleBox(pk.L, pk.x, pk.y, pk.phi, pk.A, pk.R, 0.1)
    b.set_params(gamma, 0.0, 0.0, 0.0, 0.05)
    b.rotate_by_gamma()
    b.calc_se()
    o, p = b.scalars(), b.particles()
    assert b0["Lx"] == pytest.approx(o["Lx"], rel=1e-15) and b0["gamma"] == -gamma
    assert _wrapdiff(x0, p["x"], o["Lx"]) < 1e-12 and _wrapdiff(y0, p["y"], o["Ly"]) < 1e-12
    assert _wrapdiff(p0, p["phi"], 2 * math.pi) < 1e-12
    sc = b.force_scale()
    for a_, r_, c in zip(f0, (p["fx"], p["fy"], p["tau"]), sc):
        assert np.all(np.abs(a_ - r_) <= TOL * 2 * c + 1e-15)
    for k in ("E", "cPxx", "cPyy", "cPxy"):
        assert b0[k] == pytest.approx(o[k], rel=1e-9, abs=1e-12 * abs(o["cPxx"])), k
    b.set_params(g1, 0.0, rate, 0.0, dt)
    b.find_neighbors()
    b.calc_forces()
    b.run_steps(nsteps, 0)
    b.calc_se()
    o, p = b.scalars(), b.particles()
    assert o["gamma"] < 0 and b1["gamma"] == pytest.approx(o["gamma"], abs=1e-14)  # remapped
    assert _wrapdiff(x1, p["x"], o["Lx"]) < TOL and _wrapdiff(y1, p["y"], o["Ly"]) < TOL
    assert _wrapdiff(p1, p["phi"], 2 * math.pi) < TOL
    sc = b.force_scale()
    for a_, r_, c in zip(f1, (p["fx"], p["fy"], p["tau"]), sc):
        assert np.all(np.abs(a_ - r_) <= TOL * 2 * c + 1e-15)


def test_resize_with_a_rate_per_packing():
    """Two packings with different resize rates and steps in one system: each
    box and trajectory against its own oracle run."""
    R, A = configs.bidisperse(128)
    pks = [configs.make_packing(R, A, 0.84, 31, "a"), configs.make_packing(R, A, 0.84, 32, "b")]
    rates, dts, nsteps = (2e-3, 5e-3), (0.01, 0.02), 25
    s = native.System(pks, pads=0.1)
    try:
        for k in range(2):
            s.set_dynamics(resize_rate=rates[k], step=dts[k], k=k)
        s.resize_steps(nsteps, shrink=True, corrected_image=True)
        x, y, phi = s.positions()
        boxes = [s.box(k) for k in range(2)]
    finally:
        s.close()
    for k, pk in enumerate(pks):
        ob = oracle.OracleBox(pk.L, pk.x, pk.y, pk.phi, pk.A, pk.R, 1.5)
        ob.set_params(0.0, 0.0, 0.0, rates[k], dts[k])
        ob.calc_forces()
        for _ in range(nsteps):
            ob.resize_step(shrink=True, corrected_image=True)
            ob.calc_forces()
        p, L = ob.particles(), ob.scalars()["Lx"]
        assert boxes[k]["Lx"] == L, k
        sl = slice(k * pk.n, (k + 1) * pk.n)
        assert _wrapdiff(x[sl], p["x"], L) < 1e-11 and _wrapdiff(y[sl], p["y"], L) < 1e-11, k
