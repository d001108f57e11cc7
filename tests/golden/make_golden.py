"""Write the golden vectors under tests/golden/ from the compiled reference.

Runs only where the reference tree exists (it builds oracle/_ref from the
reference's own src/cpp sources via oracle/Makefile and drives it through
oracle/ref_harness.cpp).  The outputs are data only: seeded inputs, the
reference's outputs at full precision, and the text files its run_strain /
resize_box / gradient_descent_minimize write.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from spherocyls_amd import configs  # noqa: E402


def state(box):
    p = box.particles()
    s = box.scalars()
    return {k: p[k] for k in ("x", "y", "phi", "fx", "fy", "tau")}, s


def save_case(name, pk, pad, gamma, out):
    arrays = dict(L=np.array([pk.L]), x=pk.x, y=pk.y, phi=pk.phi, A=pk.A, R=pk.R,
                  pad=np.array([pad]), gamma=np.array([gamma]))
    for k, v in out.items():
        arrays[k] = np.asarray(v)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)


def se_vec(s):
    return np.array([s["E"], s["Pxx"], s["Pyy"], s["Pxy"]])


def case_static(name, pk, pad, gamma):
    """calc_se at a fixed state: forces, torques, E, stress, neighbour set."""
    b = oracle.RefBox.from_packing(pk, pad)
    b.set_params(gamma, 0.0, 0.0, 0.0, 0.05)
    b.find_neighbors()  # neighbour criterion depends on gamma
    b.calc_se()
    st, sc = state(b)
    p, q = b.neighbors()
    save_case(name, pk, pad, gamma,
              dict(fx=st["fx"], fy=st["fy"], tau=st["tau"], se=se_vec(sc),
                   rows=np.array([sc["rows"]]), cols=np.array([sc["cols"]]),
                   eff_pad=np.array([sc["pad"]]), nbr_p=p, nbr_q=q,
                   wx=st["x"], wy=st["y"], wphi=st["phi"]))


def case_heun(name, pk, pad, gamma, strain_rate, dt, nsteps, se_every):
    """One Heun step, then nsteps more with SE records every se_every."""
    b = oracle.RefBox.from_packing(pk, pad)
    b.set_params(gamma, 0.0, strain_rate, 0.0, dt)
    b.calc_forces()
    b.strain_step()
    s1, sc1 = state(b)
    b.calc_forces()
    se = b.run_steps(nsteps, se_every)
    sN, scN = state(b)
    save_case(name, pk, pad, gamma,
              dict(strain_rate=np.array([strain_rate]), dt=np.array([dt]),
                   nsteps=np.array([nsteps]), se_every=np.array([se_every]),
                   x1=s1["x"], y1=s1["y"], phi1=s1["phi"], gamma1=np.array([sc1["gamma"]]),
                   se=se, xN=sN["x"], yN=sN["y"], phiN=sN["phi"], fxN=sN["fx"], fyN=sN["fy"],
                   tauN=sN["tau"], gammaN=np.array([scN["gamma"]]),
                   gamma_totalN=np.array([scN["gamma_total"]])))


def case_resize(name, pk, pad, rate, dt, nsteps):
    b = oracle.RefBox.from_packing(pk, pad)
    b.set_params(0.0, 0.0, 0.0, rate, dt)
    b.calc_forces()
    outs = {}
    for k in range(nsteps):
        b.resize_step(True)
        b.calc_forces()
    s, sc = state(b)
    outs = dict(rate=np.array([rate]), dt=np.array([dt]), nsteps=np.array([nsteps]),
                xN=s["x"], yN=s["y"], phiN=s["phi"], fxN=s["fx"], fyN=s["fy"], tauN=s["tau"],
                LN=np.array([sc["Lx"]]))
    save_case(name, pk, pad, 0.0, outs)


def case_gd(name, pk, pad, dt, nsteps):
    b = oracle.RefBox.from_packing(pk, pad)
    b.set_params(0.0, 0.0, 0.0, 0.0, dt)
    energies = []
    for k in range(nsteps):
        b.calc_se()
        energies.append(b.scalars()["E"])
        b.gd_step()
    b.calc_se()
    s, sc = state(b)
    save_case(name, pk, pad, 0.0,
              dict(dt=np.array([dt]), nsteps=np.array([nsteps]), energies=np.array(energies),
                   xN=s["x"], yN=s["y"], phiN=s["phi"], fxN=s["fx"], fyN=s["fy"], tauN=s["tau"],
                   se=se_vec(sc)))


def case_files():
    """Text output of the three protocol entry points on a small box."""
    R, A = configs.bidisperse(64)
    pk = configs.make_packing(R, A, 0.5, 11, "files N=64")
    files = {}
    with tempfile.TemporaryDirectory() as d:
        b = oracle.RefBox.from_packing(pk, 0.1)
        b.set_params(0.0, 0.0, 1e-4, 0.0, 0.1)
        b.calc_se()
        t = b.run_strain(2e-4, 0, d, "sp_strain_se.dat")
        for f in sorted(os.listdir(d)):
            files["strain/" + f] = open(os.path.join(d, f)).read()
        files["strain_ntime"] = t
    with tempfile.TemporaryDirectory() as d:
        b = oracle.RefBox.from_packing(pk, 0.1)
        # rate * dt must not exceed the default SE save step (1e-5), or the
        # reference's nSESaveT rounds to 0 (src/cpp/resize_dynamics.cpp:286, 299)
        b.set_params(0.0, 0.0, 0.0, 1e-4, 0.1)
        b.calc_se()
        t = b.resize_box(0.5004, 0, d, "sp_resize_se.dat")
        for f in sorted(os.listdir(d)):
            files["resize/" + f] = open(os.path.join(d, f)).read()
        files["resize_ntime"] = t
    with tempfile.TemporaryDirectory() as d:
        b = oracle.RefBox.from_packing(pk, 0.1)
        b.set_params(0.0, 0.0, 0.0, 0.0, 0.1)
        b.calc_se()
        b.gd_minimize(5, 0.0, d, "sr_strain_se.dat", pos_step=2.0)
        for f in sorted(os.listdir(d)):
            files["relax/" + f] = open(os.path.join(d, f)).read()
    inputs = dict(L=pk.L, x=pk.x.tolist(), y=pk.y.tolist(), phi=pk.phi.tolist(),
                  A=pk.A.tolist(), R=pk.R.tolist(), pad=0.1)
    with open(os.path.join(HERE, "files.json"), "w") as f:
        json.dump(dict(inputs=inputs, files=files), f, indent=0)


def main():
    oracle.build(ref=True)
    if not oracle.ref_available():
        raise SystemExit("reference library not built")
    shear = configs.shear_small(1)
    case_static("static_shear512", shear, 0.1, 0.0)
    R, A = configs.bidisperse(512)
    case_static("static_shear512_gamma", configs.make_packing(R, A, 0.85, 2, gamma=0.37), 0.1, 0.37)
    R, A = configs.bidisperse(96, aspect=4.0)
    case_static("static_aspect4", configs.make_packing(R, A, 0.82, 3, "aspect 4", gamma=-0.21), 0.1, -0.21)
    R = np.where(np.arange(48) % 2 == 0, 0.5, 0.7)
    case_static("static_discs", configs.make_packing(R, 0.0 * R, 0.8, 5, "discs", gamma=0.1), 0.1, 0.1)
    R, A = configs.bidisperse(200, aspect=1.2)
    case_static("static_aspect1p2", configs.make_packing(R, A, 0.86, 7, "aspect 1.2", gamma=0.45), 0.1, 0.45)
    case_heun("heun_shear512", shear, 0.1, 0.0, 1e-5, 0.01, 100, 10)
    # large strain rate so the Lees-Edwards remap happens inside the run
    R, A = configs.bidisperse(128)
    small = configs.make_packing(R, A, 0.84, 9, "heun LE", gamma=0.45)
    case_heun("heun_lees_edwards", small, 0.1, 0.45, 0.1, 0.01, 80, 5)
    R, A = configs.bidisperse(128)
    case_resize("resize128", configs.make_packing(R, A, 0.80, 13, "resize"), 0.1, 1e-3, 0.05, 20)
    case_gd("gd128", configs.make_packing(R, A, 0.85, 17, "gd"), 0.1, 0.05, 25)
    case_files()
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
