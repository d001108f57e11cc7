"""Text formats and output protocols against the reference's own files.

The golden files (tests/golden/files.json) were written by the reference's
run_strain / resize_box / gradient_descent_minimize on a 64-particle box
(tests/golden/make_golden.py).  The oracle reproduces the numbers bit for bit,
so writing them through the library's format entry points must reproduce the
files byte for byte; this pins the writers (sc_write_positions,
sc_format_se_line) that SpherocylBox uses, and the protocol's line cadence
(src/cpp/strain_dynamics.cpp:285-299, resize_dynamics.cpp:293-306,
relaxation_dynamics.cpp:41-83) as SpherocylBox implements it.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
import oracle
from spherocyls_amd import native

REF_PI = 3.14159265358979


@pytest.fixture(scope="module")
def files():
    with open(os.path.join(GOLDEN, "files.json")) as f:
        return json.load(f)


def _box(inp):
    return oracle.OracleBox(inp["L"], inp["x"], inp["y"], inp["phi"], inp["A"], inp["R"], inp["pad"])


def _write(tmp_path, name, b, strain_rate, step):
    s = b.scalars()
    p = b.particles()
    path = str(tmp_path / name)
    native.write_positions(path, s["Lx"], s["packing"], s["gamma"], s["gamma_total"], strain_rate,
                           step, p["x"], p["y"], p["phi"], p["R"], p["A"])
    return open(path).read()


def _se(vals, t, precision=6):
    return native.format_se_line(t, vals, precision)


def test_positions_file_bytes(files, tmp_path):
    inp = files["inputs"]
    b = _box(inp)
    b.set_params(0.0, 0.0, 1e-4, 0.0, 0.1)
    assert _write(tmp_path, "a.dat", b, 1e-4, 0.1) == files["files"]["strain/sp0000000000.dat"]
    b.calc_se()
    b.run_steps(20, 0)
    assert _write(tmp_path, "b.dat", b, 1e-4, 0.1) == files["files"]["strain/sp0000000002.dat"]


def test_read_positions_round_trip(files, tmp_path):
    path = tmp_path / "sp0000000002.dat"
    path.write_text(files["files"]["strain/sp0000000002.dat"])
    head, arr = native.read_positions(str(path))
    rows = [list(map(float, l.split())) for l in files["files"]["strain/sp0000000002.dat"].splitlines()]
    assert head["L"] == rows[0][1] and head["gamma"] == rows[0][3] and head["step"] == rows[0][6]
    body = np.array(rows[1:])
    for i, k in enumerate(("x", "y", "phi", "R", "A")):
        np.testing.assert_array_equal(arr[k], body[:, i])
    # writing what was read gives the same bytes
    out = tmp_path / "copy.dat"
    native.write_positions(str(out), head["L"], head["packing"], head["gamma"], head["gamma_total"],
                           head["strain_rate"], head["step"], arr["x"], arr["y"], arr["phi"],
                           arr["R"], arr["A"])
    assert out.read_text() == path.read_text()


def test_strain_se_file(files):
    """run_strain(2e-4, 0) at strain rate 1e-4, dt 0.1: 2 outer steps of 10
    sub-steps, an SE line after every sub-step (nSESaveT = 1) and the repeated
    line at the end of each outer step."""
    b = _box(files["inputs"])
    b.set_params(0.0, 0.0, 1e-4, 0.0, 0.1)
    b.calc_se()
    s = b.scalars()
    last = [s["E"], s["Pxx"], s["Pyy"], s["Pxy"]]
    lines = [_se(last, 0)]
    for t in (1, 2):
        rec = b.run_steps(10, 1)
        for r in rec:
            lines.append(_se(r, t))
        lines.append(_se(rec[-1], t))
    assert "".join(lines) == files["files"]["strain/sp_strain_se.dat"]


def test_resize_se_file(files):
    """resize_box(0.5004, 0) at rate 1e-4, dt 0.1: the reference's predictor
    (old box length, src/cpp/resize_dynamics.cpp:101), packing written first."""
    b = _box(files["inputs"])
    b.set_params(0.0, 0.0, 0.0, 1e-4, 0.1)
    b.calc_se()
    s = b.scalars()
    parea = s["packing"] * s["Lx"] * s["Ly"]
    lines = [_se([s["packing"], s["E"], s["Pxx"], s["Pyy"], s["Pxy"]], 0)]
    packing, t = s["packing"], 0
    while 0.5004 > packing:
        t += 1
        for _ in range(10):
            b.resize_step(shrink=True, corrected_image=False)
            b.calc_se()
            s = b.scalars()
            packing = parea / (s["Lx"] * s["Ly"])
            last = [packing, s["E"], s["Pxx"], s["Pyy"], s["Pxy"]]
            lines.append(_se(last, t))
        lines.append(_se(last, t))
    assert t == files["files"]["resize_ntime"]
    assert "".join(lines) == files["files"]["resize/sp_resize_se.dat"]


def test_relax_se_file(files):
    """gradient_descent_minimize(5, 0) with dt 0.1: a line per iteration before
    the step, then the final state as line nMax."""
    b = _box(files["inputs"])
    b.set_params(0.0, 0.0, 0.0, 0.0, 0.1)
    lines = []
    for s_ in range(5):
        b.calc_se()
        s = b.scalars()
        lines.append(_se([s["E"], s["Pxx"], s["Pyy"], s["Pxy"]], s_))
        b.gd_step()
    b.calc_se()
    s = b.scalars()
    lines.append(_se([s["E"], s["Pxx"], s["Pyy"], s["Pxy"]], 5))
    assert "".join(lines) == files["files"]["relax/sr_strain_se.dat"]
    assert sorted(k for k in files["files"] if k.startswith("relax/sr0")) == [
        "relax/sr0000000000.dat", "relax/sr0000000002.dat", "relax/sr0000000004.dat",
        "relax/sr0000000005.dat"]


def test_se_line_precision():
    assert _se([0.1234567891234, -2.5e-7], 3) == "3 0.123457 -2.5e-07\n"
    assert _se([0.1234567891234], 3, precision=14) == "3 0.1234567891234\n"
    assert _se([1.0, 0.0], 0) == "0 1 0\n"


@pytest.mark.parametrize("config", [0, 1, 2, 3, 4, 5])
def test_configuration_modes(config):
    n, R, A = 100, 0.5, 0.5
    L = np.sqrt(n * (4 * A + REF_PI * R) * R / 0.3)
    Ra, Aa = np.full(n, R), np.full(n, A)
    x, y, phi = native.random_configuration(L, Aa, Ra, seed=7, config=config)
    x2, _, _ = native.random_configuration(L, Aa, Ra, seed=7, config=config)
    np.testing.assert_array_equal(x, x2)
    assert np.all((x >= 0) & (x < L) & (y >= 0) & (y < L))
    if config % 2 == 1:
        assert np.all(phi == 0)
    b = oracle.OracleBox(L, x, y, phi, Aa, Ra, 0.1)
    cp, cq = b.brute_contacts()
    if config >= 2:  # zero-energy and grid modes: no contacts at all
        assert len(cp) == 0
    b.calc_se()
    if config >= 2:
        assert b.scalars()["E"] == 0.0


def test_random_packing_is_random_mode():
    R, A = np.full(50, 0.5), np.full(50, 0.5)
    a = native.random_packing(30.0, A, R, seed=3)
    b = native.random_configuration(30.0, A, R, seed=3, config=0)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)
