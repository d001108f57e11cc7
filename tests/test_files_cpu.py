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


def _state(n=37, seed=4):
    rng = np.random.default_rng(seed)
    return (rng.uniform(0, 20, n), rng.uniform(0, 20, n), rng.uniform(-3, 3, n),
            np.where(np.arange(n) % 2 == 0, 0.5, 0.7), rng.uniform(0, 1.5, n))


def test_gpu_code_text_positions_layout(tmp_path):
    """sd%010lu.dat of the GPU code's quasi-static protocols
    (src/cuda/strain_dynamics.cu:831-838): the printf formats of that writer,
    R at two significant digits."""
    x, y, phi, R, A = _state()
    head = (20.25, 20.25, 0.8123456789, 0.1, 1.25, 1e-5, 0.01)
    path = tmp_path / "sd0000000007.dat"
    native.write_positions_sd(str(path), *head, x, y, phi, R, A)
    want = "%d %.17g %.17g %f %.17g %.14g %g %g\n" % ((x.shape[0],) + head)
    want += "".join("%.17g %.17g %.17g %.2g %.17g\n" % v for v in zip(x, y, phi, R, A))
    assert path.read_text() == want
    # .17g round-trips doubles: the columns read back exactly (R is 0.5 / 0.7)
    rows = np.loadtxt(path, skiprows=1)
    for k, a in enumerate((x, y, phi, R, A)):
        np.testing.assert_array_equal(rows[:, k], a)


def test_binary_positions_layout_and_round_trip(tmp_path):
    """sd%010lu.bin (src/cuda/strain_dynamics.cu:845-875): a native int, seven
    doubles, then x y phi R A per particle."""
    x, y, phi, R, A = _state()
    head = (20.25, 20.5, 0.81, 0.1, 1.25, 1e-5, 0.01)
    path = str(tmp_path / "sd0000000007.bin")
    native.write_positions_bin(path, *head, x, y, phi, R, A)
    raw = open(path, "rb").read()
    assert len(raw) == 4 + 8 * (7 + 5 * x.shape[0])
    assert np.frombuffer(raw[:4], dtype=np.int32)[0] == x.shape[0]
    np.testing.assert_array_equal(np.frombuffer(raw[4:60], dtype=np.float64), head)
    recs = np.frombuffer(raw[60:], dtype=np.float64).reshape(-1, 5)
    np.testing.assert_array_equal(recs, np.column_stack([x, y, phi, R, A]))
    h, a = native.read_positions_bin(path)
    assert tuple(h[k] for k in native.SD_HEADER) == head
    for k, v in zip(("x", "y", "phi", "R", "A"), (x, y, phi, R, A)):
        np.testing.assert_array_equal(a[k], v)


def test_binary_positions_reader_rejects_bad_files(tmp_path):
    x, y, phi, R, A = _state(5)
    path = tmp_path / "a.bin"
    native.write_positions_bin(str(path), 10, 10, 0.5, 0, 0, 0, 0.01, x, y, phi, R, A)
    raw = path.read_bytes()
    for bad in (raw[:-8], raw + b"\0" * 8, raw[:30], b""):
        path.write_bytes(bad)
        with pytest.raises(native.SpherocylsError):
            native.read_positions_bin(str(path))
    with pytest.raises(native.SpherocylsError):
        native.read_positions_bin(str(tmp_path / "missing.bin"))
