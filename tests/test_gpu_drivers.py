"""The command-line drivers (SpherocylBox on the GPU) against the reference's
own output files (tests/golden/files.json, written by the reference's
run_strain / resize_box / gradient_descent_minimize on the same 64-particle
start).

Positions agree to the FP64 bar (1e-10, angles modulo 2 pi: the reference
re-wraps angles only when it rebuilds its neighbour list); stress/energy lines
are written with 6 significant digits, so their values agree to the printed
precision (relative 1.1e-5), and the line structure -- time stamps, repeated
end-of-step lines, file names -- is identical.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, wrap_diff
import oracle
from spherocyls_amd import native

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "spherocyls_amd", "bin")


@pytest.fixture(scope="module")
def files():
    with open(os.path.join(GOLDEN, "files.json")) as f:
        return json.load(f)


def _start_file(files, d, strain_rate, step):
    inp = files["inputs"]
    b = oracle.OracleBox(inp["L"], inp["x"], inp["y"], inp["phi"], inp["A"], inp["R"], inp["pad"])
    s = b.scalars()
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, "sp0000000000.dat")
    native.write_positions(path, s["Lx"], s["packing"], 0.0, 0.0, strain_rate, step,
                           np.array(inp["x"]), np.array(inp["y"]), np.array(inp["phi"]),
                           np.array(inp["R"]), np.array(inp["A"]))
    return path


def _run(exe, args, stdin=None):
    print("driver:", exe, " ".join(args), flush=True)
    r = subprocess.run([os.path.join(BIN, exe)] + ([] if stdin else args), input=stdin,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r


def _num(text):
    return [list(map(float, l.split())) for l in text.splitlines() if l.strip()]


def _cmp_se(ours, ref):
    a, b = _num(ours), _num(ref)
    assert len(a) == len(b)
    for la, lb in zip(a, b):
        assert la[0] == lb[0] and len(la) == len(lb)
        np.testing.assert_allclose(la[1:], lb[1:], rtol=1.1e-5, atol=1e-9)


def _cmp_pos(ours, ref):
    a, b = _num(ours), _num(ref)
    assert len(a) == len(b)
    np.testing.assert_allclose(a[0], b[0], rtol=1e-12, atol=1e-14)
    A, B = np.array(a[1:]), np.array(b[1:])
    L = b[0][1]
    assert np.max(np.abs(wrap_diff(A[:, 0], B[:, 0], L))) < 1e-10
    assert np.max(np.abs(wrap_diff(A[:, 1], B[:, 1], L))) < 1e-10
    assert np.max(np.abs(wrap_diff(A[:, 2], B[:, 2], 2 * 3.14159265358979))) < 1e-10
    np.testing.assert_array_equal(A[:, 3:], B[:, 3:])


def _compare_dir(d, files, prefix, skip=()):
    want = {k.split("/", 1)[1]: v for k, v in files["files"].items() if k.startswith(prefix + "/")}
    got = sorted(f for f in os.listdir(d) if f not in skip)
    assert got == sorted(want)
    for name, ref in want.items():
        ours = open(os.path.join(d, name)).read()
        if name.endswith("_se.dat"):
            _cmp_se(ours, ref)
        else:
            _cmp_pos(ours, ref)


def test_strain_driver(files, tmp_path):
    d = str(tmp_path / "out")
    start = _start_file(files, str(tmp_path / "in"), 1e-4, 0.1)
    os.makedirs(d)
    _run("spherocyl_strain", [start, d, "64", "1e-4", "0.1", "2e-4", "0.01", "1e-5", "0.1"])
    _compare_dir(d, files, "strain")


def test_strain_driver_restart(files, tmp_path):
    """Restart from sp0000000002.dat: the SE file keeps everything up to the
    first line of time 2 and continues with time 3 (the reference's seek,
    src/cpp/strain_dynamics.cpp:248-279)."""
    d = str(tmp_path / "out")
    start = _start_file(files, str(tmp_path / "in"), 1e-4, 0.1)
    os.makedirs(d)
    _run("spherocyl_strain", [start, d, "64", "1e-4", "0.1", "2e-4", "0.01", "1e-5", "0.1"])
    se_before = open(os.path.join(d, "sp_strain_se.dat")).read().splitlines()
    _run("spherocyl_strain", [os.path.join(d, "sp0000000002.dat"), d, "64", "1e-4", "0.1", "1e-4",
                              "0.01", "1e-5", "0.1"])
    se_after = open(os.path.join(d, "sp_strain_se.dat")).read().splitlines()
    first2 = next(i for i, l in enumerate(se_before) if l.startswith("2 "))
    assert se_after[:first2 + 1] == se_before[:first2 + 1]
    assert se_after[first2 + 1].startswith("3 ")
    assert os.path.exists(os.path.join(d, "sp0000000003.dat"))


def test_resize_driver_from_stdin(files, tmp_path):
    """Reference-compatible resize (predictor minimum image with the previous
    sub-step's box length, src/cpp/resize_dynamics.cpp:101).  That predictor
    depends on whether a coordinate was last wrapped before or after the
    rescale, an image shift of L (1 - (1-rate)^dt) = 1.8e-4 here, and the
    reference wraps only when it rebuilds its neighbour list
    (src/cpp/resize_dynamics.cpp:220-227): so the files agree in structure
    exactly and in value to that shift's effect.  The exact comparison is
    test_resize_driver_current_image."""
    d = str(tmp_path / "out")
    start = _start_file(files, str(tmp_path / "in"), 0.0, 0.1)
    os.makedirs(d)
    args = [start, d, "64", "1e-4", "0.1", "0.5004", "0.01", "1e-5", "0.1"]
    _run("spherocyl_resize", args, stdin="\n".join(args) + "\n")
    want = {k.split("/", 1)[1]: v for k, v in files["files"].items() if k.startswith("resize/")}
    assert sorted(os.listdir(d)) == sorted(want)
    ours = _num(open(os.path.join(d, "sp_resize_se.dat")).read())
    ref = _num(want["sp_resize_se.dat"])
    assert [l[0] for l in ours] == [l[0] for l in ref]
    o, r = np.array(ours), np.array(ref)
    np.testing.assert_allclose(o[:, 1], r[:, 1], rtol=1.1e-5)  # packing: box length only
    np.testing.assert_allclose(o[:, 2:], r[:, 2:], rtol=2e-2, atol=1e-4)
    a = np.array(_num(open(os.path.join(d, "sp0000000004.dat")).read())[1:])
    b = np.array(_num(want["sp0000000004.dat"])[1:])
    L = _num(want["sp0000000004.dat"])[0][1]
    assert np.max(np.abs(wrap_diff(a[:, 0], b[:, 0], L))) < 2e-3
    assert np.max(np.abs(wrap_diff(a[:, 1], b[:, 1], L))) < 2e-3


def _oracle_resize_rows(inp, pad):
    b = oracle.OracleBox(inp["L"], inp["x"], inp["y"], inp["phi"], inp["A"], inp["R"], pad)
    b.set_params(0.0, 0.0, 0.0, 1e-4, 0.1)
    b.calc_se()
    s = b.scalars()
    parea = s["packing"] * s["Lx"] * s["Ly"]
    rows = [[0, s["packing"], s["E"], s["Pxx"], s["Pyy"], s["Pxy"]]]
    packing, t = s["packing"], 0
    while 0.5004 > packing:
        t += 1
        for _ in range(10):
            b.resize_step(shrink=True, corrected_image=True)
            b.calc_se()
            s = b.scalars()
            packing = parea / (s["Lx"] * s["Ly"])
            rows.append([t, packing, s["E"], s["Pxx"], s["Pyy"], s["Pxy"]])
        rows.append(list(rows[-1]))
    return b, s, rows, t


def test_resize_driver_current_image(files, tmp_path):
    """Resize with the current box length in the predictor, against the
    oracle's same protocol, to the FP64 bar.  The oracle runs with a cell
    padding (2.5, 3 x 3 cells) whose rebuild threshold is never reached in
    these 40 sub-steps: a rebuild inside the reference's resize corrector loop
    wraps already-rescaled coordinates with the previous box length
    (src/cpp/resize_dynamics.cpp:220-227 calling find_cells,
    src/cpp/spherocyl_box.cpp:279-286), shifting them by L (1 - (1-rate)^dt),
    a trajectory-level quirk the GPU path does not have.  A second padding
    (2.0) must give the same oracle trajectory, which shows no rebuild
    happened in either."""
    inp = files["inputs"]
    b, s, rows, t = _oracle_resize_rows(inp, 2.5)
    _, _, rows2, _ = _oracle_resize_rows(inp, 2.0)
    assert rows2 == rows
    d = str(tmp_path / "out")
    start = _start_file(files, str(tmp_path / "in"), 0.0, 0.1)
    os.makedirs(d)
    env = dict(os.environ, SPHEROCYL_RESIZE_IMAGE="current")
    print("driver: spherocyl_resize current image", flush=True)
    r = subprocess.run([os.path.join(BIN, "spherocyl_resize"), start, d, "64", "1e-4", "0.1",
                        "0.5004", "0.01", "1e-5", "0.1"], capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    ours = _num(open(os.path.join(d, "sp_resize_se.dat")).read())
    assert [l[0] for l in ours] == [l[0] for l in rows]
    np.testing.assert_allclose(np.array(ours)[:, 1:], np.array(rows)[:, 1:], rtol=1.1e-5,
                               atol=1e-12)
    a = np.array(_num(open(os.path.join(d, "sp%010d.dat" % t)).read())[1:])
    p = b.particles()
    L = s["Lx"]
    assert np.max(np.abs(wrap_diff(a[:, 0], p["x"], L))) < 1e-10
    assert np.max(np.abs(wrap_diff(a[:, 1], p["y"], L))) < 1e-10
    assert np.max(np.abs(wrap_diff(a[:, 2], p["phi"], 2 * 3.14159265358979))) < 1e-10


def test_minimize_driver(files, tmp_path):
    d = str(tmp_path / "out")
    start = _start_file(files, str(tmp_path / "in"), 0.0, 0.1)
    os.makedirs(d)
    _run("spherocyl_minimize", [start, d, "64", "0.1", "5", "2", "1", "0.1"])
    _compare_dir(d, files, "relax")


def test_minimize_driver_fire(tmp_path):
    """FIRE through the driver: the final state's residual, evaluated by the
    oracle from the written file, is below 1e-10."""
    from spherocyls_amd import configs
    R, A = configs.bidisperse(256)
    pk = configs.make_packing(R, A, 0.70, 21, "fire driver")
    d = str(tmp_path / "out")
    os.makedirs(d)
    start = os.path.join(str(tmp_path), "sp0000000000.dat")
    native.write_positions(start, pk.L, 0.70, 0.0, 0.0, 0.0, 0.1, pk.x, pk.y, pk.phi, pk.R, pk.A)
    env = dict(os.environ, SPHEROCYL_MINIMIZER="fire")
    print("driver: spherocyl_minimize fire", flush=True)
    r = subprocess.run([os.path.join(BIN, "spherocyl_minimize"), start, d, "256", "0.1", "100", "1",
                        "1", "0.1"], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    out = [f for f in os.listdir(d) if f.startswith("sr") and f != "sr_strain_se.dat"]
    assert len(out) == 1
    head, arr = native.read_positions(os.path.join(d, out[0]))
    b = oracle.OracleBox(head["L"], arr["x"], arr["y"], arr["phi"], arr["A"], arr["R"], 0.1)
    b.calc_forces()
    # 14 printed digits perturb a relaxed state by ~1e-13 relative
    assert b.frms() < 1e-10 * 10


def _sd_file(path):
    with open(path) as f:
        head = f.readline().split()
    rows = np.loadtxt(path, skiprows=1, ndmin=2)
    return [float(v) for v in head], rows


@pytest.mark.parametrize("change", [1, 2])
def test_gpu_order_strain_driver(tmp_path, change):
    """spherocyl_box (argument order of the GPU code's src/cuda/main.cpp): a
    random start with gamma-distributed aspect ratios, the shear direction
    reversed (1) or rotated by gamma (2), then the GPU class's run_strain:
    SE lines every 10 sub-steps, sd position files every 10 time units, the
    recorded energy the oracle's on each saved state.  A restart from the
    file at t = 30 keeps the SE file up to its first line of time 30 (the
    GPU class's seek, src/cuda/strain_dynamics.cu:1018-1047) and continues
    from time 30 again with the values of the first run: to the printed
    digits at first (the restarted lists and storage order sum in another
    order), the same time stamps throughout."""
    d1, d2 = tmp_path / "a", tmp_path / "b"
    d1.mkdir()
    d2.mkdir()
    common = ["1e-3", "0.01"]
    _run("spherocyl_box", ["r", str(d1), "256", *common, "0.05", "0.01", "1e-4", "0.1",
                           str(change), "0.80", "1.0", "1.4", "0", "0.3"])
    lines = open(d1 / "sd_stress_energy.dat").read().splitlines()
    t = [int(l.split()[0]) for l in lines]
    assert t[0] == 0 and t[-1] == 50 and len(lines) == 50 * 10 + 1
    files = sorted(p.name for p in d1.glob("sd*.dat") if p.name != "sd_stress_energy.dat")
    assert files == ["sd%010d.dat" % k for k in range(0, 51, 10)]
    head, rows = _sd_file(d1 / "sd0000000030.dat")
    assert len(set(np.round(rows[:, 4] / rows[:, 3], 9))) > 100  # gamma-distributed aspects
    for k in (0, 30, 50):
        head, rows = _sd_file(d1 / ("sd%010d.dat" % k))
        b = oracle.OracleBox(head[1], rows[:, 0], rows[:, 1], rows[:, 2], rows[:, 4], rows[:, 3], 0.1)
        b.set_params(head[4], 0.0, 0.0, 0.0, 0.01)
        b.find_neighbors()
        b.calc_se()
        line = next(l for l in lines if int(l.split()[0]) == k)
        assert float(line.split()[1]) == pytest.approx(b.scalars()["E"], rel=2e-6), k
    # restart at t = 30 from the saved file into a copy of the SE file
    (d2 / "sd_stress_energy.dat").write_text("\n".join(lines) + "\n")
    _run("spherocyl_box", [str(d1 / "sd0000000030.dat"), str(d2), "256", *common, "0.02", "0.01",
                           "1e-4", "0.1", "0"])
    lines2 = open(d2 / "sd_stress_energy.dat").read().splitlines()
    i30 = next(i for i, l in enumerate(lines) if l.startswith("30 "))
    assert lines2[:i30 + 1] == lines[:i30 + 1]
    rest, rest2 = _num("\n".join(lines[i30:])), _num("\n".join(lines2[i30 + 1:]))
    assert len(rest2) == len(rest)
    assert [a[0] for a in rest] == [b[0] for b in rest2]
    # the first two time units agree to the printed digits; later the sheared
    # dynamics (strain rate 1e-3) amplifies the rounding of the other order
    for a, b in zip(rest[:20], rest2[:20]):
        np.testing.assert_allclose(b[1:], a[1:], rtol=0, atol=1e-6 * max(abs(v) for v in a[1:]))


def test_pure_shear_driver(tmp_path):
    """spherocyl_pureshear (argument order of src/cuda/main_pure_shear.cpp):
    pure shear with the two-point model from a random start until Lx / Ly
    reaches 1.05; the box aspect of every saved state grows, the last one is
    at the stop, and each recorded energy is the oracle's on the saved
    (non-square) state."""
    d = tmp_path / "p"
    d.mkdir()
    _run("spherocyl_pureshear", ["r", str(d), "128", "1e-3", "0.01", "1.05", "0.01", "1e-4", "0.1",
                                 "0", "0.80", "1.0", "1.4"])
    lines = open(d / "sd_stress_energy.dat").read().splitlines()
    files = sorted(p for p in d.glob("sd0*.dat"))
    assert len(files) >= 3
    aspects = []
    for path in files:
        head, rows = _sd_file(path)
        aspects.append(head[1] / head[2])
        k = int(path.name[2:12])
        b = oracle.OracleBox(head[1], rows[:, 0], rows[:, 1], rows[:, 2], rows[:, 4], rows[:, 3], 0.1)
        b.set_box(head[1], head[2])
        b.set_contact_model(1)
        b.set_params(head[4], 0.0, 0.0, 0.0, 0.01)
        b.find_neighbors()
        b.calc_se()
        line = next(l for l in lines if int(l.split()[0]) == k)
        e = b.scalars()["E"]
        assert float(line.split()[1]) == pytest.approx(e, rel=2e-6, abs=1e-12), k
    assert all(a2 > a1 for a1, a2 in zip(aspects, aspects[1:]))
    assert aspects[-1] >= 1.05 and aspects[-2] < 1.05
