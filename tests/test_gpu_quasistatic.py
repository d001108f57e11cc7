"""Quasi-static protocols (compress_qs2p, find_jam_2p of the reference's GPU
code, src/cuda/2point_dynamics.cu:4300-4606) through the command-line drivers.

The reference's GPU code cannot run (it reads an unallocated buffer), so the
files are checked against what pins them: the affine step rule fixes every
packing fraction; the oracle, evaluated at the state in each step's position
file (written at 17 digits), must find a minimum (F_rms < 1e-10) with the
energy and stress printed on that step's SE line; and find_jam_2p must end
on the target energy.
"""
import glob
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
import oracle
from spherocyls_amd import native

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "spherocyls_amd", "bin")
REF_PI = 3.14159265358979


def _run(exe, args):
    print("driver:", exe, " ".join(args), flush=True)
    r = subprocess.run([os.path.join(BIN, exe)] + args, capture_output=True, text=True,
                       timeout=300, env=dict(os.environ, SPHEROCYL_SEED="3"))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r


def _sd(path):
    with open(path) as f:
        head = f.readline().split()
    rows = np.loadtxt(path, skiprows=1, ndmin=2)
    return int(head[0]), float(head[1]), float(head[2]), float(head[3]), rows


def _oracle_at(path):
    n, Lx, Ly, packing, rows = _sd(path)
    assert rows.shape == (n, 5) and Lx == Ly
    b = oracle.OracleBox(Lx, rows[:, 0], rows[:, 1], rows[:, 2], rows[:, 4], rows[:, 3], 0.1)
    b.set_contact_model(1)
    b.find_neighbors()
    b.calc_se()
    return b, packing, rows


def _check_line(line, b, packing, digits):
    v = [float(t) for t in line.split()]
    s = b.scalars()
    assert v[1] == pytest.approx(packing, abs=10.0 ** -digits)
    if s["E"] < 1e-14:
        # unjammed: energy and stress are rounding-level remainders of forces
        # that cancel; both sides must be at that level
        assert abs(v[2]) < 1e-14 and max(abs(x) for x in v[3:]) < 1e-7
        return
    for got, key in zip(v[2:], ("E", "cPxx", "cPyy", "cPxy", "cPyx")):
        scale = max(abs(s["cPxx"]), abs(s["cPyy"])) if key != "E" else abs(s["E"])
        # %g prints 6 significant digits
        assert got == pytest.approx(s[key], rel=1e-5, abs=1e-5 * scale), key


def test_compress_qs2p_files(tmp_path):
    eps, phi0, phi1, n = 1e-3, 0.84, 0.86, 128
    _run("spherocyl_qscompress", ["r", str(tmp_path), str(n), str(eps), "0.01", str(phi1), "0.01",
                                  "0.01", "0.1", str(phi0), "1", "1.4"])
    lines = open(tmp_path / "sd_stress_energy.dat").read().splitlines()
    files = sorted(glob.glob(str(tmp_path / "sd*.dat")))
    files = [f for f in files if not f.endswith("sd_stress_energy.dat")]
    steps = len(lines) - 1
    # packing grows by (1 - eps)^-2 per step until it reaches phi1
    assert steps == int(np.ceil(np.log(phi1 / phi0) / (-2 * np.log1p(-eps)) - 1e-9))
    assert len(files) == steps + 1
    assert lines[0].split()[0] == "0" and lines[-1].split()[0] == str(steps)
    pressures = []
    for t, (line, path) in enumerate(zip(lines, files)):
        assert os.path.basename(path) == "sd%010d.dat" % t
        b, packing, rows = _oracle_at(path)
        want = phi0 / (1 - eps) ** (2 * t)
        assert packing == pytest.approx(want, abs=1e-6)  # printed with %f
        assert b.frms() < 1e-10, t  # every state is a minimum, judged by the oracle
        _check_line(line, b, want, 6 if t == 0 else 9)
        s = b.scalars()
        pressures.append(0.5 * (s["cPxx"] + s["cPyy"]))
    print("compress_qs2p: steps", steps, "pressure", pressures[0], "->", pressures[-1], flush=True)
    # the first line prints the packing with %f, later ones with %.9f
    assert len(lines[0].split()[1].split(".")[1]) == 6
    assert len(lines[1].split()[1].split(".")[1]) == 9


def _check_search(lines, phi0, target, s0, factor, grow_after):
    """Replay the bracketing search of resize_to_energy_{1p,2p} on the SE
    lines and check every line's packing.  A search compresses while
    E < target (expands while E > target) with steps s, 1.1 s, ...; the next
    search starts from s 1.1^(steps + grow_after) / 2 in the other direction.
    factor(compress, s) is one step's packing ratio; each line is checked
    against the line before it, so the tolerance is the printed precision.
    Searches are told apart by the direction of the packing's change, so the
    replay stops where a step no longer shows in the printed digits; the
    energies (printed with 6 digits) must lie on the search's side of the
    target until its last line, which crosses it."""
    digits = len(lines[0].split()[1].split(".")[1])
    res = 10.0 ** -digits
    t = [int(l.split()[0]) for l in lines]
    phi = [float(l.split()[1]) for l in lines]
    E = [float(l.split()[2]) for l in lines]
    assert t == list(range(1, len(lines) + 1))
    compress = phi[0] > phi0  # the random start overlaps: the first search expands
    i, start, searches = 0, s0, 0
    while i < len(lines):
        prev = phi0 if i == 0 else phi[i - 1]
        if 2 * start * prev < 40 * res:
            break  # the steps no longer show in the printed packing
        s, steps = start, 0
        while i < len(lines):
            prev = phi0 if i == 0 else phi[i - 1]
            if steps and (phi[i] > prev) != compress:
                break  # the next search
            assert phi[i] == pytest.approx(prev * factor(compress, s), abs=2.5 * res), (i, s)
            i += 1
            steps += 1
            s *= 1.1
        # every line but the last stays on the search's side of the target
        side = E[i - steps:i - 1]
        if compress:
            assert all(e < target * (1 + 1e-5) for e in side) and E[i - 1] >= target * (1 - 1e-5)
        else:
            assert all(e > target * (1 - 1e-5) for e in side) and E[i - 1] <= target * (1 + 1e-5)
        searches += 1
        start = start * 1.1 ** (steps + grow_after) / 2
        compress = not compress
    return searches


def test_find_jam_2p_ends_on_the_target_energy(tmp_path):
    target, n = 1e-6, 64
    _run("spherocyl_findjam", ["r", str(tmp_path), str(n), "1e-3", "0.01", str(target), "0.01",
                               "0.01", "0.1", "0.80", "1", "1.4"])
    lines = open(tmp_path / "sd_stress_energy.dat").read().splitlines()
    # two-point steps: x, y, L times 1 - s, or 1 / (1 - s) when expanding
    # (resize_step_2p), so the packing changes by (1 - s)^-2 or (1 - s)^2;
    # the next search starts from s 1.1^(steps - 1) / 2
    phi0 = _sd(str(tmp_path / "sd0000000000.dat"))[3]
    assert _check_search(lines, phi0, target, 1e-3,
                         lambda c, s: (1 - s) ** -2 if c else (1 - s) ** 2, -1) >= 5
    bins = glob.glob(str(tmp_path / "sd*.bin"))
    assert len(bins) == 1
    h, a = native.read_positions_bin(bins[0])
    t_end = int(os.path.basename(bins[0])[2:12])
    assert int(lines[-1].split()[0]) == t_end and a["x"].shape[0] == n
    energies = np.array([float(l.split()[2]) for l in lines])
    packings = np.array([float(l.split()[1]) for l in lines])
    # the search brackets the target ever more tightly: the last steps straddle it
    crossings = np.nonzero(np.diff(np.sign(energies - target)))[0]
    assert len(crossings) >= 5
    assert abs(packings[-1] - packings[-2]) < 1e-9
    # the final state, in the binary file, is a minimum whose energy the oracle
    # puts next to the target (within the last step's change)
    b = oracle.OracleBox(h["Lx"], a["x"], a["y"], a["phi"], a["A"], a["R"], 0.1)
    b.set_contact_model(1)
    b.find_neighbors()
    b.calc_se()
    assert b.frms() < 1e-10
    e = b.scalars()["E"]
    print("find_jam_2p: steps", t_end, "phi_J", packings[-1], "E", e, flush=True)
    assert e == pytest.approx(energies[-1], rel=1e-5)
    assert abs(e - target) < 0.1 * target


def test_y_compress_energy_2p_moves_only_y(tmp_path):
    """y_compress_energy_2p (src/cuda/2point_dynamics.cu:4458-4632) through
    spherocyl_ycompress (argument order of main_2point_ycompress.cpp): the
    bracketing search of find_jam_2p with steps that scale y and Ly only: Lx
    stays, every step changes the packing by the factor of its Ly change, and
    the final state is a minimum on the target energy."""
    target, n, s0 = 1e-6, 64, 1e-3
    _run("spherocyl_ycompress", ["r", str(tmp_path), str(n), str(s0), "0.01", str(target), "0.01",
                                 "0.01", "0.1", "0.80", "1", "1.4"])
    n0, Lx0, Ly0, phi0, _ = _sd(str(tmp_path / "sd0000000000.dat"))
    assert Lx0 == Ly0
    lines = open(tmp_path / "sd_stress_energy.dat").read().splitlines()
    phi = np.array([float(l.split()[1]) for l in lines])
    E = np.array([float(l.split()[2]) for l in lines])
    # y-only steps: the packing changes by (1 - s)^-1 or (1 - s)
    assert _check_search(lines, phi0, target, s0,
                         lambda c, s: 1 / (1 - s) if c else 1 - s, -1) >= 5
    crossings = np.nonzero(np.diff(np.sign(E - target)))[0]
    assert len(crossings) >= 5
    bins = glob.glob(str(tmp_path / "sd*.bin"))
    assert len(bins) == 1
    h, a = native.read_positions_bin(bins[0])
    assert h["Lx"] == pytest.approx(Lx0, rel=1e-13) and abs(h["Ly"] / Ly0 - 1) > 1e-3
    assert h["Ly"] == pytest.approx(Ly0 * phi0 / phi[-1], rel=1e-9)
    b = oracle.OracleBox(h["Lx"], a["x"], a["y"], a["phi"], a["A"], a["R"], 0.1)
    b.set_box(h["Lx"], h["Ly"])
    b.set_contact_model(1)
    b.find_neighbors()
    b.calc_se()
    e = b.scalars()["E"]
    print("y_compress_energy_2p: steps", len(lines), "Ly/Lx", h["Ly"] / h["Lx"], "E", e, flush=True)
    assert b.frms() < 1e-10
    assert e == pytest.approx(E[-1], rel=1e-5)
    assert abs(e - target) < 0.1 * target


def _oracle_1p(path):
    n, Lx, Ly, packing, rows = _sd(path)
    b = oracle.OracleBox(Lx, rows[:, 0], rows[:, 1], rows[:, 2], rows[:, 4], rows[:, 3], 0.1)
    b.find_neighbors()
    b.calc_se()
    return b


def test_find_jam_1p_ends_on_the_target_energy(tmp_path):
    """find_jam_1p (src/cuda/energy_minimize.cu:1882-1961): the bracketing
    search ends next to the target energy, and the oracle's energy of the
    final state (binary file) and of the saved states equals the SE lines'.
    The one-point step rule differs from the two-point one: an expansion
    multiplies x, y, L by 1 + s (packing / (1 + s)^2), and the next search
    starts from s 1.1^(steps) / 2.  Bidisperse discs (aspect 0), so that no
    minimisation stalls on the closest-point model's alignment kink."""
    target, n, s0 = 1e-5, 64, 1e-3
    r = subprocess.run([os.path.join(BIN, "spherocyl_findjam_1p"), "r", str(tmp_path), str(n),
                        str(s0), "0.01", str(target), "0.01", "0.01", "0.1", "0.80", "0", "1.4"],
                       capture_output=True, text=True, timeout=140,
                       env=dict(os.environ, SPHEROCYL_SEED="4"))
    print("driver: spherocyl_findjam_1p", flush=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    lines = open(tmp_path / "sd_stress_energy.dat").read().splitlines()
    bins = glob.glob(str(tmp_path / "sd*.bin"))
    assert len(bins) == 1
    h, a = native.read_positions_bin(bins[0])
    t_end = int(os.path.basename(bins[0])[2:12])
    assert int(lines[-1].split()[0]) == t_end
    energies = np.array([float(l.split()[2]) for l in lines])
    phi = np.array([float(l.split()[1]) for l in lines])
    phi0 = _sd(str(tmp_path / "sd0000000000.dat"))[3]
    assert _check_search(lines, phi0, target, s0,
                         lambda c, s: (1 - s) ** -2 if c else (1 + s) ** -2, 0) >= 5
    crossings = np.nonzero(np.diff(np.sign(energies - target)))[0]
    assert len(crossings) >= 5
    # every line prints the packing with %.9f (resize_to_energy_1p's format)
    assert all(len(l.split()[1].split(".")[1]) == 9 for l in lines)
    b = oracle.OracleBox(h["Lx"], a["x"], a["y"], a["phi"], a["A"], a["R"], 0.1)
    b.find_neighbors()
    b.calc_se()
    e = b.scalars()["E"]
    print("find_jam_1p: steps", t_end, "E", e, "F_rms", b.frms(), flush=True)
    assert e == pytest.approx(energies[-1], rel=1e-5)
    assert abs(e - target) < 0.3 * target
    for path in sorted(glob.glob(str(tmp_path / "sd0*.dat")))[1::7]:
        t = int(os.path.basename(path)[2:12])
        ob = _oracle_1p(path)
        line = next(l for l in lines if int(l.split()[0]) == t)
        assert float(line.split()[2]) == pytest.approx(ob.scalars()["E"], rel=1e-5, abs=1e-14), t


def test_quasistatic_find_jam_cycles_and_compress(tmp_path):
    """quasistatic_find_jam and quasistatic_compress (src/cuda/energy_minimize.cu:
    1731-1880) through SpherocylBox in a small C++ program
    (tests/cpp/box_qs_1p.cpp), closest-point model.  Each cycle expands until
    the energy vanishes, then compresses until it reaches that cycle's
    maximum, which halves from cycle to cycle together with the step; the
    compression then expands to zero energy and compresses to the target
    packing.  Saved states agree with their SE lines (oracle energy)."""
    pkg = os.path.join(ROOT, "spherocyls_amd")
    exe = str(tmp_path / "box_qs_1p")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "box_qs_1p.cpp"), "-L", pkg,
                    "-lspherocyls", f"-Wl,-rpath,{pkg}", "-o", exe], check=True)
    n, step, emax, emin, maxpack = 64, 2e-3, 1e-4, 1e-6, 0.90
    print("box_qs_1p", n, flush=True)
    (tmp_path / "c").mkdir()
    r = subprocess.run([exe, str(tmp_path), str(n), "0.84", "5", str(step), str(emax), str(emin),
                        str(maxpack), str(tmp_path / "c")], capture_output=True, text=True,
                       timeout=140)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = open(tmp_path / "cycle_se.dat").read().splitlines()
    t = np.array([int(l.split()[0]) for l in lines])
    phi = np.array([float(l.split()[1]) for l in lines])
    E = np.array([float(l.split()[2]) for l in lines])
    assert np.all(np.diff(t) == 1)
    # a cycle starts by minimising the state it is given, without a resize
    # step: its first line repeats the packing of the line before
    starts = list(np.nonzero(np.diff(phi) == 0)[0] + 1)
    bounds = [0] + starts + [len(t)]
    ncyc = len(bounds) - 1
    assert ncyc == int(np.ceil(np.log2(emax / emin)))  # dMaxE halves until <= dMinE
    for c in range(ncyc):
        cyc_E, cyc_phi = E[bounds[c]:bounds[c + 1]], phi[bounds[c]:bounds[c + 1]]
        emax_c = emax * 0.5 ** c
        # compression ends at the first state at or above the cycle's maximum
        assert cyc_E[-1] >= emax_c * (1 - 1e-5), c
        comp = np.nonzero(np.diff(cyc_phi) > 0)[0]
        assert np.all(cyc_E[comp] < emax_c * (1 + 1e-5)), c
        # compression starts from a state without overlaps
        assert len(comp) and cyc_E[comp[0]] == 0.0, c
        # steps of the cycle: phi changes by (1 +- step_c)^-2
        ratio = cyc_phi[1:] / cyc_phi[:-1]
        step_c = step * 0.5 ** c
        want = np.where(np.diff(cyc_phi) > 0, (1 - step_c) ** -2, (1 + step_c) ** -2)
        np.testing.assert_allclose(ratio, want, rtol=2e-6)
    comp_lines = open(tmp_path / "c" / "compress_se.dat").read().splitlines()
    phi_c = np.array([float(l.split()[1]) for l in comp_lines])
    assert phi_c[-1] >= maxpack and phi_c[-2] < maxpack
    t_end, phi1, phi2 = r.stdout.split()
    assert float(phi2) == pytest.approx(phi_c[-1], abs=1e-6)
    checked = 0
    for path in sorted(glob.glob(str(tmp_path / "sd0*.dat"))):
        ts = int(os.path.basename(path)[2:12])
        if ts % 25:
            continue
        ob = _oracle_1p(path)
        line = next(l for l in lines if int(l.split()[0]) == ts)
        assert float(line.split()[2]) == pytest.approx(ob.scalars()["E"], rel=1e-5, abs=1e-14), ts
        checked += 1
    assert checked >= 2
