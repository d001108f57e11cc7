"""CPU-only checks of the native library: it loads, exports every entry point
include/spherocyls.h declares, and its host-only parts (status strings, the
seeded placement generator) behave.  No GPU compute is called here."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from spherocyls_amd import configs, native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


C_HEADERS = ("spherocyls.h", "spherocyl_files.h")


def declared_functions():
    names = set()
    for h in C_HEADERS:
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(sc_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    assert os.path.exists(native.LIB_PATH)
    out = subprocess.run(["nm", "-D", "--defined-only", native.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (sc_\w+)$", out, flags=re.M))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing
    assert len(declared_functions()) >= 35


def test_cpp_box_class_is_exported():
    """SpherocylBox (include/spherocyl_box.h) lives in the same library."""
    out = subprocess.run(["nm", "-DC", "--defined-only", native.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for m in ("SpherocylBox::run_strain(double, unsigned long)", "SpherocylBox::resize_box(double)",
              "SpherocylBox::gradient_descent_minimize(int, double)",
              "SpherocylBox::save_positions(std::", "SpherocylBox::random_configuration("):
        assert m in out, m


def test_library_loads_and_describes_status():
    lib = native.load()
    assert lib.sc_version() >= 10000
    assert lib.sc_status_string(native.SC_OK) == b"ok"
    assert lib.sc_status_string(native.SC_ERR_CELL_OVERFLOW) == b"cell bucket overflow"
    assert lib.sc_device_count() >= 0


def test_create_without_gpu_is_an_error_not_a_fallback():
    lib = native.load()
    if lib.sc_device_count() > 0:
        pytest.skip("a GPU is visible")
    pk = configs.shear_small(1)
    with pytest.raises(native.SpherocylsError) as e:
        native.System(pk)
    assert e.value.status == native.SC_ERR_NO_DEVICE


def test_random_packing_seeded_and_crossing_free():
    R, A = configs.bidisperse(300, aspect=3.0)
    L = configs.box_side(R, A, 0.85)
    a = native.random_packing(L, A, R, 7, gamma=0.3)
    b = native.random_packing(L, A, R, 7, gamma=0.3)
    c = native.random_packing(L, A, R, 8, gamma=0.3)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)
    assert not np.array_equal(a[0], c[0])
    x, y, phi = a
    assert np.all((x >= 0) & (x < L) & (y >= 0) & (y < L) & (phi >= 0) & (phi < np.pi))
    # no crossing pair: every contact force is finite at the strain it was made for
    box = oracle.OracleBox(L, x, y, phi, A, R, 0.1)
    box.set_params(0.3, 0.0, 0.0, 0.0, 0.05)
    box.find_neighbors()
    box.calc_se()
    p = box.particles()
    assert np.all(np.isfinite(p["fx"])) and np.all(np.isfinite(p["tau"]))
    assert box.scalars()["ncontacts"] > 0  # overlaps are allowed, crossings are not


def test_random_packing_rejects_bad_arguments():
    with pytest.raises(native.SpherocylsError):
        native.random_packing(10.0, np.array([0.5]), np.array([-1.0]), 1)
    with pytest.raises(native.SpherocylsError):
        native.random_packing(10.0, np.array([0.5]), np.array([0.5]), 1, gamma=0.7)
    # an impossible density gives up instead of looping forever
    R, A = configs.bidisperse(400)
    with pytest.raises(native.SpherocylsError) as e:
        native.random_packing(5.0, A, R, 1, max_tries=50)
    assert e.value.status == native.SC_ERR_PLACEMENT


def test_workload_generators():
    pk = configs.shear_small(1)
    assert pk.n == 512
    np.testing.assert_array_equal(pk.R[:4], [0.5, 0.7, 0.5, 0.7])
    np.testing.assert_array_equal(pk.A, pk.R)
    area = np.sum((configs.REF_PI * pk.R + 4 * pk.A) * pk.R)
    assert area / pk.L ** 2 == pytest.approx(0.85, rel=1e-12)
    ens = configs.ensemble(count=24, seed=3)
    ns = [p.n for p in ens]
    assert min(ns) >= 64 and max(ns) <= 4096
    aspects = {round(float((p.A[0] + p.R[0]) / p.R[0]), 6) for p in ens}
    assert aspects <= {1.2, 2.0, 4.0}
    assert all(0.80 <= p.packing_fraction <= 0.86 for p in ens)


def test_drivers_built():
    from spherocyls_amd import build
    for exe in build.DRIVERS:
        assert os.access(os.path.join(ROOT, "spherocyls_amd", "bin", exe), os.X_OK), exe


def test_gpu_order_drivers_reject_unbuilt_options(tmp_path):
    """spherocyl_box stops with an error, before any GPU call, for the shear
    change it does not build (3: shear past the remap, then rotate) and
    spherocyl_pureshear for a change other than 0 or 1."""
    import subprocess
    common = ["r", str(tmp_path), "16", "1e-3", "0.01", "0.05", "0.01", "1e-4", "0.1"]
    for exe, change in (("spherocyl_box", "3"), ("spherocyl_box", "4"),
                        ("spherocyl_pureshear", "2")):
        r = subprocess.run([os.path.join(ROOT, "spherocyls_amd", "bin", exe), *common, change],
                           capture_output=True, text=True, timeout=60)
        assert r.returncode == 1, (exe, change, r.stderr)
        assert "not supported" in r.stderr or "must be" in r.stderr, (exe, change, r.stderr)


def test_new_entry_points_reject_bad_arguments():
    """Argument checks of the padding setter, the cell-size autotuner and the
    y-only affine step come before any device work."""
    import ctypes as C
    lib = native.load()
    assert lib.sc_set_pad(None, -1, C.c_double(0.1)) == native.SC_ERR_ARG
    pads = (C.c_double * 2)(0.1, 0.2)
    assert lib.sc_autotune_pad(None, 2, pads, 10, None, None) == native.SC_ERR_ARG
    scale = (C.c_double * 1)(0.99)
    assert lib.sc_affine_resize_y(None, scale) == native.SC_ERR_ARG


def test_driver_reports_missing_gpu_or_bad_input(tmp_path):
    """Without a readable start file the driver exits with status 1 and a message
    (the reference exits the same way), never a crash."""
    import subprocess
    r = subprocess.run([os.path.join(ROOT, "spherocyls_amd", "bin", "spherocyl_strain"),
                        str(tmp_path / "missing.dat"), str(tmp_path), "64", "1e-4", "0.1", "1e-4",
                        "0.01", "1e-5", "0.1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "cannot read" in r.stderr


def test_exceptions_map_to_status_codes(tmp_path):
    """Host code inside the library throws; no exception may leave an extern "C"
    entry point.  The shared mapping turns each kind into a status (built here
    with g++; it needs no GPU), and every entry-point guard ends in catch (...)."""
    exe = str(tmp_path / "status_map")
    subprocess.run(["g++", "-std=c++17", "-O1", os.path.join(ROOT, "tests", "cpp", "status_map.cpp"),
                    "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    csrc = os.path.join(ROOT, "spherocyls_amd", "csrc")
    api = open(os.path.join(csrc, "sc_api.hip")).read()
    mp = open(os.path.join(csrc, "sc_mpnn.hip")).read()
    gen = open(os.path.join(csrc, "sc_configgen.cpp")).read()
    for text, macro in ((api, "#define API_END"), (mp, "#define MP_END")):
        body = text[text.index(macro):].split("\n\n", 1)[0]
        assert "catch (...)" in body and "current_exception_status" in body, macro
    assert "catch (...)" in gen
    assert native.load().sc_status_string(10) == b"host memory or internal error"


def test_gamma_shapes_follow_the_gamma_distribution():
    """sc_gamma_shapes (set_shapes with an aspect width, src/cuda/spherocyl_box.cpp:
    1266-1300): A/R is gamma-distributed with mean mu and standard deviation
    sigma_rel mu (shape 1/sigma_rel^2, scale mu sigma_rel^2), for shapes above
    and below 1 (the two branches of the sampler); radii alternate
    0.5 bidispersity / 0.5; a seed gives one sequence."""
    import scipy.stats as st

    for mu, rel, seed in ((2.0, 0.5, 7), (1.5, 1.5, 8)):
        R, A = native.gamma_shapes(50000, mu, rel, 1.4, seed)
        a = A / R
        k, theta = 1.0 / rel ** 2, mu * rel ** 2
        assert abs(a.mean() - mu) < 5 * mu * rel / np.sqrt(len(a))
        assert st.kstest(a, st.gamma(a=k, scale=theta).cdf).pvalue > 1e-3
        np.testing.assert_array_equal(R[0::2], 0.7)
        np.testing.assert_array_equal(R[1::2], 0.5)
    R1, A1 = native.gamma_shapes(100, 2.0, 0.3, 1.0, 3)
    R2, A2 = native.gamma_shapes(100, 2.0, 0.3, 1.0, 3)
    np.testing.assert_array_equal(A1, A2)
    with pytest.raises(native.SpherocylsError):
        native.gamma_shapes(10, 2.0, 0.0)


REF_CPP = "/root/reference/src/cpp"


@pytest.mark.skipif(not os.path.isdir(REF_CPP), reason="reference sources not present")
@pytest.mark.parametrize("driver", ["main.cpp", "main_resize.cpp", "main_relax.cpp"])
def test_reference_drivers_build_against_the_box_header(tmp_path, driver):
    """The drop-in claim: the reference's own drivers compile unchanged against
    include/spherocyl_box.h and link against the library (the reference's box
    sources left out).  The driver and its file-input sources are compiled
    from where they lie; only the reference's own spherocyl_box.h is kept out
    of the include path, so the quoted include resolves to this header."""
    src = tmp_path / "src"
    src.mkdir()
    # a directory holding the driver's other includes but not the old box header
    for f in ("file_input.h",):
        os.symlink(os.path.join(REF_CPP, f), src / f)
    os.symlink(os.path.join(REF_CPP, driver), src / driver)
    exe = str(tmp_path / "drv")
    pkg = os.path.join(ROOT, "spherocyls_amd")
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-w", "-I", os.path.join(ROOT, "include"),
                        str(src / driver), os.path.join(REF_CPP, "dat_file_input.cpp"), "-L", pkg,
                        "-lspherocyls", f"-Wl,-rpath,{pkg}", "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    assert os.path.getsize(exe) > 0
