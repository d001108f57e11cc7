"""sc_calc_se_noalloc -- the allocation-free calc_se -- against sc_calc_se.

The inputs and outputs are device buffers the caller owns, allocated here
with the HIP runtime the library itself links (hipMalloc through ctypes, so
both sides use one runtime); the entry point's working memory is its static
device scratch array.  Two calls in a row, the second with a smaller packing
and the other contact model, each against sc_calc_se on the list path with
the same launch shape (one lane per particle, 256 threads)."""
import ctypes as C

import numpy as np
import pytest

from spherocyls_amd import configs, native

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _lib():
    lib = native.load()
    fn = lib.sc_calc_se_noalloc
    fn.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int] + [C.c_void_p] * 11
    fn.restype = C.c_int
    return fn


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libspherocyls.so links."""

    def __init__(self):
        native.load()  # loads libamdhip64 as a dependency of the library
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # H2D
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0  # D2H
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _noalloc(pk, gamma, model):
    hip = _Hip()
    try:
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        x, y, phi, A, R = (hip.upload(f64(v)) for v in (pk.x, pk.y, pk.phi, pk.A, pk.R))
        nan = np.full(pk.n, np.nan)
        fx, fy, tau = (hip.upload(nan) for _ in range(3))
        se = hip.upload(np.full(10, np.nan))
        status = hip.upload(np.full(1, -1, dtype=np.int32))
        st = _lib()(pk.n, pk.L, gamma, 0.1, model, x, y, phi, A, R, fx, fy, tau, se, status, None)
        assert hip.rt.hipDeviceSynchronize() == 0
        return (st, hip.download(fx, nan), hip.download(fy, nan), hip.download(tau, nan),
                hip.download(se, np.zeros(10)),
                int(hip.download(status, np.zeros(1, dtype=np.int32))[0]))
    finally:
        hip.free()


def _system(pk, gamma, model):
    s = native.System(pk, pads=0.1)
    try:
        s.set_path("list")
        s.set_list_block(256, 1)
        s.set_contact_model(model)
        s.set_gamma(gamma, 0.0)
        s.calc_se()
        return s.forces(), s.box()
    finally:
        s.close()


def _check(pk, gamma, model):
    code = native.CONST["SC_CONTACT_TWO_POINT"] if model == "two_point" else 0
    print("noalloc calc_se", pk.name, pk.n, model, flush=True)
    st, fx, fy, tau, se, status = _noalloc(pk, gamma, code)
    assert st == 0 and status == 0
    (rx, ry, rt), bx = _system(pk, gamma, model)
    scale = np.abs(rx).max() + np.abs(ry).max()
    assert np.max(np.abs(fx - rx)) <= TOL * scale
    assert np.max(np.abs(fy - ry)) <= TOL * scale
    assert np.max(np.abs(tau - rt)) <= TOL * (np.abs(rt).max() + 1e-300)
    keys = ("E", "Pxx", "Pyy", "Pxy", "cPxx", "cPyy", "cPxy", "cPyx", "ncontacts", "frms")
    for i, k in enumerate(keys):
        assert se[i] == pytest.approx(bx[k], rel=TOL, abs=TOL * abs(bx["cPxx"])), k
    assert se[8] > 0.5 * pk.n  # a dense packing: many contacts
    return fx


def test_noalloc_matches_calc_se_on_two_calls_in_a_row():
    R1, A1 = configs.bidisperse(4096, aspect=2.0)
    big = configs.make_packing(R1, A1, 0.85, 91, "noalloc N=4096 aspect 2", gamma=0.2)
    R2, A2 = configs.bidisperse(700, aspect=3.0)
    small = configs.make_packing(R2, A2, 0.83, 92, "noalloc N=700 aspect 3", gamma=-0.1)
    f1 = _check(big, 0.2, "two_point")
    f2 = _check(small, -0.1, "closest")
    assert f1.shape == (4096,) and f2.shape == (700,)


def test_noalloc_rejects_sizes_beyond_its_scratch():
    n = native.CONST.get("SC_NOALLOC_MAX_PARTICLES", 32768) + 1
    one = C.c_void_p(1)  # never dereferenced: the size check comes first
    st = _lib()(n, 100.0, 0.0, 0.1, 0, *([one] * 10), None)
    assert st == native.CONST["SC_ERR_ARG"]
