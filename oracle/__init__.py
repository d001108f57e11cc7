"""ORACLE -- test infrastructure only.

ctypes bindings for the two CPU references:
  * ``OracleBox``  -- liboracle.so, the project's CPU restatement (sc_oracle.cpp)
  * ``RefBox``     -- _ref/libspherocyl_ref.so, the reference's own src/cpp
                      compiled from the reference's own sources (present
                      only where those exist and ``make ref`` ran)
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only to test results against or as the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_SO may name another build of the restatement (the planted-mistake
# check in oracle/mutants.py builds its mutants elsewhere and points here)
ORACLE_SO = os.environ.get("SPHEROCYLS_ORACLE_SO") or os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libspherocyl_ref.so")
REF_SRC = "/root/reference/src/cpp"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build(ref: bool = True) -> None:
    """Compile liboracle.so (and _ref/ when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if ref and os.path.isdir(REF_SRC):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


_olib = None
_rlib = None


def _load_oracle():
    global _olib
    if _olib is None:
        if not os.path.exists(ORACLE_SO):
            build(ref=False)
        lib = C.CDLL(ORACLE_SO)
        lib.orc_new.restype = C.c_void_p
        lib.orc_new.argtypes = [C.c_double, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_double]
        for name in ("orc_free", "orc_find_neighbors", "orc_calc_forces", "orc_calc_se",
                     "orc_strain_step", "orc_gd_step", "orc_set_back_gamma", "orc_pure_step",
                     "orc_flip_shear", "orc_rotate_by_gamma"):
            getattr(lib, name).argtypes = [C.c_void_p]
            getattr(lib, name).restype = None
        lib.orc_set_params.argtypes = [C.c_void_p] + [C.c_double] * 5
        lib.orc_set_contact_model.argtypes = [C.c_void_p, C.c_int]
        lib.orc_set_contact_model.restype = None
        lib.orc_set_positions.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.orc_set_box.argtypes = [C.c_void_p, C.c_double, C.c_double]
        lib.orc_set_box.restype = None
        lib.orc_run_steps.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        lib.orc_run_steps.restype = C.c_int
        lib.orc_resize_step.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.orc_get_particles.argtypes = [C.c_void_p] + [_dp] * 10
        lib.orc_force_scale.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.orc_get_scalars.argtypes = [C.c_void_p, _dp]
        for name in ("orc_get_neighbors", "orc_get_contacts", "orc_brute_contacts"):
            getattr(lib, name).argtypes = [C.c_void_p, _ip, _ip, C.c_int]
            getattr(lib, name).restype = C.c_int
        lib.orc_fire.argtypes = [C.c_void_p, _dp, _dp]
        lib.orc_fire.restype = C.c_int
        lib.orc_fire_iter.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
        lib.orc_frms.argtypes = [C.c_void_p]
        lib.orc_frms.restype = C.c_double
        lib.orc_pair_contacts.argtypes = [C.c_int] + [C.c_double] * 12 + [_dp]
        lib.orc_pair_contacts.restype = C.c_int
        _olib = lib
    return _olib


def pair_contacts(p, q, model=1, L=1000.0, gamma=0.0):
    """Contacts of one pair, p = (x, y, phi, A, R) as A and q as B, in an open box:
    a list of dicts with s, t, D = (dx, dy), dist, the force (fx, fy) on p and
    the torques tau_p, tau_q."""
    lib = _load_oracle()
    out = np.zeros(18)
    xp, yp, pp, Ap, Rp = (float(v) for v in p)
    xq, yq, pq, Aq, Rq = (float(v) for v in q)
    k = lib.orc_pair_contacts(int(model), xp, yp, pp, Rp, Ap, xq, yq, pq, Rq, Aq, float(L),
                              float(gamma), _ptr(out))
    keys = ("s", "t", "dx", "dy", "dist", "fx", "fy", "tau_p", "tau_q")
    return [dict(zip(keys, out[9 * j:9 * j + 9])) for j in range(k)]


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def _load_ref():
    global _rlib
    if _rlib is None:
        if not os.path.exists(REF_SO):
            raise FileNotFoundError(REF_SO + " (run `make -C oracle ref` where /root/reference exists)")
        lib = C.CDLL(REF_SO)
        lib.ref_box_new.restype = C.c_void_p
        lib.ref_box_new.argtypes = [C.c_double, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_double]
        for name in ("ref_box_free", "ref_calc_se", "ref_calc_forces", "ref_find_neighbors",
                     "ref_set_back_gamma", "ref_strain_step", "ref_gd_step"):
            getattr(lib, name).argtypes = [C.c_void_p]
            getattr(lib, name).restype = None
        lib.ref_verbose.argtypes = [C.c_int]
        lib.ref_set_params.argtypes = [C.c_void_p] + [C.c_double] * 5
        lib.ref_set_save_steps.argtypes = [C.c_void_p, C.c_double, C.c_double]
        lib.ref_resize_step.argtypes = [C.c_void_p, C.c_int]
        lib.ref_get_particles.argtypes = [C.c_void_p] + [_dp] * 10
        lib.ref_get_scalars.argtypes = [C.c_void_p, _dp]
        lib.ref_get_neighbors.argtypes = [C.c_void_p, _ip, _ip, C.c_int]
        lib.ref_get_neighbors.restype = C.c_int
        lib.ref_run_steps.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        lib.ref_run_steps.restype = C.c_int
        lib.ref_run_strain.argtypes = [C.c_void_p, C.c_double, C.c_ulong, C.c_char_p, C.c_char_p]
        lib.ref_run_strain.restype = C.c_ulong
        lib.ref_resize_box.argtypes = [C.c_void_p, C.c_double, C.c_ulong, C.c_char_p, C.c_char_p]
        lib.ref_resize_box.restype = C.c_ulong
        lib.ref_gd_minimize.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_char_p, C.c_char_p]
        _rlib = lib
    return _rlib


class _BoxBase:
    _prefix = ""

    def particles(self):
        n = self.n
        arrs = [np.empty(n) for _ in range(10)]
        getattr(self.lib, self._prefix + "get_particles")(self.h, *[_ptr(a) for a in arrs])
        keys = ("x", "y", "phi", "R", "A", "fx", "fy", "tau", "I", "RC")
        return dict(zip(keys, arrs))

    def _pairs(self, fn):
        cap = max(64, 64 * self.n)
        p = np.empty(cap, dtype=np.int32)
        q = np.empty(cap, dtype=np.int32)
        k = fn(self.h, p.ctypes.data_as(_ip), q.ctypes.data_as(_ip), cap)
        if k > cap:
            raise RuntimeError("pair buffer too small")
        return p[:k].copy(), q[:k].copy()


class OracleBox(_BoxBase):
    """CPU restatement; arguments follow SpherocylBox(dL, N, x, y, phi, A*, R*, pad)."""
    _prefix = "orc_"

    def __init__(self, L, x, y, phi, A, R, pad):
        self.lib = _load_oracle()
        x, y, phi, A, R = (_f64(v) for v in (x, y, phi, A, R))
        self.n = int(x.shape[0])
        self.h = self.lib.orc_new(float(L), self.n, _ptr(x), _ptr(y), _ptr(phi), _ptr(A), _ptr(R),
                                  float(pad))

    @classmethod
    def from_packing(cls, pk, pad=0.1):
        return cls(pk.L, pk.x, pk.y, pk.phi, pk.A, pk.R, pad)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_free(self.h)
            self.h = None

    def set_params(self, gamma=0.0, gamma_total=0.0, strain_rate=0.0, resize_rate=0.0, step=0.05):
        self.lib.orc_set_params(self.h, gamma, gamma_total, strain_rate, resize_rate, step)

    def set_contact_model(self, model):
        """0: closest point (reference CPU); 1: two-point contact model."""
        self.lib.orc_set_contact_model(self.h, int(model))

    def set_positions(self, x=None, y=None, phi=None):
        """Overwrite coordinates; the neighbour list is kept (call
        find_neighbors() before evaluating a state far from the old one)."""
        x, y, phi = (None if v is None else _f64(v) for v in (x, y, phi))
        self.lib.orc_set_positions(self.h, _ptr(x), _ptr(y), _ptr(phi))

    def set_box(self, Lx, Ly):
        self.lib.orc_set_box(self.h, float(Lx), float(Ly))

    def find_neighbors(self): self.lib.orc_find_neighbors(self.h)
    def calc_forces(self): self.lib.orc_calc_forces(self.h)
    def calc_se(self): self.lib.orc_calc_se(self.h)
    def strain_step(self): self.lib.orc_strain_step(self.h)
    def gd_step(self): self.lib.orc_gd_step(self.h)
    def set_back_gamma(self): self.lib.orc_set_back_gamma(self.h)
    def pure_step(self): self.lib.orc_pure_step(self.h)
    def flip_shear(self): self.lib.orc_flip_shear(self.h)
    def rotate_by_gamma(self): self.lib.orc_rotate_by_gamma(self.h)

    def resize_step(self, shrink=True, corrected_image=False):
        self.lib.orc_resize_step(self.h, int(shrink), int(corrected_image))

    def run_steps(self, nsteps, se_every=0):
        nrec = nsteps // se_every if se_every > 0 else 0
        out = np.zeros(max(1, 4 * nrec))
        k = self.lib.orc_run_steps(self.h, nsteps, se_every, _ptr(out))
        return out[:4 * k].reshape(k, 4)

    def scalars(self):
        out = np.zeros(20)
        self.lib.orc_get_scalars(self.h, _ptr(out))
        keys = ("Lx", "Ly", "gamma", "gamma_total", "strain_rate", "resize_rate", "step",
                "packing", "pad", "E", "Pxx", "Pyy", "Pxy", "cPxx", "cPyy", "cPxy", "cPyx",
                "rows", "cols", "ncontacts")
        return dict(zip(keys, out.tolist()))

    def force_scale(self):
        sx, sy, st = np.empty(self.n), np.empty(self.n), np.empty(self.n)
        self.lib.orc_force_scale(self.h, _ptr(sx), _ptr(sy), _ptr(st))
        return sx, sy, st

    def neighbors(self): return self._pairs(self.lib.orc_get_neighbors)
    def contacts(self): return self._pairs(self.lib.orc_get_contacts)
    def brute_contacts(self): return self._pairs(self.lib.orc_brute_contacts)

    def frms(self):
        return float(self.lib.orc_frms(self.h))

    def fire(self, params):
        p = _f64(params)
        fr = np.zeros(1)
        it = self.lib.orc_fire(self.h, _ptr(p), _ptr(fr))
        return int(it), float(fr[0])

    def fire_iter(self, params, vx, vy, vpsi, state):
        self.lib.orc_fire_iter(self.h, _ptr(_f64(params)), _ptr(vx), _ptr(vy), _ptr(vpsi),
                               _ptr(state))


class RefBox(_BoxBase):
    """The compiled reference (src/cpp) behind ref_harness.cpp."""
    _prefix = "ref_"

    def __init__(self, L, x, y, phi, A, R, pad):
        self.lib = _load_ref()
        x, y, phi, A, R = (_f64(v) for v in (x, y, phi, A, R))
        self.n = int(x.shape[0])
        self.h = self.lib.ref_box_new(float(L), self.n, _ptr(x), _ptr(y), _ptr(phi), _ptr(A),
                                      _ptr(R), float(pad))

    @classmethod
    def from_packing(cls, pk, pad=0.1):
        return cls(pk.L, pk.x, pk.y, pk.phi, pk.A, pk.R, pad)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ref_box_free(self.h)
            self.h = None

    def set_params(self, gamma=0.0, gamma_total=0.0, strain_rate=0.0, resize_rate=0.0, step=0.05):
        self.lib.ref_set_params(self.h, gamma, gamma_total, strain_rate, resize_rate, step)

    def find_neighbors(self): self.lib.ref_find_neighbors(self.h)
    def calc_forces(self): self.lib.ref_calc_forces(self.h)
    def calc_se(self): self.lib.ref_calc_se(self.h)
    def strain_step(self): self.lib.ref_strain_step(self.h)
    def gd_step(self): self.lib.ref_gd_step(self.h)
    def set_back_gamma(self): self.lib.ref_set_back_gamma(self.h)

    def resize_step(self, shrink=True):
        self.lib.ref_resize_step(self.h, int(shrink))

    def run_steps(self, nsteps, se_every=0):
        nrec = nsteps // se_every if se_every > 0 else 0
        out = np.zeros(max(1, 4 * nrec))
        k = self.lib.ref_run_steps(self.h, nsteps, se_every, _ptr(out))
        return out[:4 * k].reshape(k, 4)

    def scalars(self):
        out = np.zeros(16)
        self.lib.ref_get_scalars(self.h, _ptr(out))
        keys = ("Lx", "Ly", "gamma", "gamma_total", "strain_rate", "resize_rate", "step",
                "packing", "pad", "E", "Pxx", "Pyy", "Pxy", "Pyx", "rows", "cols")
        return dict(zip(keys, out.tolist()))

    def neighbors(self): return self._pairs(self.lib.ref_get_neighbors)

    def run_strain(self, run_length, ntime, outdir, sefile, pos_step=None, se_step=None):
        if pos_step is not None:
            self.lib.ref_set_save_steps(self.h, pos_step, se_step)
        return int(self.lib.ref_run_strain(self.h, run_length, ntime, outdir.encode(),
                                           sefile.encode()))

    def resize_box(self, final_packing, ntime, outdir, sefile, pos_step=None, se_step=None):
        if pos_step is not None:
            self.lib.ref_set_save_steps(self.h, pos_step, se_step)
        return int(self.lib.ref_resize_box(self.h, final_packing, ntime, outdir.encode(),
                                           sefile.encode()))

    def gd_minimize(self, nmax, emin, outdir, sefile, pos_step=None):
        if pos_step is not None:
            self.lib.ref_set_save_steps(self.h, pos_step, 1e-5)
        self.lib.ref_gd_minimize(self.h, nmax, emin, outdir.encode(), sefile.encode())
