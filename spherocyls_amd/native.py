"""ctypes binding of the C ABI in include/spherocyls.h (libspherocyls.so).

The library is loaded from the package directory (built in-tree by
``python -m spherocyls_amd.build``).  There is no fallback: if the library is
missing, or no GPU is visible when a GPU call is made, these functions raise.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
# SPHEROCYLS_LIB names another build of the same library (for example one
# built with different compile-time options); the default is the in-tree build.
LIB_PATH = os.environ.get("SPHEROCYLS_LIB") or os.path.join(PKG, "libspherocyls.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "spherocyls.h")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_lib = None


def _constants():
    out = {}
    with open(HEADER) as f:
        for m in re.finditer(r"#define\s+(SC_\w+)\s+(-?\d+)", f.read()):
            out[m.group(1)] = int(m.group(2))
    return out


CONST = _constants()
globals().update(CONST)


class SpherocylsError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        name = next((k for k, v in CONST.items() if k.startswith("SC_ERR") and v == status), str(status))
        super().__init__(f"{name}: {msg}" if msg else name)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is not built (python -m spherocyls_amd.build)")
    lib = C.CDLL(LIB_PATH)
    sysp = C.c_void_p
    sig = {
        "sc_version": ([], C.c_int),
        "sc_status_string": ([C.c_int], C.c_char_p),
        "sc_device_count": ([], C.c_int),
        "sc_system_create": ([C.POINTER(sysp), C.c_int, C.c_int, _ip, _dp, _dp, _dp, _dp, _dp,
                              _dp, _dp], C.c_int),
        "sc_system_destroy": ([sysp], C.c_int),
        "sc_last_error": ([sysp], C.c_char_p),
        "sc_num_particles": ([sysp, _ip, _ip], C.c_int),
        "sc_set_dynamics": ([sysp, C.c_int, C.c_double, C.c_double, C.c_double], C.c_int),
        "sc_set_gamma": ([sysp, C.c_int, C.c_double, C.c_double], C.c_int),
        "sc_set_precision": ([sysp, C.c_int], C.c_int),
        "sc_set_contact_model": ([sysp, C.c_int], C.c_int),
        "sc_set_graphs": ([sysp, C.c_int], C.c_int),
        "sc_set_small_fire": ([sysp, C.c_int, C.c_int], C.c_int),
        "sc_get_contact_model": ([sysp, _ip], C.c_int),
        "sc_set_tiling": ([sysp, C.c_int, C.c_int], C.c_int),
        "sc_set_path": ([sysp, C.c_int], C.c_int),
        "sc_get_path": ([sysp, _ip], C.c_int),
        "sc_set_list_block": ([sysp, C.c_int, C.c_int], C.c_int),
        "sc_autotune": ([sysp, C.c_int, _ip, _ip, _ip, _dp], C.c_int),
        "sc_autotune_fire": ([sysp, _dp, C.c_double, C.c_int, _ip, _ip, _ip, _dp], C.c_int),
        "sc_set_pad": ([sysp, C.c_int, C.c_double], C.c_int),
        "sc_autotune_pad": ([sysp, C.c_int, _dp, C.c_int, _dp, _dp], C.c_int),
        "sc_get_tiling": ([sysp, _ip, _ip], C.c_int),
        "sc_set_positions": ([sysp, _dp, _dp, _dp], C.c_int),
        "sc_get_positions": ([sysp, _dp, _dp, _dp], C.c_int),
        "sc_get_forces": ([sysp, _dp, _dp, _dp], C.c_int),
        "sc_get_box": ([sysp, C.c_int, _dp], C.c_int),
        "sc_calc_forces": ([sysp], C.c_int),
        "sc_calc_se": ([sysp], C.c_int),
        "sc_strain_steps": ([sysp, C.c_int, C.c_int, _dp], C.c_int),
        "sc_resize_steps": ([sysp, C.c_int, C.c_int, C.c_int, C.c_int, _dp], C.c_int),
        "sc_gd_steps": ([sysp, C.c_int, C.c_double, _ip], C.c_int),
        "sc_affine_resize": ([sysp, _dp], C.c_int),
        "sc_affine_resize_y": ([sysp, _dp], C.c_int),
        "sc_pure_strain_steps": ([sysp, C.c_int, C.c_int, _dp], C.c_int),
        "sc_flip_shear": ([sysp, C.c_int], C.c_int),
        "sc_rotate_by_gamma": ([sysp, C.c_int], C.c_int),
        "sc_fire_minimize": ([sysp, _dp, C.c_double, _ip, _dp], C.c_int),
        "sc_get_neighbors": ([sysp, C.c_int, _ip, _ip, C.c_int, _ip], C.c_int),
        "sc_get_contacts": ([sysp, C.c_int, _ip, _ip, C.c_int, _ip], C.c_int),
        "sc_timing_enable": ([sysp, C.c_int], C.c_int),
        "sc_timing_get": ([sysp, C.c_int, _dp, C.POINTER(C.c_longlong)], C.c_int),
        "sc_timing_reset": ([sysp], C.c_int),
        "sc_synchronize": ([sysp], C.c_int),
        "sc_random_packing": ([C.c_int, C.c_double, C.c_double, _dp, _dp, C.c_ulonglong, C.c_int,
                               _dp, _dp, _dp], C.c_int),
        "sc_gamma_shapes": ([C.c_int, C.c_double, C.c_double, C.c_double, C.c_ulonglong, _dp,
                             _dp], C.c_int),
        "sc_random_configuration": ([C.c_int, C.c_double, C.c_double, _dp, _dp, C.c_int,
                                     C.c_ulonglong, C.c_int, _dp, _dp, _dp], C.c_int),
        "sc_write_positions": ([C.c_char_p, C.c_int, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp],
                               C.c_int),
        "sc_read_positions": ([C.c_char_p, C.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp], C.c_int),
        "sc_write_positions_sd": ([C.c_char_p, C.c_int] + [C.c_double] * 7 + [_dp] * 5, C.c_int),
        "sc_write_positions_bin": ([C.c_char_p, C.c_int] + [C.c_double] * 7 + [_dp] * 5, C.c_int),
        "sc_read_positions_bin": ([C.c_char_p, C.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp],
                                  C.c_int),
        "sc_format_se_line": ([C.c_char_p, C.c_int, C.c_ulong, C.c_int, _dp, C.c_int], C.c_int),
        "sc_mpnn_create": ([C.POINTER(sysp), C.c_int, C.c_ulonglong], C.c_int),
        "sc_mpnn_destroy": ([sysp], C.c_int),
        "sc_mpnn_last_error": ([sysp], C.c_char_p),
        "sc_mpnn_num_params": ([sysp, _ip], C.c_int),
        "sc_mpnn_get_params": ([sysp, _fp], C.c_int),
        "sc_mpnn_set_params": ([sysp, _fp], C.c_int),
        "sc_mpnn_get_grads": ([sysp, _fp], C.c_int),
        "sc_mpnn_set_output_scale": ([sysp, _fp], C.c_int),
        "sc_mpnn_get_output_scale": ([sysp, _fp], C.c_int),
        "sc_mpnn_load_graph": ([sysp, C.c_int, _fp, C.c_int, _ip, _ip, _fp], C.c_int),
        "sc_mpnn_forward": ([sysp, _fp], C.c_int),
        "sc_mpnn_backward": ([sysp, _fp, _fp], C.c_int),
        "sc_mpnn_adam": ([sysp, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int], C.c_int),
        "sc_mpnn_graph_from_system": ([sysp, sysp, _ip, _ip], C.c_int),
        "sc_mpnn_export_graph": ([sysp, sysp, _fp, _ip, _ip, _fp], C.c_int),
        "sc_mpnn_warm_start": ([sysp, sysp], C.c_int),
        "sc_mpnn_warm_start_damped": ([sysp, sysp, C.c_int, _dp, _dp, _dp, _dp], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def random_packing(L, A, R, seed, gamma=0.0, max_tries=100000):
    """Seeded non-crossing random placement in sheared coordinates at strain
    gamma (host code, no GPU needed)."""
    lib = load()
    A, R = _f64(A), _f64(R)
    n = A.shape[0]
    x, y, phi = np.empty(n), np.empty(n), np.empty(n)
    st = lib.sc_random_packing(n, float(L), float(gamma), _p(A), _p(R), C.c_ulonglong(seed),
                               max_tries, _p(x), _p(y), _p(phi))
    if st != 0:
        raise SpherocylsError(st, "random placement")
    return x, y, phi


def gamma_shapes(n, aspect_mean, aspect_rel_sigma, bidispersity=1.4, seed=1):
    """Radii and half-lengths with gamma-distributed A/R (sc_gamma_shapes)."""
    lib = load()
    R, A = np.empty(n), np.empty(n)
    st = lib.sc_gamma_shapes(n, aspect_mean, aspect_rel_sigma, bidispersity, seed, _p(R), _p(A))
    if st != 0:
        raise SpherocylsError(st, "sc_gamma_shapes")
    return R, A


def random_configuration(L, A, R, seed, config=0, gamma=0.0, max_tries=100000):
    """Seeded placement in one of the reference's initialConfig modes
    (SC_CONFIG_*); host code, no GPU needed."""
    lib = load()
    A, R = _f64(A), _f64(R)
    n = A.shape[0]
    x, y, phi = np.empty(n), np.empty(n), np.empty(n)
    st = lib.sc_random_configuration(n, float(L), float(gamma), _p(A), _p(R), int(config),
                                     C.c_ulonglong(seed), max_tries, _p(x), _p(y), _p(phi))
    if st != 0:
        raise SpherocylsError(st, "random placement")
    return x, y, phi


def write_positions(path, L, packing, gamma, gamma_total, strain_rate, step, x, y, phi, R, A):
    """save_positions file (reference layout, 14 significant digits)."""
    lib = load()
    x, y, phi, R, A = (_f64(v) for v in (x, y, phi, R, A))
    st = lib.sc_write_positions(os.fsencode(path), x.shape[0], L, packing, gamma, gamma_total,
                                strain_rate, step, _p(x), _p(y), _p(phi), _p(R), _p(A))
    if st != 0:
        raise SpherocylsError(st, f"cannot write {path}")


def read_positions(path):
    """(header dict, dict of arrays x y phi R A) of a save_positions file."""
    lib = load()
    n = C.c_int()
    h = np.zeros(6)
    st = lib.sc_read_positions(os.fsencode(path), 0, C.byref(n), _p(h), None, None, None, None,
                               None)
    if st != 0:
        raise SpherocylsError(st, f"cannot read {path}")
    arrs = [np.empty(max(n.value, 1)) for _ in range(5)]
    st = lib.sc_read_positions(os.fsencode(path), n.value, C.byref(n), _p(h),
                               *[_p(a) for a in arrs])
    if st != 0:
        raise SpherocylsError(st, f"cannot read {path}")
    head = dict(zip(("L", "packing", "gamma", "gamma_total", "strain_rate", "step"), h.tolist()))
    return head, dict(zip(("x", "y", "phi", "R", "A"), (a[:n.value] for a in arrs)))


SD_HEADER = ("Lx", "Ly", "packing", "gamma", "gamma_total", "strain_rate", "step")


def write_positions_sd(path, Lx, Ly, packing, gamma, gamma_total, strain_rate, step, x, y, phi,
                       R, A):
    """The GPU code's text position file (sd%010lu.dat)."""
    lib = load()
    x, y, phi, R, A = (_f64(v) for v in (x, y, phi, R, A))
    st = lib.sc_write_positions_sd(os.fsencode(path), x.shape[0], Lx, Ly, packing, gamma,
                                   gamma_total, strain_rate, step, _p(x), _p(y), _p(phi), _p(R),
                                   _p(A))
    if st != 0:
        raise SpherocylsError(st, f"cannot write {path}")


def write_positions_bin(path, Lx, Ly, packing, gamma, gamma_total, strain_rate, step, x, y, phi,
                        R, A):
    """save_positions_bin file (sd%010lu.bin)."""
    lib = load()
    x, y, phi, R, A = (_f64(v) for v in (x, y, phi, R, A))
    st = lib.sc_write_positions_bin(os.fsencode(path), x.shape[0], Lx, Ly, packing, gamma,
                                    gamma_total, strain_rate, step, _p(x), _p(y), _p(phi), _p(R),
                                    _p(A))
    if st != 0:
        raise SpherocylsError(st, f"cannot write {path}")


def read_positions_bin(path):
    """(header dict, dict of arrays x y phi R A) of a save_positions_bin file."""
    lib = load()
    n = C.c_int()
    h = np.zeros(7)
    st = lib.sc_read_positions_bin(os.fsencode(path), 0, C.byref(n), _p(h), None, None, None,
                                   None, None)
    if st != 0:
        raise SpherocylsError(st, f"cannot read {path}")
    arrs = [np.empty(max(n.value, 1)) for _ in range(5)]
    st = lib.sc_read_positions_bin(os.fsencode(path), n.value, C.byref(n), _p(h),
                                   *[_p(a) for a in arrs])
    if st != 0:
        raise SpherocylsError(st, f"cannot read {path}")
    return dict(zip(SD_HEADER, h.tolist())), dict(zip(("x", "y", "phi", "R", "A"),
                                                      (a[:n.value] for a in arrs)))


def format_se_line(t, values, precision=6):
    """One stress/energy line as the reference's writers format it."""
    lib = load()
    v = _f64(values)
    buf = C.create_string_buffer(64 + 32 * v.shape[0])
    k = lib.sc_format_se_line(buf, len(buf), t, v.shape[0], _p(v), precision)
    if k < 0:
        raise SpherocylsError(1, "SE line buffer")
    return buf.value.decode()


FIRE_DEFAULTS = dict(dt0=0.05, dt_max=0.5, n_min=5, f_inc=1.1, f_dec=0.5, alpha0=0.1,
                     f_alpha=0.99, max_move=0.02, ftol=1e-10, max_iter=200000)


def fire_params(**kw):
    p = dict(FIRE_DEFAULTS)
    p.update(kw)
    return np.array([p[k] for k in ("dt0", "dt_max", "n_min", "f_inc", "f_dec", "alpha0",
                                     "f_alpha", "max_move", "ftol", "max_iter")], dtype=np.float64)


BOX_FIELDS = ("Lx", "Ly", "gamma", "gamma_total", "E", "Pxx", "Pyy", "Pxy", "cPxx", "cPyy",
              "cPxy", "cPyx", "ncontacts", "frms", "packing", "rows", "cols", "pad", "iters",
              "fire_dt", "fire_alpha", "pair_fp32")
SE_FIELDS = ("E", "Pxx", "Pyy", "Pxy", "cPxx", "cPyy", "cPxy", "cPyx", "ncontacts")


class System:
    """One or more packings on one GPU (thin wrapper over sc_system)."""

    def __init__(self, packings, pads=0.1, device=0):
        lib = load()
        self.lib = lib
        if not isinstance(packings, (list, tuple)):
            packings = [packings]
        self.npack = len(packings)
        self.sizes = [int(p.x.shape[0]) for p in packings]
        self.offsets = np.zeros(self.npack + 1, dtype=np.int32)
        self.offsets[1:] = np.cumsum(self.sizes)
        cat = lambda k: _f64(np.concatenate([getattr(p, k) for p in packings]))
        x, y, phi, A, R = cat("x"), cat("y"), cat("phi"), cat("A"), cat("R")
        L = _f64([p.L for p in packings])
        pad = _f64(np.broadcast_to(np.asarray(pads, dtype=np.float64), (self.npack,)))
        self.ntot = int(self.offsets[-1])
        h = C.c_void_p()
        st = lib.sc_system_create(C.byref(h), device, self.npack, self.offsets.ctypes.data_as(_ip),
                                  _p(L), _p(x), _p(y), _p(phi), _p(A), _p(R), _p(pad))
        if st != 0:
            raise SpherocylsError(st, "sc_system_create")
        self.h = h
        self.precision = "fp64"

    def close(self):
        if getattr(self, "h", None):
            self.lib.sc_system_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _ck(self, st):
        if st != 0:
            msg = self.lib.sc_last_error(self.h)
            raise SpherocylsError(st, msg.decode() if msg else "")
        return st

    def set_dynamics(self, strain_rate=0.0, resize_rate=0.0, step=0.05, k=-1):
        self._ck(self.lib.sc_set_dynamics(self.h, k, strain_rate, resize_rate, step))

    def set_gamma(self, gamma, gamma_total=0.0, k=-1):
        self._ck(self.lib.sc_set_gamma(self.h, k, gamma, gamma_total))

    def set_precision(self, mode):
        self._ck(self.lib.sc_set_precision(self.h, {"fp64": 0, "mixed": 1}.get(mode, mode)))
        self.precision = {0: "fp64", 1: "mixed"}.get(mode, mode)

    def set_contact_model(self, model):
        """"closest" (the CPU reference) or "two_point"."""
        m = {"closest": SC_CONTACT_CLOSEST, "two_point": SC_CONTACT_TWO_POINT}.get(model, model)
        self._ck(self.lib.sc_set_contact_model(self.h, m))

    def set_graphs(self, enable):
        self._ck(self.lib.sc_set_graphs(self.h, int(bool(enable))))

    def set_small_fire(self, max_particles, steps_per_launch=0):
        """FIRE of packings of at most max_particles (<= 256; 0: never) in one
        launch once every running packing is that small, at most
        steps_per_launch iterations per launch (0: the default, 50000)."""
        self._ck(self.lib.sc_set_small_fire(self.h, int(max_particles), int(steps_per_launch)))

    def contact_model(self):
        m = C.c_int()
        self._ck(self.lib.sc_get_contact_model(self.h, C.byref(m)))
        return "two_point" if m.value == SC_CONTACT_TWO_POINT else "closest"

    def set_tiling(self, tile_cols, block):
        self._ck(self.lib.sc_set_tiling(self.h, tile_cols, block))

    def set_path(self, path):
        self._ck(self.lib.sc_set_path(self.h, {"tile": 0, "list": 1}.get(path, path)))

    def path(self):
        v = C.c_int()
        self._ck(self.lib.sc_get_path(self.h, C.byref(v)))
        return ("tile", "list")[v.value]

    def set_list_block(self, block, lanes=1):
        self._ck(self.lib.sc_set_list_block(self.h, block, lanes))

    def autotune(self, reps=5):
        p, t, b, us = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        self._ck(self.lib.sc_autotune(self.h, reps, C.byref(p), C.byref(t), C.byref(b),
                                      C.byref(us)))
        path = ("tile", "list")[p.value]
        return dict(path=path, shape=t.value, block=b.value, us_per_pass=us.value)

    def autotune_fire(self, params=None, iterations=60, switch_frms=1e-6):
        """Launch-shape autotuner for minimisation: `iterations` FIRE
        iterations per candidate path and shape from the current positions
        (restored after every trial); keeps the fastest."""
        p = fire_params() if params is None else _f64(params)
        pa, t, b, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        self._ck(self.lib.sc_autotune_fire(self.h, _p(p), switch_frms, int(iterations),
                                           C.byref(pa), C.byref(t), C.byref(b), C.byref(ms)))
        return dict(path=("tile", "list")[pa.value], shape=t.value, block=b.value,
                    ms_per_trial=ms.value)

    def set_pad(self, pad, k=-1):
        self._ck(self.lib.sc_set_pad(self.h, k, float(pad)))

    def autotune_pad(self, pads=(0.05, 0.1, 0.2, 0.3), steps=200):
        """Cell-size autotuner: times `steps` Heun steps (the system's own
        dynamics, state restored after each trial) per candidate padding and
        keeps the fastest.  Returns (chosen padding, {padding: ms})."""
        c = _f64(pads)
        ms = np.zeros(len(c))
        chosen = C.c_double()
        self._ck(self.lib.sc_autotune_pad(self.h, len(c), _p(c), int(steps), _p(ms),
                                          C.byref(chosen)))
        return chosen.value, dict(zip(c.tolist(), ms.tolist()))

    def tiling(self):
        a, b = C.c_int(), C.c_int()
        self._ck(self.lib.sc_get_tiling(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_positions(self, x, y, phi):
        x, y, phi = _f64(x), _f64(y), _f64(phi)
        self._ck(self.lib.sc_set_positions(self.h, _p(x), _p(y), _p(phi)))

    def positions(self):
        x, y, phi = np.empty(self.ntot), np.empty(self.ntot), np.empty(self.ntot)
        self._ck(self.lib.sc_get_positions(self.h, _p(x), _p(y), _p(phi)))
        return x, y, phi

    def forces(self):
        x, y, t = np.empty(self.ntot), np.empty(self.ntot), np.empty(self.ntot)
        self._ck(self.lib.sc_get_forces(self.h, _p(x), _p(y), _p(t)))
        return x, y, t

    def box(self, k=0):
        out = np.zeros(CONST["SC_BOX_NFIELDS"])
        self._ck(self.lib.sc_get_box(self.h, k, _p(out)))
        return dict(zip(BOX_FIELDS, out.tolist()))

    def calc_forces(self):
        self._ck(self.lib.sc_calc_forces(self.h))

    def calc_se(self):
        self._ck(self.lib.sc_calc_se(self.h))

    def strain_steps(self, nsteps, se_every=0):
        nrec = nsteps // se_every if se_every > 0 else 0
        out = np.zeros(max(1, nrec * self.npack * 9))
        self._ck(self.lib.sc_strain_steps(self.h, nsteps, se_every, _p(out)))
        return out[:nrec * self.npack * 9].reshape(nrec, self.npack, 9)

    def resize_steps(self, nsteps, shrink=True, corrected_image=False, se_every=0):
        nrec = nsteps // se_every if se_every > 0 else 0
        out = np.zeros(max(1, nrec * self.npack * 9))
        self._ck(self.lib.sc_resize_steps(self.h, nsteps, int(shrink), int(corrected_image),
                                          se_every, _p(out)))
        return out[:nrec * self.npack * 9].reshape(nrec, self.npack, 9)

    def pure_strain_steps(self, nsteps, se_every=0):
        nrec = nsteps // se_every if se_every > 0 else 0
        out = np.zeros(max(1, nrec * self.npack * 9))
        self._ck(self.lib.sc_pure_strain_steps(self.h, nsteps, se_every, _p(out)))
        return out[:nrec * self.npack * 9].reshape(nrec, self.npack, 9)

    def flip_shear(self, k=-1):
        self._ck(self.lib.sc_flip_shear(self.h, k))

    def rotate_by_gamma(self, k=-1):
        self._ck(self.lib.sc_rotate_by_gamma(self.h, k))

    def gd_steps(self, nmax, emin=0.0):
        done = C.c_int()
        self._ck(self.lib.sc_gd_steps(self.h, nmax, emin, C.byref(done)))
        return done.value

    def affine_resize(self, scale, y_only=False):
        s = _f64(np.broadcast_to(np.asarray(scale, dtype=np.float64), (self.npack,)))
        fn = self.lib.sc_affine_resize_y if y_only else self.lib.sc_affine_resize
        self._ck(fn(self.h, _p(s)))

    def fire(self, params=None, switch_frms=1e-6, check=True):
        p = fire_params() if params is None else _f64(params)
        it = np.zeros(self.npack, dtype=np.int32)
        fr = np.zeros(self.npack)
        st = self.lib.sc_fire_minimize(self.h, _p(p), switch_frms, it.ctypes.data_as(_ip), _p(fr))
        if check:
            self._ck(st)
        return it, fr, st

    def _pairs(self, fn, k):
        cnt = C.c_int()
        self._ck(fn(self.h, k, None, None, 0, C.byref(cnt)))
        n = cnt.value
        p = np.empty(max(n, 1), dtype=np.int32)
        q = np.empty(max(n, 1), dtype=np.int32)
        self._ck(fn(self.h, k, p.ctypes.data_as(_ip), q.ctypes.data_as(_ip), n, C.byref(cnt)))
        return p[:n], q[:n]

    def neighbors(self, k=0):
        return self._pairs(self.lib.sc_get_neighbors, k)

    def contacts(self, k=0):
        return self._pairs(self.lib.sc_get_contacts, k)

    def timing(self, enable=True):
        self._ck(self.lib.sc_timing_enable(self.h, int(enable)))

    def kernel_time(self, kernel):
        ms, n = C.c_double(), C.c_longlong()
        self._ck(self.lib.sc_timing_get(self.h, kernel, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def timing_reset(self):
        self._ck(self.lib.sc_timing_reset(self.h))

    def synchronize(self):
        self._ck(self.lib.sc_synchronize(self.h))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _pf(a):
    return a.ctypes.data_as(_fp)


def _pi(a):
    return a.ctypes.data_as(_ip)


class Graph:
    """Contact graph in CSR over receivers (host arrays)."""

    def __init__(self, node_feat, offsets, src, edge_feat):
        self.node_feat = _f32(node_feat).reshape(-1, 4)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int32)
        self.src = np.ascontiguousarray(src, dtype=np.int32)
        self.edge_feat = _f32(edge_feat).reshape(-1, 5)

    @property
    def n(self):
        return self.node_feat.shape[0]

    @property
    def e(self):
        return self.src.shape[0]

    @staticmethod
    def concat(graphs):
        nf = np.concatenate([g.node_feat for g in graphs])
        ef = np.concatenate([g.edge_feat for g in graphs])
        src, offs, base, ebase = [], [np.zeros(1, dtype=np.int32)], 0, 0
        for g in graphs:
            src.append(g.src + base)
            offs.append(g.offsets[1:] + ebase)
            base += g.n
            ebase += g.e
        return Graph(nf, np.concatenate(offs), np.concatenate(src), ef)


class Mpnn:
    """The message-passing warm-start network (sc_mpnn)."""

    def __init__(self, device=0, seed=1):
        self.lib = load()
        h = C.c_void_p()
        st = self.lib.sc_mpnn_create(C.byref(h), device, C.c_ulonglong(seed))
        if st != 0:
            raise SpherocylsError(st, "sc_mpnn_create")
        self.h = h
        n = C.c_int()
        self._ck(self.lib.sc_mpnn_num_params(self.h, C.byref(n)))
        self.num_params = n.value

    def _ck(self, st):
        if st != 0:
            msg = self.lib.sc_mpnn_last_error(self.h) if getattr(self, "h", None) else b""
            raise SpherocylsError(st, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            self.lib.sc_mpnn_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def params(self):
        out = np.empty(self.num_params, dtype=np.float32)
        self._ck(self.lib.sc_mpnn_get_params(self.h, _pf(out)))
        return out

    def set_params(self, p):
        p = _f32(p)
        assert p.shape == (self.num_params,)
        self._ck(self.lib.sc_mpnn_set_params(self.h, _pf(p)))

    def grads(self):
        out = np.empty(self.num_params, dtype=np.float32)
        self._ck(self.lib.sc_mpnn_get_grads(self.h, _pf(out)))
        return out

    def set_output_scale(self, scale):
        s = _f32(scale)
        self._ck(self.lib.sc_mpnn_set_output_scale(self.h, _pf(s)))

    def output_scale(self):
        s = np.zeros(3, dtype=np.float32)
        self._ck(self.lib.sc_mpnn_get_output_scale(self.h, _pf(s)))
        return s

    def load_graph(self, g):
        self._ck(self.lib.sc_mpnn_load_graph(self.h, g.n, _pf(g.node_feat), g.e, _pi(g.offsets),
                                             _pi(g.src), _pf(g.edge_feat)))
        self._n = g.n

    def forward(self):
        out = np.empty((self._n, 3), dtype=np.float32)
        self._ck(self.lib.sc_mpnn_forward(self.h, _pf(out)))
        return out

    def backward(self, target):
        t = _f32(target).reshape(self._n, 3)
        loss = C.c_float()
        self._ck(self.lib.sc_mpnn_backward(self.h, _pf(t), C.byref(loss)))
        return loss.value

    def adam(self, lr, step, beta1=0.9, beta2=0.999, eps=1e-8):
        self._ck(self.lib.sc_mpnn_adam(self.h, lr, beta1, beta2, eps, step))

    def graph_of(self, system):
        """Contact graph of a System's current state, in caller order."""
        n, e = C.c_int(), C.c_int()
        self._ck(self.lib.sc_mpnn_graph_from_system(self.h, system.h, C.byref(n), C.byref(e)))
        nf = np.empty((n.value, 4), dtype=np.float32)
        off = np.empty(n.value + 1, dtype=np.int32)
        src = np.empty(max(e.value, 1), dtype=np.int32)
        ef = np.empty((max(e.value, 1), 5), dtype=np.float32)
        self._ck(self.lib.sc_mpnn_export_graph(self.h, system.h, _pf(nf), _pi(off), _pi(src),
                                               _pf(ef)))
        self._n = n.value
        return Graph(nf, off, src[:e.value], ef[:e.value])

    def warm_start(self, system):
        self._ck(self.lib.sc_mpnn_warm_start(self.h, system.h))

    def warm_start_damped(self, system, lambdas=(1.0, 0.5, 0.25)):
        """Damped proposal chosen per packing on the device; returns (energy
        before, energy after, factor kept) per packing."""
        lam = _f64(lambdas)
        e0, e1, lk = (np.zeros(system.npack) for _ in range(3))
        self._ck(self.lib.sc_mpnn_warm_start_damped(self.h, system.h, len(lam), _p(lam), _p(e0),
                                                    _p(e1), _p(lk)))
        return e0, e1, lk
