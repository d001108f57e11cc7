"""Seeded synthetic packings for the benchmark and parity workloads.

The reference only offers an unseeded ``random_configuration``
(src/cpp/spherocyl_box.cpp:694-790, ``srand(time(0))`` at :739), so every
workload here is generated from an explicit numpy PCG64 seed instead.

Conventions (kept identical to the reference's input handling):
  * ``A`` is the half-shaft length, ``R`` the radius; the aspect ratio quoted
    for a workload is (R + A) / R, so "aspect 2" means A = R.
  * The box side follows src/cpp/main.cpp:121-122: L = sqrt(area / phi) with the
    particle area R (pi R + 4 A) summed over particles, using the reference's
    truncated pi (src/cpp/spherocyl_box.cpp:12).
  * Centres are uniform in [0, L)^2, angles uniform in [0, pi), placed one by
    one and redrawn while a segment comes within sqrt(1e-5)(R_p + R_q) of an
    earlier one (the reference's RANDOM mode, check_particle_cross at
    src/cpp/spherocyl_box.cpp:597-645); crossing segments would give a
    zero-length contact vector and a 0/0 force.  The placement itself runs in
    libspherocyls.so (sc_random_packing, host code).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

REF_PI = 3.14159265358979


@dataclasses.dataclass
class Packing:
    L: float
    x: np.ndarray
    y: np.ndarray
    phi: np.ndarray
    A: np.ndarray
    R: np.ndarray
    packing_fraction: float
    name: str = ""

    @property
    def n(self) -> int:
        return int(self.x.shape[0])


def box_side(R: np.ndarray, A: np.ndarray, phi_target: float) -> float:
    area = float(np.sum((REF_PI * R + 4.0 * A) * R))
    return math.sqrt(area / phi_target)


def make_packing(R, A, phi_target, seed, name="", gamma=0.0) -> Packing:
    """Random non-crossing packing; positions are sheared coordinates at strain
    ``gamma`` (the Lees-Edwards images at that strain are crossing-free too)."""
    from . import native
    R = np.ascontiguousarray(R, dtype=np.float64)
    A = np.ascontiguousarray(A, dtype=np.float64)
    L = box_side(R, A, phi_target)
    x, y, ang = native.random_packing(L, A, R, seed, gamma=gamma)
    return Packing(L, x, y, ang, A, R, phi_target, name)


def bidisperse(n: int, r_small=0.5, r_large=0.7, aspect=2.0):
    """50/50 radii by index parity; A = (aspect - 1) R."""
    R = np.where(np.arange(n) % 2 == 0, r_small, r_large).astype(np.float64)
    A = (aspect - 1.0) * R
    return R, A


def shear_small(seed=1) -> Packing:
    """N=512 bidisperse (R=0.5/0.7), aspect 2, phi=0.85 -- sheared Heun workload."""
    R, A = bidisperse(512)
    return make_packing(R, A, 0.85, seed, "shear N=512 phi=0.85 aspect 2")


def compress_start(seed=1, n=16384) -> Packing:
    """N=16384 bidisperse aspect-2 packing at phi=0.80 (quasi-static compression start)."""
    R, A = bidisperse(n)
    return make_packing(R, A, 0.80, seed, f"compress N={n} phi=0.80 aspect 2")


def shear_large(seed=1, n=262144) -> Packing:
    """N=262144, R uniform in [0.5, 0.7], aspect 1.5 (A = 0.5 R), phi=0.845."""
    rng = np.random.Generator(np.random.PCG64(seed + 1_000_003))
    R = 0.5 + 0.2 * rng.random(n)
    A = 0.5 * R
    return make_packing(R, A, 0.845, seed, f"shear N={n} phi=0.845 aspect 1.5")


def ensemble(count=256, seed=1, n_min=64, n_max=4096):
    """Ragged ensemble: N log-uniform in [n_min, n_max], phi U[0.80, 0.86],
    aspect drawn from {1.2, 2, 4}, radii bidisperse 1:1.4 (0.5 / 0.7)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(count):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        phi = float(rng.uniform(0.80, 0.86))
        aspect = float(rng.choice([1.2, 2.0, 4.0]))
        R, A = bidisperse(n, aspect=aspect)
        out.append(make_packing(R, A, phi, seed * 100003 + k,
                                f"ensemble[{k}] N={n} phi={phi:.4f} aspect {aspect}"))
    return out


def gamma_aspect(n, aspect_mean=2.0, rel_sigma=0.5, phi=0.84, seed=1, bidispersity=1.4):
    """Packing of gamma-distributed aspect ratios A/R (mean aspect_mean, relative
    width rel_sigma; the reference's set_shapes with an aspect width), radii
    0.5 bidispersity / 0.5 by index parity, box sized for phi with the drawn
    shapes."""
    from . import native
    R, A = native.gamma_shapes(n, aspect_mean, rel_sigma, bidispersity, seed)
    return make_packing(R, A, phi, seed, f"gamma aspect N={n} mean {aspect_mean} "
                                       f"sigma {rel_sigma} phi={phi}")
