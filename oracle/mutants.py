"""Planted-mistake check of the oracle -- test infrastructure only.

Each mutant is the restatement (sc_oracle.cpp) with one plausible mistake
planted in it: a flipped sign of a second-contact parameter of the two-point
model, its branch always or never taken, FIRE without its half-step retreat or
without its displacement clamp, a strict comparison made non-strict.  Each is
built on the CPU into a temporary directory and the CPU test suite runs
against it (SPHEROCYLS_ORACLE_SO); a mutant that passes every test is a
function of the oracle that its checks do not pin.

    python -m oracle.mutants        # prints one line per mutant, exits 1 if any survives
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

# (name, text in sc_oracle.cpp, replacement)
MUTANTS = [
    ("flip_s3", "s3 = b < 0 ? -t1 : t1;", "s3 = b < 0 ? t1 : -t1;"),
    ("flip_t3", "t3 = b < 0 ? -s1 : s1;", "t3 = b < 0 ? s1 : -s1;"),
    ("always_third", "if (dx3 * dx3 + dy3 * dy3 < d22) {", "if (true) {"),
    ("never_third", "if (dx3 * dx3 + dy3 * dy3 < d22) {", "if (false) {"),
    ("no_retreat", "      b->x[i] -= 0.5 * dt * vx[i];\n      b->y[i] -= 0.5 * dt * vy[i];\n"
                   "      b->phi[i] -= 0.5 * dt * vp[i] / sq;", ""),
    ("no_clamp", "if (m2 > fp.max_move * fp.max_move) {", "if (false) {"),
]

# Strict / non-strict at a projection's end: the two sides differ only where a
# projected parameter is exactly +-1 in floating point.  The geometries that
# give that by construction (axis-aligned rods, parallel or perpendicular)
# give the same contacts on both sides; others reach it only by a rounding
# coincidence.  Reported, not required to fail.
BOUNDARY = [
    ("targ_ge", "if (targ > 1) {", "if (targ >= 1) {"),
    ("sarg_ge", "if (sarg > 1) {\n        t2 = clamp1", "if (sarg >= 1) {\n        t2 = clamp1"),
]


def main() -> int:
    src = open(os.path.join(HERE, "sc_oracle.cpp")).read()
    survivors = []
    with tempfile.TemporaryDirectory() as tmp:
        for name, old, new in MUTANTS + BOUNDARY:
            boundary = (name, old, new) in BOUNDARY
            if src.count(old) < 1:  # the restatement changed: update the mutant
                print(f"{name}: pattern not found in sc_oracle.cpp", flush=True)
                survivors.append(name)
                continue
            path = os.path.join(tmp, f"{name}.cpp")
            with open(path, "w") as f:
                f.write(src.replace(old, new))
            lib = os.path.join(tmp, f"lib{name}.so")
            subprocess.run(["g++", "-O3", "-fPIC", "-ffp-contract=off", "-shared", "-I", HERE,
                            "-o", lib, path], check=True)
            r = subprocess.run([sys.executable, "-m", "pytest", "tests", "-m", "not gpu", "-x", "-q",
                                "-p", "no:cacheprovider"], cwd=ROOT, capture_output=True, text=True,
                               env=dict(os.environ, SPHEROCYLS_ORACLE_SO=lib))
            failed = [l for l in r.stdout.splitlines() if l.startswith("FAILED")]
            if r.returncode == 0:
                if not boundary:
                    survivors.append(name)
                print(f"{name}: {'survived (boundary)' if boundary else 'SURVIVED'}", flush=True)
            else:
                print(f"{name}: killed by {failed[0][7:] if failed else r.stdout[-200:]}", flush=True)
    return 1 if survivors else 0


if __name__ == "__main__":
    sys.exit(main())
