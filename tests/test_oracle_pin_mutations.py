"""The oracle's pins reject plausible mistakes.

Each case rebuilds the oracle from its own sources with one deliberate change (a swapped cross
product, a gradient stencil of the wrong sign, a one-sided difference reversed, the black colour
swept first, omega on the wrong term, an open face not counted), loads that build in place of the
real one, and runs the pin that must catch it: the pin has to fail.  The unchanged build passes the
same pins in test_oracle_force.py / test_oracle_project.py.  CPU only (g++)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import zq_oracle as O
import test_oracle_force as F
import test_oracle_project as P

ORACLE = os.path.dirname(os.path.abspath(O.__file__))
SOURCES = ("zq_oracle.cpp", "zq_oracle_coarse.cpp", "zq_oracle_octree.cpp")

# (name, exact text in zq_oracle.cpp, replacement, pin that must fail)
MUTANTS = [
    ("cross_product_w_x_N",
     "float fx = ch * (gy * wz - gz * wy);\n                float fy = ch * (gz * wx - gx * wz);\n"
     "                float fz = ch * (gx * wy - gy * wx);",
     "float fx = ch * (wy * gz - wz * gy);\n                float fy = ch * (wz * gx - wx * gz);\n"
     "                float fz = ch * (wx * gy - wy * gx);",
     "confinement"),
    ("central_gradient_sign",
     "else gx = 0.5f * (vs[o + 1] - vs[o - 1]);", "else gx = 0.5f * (vs[o - 1] - vs[o + 1]);", "confinement"),
    ("one_sided_edge_reversed",
     "if (i == 0) gx = vs[o + 1] - vs[o];", "if (i == 0) gx = vs[o] - vs[o + 1];", "confinement"),
    ("black_first",
     "for (int i = (color + j + k) & 1; i < W; i += 2) {", "for (int i = (1 + color + j + k) & 1; i < W; i += 2) {",
     "sor"),
    ("omega_on_wrong_term",
     "p[q] = omw * p[q] + omega * pg;", "p[q] = omega * p[q] + omw * pg;", "sor"),
    ("open_face_not_counted",
     "n += (j < H - 1) ? !occ[o + W] : ((open_mask & ORC_OPEN_YHI) != 0);", "n += (j < H - 1) ? !occ[o + W] : 0;",
     "sor"),
]


def _build_mutant(tmp, old, new):
    for f in SOURCES + ("zq_oracle.h",):
        shutil.copy(os.path.join(ORACLE, f), tmp)
    src = os.path.join(tmp, "zq_oracle.cpp")
    text = open(src).read()
    assert text.count(old) == 1, "mutation site not found exactly once: %r" % old
    open(src, "w").write(text.replace(old, new))
    out = os.path.join(tmp, "libmutant.so")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fopenmp", "-shared", "-o", out]
                          + [os.path.join(tmp, f) for f in SOURCES])
    return out


def _confinement_pins():
    for comp, axis in F._CASES:
        F.test_confinement_closed_form(comp, axis)


def _sor_pins():
    rng = np.random.default_rng(7)
    for shape in ((4, 4, 4), (5, 3, 4)):
        for open_mask, frac in ((0, 0.0), (63, 0.15), (O.OPEN_YHI | O.OPEN_XLO, 0.2)):
            for omega in (1.0, 1.8):
                P.test_sor_iteration_equals_dense_red_black_sor(rng, shape, open_mask, frac, omega, 1)


PINS = {"confinement": _confinement_pins, "sor": _sor_pins}


def test_unchanged_oracle_passes_the_pins():
    for pin in PINS.values():
        pin()


@pytest.mark.parametrize("name,old,new,pin", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_pin_rejects_mutant(tmp_path, monkeypatch, name, old, new, pin):
    path = _build_mutant(str(tmp_path), old, new)
    monkeypatch.setattr(O, "_lib", O.load(path))
    with pytest.raises(AssertionError):
        PINS[pin]()
