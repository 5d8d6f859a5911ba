"""Pins of the VTK restatement (oracle/vtk.py) on the CPU.

* fixed6 equals the exact decimal rounding of the binary value, computed with
  rational arithmetic (round half to even at the sixth decimal, sign kept),
  for the adversarial values of tests/vtk_values.py in float and double;
* the text of a small group assembled by hand from save_vtk.cpp:57-129;
* the C++ interface's SaveVTK (libstdc++'s std::fixed, the library the
  reference prints with) writes the same bytes, non-finite values included.
"""

import os
import subprocess
from fractions import Fraction

import numpy as np
import pytest

from oracle import vtk as V
from vtk_values import adversarial


def exact_fixed6(v) -> str:
    v = float(v)
    neg = np.signbit(v)
    q = Fraction(abs(v)) * 10 ** 6
    n = q.numerator // q.denominator
    r = q - n
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and n % 2 == 1):
        n += 1
    return ("-" if neg else "") + f"{n // 10 ** 6}.{n % 10 ** 6:06d}"


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fixed6_is_the_exact_rounding(dtype):
    for v in adversarial(3000, dtype, seed=1):
        assert V.fixed6(v) == exact_fixed6(v), repr(v)


def test_fixed6_known_answers():
    assert V.fixed6(1 / 128) == "0.007812"  # 7812.5 -> even
    assert V.fixed6(3 / 128) == "0.023438"  # 23437.5 -> even
    assert V.fixed6(np.float32(0.1)) == "0.100000"
    assert V.fixed6(-0.0) == "-0.000000"
    assert V.fixed6(-1e-9) == "-0.000000"
    assert V.fixed6(np.float32(0.9999995)) == "1.000000"
    assert V.fixed6(2.0 ** 70) == str(2 ** 70) + ".000000"
    assert V.fixed6(np.finfo(np.float32).max) == str(int(np.finfo(np.float32).max)) + ".000000"
    assert V.fixed6(float("nan")) == "nan" and V.fixed6(-float("nan")) == "-nan"
    assert V.fixed6(float("-inf")) == "-inf"


def test_small_group_by_hand():
    f = {"position": np.array([[0.5, 1, 2], [0.25, -1, 3]], np.float32),
         "velocity": np.zeros((2, 3), np.float32),
         "mass": np.array([1, 2], np.float32),
         "time_of_birth": np.array([-1, 0], np.float32)}
    expect = ("# vtk DataFile Version 2.0\nf fluid\nASCII\nDATASET POLYDATA\n"
              "POINTS 2 float\n0.500000 1.000000 2.000000\n0.250000 -1.000000 3.000000\n"
              "POINT_DATA 2\n"
              "VECTORS velocity float\n0.000000 0.000000 0.000000\n0.000000 0.000000 0.000000\n"
              "SCALARS _index float 1\nLOOKUP_TABLE default\n0\n1\n"
              "SCALARS mass float 1\nLOOKUP_TABLE default\n1.000000\n2.000000\n"
              "SCALARS time_of_birth float 1\nLOOKUP_TABLE default\n-1.000000\n0.000000\n")
    assert V.write_vtk("f", "fluid", f) == expect.encode()


FIELDS = (("position", 3), ("velocity", 3), ("acceleration", 3), ("mass", 1), ("density", 1),
          ("pressure", 1), ("time_of_birth", 1))


def vtk_fields(n, dtype, seed, nonfinite):
    out = {}
    for k, (name, c) in enumerate(FIELDS):
        out[name] = adversarial(n * c, dtype, seed + k, nonfinite).reshape(n, c) if c == 3 \
            else adversarial(n, dtype, seed + k, nonfinite)
    return out


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_compat_savevtk_matches_restatement(tmp_path, dtype):
    from test_compat import build_check
    n = 700
    f = vtk_fields(n, dtype, 11, nonfinite=True)
    src = tmp_path / "in.bin"
    with open(src, "wb") as fh:
        np.array([1 if dtype == np.float64 else 0, n], np.int64).tofile(fh)
        for name, _ in FIELDS:
            f[name].tofile(fh)
    exe = build_check(tmp_path)
    out = tmp_path / "out.vtk"
    r = subprocess.run([exe, "vtk", str(src), str(out)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert out.read_bytes() == V.write_vtk("f", "fluid", f)
