"""ORACLE — test infrastructure only.

Restatement of the reference's legacy-VTK writer, SaveVTK
(src/prtcl/util/save_vtk.cpp:57-129), used by tests/ as the checker of the
text written by the C++ interface's SaveVTK and by the C ABI's
prtcl_hip_scene_save_vtk.

The text, in the order the reference writes it:

  # vtk DataFile Version 2.0        (:58)
  <group name> <group type>         (:59)
  ASCII / DATASET POLYDATA          (:60-61)
  POINTS <n> float, then one line per item of "position"        (:67-68)
  POINT_DATA <n>                                                (:70)
  VECTORS velocity float / acceleration float  if present       (:72-80)
  SCALARS _index float 1, LOOKUP_TABLE default, the item index  (:82-85)
  SCALARS <name> float 1, LOOKUP_TABLE default for mass, volume, density,
    pressure, time_of_birth, implicit_function if present       (:87-121)
  VECTORS implicit_function_gradient float if present           (:123-126)

A line holds an item's components separated by one space (:18-31). Reals
are printed with std::fixed at the default precision 6 (:14, :27), the C
library's "%.6f": the exact binary value rounded half to even, the sign kept
for -0 and for values that round to zero, and (glibc) "nan", "-nan", "inf",
"-inf" for the non-finite values. Python's "%.6f" is the same correctly
rounded conversion for finite values; the non-finite spellings are set here.
"""

from __future__ import annotations

import math

import numpy as np

SCALAR_ORDER = ("mass", "volume", "density", "pressure", "time_of_birth", "implicit_function")


def fixed6(v) -> str:
    v = float(v)
    if math.isnan(v):
        return "-nan" if math.copysign(1.0, v) < 0 else "nan"
    if math.isinf(v):
        return "-inf" if v < 0 else "inf"
    return "%.6f" % v


def _lines(a) -> str:
    a = np.asarray(a)
    a = a.reshape(a.shape[0], -1) if a.ndim > 1 else a.reshape(-1, 1)
    return "".join(" ".join(fixed6(x) for x in row) + "\n" for row in a)


def write_vtk(name: str, gtype: str, fields: dict) -> bytes:
    """fields: name -> array (n,) or (n, k) in host order; "position" is
    required."""
    n = len(fields["position"])
    out = ["# vtk DataFile Version 2.0\n", f"{name} {gtype}\n", "ASCII\n",
           "DATASET POLYDATA\n", f"POINTS {n} float\n", _lines(fields["position"]),
           f"POINT_DATA {n}\n"]
    for vec in ("velocity", "acceleration"):
        if vec in fields:
            out += [f"VECTORS {vec} float\n", _lines(fields[vec])]
    out += ["SCALARS _index float 1\n", "LOOKUP_TABLE default\n",
            "".join(f"{i}\n" for i in range(n))]
    for sc in SCALAR_ORDER:
        if sc in fields:
            out += [f"SCALARS {sc} float 1\n", "LOOKUP_TABLE default\n", _lines(fields[sc])]
    if "implicit_function_gradient" in fields:
        out += ["VECTORS implicit_function_gradient float\n",
                _lines(fields["implicit_function_gradient"])]
    return "".join(out).encode()
