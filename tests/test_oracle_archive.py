"""Pins of the archive restatement (oracle/archive.py) on the CPU.

* Byte layouts assembled by hand from the format of util/archive.cpp:4-83,
  model.cpp:54-75, group.cpp:31-59, varying_manager.cpp:21-45,
  uniform_manager.cpp:22-41, component_type.cpp:23-38, shape.cpp:9-20 for
  the model of model.test.cpp:70-103 and for one field of each kind: a
  dropped count, a swapped order or a wrong type string fails them.
* The reference's round-trip tests with their known answers:
  varying_manager.test.cpp:87-157, varying_field.test.cpp:10-66,
  uniform_field.test.cpp:8-31, group.test.cpp:128-160 (the matrix entry
  (r, c) of an R x C item is component r + R c, Eigen's column-major
  storage).
* Every proper prefix of an archive is rejected, and names are written in
  ascending byte order whatever order they are given in.
"""

import struct

import numpy as np
import pytest

from oracle import archive as A


def sz(v):
    return struct.pack("=Q", v)


def st(s):
    b = s.encode()
    return sz(len(b)) + b


def ty(ctype, shape):
    return st(ctype) + sz(len(shape)) + b"".join(sz(e) for e in shape)


def test_model_test_bytes_by_hand():
    # model.test.cpp:70-85: groups a (a_type), b (b_type), global bool[1][2]
    # bx1x2 (new fields are zero-filled)
    model = {"globals": {"bx1x2": A.field("b", [1, 2], [False, False])},
             "groups": {"b": {"type": "b_type", "tags": [], "items": 0},
                        "a": {"type": "a_type", "tags": [], "items": 0}}}
    expect = (sz(1) + st("bx1x2") + ty("b", [1, 2]) + b"\x00\x00"
              + sz(2)
              + st("a") + st("a_type") + sz(0) + sz(0) + sz(0) + sz(0)
              + st("b") + st("b_type") + sz(0) + sz(0) + sz(0) + sz(0))
    got = A.write_model(model)
    assert got == expect
    back = A.read_model(expect)
    assert back["groups"]["a"]["type"] == "a_type"
    assert back["groups"]["b"]["type"] == "b_type"
    assert list(back["globals"]) == ["bx1x2"]
    assert back["globals"]["bx1x2"]["shape"] == [1, 2]
    assert back["globals"]["bx1x2"]["data"].dtype == np.bool_


def test_group_field_bytes_by_hand():
    model = {"globals": {},
             "groups": {"f": {"type": "fluid", "tags": ["visible", "dynamic"], "items": 2,
                              "varying": {"x": A.field("f32", [3], [[1, 2, 3], [4, 5, 6]]),
                                          "m": A.field("f64", [], [0.5, 0.25])},
                              "uniform": {"rho": A.field("s32", [], [1000])}}}}
    expect = (sz(0) + sz(1) + st("f") + st("fluid")
              + sz(2) + st("dynamic") + st("visible")
              + sz(2) + sz(2)
              + st("m") + ty("f64", []) + sz(2) + struct.pack("=2d", 0.5, 0.25)
              + st("x") + ty("f32", [3]) + sz(2) + struct.pack("=6f", 1, 2, 3, 4, 5, 6)
              + sz(1) + st("rho") + ty("s32", []) + struct.pack("=i", 1000))
    assert A.write_model(model) == expect
    back = A.read_model(expect)["groups"]["f"]
    assert back["tags"] == ["dynamic", "visible"] and back["items"] == 2
    assert np.array_equal(back["varying"]["x"]["data"], [[1, 2, 3], [4, 5, 6]])
    assert back["varying"]["m"]["data"].dtype == np.float64
    assert int(back["uniform"]["rho"]["data"][0]) == 1000


def _item(R, C, r, c):
    return r + R * c  # Eigen column-major


def test_varying_manager_known_answers():
    # varying_manager.test.cpp:87-157
    n = 5
    b = np.zeros((n, 2), bool)
    s = np.zeros((n, 3), np.int32)
    d = np.zeros((n, 6), np.float64)
    for i, (bv, sv) in enumerate(zip([True, False, True, False, True], [12, 34, 56, 78, 90])):
        b[i, _item(1, 2, 0, 1)] = bv
        s[i, _item(3, 1, 2, 0)] = sv
        d[i, _item(2, 3, 1, 1)] = float(sv)
    model = {"groups": {"g": {"type": "t", "items": n, "varying": {
        "bx1x2": A.field("b", [1, 2], b), "s32x3x1": A.field("s32", [3, 1], s),
        "f64x2x3": A.field("f64", [2, 3], d)}}}}
    back = A.read_model(A.write_model(model))["groups"]["g"]
    assert back["items"] == 5 and len(back["varying"]) == 3
    assert list(back["varying"]["bx1x2"]["data"][:, 1]) == [True, False, True, False, True]
    assert list(back["varying"]["s32x3x1"]["data"][:, 2]) == [12, 34, 56, 78, 90]
    assert list(back["varying"]["f64x2x3"]["data"][:, 3]) == [12.0, 34.0, 56.0, 78.0, 90.0]


def test_varying_field_known_answers():
    # varying_field.test.cpp:10-66: an empty bool field, then int64 [3][2]
    model = {"groups": {"g": {"type": "t", "items": 0,
                              "varying": {"e": A.field("b", [], np.zeros(0, bool))}}}}
    back = A.read_model(A.write_model(model))["groups"]["g"]["varying"]["e"]
    assert back["data"].shape == (0, 1)
    v = np.zeros((3, 6), np.int64)
    v[0, _item(3, 2, 0, 0)] = 123
    v[1, _item(3, 2, 1, 1)] = 456
    v[2, _item(3, 2, 2, 0)] = 789
    model = {"groups": {"g": {"type": "t", "items": 3,
                              "varying": {"v": A.field("s64", [3, 2], v)}}}}
    back = A.read_model(A.write_model(model))["groups"]["g"]["varying"]["v"]["data"]
    assert back.shape == (3, 6)
    assert back[0, 0] == 123 and back[1, 4] == 456 and back[2, 2] == 789


def test_uniform_and_group_known_answers():
    # uniform_field.test.cpp:8-31 and group.test.cpp:128-160
    model = {"globals": {"u": A.field("s64", [], [1234])},
             "groups": {"g": {"type": "t", "items": 102,
                              "varying": {"a": A.field("f32", [1, 2], np.zeros((102, 2)))},
                              "uniform": {"b": A.field("b", [1, 2], [False, False])}}}}
    back = A.read_model(A.write_model(model))
    assert int(back["globals"]["u"]["data"][0]) == 1234
    g = back["groups"]["g"]
    assert g["items"] == 102
    assert g["varying"]["a"]["ctype"] == "f32" and g["varying"]["a"]["shape"] == [1, 2]
    assert g["uniform"]["b"]["ctype"] == "b" and g["uniform"]["b"]["shape"] == [1, 2]


def test_every_prefix_is_rejected_and_trailing_bytes_too():
    model = {"globals": {"t": A.field("f32", [], [0.5])},
             "groups": {"f": {"type": "fluid", "tags": ["dynamic"], "items": 2,
                              "varying": {"x": A.field("f32", [3], np.ones((2, 3)))},
                              "uniform": {"r": A.field("f32", [], [1.0])}}}}
    buf = A.write_model(model)
    for k in range(len(buf)):
        with pytest.raises((EOFError, ValueError)):
            A.read_model(buf[:k])
    with pytest.raises(ValueError):
        A.read_model(buf + b"\x00")


def test_names_are_written_in_ascending_byte_order():
    fields = {n: A.field("s32", [], [i]) for i, n in enumerate(["z", "B", "a", "_x", "A1"])}
    buf = A.write_model({"globals": fields, "groups": {}})
    names = list(A.read_model(buf)["globals"])
    assert names == sorted(fields, key=lambda s: s.encode())
    assert names == ["A1", "B", "_x", "a", "z"]
