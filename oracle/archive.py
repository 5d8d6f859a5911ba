"""ORACLE — test infrastructure only.

numpy restatement of the reference's native binary model archive, the format
of Model::Save / Model::Load over NativeBinaryArchiveWriter / Reader. Used by
tests/ only, as the checker of the archives written by the C++ host
interface and by the C ABI's scene checkpoint.

The layout, in the order the reference writes it:

  size      8 bytes, a size_t in the machine's byte order
            (archive.cpp:4-13 SaveSize / LoadSize)
  string    size, then the bytes (archive.cpp:15-25)
  values    each value's native bytes, one after the other; bool is one
            byte (archive.cpp:27-83)
  type      string component type "b" / "s32" / "s64" / "f32" / "f64"
            (component_type.cpp:15-38), then size rank and rank sizes of
            extents (shape.cpp:9-20; tensor_type.cpp:7-15)
  uniform manager   size field count, then per field in ascending name
            order: string name, type, the item's components
            (uniform_manager.cpp:22-41; uniform_field.hpp Save/Load)
  varying manager   size item count, size field count, then per field in
            ascending name order: string name, type, size count, count x
            components (varying_manager.cpp:21-45; varying_field.hpp:224-250)
  group     size tag count, tags in ascending order, varying manager,
            uniform manager (group.cpp:31-59)
  model     global uniform manager, size group count, then per group in
            ascending name order: string name, string type, group
            (model.cpp:54-75)

The components of a matrix item are Eigen's storage, column-major: entry
(r, c) of an R x C item is component r + R c. The reference's reader
(share/python/prtcl_archive.py:8-112) walks the same layout; it maps "b" to
int8, this restatement to bool (same bytes).

A model here is a dict:
  {"globals": {name: Field}, "groups": {name: Group}}
  Field = {"ctype": str, "shape": [int], "data": ndarray}
      uniform: data has the item's components (flat, column-major);
      varying: data has shape (count, extent).
  Group = {"type": str, "tags": [str], "items": int,
           "varying": {name: Field}, "uniform": {name: Field}}
"""

from __future__ import annotations

import struct

import numpy as np

DTYPES = {
    "b": np.dtype(np.bool_),
    "s32": np.dtype(np.int32),
    "s64": np.dtype(np.int64),
    "f32": np.dtype(np.float32),
    "f64": np.dtype(np.float64),
}


def ctype_of(dtype) -> str:
    dtype = np.dtype(dtype)
    for k, v in DTYPES.items():
        if v == dtype:
            return k
    raise ValueError(f"no archive component type for {dtype}")


def extent(shape) -> int:
    n = 1
    for e in shape:
        n *= int(e)
    return n


class _Reader:
    def __init__(self, buf: bytes):
        self.buf = memoryview(buf)
        self.pos = 0

    def take(self, n: int) -> memoryview:
        if self.pos + n > len(self.buf):
            raise EOFError("archive: unexpected end of data")
        out = self.buf[self.pos:self.pos + n]
        self.pos += n
        return out

    def size(self) -> int:  # archive.cpp:10-13
        return struct.unpack("=Q", self.take(8))[0]

    def string(self) -> str:  # archive.cpp:20-25
        n = self.size()
        return bytes(self.take(n)).decode("utf-8")

    def type(self):  # tensor_type.cpp:12-15, shape.cpp:15-20
        ctype = self.string()
        if ctype not in DTYPES:
            raise ValueError(f"archive: invalid component type {ctype!r}")
        rank = self.size()
        return ctype, [self.size() for _ in range(rank)]

    def values(self, ctype: str, count: int) -> np.ndarray:
        dt = DTYPES[ctype]
        return np.frombuffer(bytes(self.take(count * dt.itemsize)), dtype=dt).copy()


def _read_uniforms(r: _Reader) -> dict:  # uniform_manager.cpp:31-41
    out = {}
    for _ in range(r.size()):
        name = r.string()
        ctype, shape = r.type()
        out[name] = {"ctype": ctype, "shape": shape, "data": r.values(ctype, extent(shape))}
    return out


def _read_varyings(r: _Reader):  # varying_manager.cpp:32-45
    items = r.size()
    out = {}
    for _ in range(r.size()):
        name = r.string()
        ctype, shape = r.type()
        count = r.size()  # varying_field.hpp:238-250
        data = r.values(ctype, count * extent(shape)).reshape(count, extent(shape))
        out[name] = {"ctype": ctype, "shape": shape, "data": data}
    return items, out


def read_model(buf: bytes) -> dict:
    """Parses a whole model archive (model.cpp:65-75); raises EOFError on
    truncated input and ValueError on bytes left over."""
    r = _Reader(buf)
    model = {"globals": _read_uniforms(r), "groups": {}}
    for _ in range(r.size()):
        name = r.string()
        gtype = r.string()
        tags = [r.string() for _ in range(r.size())]  # group.cpp:50-55
        items, varying = _read_varyings(r)
        uniform = _read_uniforms(r)
        model["groups"][name] = {"type": gtype, "tags": tags, "items": items,
                                 "varying": varying, "uniform": uniform}
    if r.pos != len(r.buf):
        raise ValueError(f"archive: {len(r.buf) - r.pos} bytes after the model")
    return model


class _Writer:
    def __init__(self):
        self.parts = []

    def size(self, v: int):  # archive.cpp:4-8
        self.parts.append(struct.pack("=Q", int(v)))

    def string(self, s: str):  # archive.cpp:15-18
        b = s.encode("utf-8")
        self.size(len(b))
        self.parts.append(b)

    def type(self, ctype: str, shape):  # tensor_type.cpp:7-10, shape.cpp:9-13
        self.string(ctype)
        self.size(len(shape))
        for e in shape:
            self.size(e)

    def values(self, ctype: str, a):
        self.parts.append(np.ascontiguousarray(a, dtype=DTYPES[ctype]).tobytes())


def _write_uniforms(w: _Writer, fields: dict):  # uniform_manager.cpp:22-29
    w.size(len(fields))
    for name in sorted(fields, key=lambda s: s.encode("utf-8")):
        f = fields[name]
        w.string(name)
        w.type(f["ctype"], f["shape"])
        data = np.asarray(f["data"]).reshape(-1)
        assert data.size == extent(f["shape"]), name
        w.values(f["ctype"], data)


def _write_varyings(w: _Writer, items: int, fields: dict):  # varying_manager.cpp:21-30
    w.size(items)
    w.size(len(fields))
    for name in sorted(fields, key=lambda s: s.encode("utf-8")):
        f = fields[name]
        w.string(name)
        w.type(f["ctype"], f["shape"])
        data = np.asarray(f["data"]).reshape(-1, extent(f["shape"]))
        w.size(data.shape[0])  # varying_field.hpp:224-236
        w.values(f["ctype"], data)


def write_model(model: dict) -> bytes:
    """Serialises a model dict (model.cpp:54-63): names, tags and fields in
    ascending byte order, as the reference's sorted containers iterate."""
    w = _Writer()
    _write_uniforms(w, model.get("globals", {}))
    groups = model.get("groups", {})
    w.size(len(groups))
    for name in sorted(groups, key=lambda s: s.encode("utf-8")):
        g = groups[name]
        w.string(name)
        w.string(g["type"])
        tags = sorted(set(g.get("tags", [])), key=lambda s: s.encode("utf-8"))
        w.size(len(tags))
        for t in tags:
            w.string(t)
        _write_varyings(w, g["items"], g.get("varying", {}))
        _write_uniforms(w, g.get("uniform", {}))
    return b"".join(w.parts)


def field(ctype: str, shape, data) -> dict:
    """A field record; data is flattened per item (column-major for
    matrices, as the archive stores it)."""
    return {"ctype": ctype, "shape": list(shape), "data": np.asarray(data, DTYPES[ctype])}
