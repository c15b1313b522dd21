"""NRRD volume files: the dose, CT and WET exports of the plan layer.

The reference writes them with SimpleITK's ImageFileWriter
(DoseCUDA/plan.py:165-204, plan_impt.py:213-242) and reads CTs with its
ImageFileReader (plan.py:25-33).  SimpleITK is not a dependency here, so this
module implements the format itself, from the NRRD0004 specification
(teem.sourceforge.net/nrrd/format.html):

* a text header: the magic line, ``field: value`` lines, ``key:=value``
  lines and ``#`` comments, ended by an empty line, then the data;
* ``sizes`` lists the fastest axis first, so a C-ordered numpy array of shape
  (nz, ny, nx) is written as ``sizes: nx ny nz`` with its bytes unchanged;
* geometry as ``space: left-posterior-superior``, ``space directions`` (one
  vector per axis, the voxel step) and ``space origin`` (the first voxel's
  centre), which is how SimpleITK stores an image's origin and spacing.

The writer emits raw (default) or gzip encoding in little-endian order.  The
reader accepts what the writer emits plus the common variants of the format:
any of the scalar types, raw/gzip encoding, either endianness, ``spacings``
instead of ``space directions``, and the ``byte skip``/``line skip`` fields.
Detached data files are rejected with an error.
"""
from __future__ import annotations

import gzip
import re

import numpy as np

_TYPES = {
    np.dtype(np.int8): "signed char", np.dtype(np.uint8): "uchar",
    np.dtype(np.int16): "short", np.dtype(np.uint16): "ushort",
    np.dtype(np.int32): "int", np.dtype(np.uint32): "uint",
    np.dtype(np.int64): "longlong", np.dtype(np.uint64): "ulonglong",
    np.dtype(np.float32): "float", np.dtype(np.float64): "double",
}

# every spelling of the scalar types that the specification allows
_READ_TYPES = {}
for _names, _dt in (
        (("signed char", "int8", "int8_t"), np.int8),
        (("uchar", "unsigned char", "uint8", "uint8_t"), np.uint8),
        (("short", "short int", "signed short", "signed short int", "int16", "int16_t"), np.int16),
        (("ushort", "unsigned short", "unsigned short int", "uint16", "uint16_t"), np.uint16),
        (("int", "signed int", "int32", "int32_t"), np.int32),
        (("uint", "unsigned int", "uint32", "uint32_t"), np.uint32),
        (("longlong", "long long", "long long int", "signed long long", "signed long long int",
          "int64", "int64_t"), np.int64),
        (("ulonglong", "unsigned long long", "unsigned long long int", "uint64", "uint64_t"),
         np.uint64),
        (("float",), np.float32),
        (("double",), np.float64)):
    for _n in _names:
        _READ_TYPES[_n] = np.dtype(_dt)


def _num(v: float) -> str:
    """Shortest decimal that reads back as the same float64."""
    r = repr(float(v))
    return r[:-2] if r.endswith(".0") else r


def write(path: str, array: np.ndarray, origin, spacing, compress: bool = False) -> None:
    """Write a 3-D array of shape (nz, ny, nx) with the grid's origin and
    spacing, both given in (x, y, z) order as for the engine's volumes."""
    a = np.ascontiguousarray(array)
    if a.ndim != 3:
        raise ValueError("NRRD volumes here are 3-dimensional")
    if a.dtype not in _TYPES:
        raise ValueError("unsupported NRRD element type %s" % a.dtype)
    o = [float(v) for v in np.asarray(origin, dtype=np.float64).reshape(-1)[:3]]
    s = [float(v) for v in np.asarray(spacing, dtype=np.float64).reshape(-1)[:3]]
    if len(o) != 3 or len(s) != 3:
        raise ValueError("origin and spacing must hold 3 values")
    sizes = (a.shape[2], a.shape[1], a.shape[0])
    dirs = " ".join("(%s)" % ",".join(_num(s[i] if j == i else 0.0) for j in range(3))
                    for i in range(3))
    head = ["NRRD0004",
            "type: %s" % _TYPES[a.dtype],
            "dimension: 3",
            "space: left-posterior-superior",
            "sizes: %d %d %d" % sizes,
            "space directions: %s" % dirs,
            "kinds: domain domain domain",
            "endian: little",
            "encoding: %s" % ("gzip" if compress else "raw"),
            "space origin: (%s)" % ",".join(_num(v) for v in o),
            "", ""]
    data = a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes()
    if compress:
        data = gzip.compress(data)
    with open(path, "wb") as f:
        f.write("\n".join(head).encode("ascii"))
        f.write(data)


def _vector(text: str):
    m = re.fullmatch(r"\s*\(([^)]*)\)\s*", text)
    if not m:
        raise ValueError("malformed NRRD vector %r" % text)
    return [float(v) for v in m.group(1).split(",")]


def _vectors(text: str):
    out = []
    for tok in re.findall(r"\([^)]*\)|none", text):
        out.append(None if tok == "none" else _vector(tok))
    return out


def read_header(f):
    """Parse the header from an open binary file; returns (fields, keys)."""
    magic = f.readline().decode("ascii", "replace").strip()
    if not re.fullmatch(r"NRRD000[1-5]", magic):
        raise ValueError("not an NRRD file (magic %r)" % magic)
    fields, keys = {}, {}
    while True:
        raw = f.readline()
        if not raw:
            raise ValueError("NRRD header has no end")
        line = raw.decode("latin-1").rstrip("\r\n")
        if line == "":
            break
        if line.startswith("#"):
            continue
        if ":=" in line:
            k, v = line.split(":=", 1)
            keys[k] = v
            continue
        if ": " not in line:
            raise ValueError("malformed NRRD header line %r" % line)
        k, v = line.split(": ", 1)
        fields[k.strip().lower()] = v.strip()
    return fields, keys


def read(path: str):
    """Read a 3-D NRRD volume; returns (array (nz, ny, nx), origin (x, y, z),
    spacing (x, y, z)) with origin and spacing as float64 arrays.  The
    spacing of an axis is the length of its ``space directions`` vector."""
    with open(path, "rb") as f:
        fields, _ = read_header(f)
        if "data file" in fields or "datafile" in fields:
            raise ValueError("detached NRRD data files are not supported")
        body = f.read()
    dim = int(fields.get("dimension", "0"))
    if dim != 3:
        raise ValueError("expected a 3-dimensional NRRD volume, got dimension %d" % dim)
    tname = fields.get("type", "").lower()
    if tname not in _READ_TYPES:
        raise ValueError("unsupported NRRD type %r" % tname)
    dt = _READ_TYPES[tname]
    sizes = [int(v) for v in fields["sizes"].split()]
    if len(sizes) != 3:
        raise ValueError("NRRD sizes must list 3 axes")
    enc = fields.get("encoding", "raw").lower()
    if enc in ("gzip", "gz"):
        body = gzip.decompress(body)
    elif enc != "raw":
        raise ValueError("unsupported NRRD encoding %r" % enc)
    # byte skip / line skip apply to the decoded data (specification, "line
    # skip" and "byte skip"); -1 means the data are the file's last bytes
    lskip = int(fields.get("line skip", fields.get("lineskip", "0")))
    for _ in range(lskip):
        body = body[body.index(b"\n") + 1:]
    n = sizes[0] * sizes[1] * sizes[2]
    bskip = int(fields.get("byte skip", fields.get("byteskip", "0")))
    if bskip == -1:
        body = body[len(body) - n * dt.itemsize:]
    else:
        body = body[bskip:]
    if len(body) < n * dt.itemsize:
        raise ValueError("NRRD data are shorter than sizes say")
    order = "<" if fields.get("endian", "little").lower() == "little" else ">"
    a = np.frombuffer(body[:n * dt.itemsize], dtype=dt.newbyteorder(order))
    a = a.astype(dt).reshape(sizes[2], sizes[1], sizes[0])
    if "space directions" in fields:
        vecs = [v for v in _vectors(fields["space directions"]) if v is not None]
        if len(vecs) != 3:
            raise ValueError("NRRD space directions must give 3 vectors")
        spacing = np.array([float(np.linalg.norm(v)) for v in vecs])
    elif "spacings" in fields:
        spacing = np.array([float(v) for v in fields["spacings"].split()])
    else:
        spacing = np.ones(3)
    if "space origin" in fields:
        origin = np.array(_vector(fields["space origin"]), dtype=np.float64)
    else:
        origin = np.zeros(3)
    return a, origin, spacing


def beam_path(path: str, index: int) -> str:
    """Per-beam file name of the reference's exports: ``x.nrrd`` ->
    ``x_beam01.nrrd`` for the first beam (plan.py:185, plan_impt.py:240)."""
    return path.replace(".nrrd", "_beam%02i.nrrd" % (index + 1))
