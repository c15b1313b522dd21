"""NIfTI-1 single-file images (.nii, .nii.gz) for the plan layer's CT export
(the reference writes them with SimpleITK, plan.py:206-216).

Arrays are (z, y, x) C-order float32, origin and spacing in the patient
(LPS) frame with identity direction, as the plan layer keeps them.  The
header carries the LPS -> RAS flip the way ITK writes it: x and y of the
origin and of the axes are negated (qform and sform, code 1 = scanner), so
other NIfTI readers place the volume where the DICOM CT was.
"""
from __future__ import annotations

import gzip
import struct

import numpy as np

_HDR = 348


def _header(shape_zyx, spacing, origin) -> bytes:
    nz, ny, nx = (int(n) for n in shape_zyx)
    sx, sy, sz = (float(s) for s in spacing)
    ox, oy, oz = (float(o) for o in origin)
    h = bytearray(_HDR)
    struct.pack_into("<i", h, 0, _HDR)                                   # sizeof_hdr
    struct.pack_into("<8h", h, 40, 3, nx, ny, nz, 1, 1, 1, 1)            # dim
    struct.pack_into("<h", h, 70, 16)                                    # datatype: float32
    struct.pack_into("<h", h, 72, 32)                                    # bitpix
    struct.pack_into("<8f", h, 76, 1.0, sx, sy, sz, 0.0, 0.0, 0.0, 0.0)  # pixdim (qfac = 1)
    struct.pack_into("<f", h, 108, 352.0)                                # vox_offset
    struct.pack_into("<f", h, 112, 1.0)                                  # scl_slope
    struct.pack_into("<f", h, 116, 0.0)                                  # scl_inter
    struct.pack_into("<B", h, 123, 2)                                    # xyzt_units: mm
    struct.pack_into("<h", h, 252, 1)                                    # qform_code
    struct.pack_into("<h", h, 254, 1)                                    # sform_code
    # diag(-1, -1, 1) is a half turn about z: quaternion (a, b, c, d) = (0, 0, 0, 1)
    struct.pack_into("<6f", h, 256, 0.0, 0.0, 1.0, -ox, -oy, oz)         # quatern_b..d, qoffset
    struct.pack_into("<4f", h, 280, -sx, 0.0, 0.0, -ox)                  # srow_x
    struct.pack_into("<4f", h, 296, 0.0, -sy, 0.0, -oy)                  # srow_y
    struct.pack_into("<4f", h, 312, 0.0, 0.0, sz, oz)                    # srow_z
    h[344:348] = b"n+1\x00"
    return bytes(h) + b"\x00" * 4                                        # extension flag


def write(path: str, array, origin, spacing):
    a = np.ascontiguousarray(np.asarray(array, dtype="<f4"))
    if a.ndim != 3:
        raise ValueError("NIfTI export takes a 3-D (z, y, x) volume")
    data = _header(a.shape, spacing, origin) + a.tobytes()
    if path.endswith(".gz"):
        with gzip.open(path, "wb") as f:
            f.write(data)
    else:
        with open(path, "wb") as f:
            f.write(data)


def read(path: str):
    """(array (z, y, x) float32, origin (LPS), spacing) of a file written by
    ``write`` (float32, sform code 1, axis-aligned)."""
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rb") as f:
        buf = f.read()
    if struct.unpack_from("<i", buf, 0)[0] != _HDR or buf[344:347] != b"n+1":
        raise ValueError("%s is not a single-file NIfTI-1 image" % path)
    dim = struct.unpack_from("<8h", buf, 40)
    if dim[0] != 3 or struct.unpack_from("<h", buf, 70)[0] != 16:
        raise NotImplementedError("only 3-D float32 NIfTI images are read")
    nx, ny, nz = dim[1:4]
    off = int(struct.unpack_from("<f", buf, 108)[0])
    srow = [struct.unpack_from("<4f", buf, 280 + 16 * r) for r in range(3)]
    spacing = np.array([-srow[0][0], -srow[1][1], srow[2][2]])
    origin = np.array([-srow[0][3], -srow[1][3], srow[2][3]])
    a = np.frombuffer(buf, dtype="<f4", count=nx * ny * nz, offset=off).reshape(nz, ny, nx)
    return a.astype(np.float32), origin, spacing
