"""Reader for the binary phase dumps written by oracle/ref_dump.c.

Each record is: b"CDMP", char name[24], char type ('d' float64 | 'i' int32),
int32 ndim, int32 dims[ndim] (Fortran order), raw little-endian data.
Arrays are returned as numpy arrays in Fortran order with the reference's own
shape, e.g. UU(NUMI,NUMJ,NUMK) comes back with shape (NUMI, NUMJ, NUMK).
"""
import struct

import numpy as np


def read_dump(path):
    out = {}
    with open(path, "rb") as fh:
        buf = fh.read()
    pos = 0
    while pos < len(buf):
        if buf[pos:pos + 4] != b"CDMP":
            raise ValueError(f"{path}: bad record marker at byte {pos}")
        pos += 4
        name = buf[pos:pos + 24].split(b"\0", 1)[0].decode()
        pos += 24
        typ = chr(buf[pos])
        pos += 1
        (ndim,) = struct.unpack_from("<i", buf, pos)
        pos += 4
        dims = struct.unpack_from(f"<{ndim}i", buf, pos)
        pos += 4 * ndim
        dt = np.float64 if typ == "d" else np.int32
        n = int(np.prod(dims))
        arr = np.frombuffer(buf, dtype=dt, count=n, offset=pos).reshape(dims, order="F").copy()
        pos += n * arr.itemsize
        out[name] = arr
    return out
