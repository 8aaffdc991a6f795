"""Fortran sequential-unformatted records, as read and written by
src/fortran/seidartio.f08 (material_rw2/3 225-292, loadsource 161-183,
write_image2/3 366-426, write_array 186-200).

A record is a 4-byte little-endian length marker, the payload, and the same
marker again.  Payloads larger than 2**31 - 9 bytes are split into
subrecords whose markers are negative on every subrecord but the last (the
gfortran convention).
"""
from __future__ import annotations

import os
import struct

import numpy as np

_MAX_SUB = 2**31 - 9


def read_record(path, dtype=np.float64, count=None):
    """Payload of the first record of `path` as a 1-D array."""
    chunks = []
    with open(path, "rb") as f:
        while True:
            head = f.read(4)
            if len(head) < 4:
                raise ValueError(f"{path}: truncated record marker")
            n = struct.unpack("<i", head)[0]
            more = n < 0
            n = abs(n)
            chunks.append(f.read(n))
            tail = struct.unpack("<i", f.read(4))[0]
            if abs(tail) != n:
                raise ValueError(f"{path}: record markers disagree ({n} vs {tail})")
            if not more:
                break
    data = np.frombuffer(b"".join(chunks), dtype=np.dtype(dtype).newbyteorder("<"))
    if count is not None and data.size != count:
        raise ValueError(f"{path}: expected {count} values, found {data.size}")
    return data


def read_array(path, shape, dtype=np.float64):
    """Column-major (x fastest) array of `shape` from one record."""
    n = int(np.prod(shape))
    return read_record(path, dtype, n).reshape(shape, order="F").astype(np.float64)


def write_record(path, array, dtype=None):
    """Write one record holding `array` in column-major order."""
    a = np.asarray(array)
    if dtype is not None:
        a = a.astype(dtype)
    payload = np.asfortranarray(a).ravel(order="F").astype(a.dtype.newbyteorder("<")).tobytes()
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        pos, n = 0, len(payload)
        while True:
            size = min(n - pos, _MAX_SUB)
            last = pos + size >= n
            marker = size if last else -size
            f.write(struct.pack("<i", marker))
            f.write(payload[pos:pos + size])
            f.write(struct.pack("<i", marker))
            pos += size
            if last:
                break
    os.replace(tmp, path)


def snapshot_name(channel, it, xind, zind, yind=None):
    """`(a2,'.',i6.6,'.',i0,'.',i0[,'.',i0],'.dat')` of write_image2/3
    (seidartio.f08:378-379, 413-414)."""
    if yind is None:
        return f"{channel}.{it:06d}.{xind}.{zind}.dat"
    return f"{channel}.{it:06d}.{xind}.{yind}.{zind}.dat"
