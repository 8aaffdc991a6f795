"""Fortran sequential-unformatted records, as read and written by
src/fortran/seidartio.f08 (material_rw2/3 225-292, loadsource 161-183,
write_image2/3 366-426, write_array 186-200).

A record is a 4-byte little-endian length marker, the payload, and the same
marker again.  Payloads larger than 2**31 - 9 bytes are split into
subrecords (the gfortran convention): a subrecord's head marker is negative
when more subrecords follow it, its tail marker is negative when it
continues an earlier subrecord, so the head of the first and the tail of the
last are the only positive markers of a split record.
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
    """Column-major (x fastest) array of `shape` from one record (complex128
    records, as read by material_rw2c at seidartio.f08:249-270, stay complex)."""
    n = int(np.prod(shape))
    out = np.dtype(np.complex128) if np.dtype(dtype).kind == "c" else np.dtype(np.float64)
    return read_record(path, dtype, n).reshape(shape, order="F").astype(out)


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
            f.write(struct.pack("<i", size if last else -size))
            f.write(payload[pos:pos + size])
            f.write(struct.pack("<i", -size if pos > 0 else size))
            pos += size
            if last:
                break
    os.replace(tmp, path)


def write_record_raw(path, payload):
    """Write one record whose payload is already in file order: `payload` is
    a contiguous array (or buffer) whose bytes are written as they are, with
    no copy (snapshots packed on the device)."""
    mv = memoryview(payload).cast("B")
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        pos, n = 0, len(mv)
        while True:
            size = min(n - pos, _MAX_SUB)
            last = pos + size >= n
            f.write(struct.pack("<i", size if last else -size))
            f.write(mv[pos:pos + size])
            f.write(struct.pack("<i", -size if pos > 0 else size))
            pos += size
            if last:
                break
    os.replace(tmp, path)


def record_layout(n):
    """Subrecords of an n-byte payload: (file offset of the head marker,
    head marker, tail marker, payload offset, size) each, and the file size."""
    out, pos, off = [], 0, 0
    while True:
        size = min(n - pos, _MAX_SUB)
        last = pos + size >= n
        out.append((off, size if last else -size, -size if pos > 0 else size, pos, size))
        off += size + 8
        pos += size
        if last:
            return out, off


def _pwrite_all(fd, mv, off):
    while len(mv):
        k = os.pwrite(fd, mv, off)
        mv, off = mv[k:], off + k


def start_record(path, payload, pool, chunk=16 << 20):
    """Queue the write of one record (payload already in file order, kept
    alive by the caller until the write ends) as parallel chunk writes on the
    thread pool `pool`.  Returns a callable that waits for them and then
    moves the finished file to `path`; the bytes equal write_record_raw's."""
    mv = memoryview(payload).cast("B")
    subs, total = record_layout(len(mv))
    tmp = path + ".tmp"
    fd = os.open(tmp, os.O_WRONLY | os.O_CREAT | os.O_TRUNC, 0o666)
    futs = []
    try:
        os.ftruncate(fd, total)
        for off, head, tail, ppos, size in subs:
            _pwrite_all(fd, memoryview(struct.pack("<i", head)), off)
            _pwrite_all(fd, memoryview(struct.pack("<i", tail)), off + 4 + size)
            for c in range(0, size, chunk):
                m = min(chunk, size - c)
                futs.append(pool.submit(_pwrite_all, fd, mv[ppos + c:ppos + c + m], off + 4 + c))
    except BaseException:
        for f in futs:
            f.cancel()
        os.close(fd)
        raise

    def finish():
        try:
            for f in futs:
                f.result()
        finally:
            os.close(fd)
        os.replace(tmp, path)
    return finish


def snapshot_name(channel, it, xind, zind, yind=None):
    """`(a2,'.',i6.6,'.',i0,'.',i0[,'.',i0],'.dat')` of write_image2/3
    (seidartio.f08:378-379, 413-414)."""
    if yind is None:
        return f"{channel}.{it:06d}.{xind}.{zind}.dat"
    return f"{channel}.{it:06d}.{xind}.{yind}.{zind}.dat"


# --------------------------------------------------------------- zstd blocks ----
_ZSTD_MAGIC = 0xFD2FB528
_ZSTD_RAW_BLOCK = 128 * 1024


def zstd_raw_frame(payload):
    """A zstd frame (RFC 8878) of raw (stored) blocks: single-segment header
    with the 8-byte content size, blocks of at most 128 KiB, no checksum.
    Every zstd decoder reads it; it is used when no zstd compressor is
    importable."""
    payload = bytes(payload)
    out = bytearray(struct.pack("<I", _ZSTD_MAGIC))
    out.append((3 << 6) | (1 << 5))          # FCS 8 bytes, single segment
    out += struct.pack("<Q", len(payload))
    pos, n = 0, len(payload)
    while True:
        size = min(n - pos, _ZSTD_RAW_BLOCK)
        last = pos + size >= n
        header = (size << 3) | (0 << 1) | (1 if last else 0)   # block type 0 = raw
        out += struct.pack("<I", header)[:3]
        out += payload[pos:pos + size]
        pos += size
        if last:
            break
    return bytes(out)


def zstd_compress(payload, level=3):
    """zstd frame of `payload` at `level` (seidartio.f08:559-560 uses 3) via
    pyarrow's bundled libzstd, else a frame of raw blocks."""
    try:
        import pyarrow as pa
        return pa.Codec("zstd", compression_level=level).compress(payload, asbytes=True)
    except ImportError:
        return zstd_raw_frame(payload)


def zstd_decompress(frame, size):
    """Inverse of zstd_compress for a frame of known decompressed size."""
    import pyarrow as pa
    return pa.Codec("zstd").decompress(frame, decompressed_size=size, asbytes=True)


def block_name(prefix, first, last, xind, yind, zind):
    """`(a, i6.6, '-', i6.6, '.', i0, '.', i0, '.', i0, '.blk.zst')` of the
    2.5D block files (cpmlfdtd.f08:1299-1302 'S25.', 3317-3320 'EM25.')."""
    return f"{prefix}.{first:06d}-{last:06d}.{xind}.{yind}.{zind}.blk.zst"


class BlockWriter2D:
    """Block output of the 2.5D solvers (seidartio.f08:614-734): each step's
    real parts of three components, CPML stripped, as FP32 in a Fortran-order
    buffer (nx - 2c, 1, nz - 2c, 3, steps); a full (or the final, partial)
    block is written as one zstd frame of its populated leading part."""

    def __init__(self, nx, nz, cpml, steps_per_block, time_steps, prefix, xind, yind, zind, workdir="."):
        if steps_per_block < 1:
            raise ValueError("steps per zstd block must be at least 1")
        self.onx, self.onz = nx - 2 * cpml, nz - 2 * cpml
        if self.onx < 1 or self.onz < 1:
            raise ValueError("cpml removes the entire zstd block output domain")
        self.cpml, self.spb, self.nsteps = cpml, steps_per_block, time_steps
        self.buf = np.zeros((self.onx, 1, self.onz, 3, steps_per_block), dtype=np.float32, order="F")
        self.cur, self.block = 0, 0
        self.prefix, self.ind, self.workdir = prefix, (xind, yind, zind), workdir
        self.written = []

    def add(self, a, b, c):
        if self.cur == 0:
            self.block += 1
            first = (self.block - 1) * self.spb + 1
            last = min(self.block * self.spb, self.nsteps)
            self.name = block_name(self.prefix, first, last, *self.ind)
        cp = self.cpml
        for q, f in enumerate((a, b, c)):
            f = np.asarray(f)
            self.buf[:, 0, :, q, self.cur] = np.real(f[cp:f.shape[0] - cp, cp:f.shape[1] - cp]).astype(np.float32)
        self.cur += 1
        if self.cur == self.spb:
            self.flush()

    def flush(self):
        if self.cur < 1:
            return
        n = self.onx * self.onz * 3 * self.cur
        payload = self.buf.ravel(order="F")[:n].astype("<f4").tobytes()
        path = os.path.join(self.workdir, self.name)
        with open(path, "wb") as f:
            f.write(zstd_compress(payload))
        self.written.append((path, len(payload)))
        self.cur = 0


def steps_per_block_2d(nx, nz, cpml, env=None):
    """FDTD_STEPS_PER_BLOCK, else about 2 GiB of FP32 per block capped at 500
    steps (seidartio.f08:685-734)."""
    env = os.environ if env is None else env
    if "FDTD_STEPS_PER_BLOCK" in env:
        return int(env["FDTD_STEPS_PER_BLOCK"].strip())
    bytes_per_step = float(nx - 2 * cpml) * (nz - 2 * cpml) * 3 * 4
    return min(500, max(1, int(2147483648.0 / bytes_per_step)))
