"""Fortran sequential-unformatted records (seidartio.f08 formats) and the
reference's snapshot names; CPU only."""
import struct

import numpy as np

from seidart_hip import fortran_io as fio


def test_record_round_trip_and_markers(tmp_path):
    a = np.arange(12.0).reshape(3, 4, order="F") * 1.5
    p = str(tmp_path / "c11.dat")
    fio.write_record(p, a)
    raw = open(p, "rb").read()
    assert struct.unpack("<i", raw[:4])[0] == 12 * 8
    assert struct.unpack("<i", raw[-4:])[0] == 12 * 8
    assert len(raw) == 12 * 8 + 8
    # payload is column-major: first value a[0,0], second a[1,0]
    assert np.frombuffer(raw[4:20], "<f8").tolist() == [a[0, 0], a[1, 0]]
    np.testing.assert_array_equal(fio.read_array(p, (3, 4)), a)


def test_single_precision_snapshot(tmp_path):
    a = np.random.default_rng(0).standard_normal((5, 6))
    p = str(tmp_path / "Vx.000001.10.40.dat")
    fio.write_record(p, a, np.float32)
    assert len(open(p, "rb").read()) == 30 * 4 + 8
    np.testing.assert_array_equal(fio.read_array(p, (5, 6), np.float32), a.astype(np.float32))


def test_subrecords(tmp_path, monkeypatch):
    monkeypatch.setattr(fio, "_MAX_SUB", 16)
    a = np.arange(7.0)
    p = str(tmp_path / "big.dat")
    fio.write_record(p, a)
    raw = open(p, "rb").read()
    # 56 bytes in subrecords of 16, 16, 16, 8: (head, tail) markers are
    # (-16, 16), (-16, -16), (-16, -16), (8, -8) -- head negative when more
    # follow, tail negative when the subrecord continues an earlier one
    markers, pos = [], 0
    while pos < len(raw):
        head = struct.unpack("<i", raw[pos:pos + 4])[0]
        size = abs(head)
        tail = struct.unpack("<i", raw[pos + 4 + size:pos + 8 + size])[0]
        markers.append((head, tail))
        pos += size + 8
    assert markers == [(-16, 16), (-16, -16), (-16, -16), (8, -8)]
    np.testing.assert_array_equal(fio.read_record(p, np.float64, 7), a)


def test_snapshot_names():
    assert fio.snapshot_name("Vx", 1, 100, 40) == "Vx.000001.100.40.dat"
    assert fio.snapshot_name("Ez", 123456, 7, 9, yind=8) == "Ez.123456.7.8.9.dat"
