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


def test_raw_record_writer_matches_record_writer(tmp_path, monkeypatch):
    """write_record_raw (payload already in file order, as device snapshots
    are packed) writes the same bytes as write_record, with and without
    subrecords."""
    a = np.random.default_rng(1).standard_normal((5, 4, 3))
    for dt in (np.float32, np.float64):
        p1, p2 = str(tmp_path / "a.dat"), str(tmp_path / "b.dat")
        fio.write_record(p1, a, dt)
        fio.write_record_raw(p2, np.asfortranarray(a).ravel(order="F").astype(dt))
        assert open(p1, "rb").read() == open(p2, "rb").read()
    monkeypatch.setattr(fio, "_MAX_SUB", 40)
    fio.write_record(p1, a, np.float64)
    fio.write_record_raw(p2, np.asfortranarray(a).ravel(order="F"))
    assert open(p1, "rb").read() == open(p2, "rb").read()


def test_parallel_chunked_record_matches(tmp_path, monkeypatch):
    """start_record (chunks written by a thread pool at their offsets)
    produces the bytes of write_record_raw, also across subrecords."""
    from concurrent.futures import ThreadPoolExecutor
    a = np.random.default_rng(2).standard_normal(1000).astype(np.float32)
    with ThreadPoolExecutor(4) as pool:
        for max_sub in (fio._MAX_SUB, 1000):
            monkeypatch.setattr(fio, "_MAX_SUB", max_sub)
            p1, p2 = str(tmp_path / "a.dat"), str(tmp_path / "b.dat")
            fio.write_record_raw(p1, a)
            fio.start_record(p2, a, pool, chunk=96)()
            assert open(p1, "rb").read() == open(p2, "rb").read()
            assert not (tmp_path / "b.dat.tmp").exists()
