"""sdt_snapshot_noalloc (the allocation-free snapshot, fixed device scratch)
against sdt_snapshot (staging buffer allocated by the handle), through the
C-ABI: bit-identical output on two calls in a row, the second with a smaller
grid, in FP64 and FP32, across several scratch chunks; a handle that was
never set up is refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _both(s, name, single):
    from seidart_hip import _lib as L
    from seidart_hip.solver import PinnedArray
    dt = np.float32 if single else np.float64
    a, b = PinnedArray(s.ncell, dt), PinnedArray(s.ncell, dt)
    a.array[:] = -1
    b.array[:] = -2
    fid = s.fields[name]
    L.check(L.lib.sdt_snapshot(s._h, fid, int(single), a.array.ctypes.data_as(C.c_void_p), s.ncell), s._h)
    L.check(L.lib.sdt_synchronize(s._h), s._h)
    L.check(L.lib.sdt_snapshot_noalloc(s._h, fid, int(single), b.array.ctypes.data_as(C.c_void_p), s.ncell),
            s._h)
    return a.array.copy(), b.array.copy()


def test_noalloc_snapshot_equals_snapshot_twice_second_smaller():
    from seidart_hip import models
    from seidart_hip.solver import make_solver
    big = models.elastic3(nx=150, ny=150, nz=150, cpml=10, steps=8, src=(75, 75, 20), nrec=4)
    small = models.em2(nx=60, nz=44, steps=12, src=(30, 15), nrec=4)
    cases = [(big, ("vx", "vz", "syz")), (small, ("ex", "hy"))]
    assert big["shape"] == (170, 170, 170)   # FP64 39 MB, FP32 20 MB: several 16 MiB chunks
    for cfg, names in cases:
        for precision in ("fp64", "fp32"):
            print(f"[noalloc] {cfg['physics']} {cfg['shape']} {precision}", flush=True)
            s = make_solver(cfg, precision=precision)
            try:
                s.run(cfg["time_steps"])
                for name in names:
                    for single in (False, True):
                        a, b = _both(s, name, single)
                        assert np.any(a != 0), (name, single)
                        assert np.array_equal(a, b), (cfg["physics"], precision, name, single)
            finally:
                s.close()


def test_noalloc_snapshot_refuses_a_handle_not_set_up():
    from seidart_hip import _lib as L, models
    from seidart_hip.solver import PinnedArray, make_solver
    cfg = models.em2(nx=40, nz=30, steps=5, src=(20, 10), nrec=2)
    s = make_solver(cfg)
    try:
        buf = PinnedArray(s.ncell, np.float32)
        rc = L.lib.sdt_snapshot_noalloc(s._h, s.fields["ex"], 1, buf.array.ctypes.data_as(C.c_void_p), s.ncell)
        assert rc == L.ERR_STATE
        assert "allocates nothing" in L.lib.sdt_last_error(s._h).decode()
        s.run(0)
        assert L.lib.sdt_snapshot_noalloc(s._h, s.fields["ex"], 1, buf.array.ctypes.data_as(C.c_void_p),
                                          s.ncell) == L.OK
    finally:
        s.close()
