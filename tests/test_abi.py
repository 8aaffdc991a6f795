"""C-ABI checks that need no GPU: the library loads, exports every entry point
include/seidart_hip.h declares, and rejects bad parameters before touching a
device."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "seidart_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sdt_[a-z_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for must in ("sdt_create", "sdt_run", "sdt_set_array", "sdt_get_field", "sdt_destroy"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from seidart_hip import _lib as L
    for name in declared_functions():
        assert hasattr(L.lib, name), name
    assert sorted(L.EXPORTS) == declared_functions()
    assert L.lib.sdt_version() == L.ABI_VERSION


def _params(**kw):
    from seidart_hip import _lib as L
    p = L.Params()
    p.abi_version = L.ABI_VERSION
    p.physics = L.ELASTIC2
    p.precision = L.FP64
    p.nx, p.ny, p.nz = 40, 1, 40
    p.dx = p.dy = p.dz = 1.0
    p.cpml = 5
    p.dt = 1e-4
    p.time_steps = 10
    p.xind, p.yind, p.zind = 10, 1, 10
    p.density_code = 3
    p.stencil = 0
    p.device = 0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("bad,msg", [
    (dict(abi_version=99), "abi_version"),
    (dict(physics=7), "unknown physics"),
    (dict(precision=5), "unknown precision"),
    (dict(nx=3), "grid too small"),
    (dict(ny=4), "grid too small"),
    (dict(dt=0.0), "positive"),
    (dict(time_steps=0), "time_steps"),
    (dict(density_code=9), "density_code"),
    (dict(stencil=3), "stencil"),
    (dict(xind=35), "source index"),
    (dict(zind=-6), "source index"),
])
def test_create_rejects_bad_parameters(bad, msg):
    from seidart_hip import _lib as L
    h = C.c_void_p()
    rc = L.lib.sdt_create(C.byref(_params(**bad)), C.byref(h))
    assert rc == L.ERR_ARG
    assert msg in L.lib.sdt_last_error(None).decode()
    assert not h


def test_create_without_device_reports_hip_error():
    from seidart_hip import _lib as L
    n = C.c_int32(-1)
    L.lib.sdt_device_count(C.byref(n))
    if n.value > 0:
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = L.lib.sdt_create(C.byref(_params()), C.byref(h))
    assert rc == L.ERR_HIP and not h


def test_null_handle_calls_fail_cleanly():
    from seidart_hip import _lib as L
    assert L.lib.sdt_run(None, 1, None) == L.ERR_ARG
    assert L.lib.sdt_synchronize(None) == L.ERR_ARG
    assert L.lib.sdt_destroy(None) == L.OK


def test_oracle_library_builds_and_loads():
    from oracle import oracle_ctypes as O
    lib = O.lib()
    assert lib.oracle_scalar_mean(2.0, 2.0, 3) == 2.0


def test_batch_rejects_repeated_handle_before_touching_it():
    """A handle listed twice in one batch would be stepped twice per launch;
    sdt_batch_create refuses it by comparing the pointers alone, so the
    stand-in addresses below are never dereferenced."""
    from seidart_hip import _lib as L
    b = C.c_void_p()
    hs = (C.c_void_p * 3)(0x1000, 0x2000, 0x1000)
    assert L.lib.sdt_batch_create(hs, 3, C.byref(b)) == L.ERR_ARG
    assert "repeats handle 0" in L.lib.sdt_last_error(None).decode()
    assert not b
    hs = (C.c_void_p * 2)(0x1000, None)
    assert L.lib.sdt_batch_create(hs, 2, C.byref(b)) == L.ERR_ARG
    assert "null handle" in L.lib.sdt_last_error(None).decode()
