"""ctypes binding of libseidart_hip.so (declarations: include/seidart_hip.h).

The shared library is built in-tree next to this file.  Importing this module
raises if it is missing: there is no CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libseidart_hip.so")

ABI_VERSION = 2

# status codes
OK, ERR_UNSTABLE, ERR_ARG, ERR_HIP, ERR_STATE = 0, 1, 2, 3, 4
# solvers
ELASTIC2, ELASTIC25, ELASTIC3, EM2, EM25, EM3 = 0, 1, 2, 3, 4, 5
# precision
FP64, FP32 = 0, 1
# density codes, averaging.f08:5-25
DENSITY_CODES = {"harmonic": 1, "geometric": 2, "arithmetic": 3, "none": 4}
# stencils
STENCIL_REFERENCE, STENCIL_CORRECTED = 0, 1

# input arrays, named after the reference's input files (without ".dat")
ARRAY_IDS = {}
_C_NAMES = ["c11", "c12", "c13", "c14", "c15", "c16", "c22", "c23", "c24", "c25", "c26",
            "c33", "c34", "c35", "c36", "c44", "c45", "c46", "c55", "c56", "c66"]
for _q, _n in enumerate(_C_NAMES):
    ARRAY_IDS[_n] = _q
ARRAY_IDS.update({
    "rho": 21, "gamma_x": 22, "gamma_y": 23, "gamma_z": 24, "gamma_xy": 25, "gamma_xz": 26,
    "gamma_yz": 27, "kappa_cpml": 28, "alpha_cpml": 29, "acoef_cpml": 30, "bcoef_cpml": 31,
    "kappa_half_cpml": 32, "alpha_half_cpml": 33, "acoef_half_cpml": 34, "bcoef_half_cpml": 35,
    "e11": 36, "e12": 37, "e13": 38, "e22": 39, "e23": 40, "e33": 41,
    "s11": 42, "s12": 43, "s13": 44, "s22": 45, "s23": 46, "s33": 47,
    "ic_x": 48, "ic_y": 49, "ic_z": 50,
})
NUM_ARRAYS = 51

SOURCE_IDS = {"x": 0, "y": 1, "z": 2, "xx": 3, "xy": 4, "xz": 5, "yy": 6, "yz": 7, "zz": 8}

FIELD_IDS_ELASTIC = {"vx": 0, "vy": 1, "vz": 2, "sxx": 3, "syy": 4, "szz": 5, "sxy": 6, "sxz": 7,
                     "syz": 8}
FIELD_IDS_EM = {"ex": 0, "ey": 1, "ez": 2, "hx": 3, "hy": 4, "hz": 5}


class Params(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("physics", C.c_int32), ("precision", C.c_int32),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("cpml", C.c_int32), ("dt", C.c_double), ("time_steps", C.c_int32),
        ("xind", C.c_int32), ("yind", C.c_int32), ("zind", C.c_int32),
        ("density_code", C.c_int32), ("stencil", C.c_int32), ("device", C.c_int32),
        ("reserved0", C.c_int32), ("ky", C.c_double), ("reserved", C.c_int32 * 6),
    ]


EXPORTS = [
    "sdt_version", "sdt_device_count", "sdt_last_error", "sdt_create", "sdt_destroy",
    "sdt_set_array", "sdt_set_source", "sdt_set_receivers", "sdt_run", "sdt_steps_done",
    "sdt_get_field", "sdt_set_field", "sdt_get_receivers", "sdt_get_norms", "sdt_set_timing",
    "sdt_get_kernel_times", "sdt_synchronize", "sdt_set_option", "sdt_get_info",
    "sdt_batch_create", "sdt_batch_run", "sdt_batch_info", "sdt_batch_destroy",
    "sdt_set_coefficients", "sdt_adjoint", "sdt_batch_adjoint", "sdt_set_plane_wave",
    "sdt_snapshot", "sdt_host_alloc", "sdt_host_free", "sdt_snapshot_noalloc",
]

OPT_ZCHUNK, OPT_GRAPH, OPT_STAGGER, OPT_TILE_ROWS, OPT_MATERIAL_TABLE, OPT_CONCURRENT, OPT_XCD_REMAP = (
    1, 2, 3, 4, 5, 6, 7)
OPT_CHECKPOINT, OPT_RAW_SOURCE, OPT_ISOTROPIC, OPT_FUSED, OPT_SOURCE_WINDOW = 8, 9, 10, 11, 12
(INFO_MATERIALS, INFO_SHELL_TILES, INFO_DEVICE_BYTES, INFO_COEFFICIENTS, INFO_ISOTROPIC, INFO_FUSED,
 INFO_CPML_TUPLES) = 1, 2, 3, 4, 5, 6, 7
NUM_SRC = 9
PW_S, PW_P = 1, 2
BATCH_INFO_TILE_ROWS, BATCH_INFO_TILES, BATCH_INFO_XCD_ORDER = 1, 2, 3


class SdtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"seidart_hip error {code}: {msg}")
        self.code = code


class UnstableError(SdtError):
    """Raised when the max velocity/E norm exceeds 1e25, where the reference
    stops with 'code became unstable and blew up' (cpmlfdtd.f08:567)."""

    def __init__(self, code, msg, step):
        super().__init__(code, msg)
        self.step = step


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `make -C seidart-backend_amd/csrc` (no CPU fallback exists)")
    lib = C.CDLL(LIB_PATH)
    P, I32, I64, D = C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_double)
    sig = {
        "sdt_version": ([], I32),
        "sdt_device_count": ([C.POINTER(I32)], I32),
        "sdt_last_error": ([P], C.c_char_p),
        "sdt_create": ([C.POINTER(Params), C.POINTER(P)], I32),
        "sdt_destroy": ([P], I32),
        "sdt_set_array": ([P, I32, D, I64], I32),
        "sdt_set_source": ([P, I32, D, I64], I32),
        "sdt_set_receivers": ([P, I32, C.POINTER(I32)], I32),
        "sdt_run": ([P, I32, C.POINTER(I32)], I32),
        "sdt_steps_done": ([P, C.POINTER(I32)], I32),
        "sdt_get_field": ([P, I32, D, I64], I32),
        "sdt_set_field": ([P, I32, D, I64], I32),
        "sdt_get_receivers": ([P, D, I64], I32),
        "sdt_get_norms": ([P, D, D, I64], I32),
        "sdt_set_timing": ([P, I32], I32),
        "sdt_get_kernel_times": ([P, D, D, C.POINTER(I64)], I32),
        "sdt_synchronize": ([P], I32),
        "sdt_set_option": ([P, I32, I64], I32),
        "sdt_get_info": ([P, I32, C.POINTER(I64)], I32),
        "sdt_batch_create": ([C.POINTER(P), I32, C.POINTER(P)], I32),
        "sdt_batch_run": ([P, I32, C.POINTER(I32), C.POINTER(I32)], I32),
        "sdt_batch_info": ([P, I32, C.POINTER(I64)], I32),
        "sdt_batch_destroy": ([P], I32),
        "sdt_set_coefficients": ([P, D, I64], I32),
        "sdt_adjoint": ([P, D, I64, D, I64, D, I64], I32),
        "sdt_batch_adjoint": ([P, C.POINTER(D), C.POINTER(D), C.POINTER(D)], I32),
        "sdt_set_plane_wave": ([P, I32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, I32], I32),
        "sdt_snapshot": ([P, I32, I32, P, I64], I32),
        "sdt_host_alloc": ([I64, C.POINTER(P)], I32),
        "sdt_host_free": ([P], I32),
        "sdt_snapshot_noalloc": ([P, I32, I32, P, I64], I32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if lib.sdt_version() != ABI_VERSION:
        raise ImportError("libseidart_hip.so ABI version mismatch; rebuild it")
    return lib


lib = _load()


def check(rc, handle=None, step=None):
    if rc == OK:
        return
    msg = lib.sdt_last_error(handle)
    msg = msg.decode() if msg else ""
    if rc == ERR_UNSTABLE:
        raise UnstableError(rc, msg, step)
    raise SdtError(rc, msg)
