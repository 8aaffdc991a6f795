"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker used by tests/, __graft_entry__.smoke()
and bench.py's CPU-baseline leg.  It mirrors seidart_hip.solver.run_config so
a configuration dict runs unchanged through either.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
# the same source built with the reference's production flags (-O3,
# fast-math, FP contraction): timed as the CPU baseline, never the checker
FAST_LIB_PATH = os.path.join(_HERE, "liboracle_fast.so")

NUM_ARRAYS, NUM_SRC, NUM_FIELDS = 51, 9, 9

_C_NAMES = ["c11", "c12", "c13", "c14", "c15", "c16", "c22", "c23", "c24", "c25", "c26",
            "c33", "c34", "c35", "c36", "c44", "c45", "c46", "c55", "c56", "c66"]
ARRAY_IDS = {n: q for q, n in enumerate(_C_NAMES)}
ARRAY_IDS.update({
    "rho": 21, "gamma_x": 22, "gamma_y": 23, "gamma_z": 24, "gamma_xy": 25, "gamma_xz": 26,
    "gamma_yz": 27, "kappa_cpml": 28, "alpha_cpml": 29, "acoef_cpml": 30, "bcoef_cpml": 31,
    "kappa_half_cpml": 32, "alpha_half_cpml": 33, "acoef_half_cpml": 34, "bcoef_half_cpml": 35,
    "e11": 36, "e12": 37, "e13": 38, "e22": 39, "e23": 40, "e33": 41,
    "s11": 42, "s12": 43, "s13": 44, "s22": 45, "s23": 46, "s33": 47,
    "ic_x": 48, "ic_y": 49, "ic_z": 50,
})
SOURCE_IDS = {"x": 0, "y": 1, "z": 2, "xx": 3, "xy": 4, "xz": 5, "yy": 6, "yz": 7, "zz": 8}
FIELD_IDS_ELASTIC = {"vx": 0, "vy": 1, "vz": 2, "sxx": 3, "syy": 4, "szz": 5, "sxy": 6, "sxz": 7,
                     "syz": 8}
FIELD_IDS_EM = {"ex": 0, "ey": 1, "ez": 2, "hx": 3, "hy": 4, "hz": 5}
DENSITY_CODES = {"harmonic": 1, "geometric": 2, "arithmetic": 3, "none": 4}

_D = C.POINTER(C.c_double)


class OracleRun(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("nsteps", C.c_int32), ("isrc", C.c_int32), ("jsrc", C.c_int32), ("ksrc", C.c_int32),
        ("density_code", C.c_int32), ("stencil", C.c_int32),
        ("arr", _D * NUM_ARRAYS), ("src", _D * NUM_SRC),
        ("nrec", C.c_int32), ("rec", C.POINTER(C.c_int32)), ("rec_out", _D),
        ("field_in", _D * NUM_FIELDS), ("field_out", _D * NUM_FIELDS),
        ("velnorm", _D), ("stressnorm", _D), ("unstable_step", C.c_int32), ("nthreads", C.c_int32),
        ("ky", C.c_double),
        ("pw_kind", C.c_int32), ("pw_wavelet", C.c_int32), ("cpml", C.c_int32),
        ("pw_angles", C.c_double * 3), ("pw_f0", C.c_double), ("pw_amp", C.c_double), ("pw_speed", C.c_double),
        ("pw_eta", C.c_double),
    ]


_VOIGT = [(0, 0), (0, 1), (0, 2), (0, 3), (0, 4), (0, 5), (1, 1), (1, 2), (1, 3), (1, 4), (1, 5), (2, 2), (2, 3),
          (2, 4), (2, 5), (3, 3), (3, 4), (3, 5), (4, 4), (4, 5), (5, 5)]


def background_speed(arrays, two_d):
    """sqrt(max over cells of the largest eigenvalue of the Kelvin-form
    stiffness / min density), as cpmlfdtd.f08:228-229 (2D), 829-832 (2.5D)
    and 1601-1605 (3D) compute it with LAPACK dsyev
    (tensor_operations.f08:57-132, 157-181, 207-236); numpy's eigvalsh is
    the same LAPACK family."""
    names = ["c11", "c12", "c13", "c14", "c15", "c16", "c22", "c23", "c24", "c25", "c26",
             "c33", "c34", "c35", "c36", "c44", "c45", "c46", "c55", "c56", "c66"]
    rho = np.asarray(arrays["rho"], dtype=np.float64).ravel()
    n = rho.size
    keep = ("c11", "c13", "c15", "c33", "c35", "c55") if two_d else names
    K = np.zeros((n, 6, 6))
    s = np.array([1, 1, 1, np.sqrt(2.0), np.sqrt(2.0), np.sqrt(2.0)])
    for q, (a, b) in enumerate(_VOIGT):
        if names[q] not in keep or names[q] not in arrays:
            continue
        v = np.asarray(arrays[names[q]], dtype=np.float64).ravel() * s[a] * s[b]
        K[:, a, b] = v
        K[:, b, a] = v
    return float(np.sqrt(np.linalg.eigvalsh(K).max() / rho.min()))


CLIGHT = 2.9979458e8                                 # constants.f08:8
MU0 = 4.0 * 3.141592653589793238462643 * 1.0e-7       # constants.f08:9
EPS0 = 8.85418782e-12                                 # constants.f08:10


def em_background(arrays, two_d):
    """(speed, impedance) of the smallest relative-permittivity eigenvalue
    over the model: clight / sqrt(eps_min), sqrt(mu0 / (eps0 eps_min))
    (cpmlfdtd.f08:2472-2475, 3577-3580).  2D takes eigenvalues 2..3 of
    [[e11, e13, 0], [e13, e33, 0], [0, 0, 0]] per cell
    (tensor_operations.f08:135-154); 3D all three of the full tensor
    (240-262)."""
    def a(name):
        return np.asarray(arrays[name], dtype=np.float64).ravel() if name in arrays else None
    e11 = a("e11")
    n = e11.size
    z = np.zeros(n)
    if two_d:
        K = np.zeros((n, 3, 3))
        K[:, 0, 0] = e11
        K[:, 0, 1] = K[:, 1, 0] = a("e13") if a("e13") is not None else z
        K[:, 1, 1] = a("e33")
        emin = np.linalg.eigvalsh(K)[:, 1:].min()
    else:
        K = np.zeros((n, 3, 3))
        for (i, j), nm in zip([(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)],
                              ["e11", "e12", "e13", "e22", "e23", "e33"]):
            v = a(nm)
            K[:, i, j] = K[:, j, i] = v if v is not None else z
        emin = np.linalg.eigvalsh(K).min()
    return float(CLIGHT / np.sqrt(emin)), float(np.sqrt(MU0 / (EPS0 * emin)))


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load(path):
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    for name in ("oracle_seismic2", "oracle_seismic25", "oracle_seismic3", "oracle_electromag2",
                 "oracle_electromag25", "oracle_electromag3"):
        fn = getattr(lib, name)
        fn.argtypes = [C.POINTER(OracleRun)]
        fn.restype = C.c_int32
    lib.oracle_scalar_mean.argtypes = [C.c_double, C.c_double, C.c_int32]
    lib.oracle_scalar_mean.restype = C.c_double
    lib.oracle_array_averaging2d.argtypes = [_D, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
    lib.oracle_array_averaging2d.restype = None
    return lib


_libs = {}


def lib(fast=False):
    """The checker library, or with fast=True the production-flag build
    used only for CPU-baseline timing."""
    path = FAST_LIB_PATH if fast else LIB_PATH
    if path not in _libs:
        _libs[path] = _load(path)
    return _libs[path]


def scalar_mean(a, b, code):
    return lib().oracle_scalar_mean(float(a), float(b), int(code))


def array_averaging2d(a, code, direction):
    """In-place sweep of averaging.f08:27-80 on a copy of the (ni, nj) array;
    direction one of '+i', '-i', '+j', '-j'."""
    out = np.asfortranarray(np.array(a, dtype=np.float64))
    ni, nj = out.shape
    d = {"+i": 0, "-i": 1, "+j": 2, "-j": 3}[direction]
    lib().oracle_array_averaging2d(out.ctypes.data_as(_D), ni, nj, int(code), d)
    return out


_ENTRY = {"elastic2": "oracle_seismic2", "elastic25": "oracle_seismic25", "em25": "oracle_electromag25",
          "elastic3": "oracle_seismic3", "em2": "oracle_electromag2",
          "em3": "oracle_electromag3"}


def run_oracle(cfg, *, steps=None, fields=(), field_in=None, nthreads=0, fast=False):
    """Run a configuration dict through the oracle.

    Returns (receivers[steps, nrec, 3], (velnorm, stressnorm), {field: array},
    unstable_step).  For "elastic25" the receivers, fields and ic_* arrays are
    complex128 and cfg["ky"] is the y wavenumber (likewise "em25").
    fast=True runs the production-flag build (CPU-baseline timing only)."""
    phys = cfg["physics"]
    is3d = phys.endswith("3")
    shape = tuple(cfg["shape"])
    nsteps = int(steps if steps is not None else cfg["time_steps"])
    r = OracleRun()
    if is3d:
        r.nx, r.ny, r.nz = shape
        r.dx, r.dy, r.dz = cfg["spacing"]
    else:
        r.nx, r.nz = shape
        r.ny = 1
        r.dx, r.dz = cfg["spacing"]
        r.dy = 1.0
    r.dt = cfg["dt"]
    r.ky = float(cfg.get("ky", 0.0))
    cplx = phys in ("elastic25", "em25")
    r.nsteps = nsteps
    cp = cfg["cpml"]
    src = cfg["src"]
    if is3d:
        r.isrc, r.jsrc, r.ksrc = src[0] + cp, src[1] + cp, src[2] + cp
    else:
        r.isrc, r.jsrc, r.ksrc = src[0] + cp, 1, src[1] + cp
    r.density_code = DENSITY_CODES[cfg.get("density_method", "arithmetic")]
    r.cpml = int(cp)
    pw = cfg.get("plane_wave")
    if pw:
        r.pw_kind = {"s": 1, "p": 2}[pw["kind"]]
        r.pw_wavelet = int(pw.get("wavelet", 0))
        for q in range(3):
            r.pw_angles[q] = float(pw["angles"][q])
        r.pw_f0 = float(pw["f0"])
        r.pw_amp = float(pw["amplitude"])
        if phys in ("em2", "em3"):
            r.pw_speed, r.pw_eta = em_background(cfg["arrays"], phys == "em2")
        else:
            r.pw_speed = background_speed(cfg["arrays"], phys == "elastic2")
    r.stencil = 1 if cfg.get("stencil", "reference") == "corrected" else 0
    r.nthreads = nthreads
    keep = []

    def ptr(a):
        a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))
        keep.append(a)
        return a.ctypes.data_as(_D)

    def cptr(a):   # complex grid array -> interleaved (re, im) doubles
        a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128).reshape(shape).ravel(order="F"))
        keep.append(a)
        return a.ctypes.data_as(_D)

    for name, a in cfg["arrays"].items():
        if cplx and name.startswith("ic_"):
            r.arr[ARRAY_IDS[name]] = cptr(a)
        else:
            r.arr[ARRAY_IDS[name]] = ptr(np.asarray(a).reshape(shape).ravel(order="F"))
    for comp, s in cfg["sources"].items():
        s = np.asarray(s, dtype=np.float64)
        if s.size < nsteps:
            raise ValueError("source series shorter than the run")
        r.src[SOURCE_IDS[comp]] = ptr(s[:nsteps])
    recs = np.asarray(cfg.get("receivers", np.zeros((0, 3))), dtype=np.int32).reshape(-1, 3)
    padded = recs.copy()
    padded[:, 0] += cp
    padded[:, 2] += cp
    if is3d:
        padded[:, 1] += cp
    else:
        padded[:, 1] = 1
    padded = np.ascontiguousarray(padded)
    keep.append(padded)
    r.nrec = padded.shape[0]
    r.rec = padded.ctypes.data_as(C.POINTER(C.c_int32))
    rec_out = np.zeros((nsteps, padded.shape[0], 6 if cplx else 3))
    r.rec_out = rec_out.ctypes.data_as(_D)
    fid = FIELD_IDS_ELASTIC if phys.startswith("elastic") else FIELD_IDS_EM
    if field_in:
        for name, a in field_in.items():
            r.field_in[fid[name]] = cptr(a) if cplx else ptr(np.asarray(a).reshape(shape).ravel(order="F"))
    outs = {}
    for name in fields:
        o = np.zeros(int(np.prod(shape)), dtype=np.complex128 if cplx else np.float64)
        outs[name] = o
        r.field_out[fid[name]] = o.ctypes.data_as(_D)
    vn = np.zeros(nsteps)
    sn = np.zeros(nsteps)
    r.velnorm = vn.ctypes.data_as(_D)
    r.stressnorm = sn.ctypes.data_as(_D)
    getattr(lib(fast), _ENTRY[phys])(C.byref(r))
    fields_out = {k: v.reshape(shape, order="F") for k, v in outs.items()}
    if cplx:
        rec_out = rec_out.view(np.complex128)
    return rec_out, (vn, sn), fields_out, r.unstable_step
