"""Pins of the oracle's reference-mode velocity stencils (CPU only).

The elastic velocity updates of cpmlfdtd.f08 take their stress differences
with 3-term weights (27 s0 - 27 s-1 + s-2 backward, -27 s0 + 27 s+1 - s+2
forward; seismic2 405-406 / 434-435, seismic25 1132-1134 / 1161-1163 /
1189-1191, seismic3 2022-2024 / 2063-2066 / 2107-2111), the consistent
4th-order difference with one term left out:

    backward:  27 (s0 - s-1) - (s+1 - s-2)  minus  (reference) = -s+1
    forward:   27 (s+1 - s0) - (s+2 - s-1)  minus  (reference) = +s-1

The corrected (4-term) mode is pinned by closed-form plane waves elsewhere.
Here one step is run from an arbitrary stress field with zero velocities,
no CPML, no sources, uniform density and no attenuation (so the stress
update leaves the stresses unchanged and every velocity is dt / rho times
its three stress differences), once in each mode; the difference of the two
runs must equal dt / rho times the omitted terms, computed below from the
stress field alone.  A mistake in a reference-mode weight, in which
neighbour a nest reads, or in the direction (T- vs T+) of a nest's stencil
changes that difference.
"""
import numpy as np

RHO = 1700.0


def _run(cfg, field_in, stencil, fields):
    from oracle.oracle_ctypes import run_oracle
    c = dict(cfg)
    c["stencil"] = stencil
    rec, _, out, us = run_oracle(c, steps=1, fields=fields, field_in=field_in)
    assert us == 0
    return out


def _prepare(cfg):
    from seidart_hip import models
    shape = cfg["shape"]
    cfg["arrays"].update(models.no_cpml(shape))
    cfg["arrays"]["rho"] = np.full(shape, RHO, order="F")
    for g in [k for k in cfg["arrays"] if k.startswith("gamma")]:
        cfg["arrays"][g] = np.zeros(shape, order="F")
    cfg["sources"] = {}
    cfg["receivers"] = np.zeros((0, 3), dtype=np.int32)
    return cfg


def _shift(a, axis, by):
    """b[i] = a[i + by] along `axis` (edges wrap; only inner cells are compared)."""
    return np.roll(a, -by, axis=axis)


def _inner(a, w=3):
    return a[tuple(slice(w, n - w) for n in a.shape)]


def test_seismic2_reference_stencil_omits_one_term():
    from seidart_hip import models
    cfg = _prepare(models.elastic2(nx=26, nz=22, cpml=4, steps=1, src=(10, 10), nrec=1))
    rng = np.random.default_rng(41)
    shape = cfg["shape"]
    s = {k: rng.uniform(-1e6, 1e6, shape) for k in ("sxx", "szz", "sxz")}
    fin = dict(s, vx=np.zeros(shape), vz=np.zeros(shape))
    ref = _run(cfg, fin, "reference", ("vx", "vz"))
    cor = _run(cfg, fin, "corrected", ("vx", "vz"))
    dx, dz = cfg["spacing"]
    dt = cfg["dt"]
    # vx: T-x sigma_xx, T-z sigma_xz; vz: T+x sigma_xz, T+z sigma_zz (axes x=0, z=1)
    want_vx = dt / RHO * (_shift(s["sxx"], 0, 1) / (24 * dx) + _shift(s["sxz"], 1, 1) / (24 * dz))
    want_vz = dt / RHO * (-_shift(s["sxz"], 0, -1) / (24 * dx) - _shift(s["szz"], 1, -1) / (24 * dz))
    for name, want in (("vx", want_vx), ("vz", want_vz)):
        got = ref[name] - cor[name]
        scale = np.abs(_inner(ref[name])).max()
        assert np.abs(_inner(want)).max() > 1e-3 * scale
        np.testing.assert_allclose(_inner(got), _inner(want), rtol=0, atol=1e-12 * scale, err_msg=name)


def test_seismic3_reference_stencil_omits_one_term():
    from seidart_hip import models
    cfg = _prepare(models.elastic3(nx=14, ny=12, nz=10, cpml=3, steps=1, src=(7, 6, 5), nrec=1))
    rng = np.random.default_rng(43)
    shape = cfg["shape"]
    names = ("sxx", "syy", "szz", "sxy", "sxz", "syz")
    s = {k: rng.uniform(-1e6, 1e6, shape) for k in names}
    fin = dict(s, vx=np.zeros(shape), vy=np.zeros(shape), vz=np.zeros(shape))
    ref = _run(cfg, fin, "reference", ("vx", "vy", "vz"))
    cor = _run(cfg, fin, "corrected", ("vx", "vy", "vz"))
    dx, dy, dz = cfg["spacing"]
    k = cfg["dt"] / RHO
    # backward (T-) differences omit -s(+1), forward (T+) ones +s(-1)
    want = {
        "vx": k * (_shift(s["sxx"], 0, 1) / (24 * dx) + _shift(s["sxy"], 1, 1) / (24 * dy)
                   + _shift(s["sxz"], 2, 1) / (24 * dz)),
        "vy": k * (-_shift(s["sxy"], 0, -1) / (24 * dx) - _shift(s["syy"], 1, -1) / (24 * dy)
                   + _shift(s["syz"], 2, 1) / (24 * dz)),
        "vz": k * (-_shift(s["sxz"], 0, -1) / (24 * dx) + _shift(s["syz"], 1, 1) / (24 * dy)
                   - _shift(s["szz"], 2, -1) / (24 * dz)),
    }
    for name, w in want.items():
        got = ref[name] - cor[name]
        scale = np.abs(_inner(ref[name])).max()
        assert np.abs(_inner(w)).max() > 1e-3 * scale
        np.testing.assert_allclose(_inner(got), _inner(w), rtol=0, atol=1e-12 * scale, err_msg=name)


def test_seismic25_reference_stencil_omits_one_term():
    """2.5D: x and z as in 3D, the y derivative is i ky s in both modes and
    cancels in the difference."""
    from seidart_hip import models
    cfg = _prepare(models.elastic25(nx=22, nz=20, cpml=4, steps=1, src=(10, 8), nrec=1))
    rng = np.random.default_rng(47)
    shape = cfg["shape"]
    names = ("sxx", "syy", "szz", "sxy", "sxz", "syz")
    s = {k: rng.uniform(-1e6, 1e6, shape) + 1j * rng.uniform(-1e6, 1e6, shape) for k in names}
    z = np.zeros(shape, dtype=np.complex128)
    fin = dict(s, vx=z, vy=z.copy(), vz=z.copy())
    ref = _run(cfg, fin, "reference", ("vx", "vy", "vz"))
    cor = _run(cfg, fin, "corrected", ("vx", "vy", "vz"))
    dx, dz = cfg["spacing"]
    k = cfg["dt"] / RHO
    want = {
        "vx": k * (_shift(s["sxx"], 0, 1) / (24 * dx) + _shift(s["sxz"], 1, 1) / (24 * dz)),
        "vy": k * (-_shift(s["sxy"], 0, -1) / (24 * dx) + _shift(s["syz"], 1, 1) / (24 * dz)),
        "vz": k * (-_shift(s["sxz"], 0, -1) / (24 * dx) - _shift(s["szz"], 1, -1) / (24 * dz)),
    }
    for name, w in want.items():
        got = ref[name] - cor[name]
        scale = np.abs(_inner(ref[name])).max()
        assert np.abs(_inner(w)).max() > 1e-3 * scale
        np.testing.assert_allclose(_inner(got), _inner(w), rtol=0, atol=1e-12 * scale, err_msg=name)
