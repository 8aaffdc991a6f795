"""Pins of the 2.5D elastic oracle (oracle_seismic25, cpmlfdtd.f08:600-1337)
and of array_averaging2d (averaging.f08:27-80) that do not restate their
formulas:

* ky = 0 with only the 2D stiffness entries non-zero, identity CPML and the
  'none' density code: the x-z part of the 2.5D scheme is the 2D scheme, so
  vx, vz, sxx, szz, sxz must equal the (separately pinned) 2D oracle's
  fields bit for bit and every other field stays zero.  Fails on a wrong
  stiffness slot or attenuation array in the x-z stresses, a wrong density
  array, wrong nest ranges or a wrong stress-source offset.
* complex linearity: the scheme is C-linear (real coefficients, i*ky), so a
  complex initial condition a + ib gives run(a) + i*run(b).  Fails when the
  real and imaginary parts are mixed in any other way (e.g. ky*conj(f)).
* the y mode: fields uniform in x and z (corrected stencil, isotropic
  medium) reduce to three decoupled leapfrog oscillators vy<->syy (c22),
  vx<->sxy (c66), vz<->syz (c44) with v_n = cos(n t) - w^2/(2 sin t) sin(n t),
  sin(t/2) = w/2, w = ky*dt*sqrt(c/rho), and syy_1 = +i*ky*dt*c22*v_0.  Fails
  on a wrong sign of i*ky, a wrong stress feeding a velocity, a wrong modulus
  or density in the y terms.
* array_averaging2d: a unit impulse smeared by '+i' / '+j' (arithmetic) decays
  as 2^-n along the sweep (the recursion reuses updated neighbours), while
  '-i' / '-j' spread it over exactly two cells (original neighbours).

The per-nest choice of full or half CPML coefficients and the 3-term velocity
stencils are restated from the cited reference lines only.
"""
import math

import numpy as np

from oracle.oracle_ctypes import array_averaging2d, run_oracle
from seidart_hip import models

VOIGT = ["c11", "c12", "c13", "c14", "c15", "c16", "c22", "c23", "c24", "c25", "c26",
         "c33", "c34", "c35", "c36", "c44", "c45", "c46", "c55", "c56", "c66"]
FIELDS = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")


def _el25_from_el2(cfg2):
    shape = cfg2["shape"]
    arrays = {k: np.zeros(shape, order="F") for k in VOIGT}
    for k in ("c11", "c13", "c15", "c33", "c35", "c55", "rho"):
        arrays[k] = cfg2["arrays"][k]
    for g in ("gamma_x", "gamma_y", "gamma_z", "gamma_xy", "gamma_xz", "gamma_yz"):
        arrays[g] = cfg2["arrays"].get(g, np.zeros(shape, order="F"))
    arrays.update(models.no_cpml(shape))
    cfg = dict(cfg2)
    cfg.update(physics="elastic25", arrays=arrays, ky=0.0)
    return cfg


def test_ky0_reduces_to_2d_scheme():
    cfg2 = models.elastic2(nx=60, nz=50, cpml=6, steps=120, src=(30, 20), rec_z=20, nrec=6,
                           density_method="none", gamma=40.0, with_cpml=False)
    s = cfg2["sources"]["z"]
    cfg2["sources"] = {"xx": 0.6 * s, "zz": -s}
    cfg2["arrays"].update(models.no_cpml(cfg2["shape"]))
    cfg25 = _el25_from_el2(cfg2)
    _, _, f2, _ = run_oracle(cfg2, fields=("vx", "vz", "sxx", "szz", "sxz"))
    _, _, f25, _ = run_oracle(cfg25, fields=FIELDS)
    assert np.abs(f2["vz"]).max() > 0
    for k in ("vx", "vz", "sxx", "szz", "sxz"):
        assert np.array_equal(f25[k].real, f2[k]), k
        assert not np.any(f25[k].imag)
    for k in ("vy", "syy", "sxy", "syz"):
        assert not np.any(f25[k]), k


def test_complex_linearity():
    cfg = models.elastic25(nx=50, nz=44, cpml=6, steps=90, src=(25, 15), nrec=4, ky=0.2)
    cfg["sources"] = {}
    shape = cfg["shape"]
    x, z = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    rng = np.random.default_rng(3)
    ics = {}
    for c in ("ic_x", "ic_y", "ic_z"):
        x0, z0 = rng.uniform(20, 40), rng.uniform(18, 36)
        g = np.exp(-((x - x0) ** 2 + (z - z0) ** 2) / 18.0)
        ics[c] = (rng.normal() * g, rng.normal() * g)
    outs = []
    for part in ("re", "im", "both"):
        c = dict(cfg)
        c["arrays"] = dict(cfg["arrays"])
        for k, (a, b) in ics.items():
            c["arrays"][k] = {"re": a + 0j, "im": b + 0j, "both": a + 1j * b}[part]
        outs.append(run_oracle(c, fields=FIELDS)[2])
    for k in FIELDS:
        ref = outs[0][k] + 1j * outs[1][k]
        err = np.linalg.norm(outs[2][k] - ref) / np.linalg.norm(ref)
        assert err < 1e-13, (k, err)


def test_y_mode_closed_form():
    nx = nz = 100
    cp = 10
    NX, NZ = nx + 2 * cp, nz + 2 * cp
    shape = (NX, NZ)
    lam, mu, rho = 7.0e9, 3.0e9, 1800.0
    arrays = {k: np.zeros(shape, order="F") for k in VOIGT}
    for k in ("c11", "c22", "c33"):
        arrays[k][:] = lam + 2 * mu
    for k in ("c12", "c13", "c23"):
        arrays[k][:] = lam
    for k in ("c44", "c55", "c66"):
        arrays[k][:] = mu
    # distinct y-coupling moduli so a swapped modulus shows
    arrays["c44"][:] = 0.8 * mu
    arrays["c66"][:] = 1.3 * mu
    arrays["rho"] = np.full(shape, rho, order="F")
    arrays.update(models.no_cpml(shape))
    v0 = {"ic_x": 1.0, "ic_y": 1.0, "ic_z": 1.0}
    for k, v in v0.items():
        arrays[k] = np.full(shape, v + 0j, order="F")
    ky, dt, nsteps = 0.9, 2.0e-4, 12
    cfg = dict(physics="elastic25", shape=shape, spacing=(1.0, 1.0), cpml=cp, dt=dt, time_steps=nsteps,
               src=(50, 50), arrays=arrays, ky=ky, sources={}, receivers=np.zeros((0, 3), np.int32),
               density_method="arithmetic", stencil="corrected")
    c = (NX // 2, NZ // 2)
    for n in (1, 5, nsteps):
        _, _, f, _ = run_oracle(cfg, steps=n, fields=FIELDS)
        for vname, sname, modulus in (("vy", "syy", lam + 2 * mu), ("vx", "sxy", 1.3 * mu),
                                      ("vz", "syz", 0.8 * mu)):
            w = ky * dt * math.sqrt(modulus / rho)
            th = 2.0 * math.asin(w / 2.0)
            expect = math.cos(n * th) - w * w / (2.0 * math.sin(th)) * math.sin(n * th)
            got = f[vname][c]
            assert abs(got.real - expect) < 1e-12 and abs(got.imag) < 1e-15, (vname, n, got, expect)
            if n == 1:
                s = f[sname][c]
                assert abs(s.imag - ky * dt * modulus) < 1e-9 * ky * dt * modulus, (sname, s)
                assert abs(s.real) < 1e-6


def test_array_averaging_recursion():
    a = np.zeros((6, 4))
    a[0, :] = 1.0
    out = array_averaging2d(a, 3, "+i")
    assert np.allclose(out[:, 1], 2.0 ** -np.arange(6), rtol=0, atol=0)
    b = np.zeros((6, 4))
    b[2, :] = 1.0
    out = array_averaging2d(b, 3, "-i")
    assert np.array_equal(out[:, 0], [0, 0.5, 0.5, 0, 0, 0])
    c = np.zeros((4, 6))
    c[:, 0] = 1.0
    out = array_averaging2d(c, 3, "+j")
    assert np.array_equal(out[1, :], 2.0 ** -np.arange(6))
    d = np.zeros((4, 6))
    d[:, 3] = 1.0
    out = array_averaging2d(d, 3, "-j")
    assert np.array_equal(out[2, :], [0, 0, 0.5, 0.5, 0, 0])
    # 'none' leaves the array unchanged; geometric/harmonic of equal values too
    r = np.random.default_rng(0).uniform(1, 2, (5, 5))
    assert np.array_equal(array_averaging2d(r, 4, "+i"), r)
    u = np.full((5, 5), 4.0)
    assert np.array_equal(array_averaging2d(u, 2, "+j"), u)
    assert np.array_equal(array_averaging2d(u, 1, "-i"), u)


# ------------------------------------------------------------- electromag25
EPS0 = 8.85418782e-12
MU0 = 4.0 * math.pi * 1e-7


def _em25_uniform(nsteps, ky, eps_r, sig, ic):
    nx = nz = 100
    cp = 10
    shape = (nx + 2 * cp, nz + 2 * cp)
    arrays = {}
    for k, v in (("e11", eps_r), ("e22", eps_r), ("e33", eps_r), ("s11", sig), ("s22", sig), ("s33", sig)):
        arrays[k] = np.full(shape, v, order="F")
    for k in ("e12", "e13", "e23", "s12", "s13", "s23"):
        arrays[k] = np.zeros(shape, order="F")
    arrays.update(models.no_cpml(shape))
    for k, v in ic.items():
        arrays[k] = np.full(shape, v + 0j, order="F")
    d = 0.05
    dt = 0.2 * d / 3e8
    cfg = dict(physics="em25", shape=shape, spacing=(d, d), cpml=cp, dt=dt, time_steps=nsteps, src=(50, 50),
               arrays=arrays, ky=ky, sources={}, receivers=np.zeros((0, 3), np.int32))
    return cfg, dt


def test_em25_y_mode_closed_form():
    """Uniform fields: Ez<->Hx and Ex<->Hz are leapfrog oscillators with
    w = ky*dt/sqrt(mu0*eps), E_n = cos(n t) - w^2/(2 sin t) sin(n t); Hx_1 =
    -i*ky*dt/mu0*Ez_0 fixes the sign of the y derivative; Ey stays put."""
    ky, eps_r = 40.0, 4.0
    for n in (1, 4, 10):
        cfg, dt = _em25_uniform(n, ky, eps_r, 0.0, {"ic_x": 1.0, "ic_y": 1.0, "ic_z": 1.0})
        _, _, f, _ = run_oracle(cfg, fields=("ex", "ey", "ez", "hx", "hz"))
        c = (60, 60)
        w = ky * dt / math.sqrt(MU0 * EPS0 * eps_r)
        th = 2.0 * math.asin(w / 2.0)
        expect = math.cos(n * th) - w * w / (2.0 * math.sin(th)) * math.sin(n * th)
        for k in ("ex", "ez"):
            assert abs(f[k][c].real - expect) < 1e-12 and abs(f[k][c].imag) < 1e-15, (k, n, f[k][c], expect)
        assert abs(f["ey"][c] - 1.0) < 1e-15
        if n == 1:
            assert abs(f["hx"][c].imag + ky * dt / MU0) < 1e-12 * ky * dt / MU0
            assert abs(f["hz"][c].imag - ky * dt / MU0) < 1e-12 * ky * dt / MU0


def test_em25_conductive_decay():
    """ky = 0, uniform E: E_n = (1 - dt*sigma/eps)^n E_0 (sigma acts on E_old)."""
    eps_r, sig, n = 6.0, 2e-3, 12
    cfg, dt = _em25_uniform(n, 0.0, eps_r, sig, {"ic_x": 0.7, "ic_y": -0.3, "ic_z": 0.2})
    _, _, f, _ = run_oracle(cfg, fields=("ex", "ey", "ez"))
    g = (1.0 - dt * sig / (EPS0 * eps_r)) ** n
    for k, v in (("ex", 0.7), ("ey", -0.3), ("ez", 0.2)):
        got = f[k][60, 60]
        assert abs(got.real - v * g) < 1e-13 and got.imag == 0.0, (k, got, v * g)


def test_em25_complex_linearity():
    shape = (70, 60)
    rng = np.random.default_rng(5)
    cfg, _ = _em25_uniform(1, 25.0, 5.0, 1e-3, {})
    cfg["arrays"] = {k: (v[:70, :60] if hasattr(v, "shape") else v) for k, v in cfg["arrays"].items()}
    cfg["shape"] = shape
    cfg["arrays"].update(models.cpml_arrays(shape, (0.05, 0.05), 8, 1.5e8, cfg["dt"], 1e8))
    cfg["cpml"] = 8
    cfg["time_steps"] = 80
    cfg["src"] = (20, 20)
    x, z = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    ics = {}
    for c in ("ic_x", "ic_y", "ic_z"):
        x0, z0 = rng.uniform(25, 45), rng.uniform(20, 40)
        gss = np.exp(-((x - x0) ** 2 + (z - z0) ** 2) / 15.0)
        ics[c] = (rng.normal() * gss, rng.normal() * gss)
    outs = []
    for part in ("re", "im", "both"):
        c = dict(cfg)
        c["arrays"] = dict(cfg["arrays"])
        for k, (a, b) in ics.items():
            c["arrays"][k] = {"re": a + 0j, "im": b + 0j, "both": a + 1j * b}[part]
        outs.append(run_oracle(c, fields=("ex", "ey", "ez", "hx", "hy", "hz"))[2])
    for k in ("ex", "ey", "ez", "hx", "hy", "hz"):
        ref = outs[0][k] + 1j * outs[1][k]
        err = np.linalg.norm(outs[2][k] - ref) / np.linalg.norm(ref)
        assert err < 1e-13, (k, err)
