"""Pins of the CPU oracle (oracle/oracle.c) that do not restate its formulas:

* linear-field patch tests: for v linear in space every 4th-order staggered
  difference is exact, so one stress update equals dt*C*strain/(1+gamma*dt)
  in closed form.  Fails on a transposed stiffness entry, a wrong Voigt
  pairing, a wrong difference scale or a wrong attenuation form.
* CPML recursion closed form: uniform a, b, kappa (different for the full and
  the half set) and a linear v field give psi_n = a d (1 - b^n)/(1 - b) per
  derivative; fails on a swapped coefficient set or a wrong recursion.
* plane waves (corrected stencil): a right-going Christoffel eigen-mode
  started with consistently staggered stresses must travel at the
  eigen-speed without splitting; fails on a wrong staggering offset, a wrong
  Christoffel entry, wrong density or a wrong time step.  Run along every
  axis of the 2D and 3D elastic solvers and the EM solvers (Yee).
* EM one-step closed forms: E_1 = dt * inv(eps0*eps_r) * curl H for a linear
  H field, E_1 = (I - dt inv(eps) sigma) E_0 for a uniform E field, and
  H_1 = -dt/mu0 curl E; the adjugate/determinant inverse is checked against
  numpy.linalg.inv.
* linearity, superposition, zero-in/zero-out and the blow-up threshold.
* scalar_mean against the textbook means.

The reference's 3-term velocity stencils and its point-source offsets are
restated from the cited lines only (transcription known-answer tests below):
nothing outside the reference fixes them.
"""
import math

import numpy as np
import pytest

from oracle.oracle_ctypes import run_oracle, scalar_mean
from seidart_hip import models

MU0 = 4.0 * math.pi * 1e-7
EPS0 = 8.85418782e-12

VOIGT = ["c11", "c12", "c13", "c14", "c15", "c16", "c22", "c23", "c24", "c25", "c26",
         "c33", "c34", "c35", "c36", "c44", "c45", "c46", "c55", "c56", "c66"]
VIJ = [(0, 0), (0, 1), (0, 2), (0, 3), (0, 4), (0, 5), (1, 1), (1, 2), (1, 3), (1, 4), (1, 5),
       (2, 2), (2, 3), (2, 4), (2, 5), (3, 3), (3, 4), (3, 5), (4, 4), (4, 5), (5, 5)]


def full(shape, v):
    return np.full(shape, float(v), order="F")


def base_cfg(physics, shape, spacing, dt, steps, src, stencil="reference", density="arithmetic"):
    srcs = {c: np.zeros(steps) for c in ("x", "y", "z")}
    return dict(physics=physics, shape=shape, spacing=spacing, cpml=0, dt=dt, time_steps=steps,
                src=src, arrays={}, sources=srcs, receivers=np.zeros((0, 3), dtype=np.int32),
                stencil=stencil, density_method=density)


def coords(shape, spacing):
    axes = [np.arange(n, dtype=float) * h for n, h in zip(shape, spacing)]
    return np.meshgrid(*axes, indexing="ij")


# ------------------------------------------------------------ scalar_mean ----
def test_scalar_mean_textbook():
    a, b = 917.0, 2000.0
    assert scalar_mean(a, b, 1) == pytest.approx(2 * a * b / (a + b), rel=1e-15)
    assert scalar_mean(a, b, 2) == pytest.approx(math.sqrt(a * b), rel=1e-15)
    assert scalar_mean(a, b, 3) == pytest.approx(0.5 * (a + b), rel=1e-15)
    assert scalar_mean(a, b, 4) == a
    # harmonic <= geometric <= arithmetic for positive inputs
    assert scalar_mean(a, b, 1) < scalar_mean(a, b, 2) < scalar_mean(a, b, 3)


# --------------------------------------------------------- patch tests ----
def test_elastic2_linear_patch():
    nx, nz, dx, dz, dt = 20, 18, 0.7, 1.1, 1e-4
    cfg = base_cfg("elastic2", (nx, nz), (dx, dz), dt, 1, (8, 8))
    c = dict(c11=1.3e10, c13=0.55e10, c15=0.21e10, c33=1.6e10, c35=-0.13e10, c55=0.31e10)
    g = dict(gamma_x=20.0, gamma_z=35.0, gamma_xz=50.0)
    for k, v in {**c, **g, "rho": 1500.0}.items():
        cfg["arrays"][k] = full((nx, nz), v)
    X, Z = coords((nx, nz), (dx, dz))
    al, be, ga, de = 0.3, -0.7, 1.1, 0.45       # vx = al x + be z, vz = ga x + de z
    cfg["arrays"]["ic_x"] = al * X + be * Z
    cfg["arrays"]["ic_z"] = ga * X + de * Z
    _, _, f, _ = run_oracle(cfg, fields=("sxx", "szz", "sxz"))
    shear = ga + be
    exp_xx = (c["c11"] * al + c["c13"] * de + c["c15"] * shear) * dt / (1 + g["gamma_x"] * dt)
    exp_zz = (c["c13"] * al + c["c33"] * de + c["c35"] * shear) * dt / (1 + g["gamma_z"] * dt)
    exp_xz = (c["c15"] * al + c["c35"] * de + c["c55"] * shear) * dt / (1 + g["gamma_xz"] * dt)
    # nest ranges (1-based): L1 i=2..nx-2, j=3..nz-1 ; L2 i=3..nx-1, j=2..nz-2
    m1 = np.zeros((nx, nz), bool); m1[1:nx - 2, 2:nz - 1] = True
    m2 = np.zeros((nx, nz), bool); m2[2:nx - 1, 1:nz - 2] = True
    np.testing.assert_allclose(f["sxx"][m1], exp_xx, rtol=1e-11)
    np.testing.assert_allclose(f["szz"][m1], exp_zz, rtol=1e-11)
    np.testing.assert_allclose(f["sxz"][m2], exp_xz, rtol=1e-11)
    assert np.all(f["sxx"][~m1] == 0) and np.all(f["sxz"][~m2] == 0)


def _random_voigt(rng):
    A = rng.uniform(-1.0, 1.0, (6, 6))
    return (A + A.T) * 1e10


def test_elastic3_linear_patch():
    rng = np.random.default_rng(7)
    nx, ny, nz = 13, 12, 11
    sp = (0.6, 0.8, 1.1)
    dt = 1e-4
    cfg = base_cfg("elastic3", (nx, ny, nz), sp, dt, 1, (5, 5, 5))
    C = _random_voigt(rng)
    for name, (i, j) in zip(VOIGT, VIJ):
        cfg["arrays"][name] = full((nx, ny, nz), C[i, j])
    gam = dict(zip(["gamma_x", "gamma_y", "gamma_z", "gamma_xy", "gamma_xz", "gamma_yz"],
                   rng.uniform(0.0, 100.0, 6)))
    for k, v in gam.items():
        cfg["arrays"][k] = full((nx, ny, nz), v)
    cfg["arrays"]["rho"] = full((nx, ny, nz), 1200.0)
    G = rng.uniform(-1.0, 1.0, (3, 3))          # v_i = sum_j G_ij x_j
    X = coords((nx, ny, nz), sp)
    for q, name in enumerate(["ic_x", "ic_y", "ic_z"]):
        cfg["arrays"][name] = G[q, 0] * X[0] + G[q, 1] * X[1] + G[q, 2] * X[2]
    names = ("sxx", "syy", "szz", "sxy", "sxz", "syz")
    _, _, f, _ = run_oracle(cfg, fields=names)
    # engineering strain rate in Voigt order xx, yy, zz, yz, xz, xy
    eps = np.array([G[0, 0], G[1, 1], G[2, 2], G[1, 2] + G[2, 1], G[0, 2] + G[2, 0], G[0, 1] + G[1, 0]])
    sig = C @ eps * dt
    def box(i0, i1, j0, j1, k0, k1):
        m = np.zeros((nx, ny, nz), bool)
        m[i0 - 1:i1, j0 - 1:j1, k0 - 1:k1] = True
        return m
    L1 = box(2, nx - 2, 3, ny - 1, 3, nz - 1)
    L2 = box(3, nx - 1, 2, ny - 2, 3, nz - 1)
    L3 = box(3, nx - 1, 3, ny - 1, 2, nz - 2)
    L4 = box(2, nx - 2, 2, ny - 2, 2, nz - 2)
    checks = [("sxx", 0, "gamma_x", L1), ("syy", 1, "gamma_y", L1), ("szz", 2, "gamma_z", L1),
              ("sxy", 5, "gamma_xy", L2), ("sxz", 4, "gamma_xz", L3), ("syz", 3, "gamma_yz", L4)]
    for name, row, gname, mask in checks:
        np.testing.assert_allclose(f[name][mask], sig[row] / (1 + gam[gname] * dt), rtol=1e-10,
                                   err_msg=name)
        assert np.all(f[name][~mask] == 0), name


def test_elastic2_cpml_recursion_closed_form():
    """Uniform coefficients, different for the full (F) and half (H) sets;
    per cpmlfdtd.f08:332-340 sigma_xx/zz use dvx_dx:H dvx_dz:H dvz_dz:F
    dvz_dx:F, and sigma_xz uses H for all four (370-378)."""
    nx = nz = 60
    dx, dz, dt, n = 1.0, 1.3, 1e-4, 4
    cfg = base_cfg("elastic2", (nx, nz), (dx, dz), dt, n, (30, 30), stencil="corrected")
    c = dict(c11=1.3e10, c13=0.55e10, c15=0.21e10, c33=1.6e10, c35=-0.13e10, c55=0.31e10)
    for k, v in {**c, "rho": 1500.0}.items():
        cfg["arrays"][k] = full((nx, nz), v)
    F = dict(a=0.3, b=0.6, k=1.5)
    H = dict(a=-0.2, b=0.8, k=2.5)
    for s, sfx in ((F, "_cpml"), (H, "_half_cpml")):
        cfg["arrays"]["acoef" + sfx] = full((nx, nz), s["a"])
        cfg["arrays"]["bcoef" + sfx] = full((nx, nz), s["b"])
        cfg["arrays"]["kappa" + sfx] = full((nx, nz), s["k"])
    X, Z = coords((nx, nz), (dx, dz))
    al, be, ga, de = 0.3, -0.7, 1.1, 0.45
    cfg["arrays"]["ic_x"] = al * X + be * Z
    cfg["arrays"]["ic_z"] = ga * X + de * Z
    _, _, f, _ = run_oracle(cfg, fields=("sxx", "szz", "sxz", "vx"))

    def corrected(d, s, m):   # derivative seen at step m (1-based)
        psi = s["a"] * d * (1 - s["b"] ** m) / (1 - s["b"])
        return d / s["k"] + psi

    sxx = szz = sxz = 0.0
    for m in range(1, n + 1):
        vx_x, vx_z, vz_z, vz_x = corrected(al, H, m), corrected(be, H, m), corrected(de, F, m), corrected(ga, F, m)
        sxx += (c["c11"] * vx_x + c["c13"] * vz_z + c["c15"] * (vz_x + vx_z)) * dt
        szz += (c["c13"] * vx_x + c["c33"] * vz_z + c["c35"] * (vz_x + vx_z)) * dt
        vx_x, vx_z, vz_z, vz_x = (corrected(v, H, m) for v in (al, be, de, ga))
        sxz += (c["c15"] * vx_x + c["c35"] * vz_z + c["c55"] * (vz_x + vx_z)) * dt
    core = (slice(24, 36), slice(24, 36))
    np.testing.assert_allclose(f["sxx"][core], sxx, rtol=1e-10)
    np.testing.assert_allclose(f["szz"][core], szz, rtol=1e-10)
    np.testing.assert_allclose(f["sxz"][core], sxz, rtol=1e-10)
    np.testing.assert_allclose(f["vx"][core], (al * X + be * Z)[core], rtol=1e-12, atol=1e-12)


# -------------------------------------------------------- plane waves ----
def _gauss(x, x0, w):
    return np.exp(-((x - x0) / w) ** 2)


def _christoffel(C, axis):
    """Rows of sigma components driven along `axis` and the Christoffel
    matrix K (u = (vx, vy, vz))."""
    # (stress Voigt index per velocity component) for propagation along x, y, z
    rows = {0: [0, 5, 4], 1: [5, 1, 3], 2: [4, 3, 2]}[axis]
    K = np.array([[C[rows[i], rows[j]] for j in range(3)] for i in range(3)])
    return rows, K


# stress staggering offsets (in cells) along each axis, from the one-sided
# differences of each nest (cpmlfdtd.f08:1761-1992): +1/2 for D+, -1/2 for D-
STAGGER3 = {  # (stress voigt index, axis) -> offset
    (0, 0): +0.5, (5, 0): -0.5, (4, 0): -0.5,
    (5, 1): +0.5, (1, 1): -0.5, (3, 1): +0.5,
    (4, 2): +0.5, (3, 2): +0.5, (2, 2): -0.5,
}
SNAME = {0: "sxx", 1: "syy", 2: "szz", 3: "syz", 4: "sxz", 5: "sxy"}


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_elastic3_plane_wave(axis):
    C = models.rotate_voigt(models.ice_voigt(), models.euler_zyz(0.4, 0.9, 1.7))
    rho = 917.0
    n_long, n_tr = 200, 11
    h, htr = 1.0, 30.0
    shape = [n_tr] * 3
    shape[axis] = n_long
    sp = [htr] * 3
    sp[axis] = h
    rows, K = _christoffel(C, axis)
    lam, vec = np.linalg.eigh(K / rho)
    c, e = math.sqrt(lam[-1]), vec[:, -1]     # quasi-P
    dt = 0.25 * h / c
    nsteps = 60
    cfg = base_cfg("elastic3", tuple(shape), tuple(sp), dt, nsteps, (5, 5, 5), stencil="corrected")
    for name, (i, j) in zip(VOIGT, VIJ):
        cfg["arrays"][name] = full(tuple(shape), C[i, j])
    cfg["arrays"]["rho"] = full(tuple(shape), rho)
    X = coords(tuple(shape), tuple(sp))[axis]
    x0, w = 60.0 * h, 8.0 * h
    f0 = _gauss(X, x0, w)
    for q, name in enumerate(["vx", "vy", "vz"]):
        cfg["arrays"]["ic_" + "xyz"[q]] = e[q] * f0
    field_in = {}
    for q in range(3):      # right-going wave: sigma = -rho c e f, at t = -dt/2
        r = rows[q]
        off = STAGGER3[(r, axis)]
        field_in[SNAME[r]] = -rho * c * e[q] * _gauss(X + off * h + c * dt / 2, x0, w)
    _, _, f, _ = run_oracle(cfg, fields=("vx", "vy", "vz"), field_in=field_in)
    sl = [n_tr // 2] * 3
    sl[axis] = slice(None)
    sl = tuple(sl)
    xl = X[sl]
    exp = _gauss(xl, x0 + c * nsteps * dt, w)
    for q, name in enumerate(["vx", "vy", "vz"]):
        got = f[name][sl]
        err = np.linalg.norm(got - e[q] * exp) / np.linalg.norm(exp)
        assert err < 1e-3, (axis, name, err)


@pytest.mark.parametrize("axis", [0, 2])
def test_elastic2_plane_wave(axis):
    C = models.rotate_voigt(models.ice_voigt(), models.rot_y(math.radians(25.0)))
    rho = 917.0
    # x-z plane: u = (vx, vz); along x the stresses are (sxx, sxz), along z (sxz, szz)
    if axis == 0:
        K = np.array([[C[0, 0], C[0, 4]], [C[0, 4], C[4, 4]]])
        snames, offs = ("sxx", "sxz"), (+0.5, -0.5)        # L1 D+x, L2 D-x
    else:
        K = np.array([[C[4, 4], C[2, 4]], [C[2, 4], C[2, 2]]])
        snames, offs = ("sxz", "szz"), (+0.5, -0.5)        # L2 D+z, L1 D-z
    lam, vec = np.linalg.eigh(K / rho)
    c, e = math.sqrt(lam[-1]), vec[:, -1]
    n_long, n_tr, h, htr = 240, 11, 1.0, 30.0
    shape = (n_long, n_tr) if axis == 0 else (n_tr, n_long)
    sp = (h, htr) if axis == 0 else (htr, h)
    dt = 0.25 * h / c
    nsteps = 60
    cfg = base_cfg("elastic2", shape, sp, dt, nsteps, (4, 4), stencil="corrected")
    for name, (i, j) in zip(["c11", "c13", "c15", "c33", "c35", "c55"],
                            [(0, 0), (0, 2), (0, 4), (2, 2), (2, 4), (4, 4)]):
        cfg["arrays"][name] = full(shape, C[i, j])
    cfg["arrays"]["rho"] = full(shape, rho)
    X = coords(shape, sp)[0 if axis == 0 else 1]
    x0, w = 70.0, 8.0
    f0 = _gauss(X, x0, w)
    cfg["arrays"]["ic_x"] = e[0] * f0
    cfg["arrays"]["ic_z"] = e[1] * f0
    field_in = {snames[q]: -rho * c * e[q] * _gauss(X + offs[q] * h + c * dt / 2, x0, w) for q in range(2)}
    _, _, f, _ = run_oracle(cfg, fields=("vx", "vz"), field_in=field_in)
    sl = (slice(None), n_tr // 2) if axis == 0 else (n_tr // 2, slice(None))
    exp = _gauss(X[sl], x0 + c * nsteps * dt, w)
    for q, name in enumerate(["vx", "vz"]):
        err = np.linalg.norm(f[name][sl] - e[q] * exp) / np.linalg.norm(exp)
        assert err < 1e-3, (axis, name, err)


def test_elastic2_wrong_offset_fails():
    """The plane-wave pin is sensitive: starting the stresses half a cell off
    (as a wrong staggering would) leaves an error far above the bar."""
    C = models.rotate_voigt(models.ice_voigt(), models.rot_y(math.radians(25.0)))
    rho = 917.0
    K = np.array([[C[0, 0], C[0, 4]], [C[0, 4], C[4, 4]]])
    lam, vec = np.linalg.eigh(K / rho)
    c, e = math.sqrt(lam[-1]), vec[:, -1]
    shape, sp, h = (240, 11), (1.0, 30.0), 1.0
    dt = 0.25 / c
    cfg = base_cfg("elastic2", shape, sp, dt, 60, (4, 4), stencil="corrected")
    for name, (i, j) in zip(["c11", "c13", "c15", "c33", "c35", "c55"],
                            [(0, 0), (0, 2), (0, 4), (2, 2), (2, 4), (4, 4)]):
        cfg["arrays"][name] = full(shape, C[i, j])
    cfg["arrays"]["rho"] = full(shape, rho)
    X = coords(shape, sp)[0]
    f0 = _gauss(X, 70.0, 8.0)
    cfg["arrays"]["ic_x"] = e[0] * f0
    cfg["arrays"]["ic_z"] = e[1] * f0
    field_in = {"sxx": -rho * c * e[0] * _gauss(X - 0.5 + c * dt / 2, 70.0, 8.0),
                "sxz": -rho * c * e[1] * _gauss(X + 0.5 + c * dt / 2, 70.0, 8.0)}
    _, _, f, _ = run_oracle(cfg, fields=("vx",), field_in=field_in)
    exp = _gauss(X[:, 5], 70.0 + c * 60 * dt, 8.0)
    err = np.linalg.norm(f["vx"][:, 5] - e[0] * exp) / np.linalg.norm(exp)
    assert err > 2e-2


def test_em2_plane_wave():
    """Yee along x: Ez/Hy right-going wave at c = 1/sqrt(mu0 eps0 eps33)."""
    n_long, n_tr, h = 300, 11, 0.05
    shape, sp = (n_long, n_tr), (h, 30 * h)
    er = 4.0
    c = 1.0 / math.sqrt(MU0 * EPS0 * er)
    dt = 0.25 * h / c
    nsteps = 120
    cfg = base_cfg("em2", shape, sp, dt, nsteps, (4, 4))
    cfg["arrays"].update(e11=full(shape, 6.0), e13=full(shape, 0.0), e33=full(shape, er),
                         s11=full(shape, 0.0), s13=full(shape, 0.0), s33=full(shape, 0.0))
    X = coords(shape, sp)[0]
    x0, w = 100 * h, 16 * h
    cfg["arrays"]["ic_z"] = _gauss(X, x0, w)
    # dHy/dt = (1/mu0) dEz/dx; Ez(i) at x_i, Hy(i) at x_i + h/2, t = -dt/2
    field_in = {"hy": -(EPS0 * er) * c * _gauss(X + 0.5 * h + c * dt / 2, x0, w)}
    _, _, f, _ = run_oracle(cfg, fields=("ez",), field_in=field_in)
    sl = (slice(None), n_tr // 2)
    exp = _gauss(X[sl], x0 + c * nsteps * dt, w)
    err = np.linalg.norm(f["ez"][sl] - exp) / np.linalg.norm(exp)
    assert err < 3e-3, err


def test_em3_plane_wave():
    """Yee along z: Ex/Hy right-going wave at c = 1/sqrt(mu0 eps0 eps11)."""
    n_long, n_tr, h = 300, 11, 0.05
    shape, sp = (n_tr, n_tr, n_long), (30 * h, 30 * h, h)
    er = 5.0
    c = 1.0 / math.sqrt(MU0 * EPS0 * er)
    dt = 0.25 * h / c
    nsteps = 120
    cfg = base_cfg("em3", shape, sp, dt, nsteps, (3, 3, 3))
    for k, v in dict(e11=er, e22=7.0, e33=9.0, e12=0.0, e13=0.0, e23=0.0).items():
        cfg["arrays"][k] = full(shape, v)
    for k in ("s11", "s12", "s13", "s22", "s23", "s33"):
        cfg["arrays"][k] = full(shape, 0.0)
    Z = coords(shape, sp)[2]
    x0, w = 100 * h, 16 * h
    cfg["arrays"]["ic_x"] = _gauss(Z, x0, w)
    # dHy/dt = (1/mu0)(dEz/dx - dEx/dz); E(k) at z_k, Hy(k) at z_k + h/2
    field_in = {"hy": (EPS0 * er) * c * _gauss(Z + 0.5 * h + c * dt / 2, x0, w)}
    _, _, f, _ = run_oracle(cfg, fields=("ex",), field_in=field_in)
    sl = (n_tr // 2, n_tr // 2, slice(None))
    exp = _gauss(Z[sl], x0 + c * nsteps * dt, w)
    err = np.linalg.norm(f["ex"][sl] - exp) / np.linalg.norm(exp)
    assert err < 3e-3, err


# ------------------------------------------------- EM one-step closed forms ----
def _spd(rng, lo, hi, off):
    return models.spd_tensor(rng, lo, hi, off)


def test_em2_inverse_permittivity_and_conductivity():
    rng = np.random.default_rng(3)
    nx, nz, dx, dz, dt = 16, 14, 0.05, 0.07, 1e-11
    eps = _spd(rng, 3.0, 12.0, 0.3)[np.ix_([0, 2], [0, 2])]
    sig = np.array([[2e-3, 5e-4], [5e-4, 7e-3]])
    shape = (nx, nz)

    def cfg_with(ic_x=None, ic_z=None):
        cfg = base_cfg("em2", shape, (dx, dz), dt, 1, (5, 5))
        cfg["arrays"].update(e11=full(shape, eps[0, 0]), e13=full(shape, eps[0, 1]),
                             e33=full(shape, eps[1, 1]), s11=full(shape, sig[0, 0]),
                             s13=full(shape, sig[0, 1]), s33=full(shape, sig[1, 1]))
        if ic_x is not None:
            cfg["arrays"]["ic_x"] = ic_x
            cfg["arrays"]["ic_z"] = ic_z
        return cfg

    inv = np.linalg.inv(EPS0 * eps)
    inner = (slice(1, nx - 1), slice(1, nz - 1))
    # linear Hy = al x + be z: curl H = (-dHy/dz, dHy/dx) = (-be, al)
    X, Z = coords(shape, (dx, dz))
    al, be = 3.0, -2.0
    _, _, f, _ = run_oracle(cfg_with(), fields=("ex", "ez"), field_in={"hy": al * X + be * Z})
    exp = dt * inv @ np.array([-be, al])
    np.testing.assert_allclose(f["ex"][inner], exp[0], rtol=1e-10)
    np.testing.assert_allclose(f["ez"][inner], exp[1], rtol=1e-10)
    # uniform E: E1 = (I - dt inv(eps) sigma) E0 away from the faces
    E0 = np.array([1.5, -0.8])
    _, _, f, _ = run_oracle(cfg_with(full(shape, E0[0]), full(shape, E0[1])), fields=("ex", "ez"))
    exp = (np.eye(2) - dt * inv @ sig) @ E0
    core = (slice(2, nx - 2), slice(2, nz - 2))
    np.testing.assert_allclose(f["ex"][core], exp[0], rtol=1e-12)
    np.testing.assert_allclose(f["ez"][core], exp[1], rtol=1e-12)


def test_em3_inverse_permittivity_conductivity_and_faraday():
    rng = np.random.default_rng(5)
    shape, sp, dt = (11, 10, 9), (0.05, 0.06, 0.07), 1e-11
    eps = _spd(rng, 3.0, 12.0, 0.3)
    sig = models.spd_tensor(rng, 1e-4, 1e-2, 0.1, log=True)
    names = ["11", "12", "13", "22", "23", "33"]
    ij = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]

    def cfg_with(ic=None):
        cfg = base_cfg("em3", shape, sp, dt, 1, (4, 4, 4))
        for n_, (i, j) in zip(names, ij):
            cfg["arrays"]["e" + n_] = full(shape, eps[i, j])
            cfg["arrays"]["s" + n_] = full(shape, sig[i, j])
        if ic is not None:
            for q, c_ in enumerate("xyz"):
                cfg["arrays"]["ic_" + c_] = ic[q]
        return cfg

    inv = np.linalg.inv(EPS0 * eps)
    X = coords(shape, sp)
    G = rng.uniform(-1, 1, (3, 3))       # H_i = sum_j G_ij x_j
    Hin = {n_: G[q, 0] * X[0] + G[q, 1] * X[1] + G[q, 2] * X[2] for q, n_ in enumerate(("hx", "hy", "hz"))}
    _, _, f, _ = run_oracle(cfg_with(), fields=("ex", "ey", "ez"), field_in=Hin)
    curl = np.array([G[2, 1] - G[1, 2], G[0, 2] - G[2, 0], G[1, 0] - G[0, 1]])
    exp = dt * inv @ curl
    inner = (slice(1, -1), slice(1, -1), slice(1, -1))
    for q, n_ in enumerate(("ex", "ey", "ez")):
        np.testing.assert_allclose(f[n_][inner], exp[q], rtol=1e-10, err_msg=n_)
    # uniform E with conductivity
    E0 = np.array([0.7, -1.2, 0.4])
    _, _, f, _ = run_oracle(cfg_with([full(shape, v) for v in E0]), fields=("ex", "ey", "ez"))
    exp = (np.eye(3) - dt * inv @ sig) @ E0
    core = (slice(2, -2), slice(2, -2), slice(2, -2))
    for q, n_ in enumerate(("ex", "ey", "ez")):
        np.testing.assert_allclose(f[n_][core], exp[q], rtol=1e-12, err_msg=n_)
    # Faraday: linear E gives H_1 = -dt/mu0 curl E on the H nests
    Ge = rng.uniform(-1, 1, (3, 3))
    ic = [Ge[q, 0] * X[0] + Ge[q, 1] * X[1] + Ge[q, 2] * X[2] for q in range(3)]
    _, _, f, _ = run_oracle(cfg_with(ic), fields=("hx", "hy", "hz"))
    curlE = np.array([Ge[2, 1] - Ge[1, 2], Ge[0, 2] - Ge[2, 0], Ge[1, 0] - Ge[0, 1]])
    expH = -dt / MU0 * curlE
    core = (slice(1, -1), slice(1, -1), slice(1, -1))
    for q, n_ in enumerate(("hx", "hy", "hz")):
        np.testing.assert_allclose(f[n_][core], expH[q], rtol=1e-9, err_msg=n_)


# ----------------------------------------------- linearity, zero, blow-up ----
def _small_el2():
    return models.elastic2(nx=50, nz=44, cpml=6, steps=120, src=(25, 12), rec_z=12, nrec=6)


def test_linearity_and_superposition():
    cfg = _small_el2()
    s = cfg["sources"]["z"]
    r1, _, _, _ = run_oracle(cfg)
    cfg3 = dict(cfg, sources={"x": np.zeros_like(s), "z": 3.0 * s})
    r3, _, _, _ = run_oracle(cfg3)
    np.testing.assert_allclose(r3, 3.0 * r1, rtol=1e-12, atol=1e-12 * np.abs(r1).max())
    cfg_b = dict(cfg, sources={"x": 0.5 * s, "z": np.zeros_like(s)})
    rb, _, _, _ = run_oracle(cfg_b)
    cfg_ab = dict(cfg, sources={"x": 0.5 * s, "z": s})
    rab, _, _, _ = run_oracle(cfg_ab)
    np.testing.assert_allclose(rab, r1 + rb, rtol=1e-10, atol=1e-12 * np.abs(rab).max())


def test_zero_in_zero_out():
    cfg = _small_el2()
    cfg["sources"] = {"x": np.zeros(cfg["time_steps"]), "z": np.zeros(cfg["time_steps"])}
    r, (vn, _), _, _ = run_oracle(cfg)
    assert np.all(r == 0) and np.all(vn == 0)


def test_blowup_forced_by_initial_condition():
    cfg = _small_el2()
    ic = np.zeros(cfg["shape"], order="F")
    ic[20, 20] = 1e26
    cfg["arrays"]["ic_x"] = ic
    _, (vn, _), _, us = run_oracle(cfg)
    assert us == 1 and vn[0] > 1e25


# ---------------------------------------------- transcription known answers ----
def test_reference_velocity_stencil_known_answer():
    """The 3-term velocity difference of cpmlfdtd.f08:405 on one step with a
    prescribed sigma_xx (all else zero): vx(i) = dt*(27 s_i - 27 s_{i-1} +
    s_{i-2})/(24 dx)/rho.  Restated from the cited line only."""
    nx, nz, dx, dz, dt = 16, 12, 1.0, 1.0, 1e-3
    cfg = base_cfg("elastic2", (nx, nz), (dx, dz), dt, 1, (6, 6))
    for k in ("c11", "c13", "c15", "c33", "c35", "c55"):
        cfg["arrays"][k] = full((nx, nz), 0.0)
    cfg["arrays"]["rho"] = full((nx, nz), 2.0)
    X, _ = coords((nx, nz), (dx, dz))
    s = 1.0 + 0.25 * X ** 2
    _, _, f, _ = run_oracle(cfg, fields=("vx",), field_in={"sxx": s})
    i = 7
    exp = dt * (27 * s[i, 5] - 27 * s[i - 1, 5] + s[i - 2, 5]) / (24 * dx) / 2.0
    assert f["vx"][i, 5] == pytest.approx(exp, rel=1e-14)


def test_point_source_offsets_known_answer():
    """Delta source at step 1: 2D stress sources land on (is, js) and
    (is+1, js+1), velocity sources on (is, js) with no dt
    (cpmlfdtd.f08:519-523).  Restated from the cited lines only."""
    nx, nz, dt = 20, 20, 1e-3
    cfg = base_cfg("elastic2", (nx, nz), (1.0, 1.0), dt, 1, (7, 9))
    for k in ("c11", "c13", "c15", "c33", "c35", "c55"):
        cfg["arrays"][k] = full((nx, nz), 1e9)
    rho = full((nx, nz), 1000.0)
    rho[7, 9] = 1500.0
    cfg["arrays"]["rho"] = rho
    cfg["sources"] = {"xx": np.array([2.0]), "xz": np.array([3.0]), "zz": np.array([5.0]),
                      "x": np.array([7.0]), "z": np.array([11.0])}
    _, _, f, _ = run_oracle(cfg, fields=("sxx", "szz", "sxz", "vx", "vz"))
    is_, js = 7 - 1, 9 - 1     # 0-based
    assert f["sxx"][is_, js] == 2.0 / 1000.0 and f["szz"][is_, js] == 5.0 / 1000.0
    assert f["sxz"][is_ + 1, js + 1] == 3.0 / 1500.0
    assert f["vx"][is_, js] == 7.0 / 1000.0 and f["vz"][is_, js] == 11.0 / 1000.0
    assert np.count_nonzero(f["sxz"]) == 1 and np.count_nonzero(f["vx"]) == 1
