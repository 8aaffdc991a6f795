"""Pins of the oracle's plane-wave source (plane_wave_source.f08 and the
injection loops of cpmlfdtd.f08) by closed forms: the Kelvin-form background
speed of an isotropic solid, the EM speed and impedance of simple
permittivity tensors, and the fields one step after injection into a zero
state for rotations whose direction and polarisation are known by hand (the
rotation matrices are filled column by column, cpmlfdtd's reshape)."""
import math

import numpy as np
import pytest

from oracle.oracle_ctypes import CLIGHT, EPS0, MU0, background_speed, em_background, run_oracle


def _iso(shape, lam, mu, rho):
    A = {n: np.zeros(shape) for n in ("c11", "c12", "c13", "c14", "c15", "c16", "c22", "c23", "c24", "c25",
                                      "c26", "c33", "c34", "c35", "c36", "c44", "c45", "c46", "c55", "c56", "c66")}
    for n in ("c11", "c22", "c33"):
        A[n][:] = lam + 2 * mu
    for n in ("c12", "c13", "c23"):
        A[n][:] = lam
    for n in ("c44", "c55", "c66"):
        A[n][:] = mu
    A["rho"] = np.full(shape, rho)
    return A


def test_background_speed_isotropic_closed_form():
    lam, mu, rho = 7.0e9, 3.0e9, 1000.0
    A = _iso((6, 5), lam, mu, rho)
    A["rho"][2, 3] = 800.0          # the smallest density sets the speed
    # 3D Kelvin eigenvalues of an isotropic tensor: 3 lam + 2 mu once, 2 mu five times
    assert background_speed(A, two_d=False) == pytest.approx(math.sqrt((3 * lam + 2 * mu) / 800.0), rel=1e-14)
    # 2D keeps the (xx, zz, xz) block: eigenvalues 2 lam + 2 mu, 2 mu, 2 mu
    assert background_speed(A, two_d=True) == pytest.approx(math.sqrt((2 * lam + 2 * mu) / 800.0), rel=1e-14)


def _gaus(t, f0, kind):
    a = math.pi * f0 * (t - 1.5 / f0)
    g = math.exp(-a * a)
    return {0: g, 1: -2 * a * g, 2: (1 - 2 * a * a) * g}[kind]


@pytest.mark.parametrize("angles,kind,face,e_vec,p_vec", [
    ((0.0, 0.0, 0.0), "s", "zmin", (1, 0, 0), (0, 0, 1)),      # R = I: p = z, e = x
    ((90.0, 0.0, 0.0), "s", "xmax", (0, 0, 1), (-1, 0, 0)),    # Ry(90): p = (-sy, 0, cy), e = (cy, 0, sy)
    ((90.0, 0.0, 0.0), "p", "xmax", (-1, 0, 0), (-1, 0, 0)),
    ((180.0, 0.0, 0.0), "p", "zmax", (0, 0, -1), (0, 0, -1)),
])
def test_el2_one_step_injection_closed_form(angles, kind, face, e_vec, p_vec):
    from seidart_hip import models
    cfg = models.elastic2(nx=30, nz=24, steps=1, src=(10, 10), nrec=2, nlayers=2)
    f0, amp, wav = 400.0, 2.5e-3, 2
    cfg["plane_wave"] = dict(kind=kind, angles=angles, f0=f0, amplitude=amp, wavelet=wav)
    fields = ("vx", "vz", "sxx", "szz", "sxz")
    _, _, F, us = run_oracle(cfg, steps=1, fields=fields)
    assert us == 0
    A = cfg["arrays"]
    nx, nz = cfg["shape"]
    cp = cfg["cpml"]
    v = background_speed(A, two_d=True)
    s = amp * _gaus(cfg["dt"], f0, wav)     # on the face p.(r - r0) = 0, so tau = t = dt
    sc = -s / v
    e, p = np.array(e_vec, float), np.array(p_vec, float)
    st = [sc * e[0] * p[0], sc * e[1] * p[1], sc * e[2] * p[2], sc * (e[1] * p[2] + e[2] * p[1]),
          sc * (e[0] * p[2] + e[2] * p[0]), sc * (e[0] * p[1] + e[1] * p[0])]
    if face in ("zmin", "zmax"):
        k = cp + 1 if face == "zmin" else nz - cp     # 1-based face row
        cells = [(i, k) for i in range(cp + 2, nx - cp)]
    else:
        i = cp + 1 if face == "xmin" else nx - cp
        cells = [(i, k) for k in range(cp + 2, nz - cp)]
    mask = np.zeros((nx, nz), bool)
    for (i, k) in cells:
        c = (i - 1, k - 1)
        mask[c] = True
        # cos(90 deg) is 6e-17 in floating point: absolute slack of 1e-12 |s|
        assert F["vx"][c] == pytest.approx(s * e[0], rel=1e-12, abs=1e-12 * abs(s))
        assert F["vz"][c] == pytest.approx(s * e[2], rel=1e-12, abs=1e-12 * abs(s))
        want = {"sxx": A["c11"][c] * st[0] + A["c13"][c] * st[2] + A["c15"][c] * st[4],
                "szz": A["c13"][c] * st[0] + A["c33"][c] * st[2] + A["c35"][c] * st[4],
                "sxz": A["c15"][c] * st[0] + A["c35"][c] * st[2] + A["c55"][c] * st[4]}
        for n, w in want.items():
            assert F[n][c] == pytest.approx(w, rel=1e-12, abs=1e-12 * abs(sc) * A["c33"][c])
    for n in fields:   # nothing else moved: the first update of a zero state is zero
        assert not np.any(F[n][~mask])


def test_em_background_closed_form():
    shape2 = (5, 4)
    A = {"e11": np.full(shape2, 4.0), "e13": np.zeros(shape2), "e33": np.full(shape2, 4.0)}
    v, eta = em_background(A, two_d=True)
    assert v == pytest.approx(CLIGHT / 2.0, rel=1e-15)
    assert eta == pytest.approx(math.sqrt(MU0 / (EPS0 * 4.0)), rel=1e-15)
    # [[3, 1], [1, 3]] has eigenvalues 2 and 4; the smallest over the model wins
    A["e11"][1, 2], A["e13"][1, 2], A["e33"][1, 2] = 3.0, 1.0, 3.0
    assert em_background(A, two_d=True)[0] == pytest.approx(CLIGHT / math.sqrt(2.0), rel=1e-14)
    shape3 = (3, 4, 2)
    B = {n: np.zeros(shape3) for n in ("e12", "e13", "e23")}
    for n in ("e11", "e22", "e33"):
        B[n] = np.full(shape3, 6.0)
    # [[4, 1, 0], [1, 4, 0], [0, 0, 9]]: eigenvalues 3, 5, 9
    B["e11"][2, 1, 0], B["e12"][2, 1, 0], B["e22"][2, 1, 0], B["e33"][2, 1, 0] = 4.0, 1.0, 4.0, 9.0
    v, eta = em_background(B, two_d=False)
    assert v == pytest.approx(CLIGHT / math.sqrt(3.0), rel=1e-14)
    assert eta == pytest.approx(math.sqrt(MU0 / (EPS0 * 3.0)), rel=1e-14)


def _em_cfg(physics, cpml=4):
    from seidart_hip import models
    if physics == "em2":
        return models.em2(nx=26, nz=22, cpml=cpml, steps=1, src=(10, 10), nrec=2, nlayers=2)
    return models.em3(nx=14, ny=12, nz=16, cpml=cpml, steps=1, src=(7, 6, 8), nrec=2, nlayers=2)


# E along e, H along p x e (boundary_em_field, plane_wave_source.f08:163-194)
@pytest.mark.parametrize("physics", ["em2", "em3"])
@pytest.mark.parametrize("angles,face,e_vec,h_vec", [
    ((0.0, 0.0, 0.0), "zmin", (1, 0, 0), (0, 1, 0)),      # p = z, e = x, z x x = y
    ((90.0, 0.0, 0.0), "xmax", (0, 0, 1), (0, 1, 0)),     # p = -x, e = z, (-x) x z = y
])
def test_em_one_step_injection_closed_form(physics, angles, face, e_vec, h_vec):
    cfg = _em_cfg(physics)
    f0, amp, wav = 200e6, 3.0, 1
    cfg["plane_wave"] = dict(kind="s", angles=angles, f0=f0, amplitude=amp, wavelet=wav)
    three = physics == "em3"
    fields = ("ex", "ey", "ez", "hx", "hy", "hz") if three else ("ex", "ez", "hy")
    _, _, F, us = run_oracle(cfg, steps=1, fields=fields)
    assert us == 0
    v, eta = em_background(cfg["arrays"], two_d=not three)
    s = amp * _gaus(cfg["dt"], f0, wav)     # on the face tau = t = dt
    n = cfg["shape"]
    cp = cfg["cpml"]
    # cells along a face: cpml+2 .. n-cpml-1 in 2D, cpml+3 .. n-cpml-2 in 3D
    lo, hi = (cp + 3, lambda m: m - cp - 2) if three else (cp + 2, lambda m: m - cp - 1)
    ax = 0 if face[0] == "x" else len(n) - 1
    fidx = cp + 1 if face.endswith("min") else n[ax] - cp
    others = [a for a in range(len(n)) if a != ax]
    mask = np.zeros(n, bool)
    sl = [None] * len(n)
    sl[ax] = fidx - 1
    for a in others:
        sl[a] = slice(lo - 1, hi(n[a]))
    mask[tuple(sl)] = True
    want = {"ex": s * e_vec[0], "ey": s * e_vec[1], "ez": s * e_vec[2],
            "hx": s / eta * h_vec[0], "hy": s / eta * h_vec[1], "hz": s / eta * h_vec[2]}
    for name in fields:
        got = F[name]
        w = want[name]
        scale = abs(s) if name[0] == "e" else abs(s) / eta
        assert np.allclose(got[mask], w, rtol=1e-12, atol=1e-12 * scale), name
        assert not np.any(got[~mask]), name    # the first update of a zero state is zero


@pytest.mark.parametrize("physics", ["em2", "em3"])
def test_em_face_on_grid_edge_is_zeroed(physics):
    """With no absorbing layer the face lies on the grid edge, which the
    Dirichlet step after the injection zeroes again (cpmlfdtd.f08:2692-2711,
    3921-3961)."""
    cfg = _em_cfg(physics, cpml=0)
    cfg["plane_wave"] = dict(kind="s", angles=(0.0, 0.0, 0.0), f0=200e6, amplitude=3.0, wavelet=1)
    fields = ("ex", "ez", "hy")
    _, _, F, us = run_oracle(cfg, steps=1, fields=fields)
    assert us == 0
    for name in fields:
        assert not np.any(F[name]), name


# R = Ry(x_z) (Rz(x_y) Rx(y_z)) with the matrices filled column by column
# (plane_wave_source.f08:67-98); p = R z, e = R x, worked out by hand.  The
# last case tells the composition order apart: Rx Ry would send p to -x.
@pytest.mark.parametrize("angles,face,e_vec,h_vec", [
    ((0.0, 0.0, 90.0), "ymin", (1, 0, 0), (0, 0, -1)),    # Rx(90): p = y, e = x, y x x = -z
    ((0.0, 90.0, 0.0), "zmin", (0, -1, 0), (1, 0, 0)),    # Rz(90): p = z, e = -y, z x -y = x
    ((90.0, 90.0, 0.0), "xmax", (0, -1, 0), (0, 0, 1)),   # Ry(90) Rz(90): p = -x, e = -y
    ((90.0, 0.0, 90.0), "ymin", (0, 0, 1), (1, 0, 0)),    # Ry(90) Rx(90): p = y, e = z
])
def test_em3_rotation_composition_closed_form(angles, face, e_vec, h_vec):
    cfg = _em_cfg("em3")
    f0, amp, wav = 200e6, 3.0, 0
    cfg["plane_wave"] = dict(kind="s", angles=angles, f0=f0, amplitude=amp, wavelet=wav)
    fields = ("ex", "ey", "ez", "hx", "hy", "hz")
    _, _, F, us = run_oracle(cfg, steps=1, fields=fields)
    assert us == 0
    _, eta = em_background(cfg["arrays"], two_d=False)
    s = amp * _gaus(cfg["dt"], f0, wav)
    n, cp = cfg["shape"], cfg["cpml"]
    ax = "xyz".index(face[0])
    mask = np.zeros(n, bool)
    sl = [slice(cp + 2, n[a] - cp - 2) for a in range(3)]     # 0-based cpml+3 .. n-cpml-2
    sl[ax] = cp if face.endswith("min") else n[ax] - cp - 1
    mask[tuple(sl)] = True
    want = {"ex": s * e_vec[0], "ey": s * e_vec[1], "ez": s * e_vec[2],
            "hx": s / eta * h_vec[0], "hy": s / eta * h_vec[1], "hz": s / eta * h_vec[2]}
    for name in fields:
        scale = abs(s) if name[0] == "e" else abs(s) / eta
        assert np.allclose(F[name][mask], want[name], rtol=1e-12, atol=1e-12 * scale), name
        assert not np.any(F[name][~mask]), name


# 3D stress injection: sigma_I += C_IJ strain_J over the Voigt columns
# (add_plane_wave_stress, cpmlfdtd.f08:16-50).  Along z, a P wave has only
# strain_zz = -s/v (column 3), an S wave polarised along x only the
# engineering strain_xz = -s/v (column 5).
@pytest.mark.parametrize("kind,col", [("p", 3), ("s", 5)])
def test_el3_one_step_stress_columns(kind, col):
    from seidart_hip import models
    cfg = models.elastic3(nx=14, ny=12, nz=16, cpml=3, steps=1, src=(7, 6, 8), nrec=2, nlayers=2)
    f0, amp, wav = 2000.0, 1e-3, 2
    cfg["plane_wave"] = dict(kind=kind, angles=(0.0, 0.0, 0.0), f0=f0, amplitude=amp, wavelet=wav)
    fields = ("vx", "vy", "vz", "sxx", "syy", "szz", "syz", "sxz", "sxy")
    _, _, F, us = run_oracle(cfg, steps=1, fields=fields)
    assert us == 0
    A = cfg["arrays"]
    v = background_speed(A, two_d=False)
    s = amp * _gaus(cfg["dt"], f0, wav)
    n, cp = cfg["shape"], cfg["cpml"]
    mask = np.zeros(n, bool)
    mask[cp + 1:n[0] - cp - 1, cp + 1:n[1] - cp - 1, cp] = True    # z face cpml+1, cells cpml+2 .. n-cpml-1
    assert np.allclose(F["vz" if kind == "p" else "vx"][mask], s, rtol=1e-12)
    row = {"sxx": 1, "syy": 2, "szz": 3, "syz": 4, "sxz": 5, "sxy": 6}
    for name, r in row.items():
        a, b = sorted((r, col))
        want = A[f"c{a}{b}"][mask] * (-s / v)
        assert np.allclose(F[name][mask], want, rtol=1e-12, atol=1e-12 * abs(s / v) * A["c33"].max()), name
    for name in fields:
        assert not np.any(F[name][~mask]), name


def test_wavelets_are_gaussian_derivatives():
    """gaus1 and gaus2 against central differences of the oracle's own gaus0
    (plane_wave_source.f08:141-161): with a = pi f0 (t - t0) and g = exp(-a^2),
    gaus1 = g' / (pi f0) and gaus2 = -g'' / (2 (pi f0)^2).  The value injected
    at a face cell one step after a zero state is amplitude * wavelet(dt), so
    varying dt samples the wavelet."""
    f0, amp = 200e6, 1.0
    cp = 4

    def sample(t, wav):
        cfg = _em_cfg("em2", cpml=cp)
        cfg["dt"] = t
        cfg["plane_wave"] = dict(kind="s", angles=(0.0, 0.0, 0.0), f0=f0, amplitude=amp, wavelet=wav)
        _, _, F, us = run_oracle(cfg, steps=1, fields=("ex",))
        assert us == 0
        return F["ex"][cp + 4, cp]     # a cell of the z face (row cpml+1)

    w = math.pi * f0
    h = 2e-3 / f0
    for t in (2e-9, 5.5e-9, 7.5e-9, 9e-9, 12e-9):
        g = [sample(t + k * h, 0) for k in (-1, 0, 1)]
        d1 = (g[2] - g[0]) / (2 * h)
        d2 = (g[2] - 2 * g[1] + g[0]) / (h * h)
        assert sample(t, 1) == pytest.approx(d1 / w, rel=1e-4, abs=1e-6)
        assert sample(t, 2) == pytest.approx(-d2 / (2 * w * w), rel=1e-4, abs=1e-5)
    assert sample(1.5 / f0, 0) == pytest.approx(amp, rel=1e-15)      # peak at t0 = 1.5 / f0
