"""Synthetic inputs of the solver configurations: layered anisotropic models,
CPML coefficient arrays, Ricker source series and receiver lines.

The reference receives all of these as files written by the SeidarT frontend
(src/fortran/cpmlfdtd.f08:159-174, 253-266); nothing under the reference
defines how the frontend builds them, so this module provides a seeded
generator for tests and benchmarks:

* stiffness: ice-Ih hexagonal tensor (C11 13.93, C12 7.08, C13 5.77, C33 15.01,
  C44 3.01 GPa, C66 = (C11 - C12)/2) with seeded jitter, Bond-rotated;
* permittivity / conductivity: symmetric positive-definite tensors;
* CPML: Komatitsch & Martin (2007) profiles d = d0 (xi)^N, kappa =
  1 + (kappa_max - 1) xi^N, alpha = alpha_max (1 - xi), b = exp(-(d/kappa +
  alpha) dt), a = d (b - 1) / (kappa (d + kappa alpha)), evaluated at integer
  and half-integer cell positions.  The reference uses one coefficient array
  for all three directions, so each cell takes the profile of the axis along
  which it lies deepest in the absorbing layer.

Arrays are padded (n + 2*cpml per axis, src/fortran/main.f08:69-71) and
indexed [x, z] or [x, y, z]; source and receiver indices are 1-based and
unpadded, as in the reference's JSON.
"""
from __future__ import annotations

import math

import numpy as np

EPS0 = 8.85418782e-12       # constants.f08:10
CLIGHT = 2.9979458e8        # constants.f08:8

VOIGT_NAMES = ["c11", "c12", "c13", "c14", "c15", "c16", "c22", "c23", "c24", "c25", "c26",
               "c33", "c34", "c35", "c36", "c44", "c45", "c46", "c55", "c56", "c66"]
VOIGT_IJ = [(0, 0), (0, 1), (0, 2), (0, 3), (0, 4), (0, 5), (1, 1), (1, 2), (1, 3), (1, 4),
            (1, 5), (2, 2), (2, 3), (2, 4), (2, 5), (3, 3), (3, 4), (3, 5), (4, 4), (4, 5), (5, 5)]


# ---------------------------------------------------------------- tensors ----
def ice_voigt(rng=None, jitter=0.05):
    """6x6 Voigt stiffness (Pa) of hexagonal ice Ih, symmetry axis along z.
    Voigt order xx, yy, zz, yz, xz, xy (the reference's c14 couples xx with
    yz, cpmlfdtd.f08:1799-1802)."""
    c11, c12, c13, c33, c44 = 13.93, 7.08, 5.77, 15.01, 3.01
    if rng is not None and jitter:
        c11, c12, c13, c33, c44 = (v * (1.0 + rng.uniform(-jitter, jitter))
                                   for v in (c11, c12, c13, c33, c44))
    c66 = 0.5 * (c11 - c12)
    C = np.zeros((6, 6))
    C[0, 0] = C[1, 1] = c11
    C[0, 1] = C[1, 0] = c12
    C[0, 2] = C[2, 0] = C[1, 2] = C[2, 1] = c13
    C[2, 2] = c33
    C[3, 3] = C[4, 4] = c44
    C[5, 5] = c66
    return C * 1e9


def bond_matrix(R):
    """6x6 Bond stress-transformation matrix of rotation R (C' = M C M^T)."""
    R = np.asarray(R, dtype=float)
    M = np.zeros((6, 6))
    pairs = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    for a, (i, j) in enumerate(pairs):
        for b, (k, l) in enumerate(pairs):
            if b < 3:
                M[a, b] = R[i, k] * R[j, l]
            else:
                M[a, b] = R[i, k] * R[j, l] + R[i, l] * R[j, k]
    return M


def rot_y(theta):
    c, s = math.cos(theta), math.sin(theta)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def rot_z(theta):
    c, s = math.cos(theta), math.sin(theta)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def euler_zyz(a, b, g):
    return rot_z(a) @ rot_y(b) @ rot_z(g)


def rotate_voigt(C, R):
    M = bond_matrix(R)
    return M @ C @ M.T


def kelvin(C):
    """Kelvin (Mandel) form of a Voigt stiffness: eigenvalues bound wave speeds."""
    w = np.array([1.0, 1.0, 1.0, math.sqrt(2), math.sqrt(2), math.sqrt(2)])
    return C * np.outer(w, w)


def max_speed(Cs, rho_min):
    lam = max(np.linalg.eigvalsh(kelvin(C)).max() for C in Cs)
    return math.sqrt(lam / rho_min)


def spd_tensor(rng, diag_lo, diag_hi, off_max, log=False):
    """Random symmetric positive-definite 3x3 tensor."""
    while True:
        if log:
            d = np.exp(rng.uniform(math.log(diag_lo), math.log(diag_hi), 3))
            off = rng.uniform(-off_max, off_max, 3) * d.min()
        else:
            d = rng.uniform(diag_lo, diag_hi, 3)
            off = rng.uniform(-off_max, off_max, 3)
        T = np.diag(d)
        T[0, 1] = T[1, 0] = off[0]
        T[0, 2] = T[2, 0] = off[1]
        T[1, 2] = T[2, 1] = off[2]
        if np.linalg.eigvalsh(T).min() > 0.05 * d.min():
            return T


# ---------------------------------------------------------------- sources ----
def ricker(f0, dt, n, amplitude=1.0):
    """(1 - 2a^2) exp(-a^2), a = pi f0 (t - 1.5/f0), t_n = n dt for n = 1..N
    (the 'gaus2' wavelet of plane_wave_source.f08:141-161)."""
    t = np.arange(1, n + 1) * dt
    a = math.pi * f0 * (t - 1.5 / f0)
    return amplitude * (1.0 - 2.0 * a * a) * np.exp(-a * a)


# ------------------------------------------------------------------- CPML ----
def _axis_xi(n, cpml, half):
    """Normalised depth into the absorbing layer for padded 1-based cells."""
    i = np.arange(1, n + 1, dtype=float) + (0.5 if half else 0.0)
    lo = (cpml + 1 - i) / max(cpml, 1)
    hi = (i - (n - cpml)) / max(cpml, 1)
    xi = np.clip(np.maximum(lo, hi), 0.0, 1.0)
    if cpml == 0:
        xi[:] = 0.0
    return xi


def cpml_arrays(shape, spacing, cpml, vmax, dt, f0, *, Rc=1e-3, N=2, kappa_max=2.0,
                alpha_max=None):
    """kappa/alpha/acoef/bcoef (and _half) arrays of the padded grid."""
    if alpha_max is None:
        alpha_max = math.pi * f0
    out = {}
    for half in (False, True):
        d_best = np.zeros(shape)
        xi_best = np.zeros(shape)
        for ax, (n, h) in enumerate(zip(shape, spacing)):
            xi = _axis_xi(n, cpml, half)
            L = max(cpml, 1) * h
            d0 = -(N + 1) * vmax * math.log(Rc) / (2.0 * L)
            d = d0 * xi ** N
            sh = [1] * len(shape)
            sh[ax] = n
            d_b = np.broadcast_to(d.reshape(sh), shape)
            xi_b = np.broadcast_to(xi.reshape(sh), shape)
            take = d_b > d_best
            d_best = np.where(take, d_b, d_best)
            xi_best = np.where(take, xi_b, xi_best)
        inside = xi_best > 0
        kappa = np.where(inside, 1.0 + (kappa_max - 1.0) * xi_best ** N, 1.0)
        alpha = np.where(inside, alpha_max * (1.0 - xi_best), 0.0)
        b = np.exp(-(d_best / kappa + alpha) * dt)
        with np.errstate(invalid="ignore", divide="ignore"):
            a = np.where(d_best > 0, d_best * (b - 1.0) / (kappa * (d_best + kappa * alpha)), 0.0)
        sfx = "_half_cpml" if half else "_cpml"
        out["kappa" + sfx] = np.asfortranarray(kappa)
        out["alpha" + sfx] = np.asfortranarray(alpha)
        out["acoef" + sfx] = np.asfortranarray(a)
        out["bcoef" + sfx] = np.asfortranarray(b)
    return out


def no_cpml(shape):
    one, zero = np.ones(shape, order="F"), np.zeros(shape, order="F")
    return {"kappa_cpml": one, "kappa_half_cpml": one.copy(), "alpha_cpml": zero,
            "alpha_half_cpml": zero.copy(), "acoef_cpml": zero.copy(),
            "acoef_half_cpml": zero.copy(), "bcoef_cpml": zero.copy(),
            "bcoef_half_cpml": zero.copy()}


# ----------------------------------------------------------------- layers ----
def layer_index(nz_pad, cpml, nlayers, rng, min_thick=4):
    """Layer id per padded z index; boundaries drawn inside the unpadded grid."""
    nz = nz_pad - 2 * cpml
    if nlayers == 1:
        return np.zeros(nz_pad, dtype=int)
    cuts = np.sort(rng.choice(np.arange(min_thick, nz - min_thick), nlayers - 1, replace=False))
    lid = np.searchsorted(cuts, np.arange(nz), side="right")
    return np.concatenate([np.full(cpml, lid[0]), lid, np.full(cpml, lid[-1])])


# ------------------------------------------------------------ configurations ----
def elastic2(nx=200, nz=200, dx=1.0, dz=1.0, cpml=10, steps=1000, seed=0, nlayers=2,
             rho=(917.0, 2000.0), f0=50.0, src=(100, 40), rec_z=40, nrec=20, cfl=0.5,
             split=None, gamma=0.0, density_method="arithmetic", with_cpml=True, tilt=(0.0, 30.0)):
    """2D elastic, two seeded ice-like layers (the 200x200, 1000-step case by
    default); each layer's symmetry axis is tilted by a uniform draw from
    `tilt` degrees."""
    rng = np.random.default_rng(seed)
    NX, NZ = nx + 2 * cpml, nz + 2 * cpml
    if split is None and nlayers == 2:
        k = int(rng.integers(60, 141)) if nz >= 200 else int(rng.integers(nz // 3, 2 * nz // 3 + 1))
        lid = np.concatenate([np.zeros(cpml + k, dtype=int), np.ones(NZ - cpml - k, dtype=int)])
    else:
        lid = layer_index(NZ, cpml, nlayers, rng)
    Cs, rhos = [], []
    for q in range(max(nlayers, 2)):
        C = rotate_voigt(ice_voigt(rng), rot_y(math.radians(rng.uniform(*tilt))))
        Cs.append(C)
        rhos.append(rho[q % len(rho)])
    arrays = {}
    for name, (i, j) in zip(["c11", "c13", "c15", "c33", "c35", "c55"],
                            [(0, 0), (0, 2), (0, 4), (2, 2), (2, 4), (4, 4)]):
        col = np.array([Cs[l][i, j] for l in lid])
        arrays[name] = np.asfortranarray(np.broadcast_to(col[None, :], (NX, NZ)))
    arrays["rho"] = np.asfortranarray(np.broadcast_to(np.array([rhos[l] for l in lid])[None, :], (NX, NZ)))
    for g in ("gamma_x", "gamma_z", "gamma_xz"):
        arrays[g] = np.full((NX, NZ), gamma, order="F")
    C2 = [np.array([[C[0, 0], C[0, 2], C[0, 4]], [C[0, 2], C[2, 2], C[2, 4]], [C[0, 4], C[2, 4], C[4, 4]]])
          for C in Cs]
    lam = max(np.linalg.eigvalsh(c * np.outer([1, 1, math.sqrt(2)], [1, 1, math.sqrt(2)])).max() for c in C2)
    vmax = math.sqrt(lam / min(rhos))
    dt = cfl * min(dx, dz) / (math.sqrt(2.0) * vmax)
    arrays.update(cpml_arrays((NX, NZ), (dx, dz), cpml, vmax, dt, f0) if with_cpml else no_cpml((NX, NZ)))
    arrays["ic_x"] = np.zeros((NX, NZ), order="F")
    arrays["ic_z"] = np.zeros((NX, NZ), order="F")
    recs = np.array([[10 + (nx - 20) * k // max(nrec - 1, 1), 1, rec_z] for k in range(nrec)], dtype=np.int32)
    if nx == 200 and nrec == 20:
        recs[:, 0] = 10 + 9 * np.arange(20)
    return dict(physics="elastic2", shape=(NX, NZ), spacing=(dx, dz), cpml=cpml, dt=dt,
                time_steps=steps, src=src, arrays=arrays,
                sources={"x": np.zeros(steps), "z": ricker(f0, dt, steps)},
                receivers=recs, density_method=density_method, vmax=vmax)


def elastic3(nx=256, ny=256, nz=256, d=0.5, cpml=12, steps=2000, seed=1, nlayers=8, f0=100.0,
             src=(128, 128, 20), cfl=0.45, nrec=256, density_method="arithmetic", with_cpml=True):
    """3D elastic, eight seeded layers of rotated ice tensors with attenuation
    (the 256^3, 2000-step case by default)."""
    rng = np.random.default_rng(seed)
    NX, NY, NZ = nx + 2 * cpml, ny + 2 * cpml, nz + 2 * cpml
    lid = layer_index(NZ, cpml, nlayers, rng, min_thick=min(4, max(1, nz // (2 * nlayers))))
    Cs, rhos, gams = [], [], []
    for _ in range(nlayers):
        a, b, g = rng.uniform(0, 2 * math.pi), math.acos(rng.uniform(-1, 1)), rng.uniform(0, 2 * math.pi)
        Cs.append(rotate_voigt(ice_voigt(rng), euler_zyz(a, b, g)))
        rhos.append(rng.uniform(900.0, 2600.0))
        gams.append(rng.uniform(0.001, 0.01, 6))
    shape = (NX, NY, NZ)
    arrays = {}

    def zcol(vals):
        col = np.array([vals[l] for l in lid])
        return np.asfortranarray(np.broadcast_to(col[None, None, :], shape))

    for name, (i, j) in zip(VOIGT_NAMES, VOIGT_IJ):
        arrays[name] = zcol([C[i, j] for C in Cs])
    arrays["rho"] = zcol(rhos)
    for q, gname in enumerate(["gamma_x", "gamma_y", "gamma_z", "gamma_xy", "gamma_xz", "gamma_yz"]):
        arrays[gname] = zcol([g[q] for g in gams])
    vmax = max_speed(Cs, min(rhos))
    dt = cfl * d / (math.sqrt(3.0) * vmax)
    arrays.update(cpml_arrays(shape, (d, d, d), cpml, vmax, dt, f0) if with_cpml else no_cpml(shape))
    for c in ("ic_x", "ic_y", "ic_z"):
        arrays[c] = np.zeros(shape, order="F")
    nrec = min(nrec, nx)
    recs = np.array([[1 + (nx - 1) * k // max(nrec - 1, 1), ny // 2, 1] for k in range(nrec)], dtype=np.int32)
    return dict(physics="elastic3", shape=shape, spacing=(d, d, d), cpml=cpml, dt=dt,
                time_steps=steps, src=src, arrays=arrays,
                sources={"x": np.zeros(steps), "y": np.zeros(steps), "z": ricker(f0, dt, steps)},
                receivers=recs, density_method=density_method, vmax=vmax)


def elastic3_iso(nx=256, ny=256, nz=256, d=0.5, cpml=12, steps=2000, seed=11, f0=100.0, src=(128, 128, 20),
                 cfl=0.45, nrec=256, density_method="arithmetic", with_cpml=True, attenuation=True):
    """3D elastic, smoothly varying isotropic model (a P speed gradient with
    depth and lateral undulations, Vp/Vs about 1.8, seeded per-cell jitter of
    0.5 %, so no two cells share a material and the per-cell arrays are used):
    c11 = c22 = c33 = lambda + 2 mu, c12 = c13 = c23 = lambda, c44 = c55 = c66
    = mu, the other twelve entries absent (0).  The repeated entries share one
    array."""
    rng = np.random.default_rng(seed)
    NX, NY, NZ = nx + 2 * cpml, ny + 2 * cpml, nz + 2 * cpml
    shape = (NX, NY, NZ)
    x = np.linspace(0.0, 1.0, NX)[:, None, None]
    y = np.linspace(0.0, 1.0, NY)[None, :, None]
    z = np.linspace(0.0, 1.0, NZ)[None, None, :]
    jit = lambda: 1.0 + 0.005 * rng.standard_normal(shape)
    vp = (3000.0 + 1200.0 * z + 150.0 * np.sin(2 * math.pi * 1.5 * x) * np.cos(2 * math.pi * y)) * jit()
    vs = vp / (1.8 + 0.1 * np.cos(2 * math.pi * z)) * jit()
    rho = (1800.0 + 500.0 * z + 50.0 * np.sin(2 * math.pi * (x + y))) * jit()
    mu = np.asfortranarray(rho * vs * vs)
    lam = np.asfortranarray(rho * (vp * vp - 2.0 * vs * vs))
    c11 = np.asfortranarray(lam + 2.0 * mu)
    arrays = {"c11": c11, "c22": c11, "c33": c11, "c12": lam, "c13": lam, "c23": lam,
              "c44": mu, "c55": mu, "c66": mu, "rho": np.asfortranarray(rho)}
    if attenuation:
        g = np.asfortranarray(0.002 + 0.006 * z * np.ones(shape))
        for gname in ("gamma_x", "gamma_y", "gamma_z", "gamma_xy", "gamma_xz", "gamma_yz"):
            arrays[gname] = g
    vmax = float(np.sqrt(c11 / rho).max())
    dt = cfl * d / (math.sqrt(3.0) * vmax)
    arrays.update(cpml_arrays(shape, (d, d, d), cpml, vmax, dt, f0) if with_cpml else no_cpml(shape))
    for c in ("ic_x", "ic_y", "ic_z"):
        arrays[c] = np.zeros(shape, order="F")
    nrec = min(nrec, nx)
    recs = np.array([[1 + (nx - 1) * k // max(nrec - 1, 1), ny // 2, 1] for k in range(nrec)], dtype=np.int32)
    return dict(physics="elastic3", shape=shape, spacing=(d, d, d), cpml=cpml, dt=dt,
                time_steps=steps, src=src, arrays=arrays,
                sources={"x": np.zeros(steps), "y": np.zeros(steps), "z": ricker(f0, dt, steps)},
                receivers=recs, density_method=density_method, vmax=vmax)


def elastic25(nx=120, nz=100, dx=1.0, dz=1.0, cpml=10, steps=400, seed=5, nlayers=3, f0=60.0,
              src=(60, 30), cfl=0.4, nrec=24, ky=0.05, density_method="harmonic", anomaly=True,
              with_cpml=True):
    """2.5D elastic (x-z grid, y handled as i*ky): seeded layers of Euler-rotated
    ice tensors with attenuation and, optionally, a rectangular anomaly of a
    different material so density varies along x as well as z."""
    rng = np.random.default_rng(seed)
    NX, NZ = nx + 2 * cpml, nz + 2 * cpml
    shape = (NX, NZ)
    lid = layer_index(NZ, cpml, nlayers, rng, min_thick=min(4, max(1, nz // (2 * nlayers))))
    nmat = nlayers + (1 if anomaly else 0)
    Cs, rhos, gams = [], [], []
    for _ in range(nmat):
        a, b, g = rng.uniform(0, 2 * math.pi), math.acos(rng.uniform(-1, 1)), rng.uniform(0, 2 * math.pi)
        Cs.append(rotate_voigt(ice_voigt(rng), euler_zyz(a, b, g)))
        rhos.append(rng.uniform(900.0, 2600.0))
        gams.append(rng.uniform(0.001, 0.01, 6))
    mid = np.broadcast_to(lid[None, :], shape).copy()
    if anomaly:
        i0, i1 = cpml + nx // 3, cpml + (2 * nx) // 3
        k0, k1 = cpml + nz // 2, cpml + (3 * nz) // 4
        mid[i0:i1, k0:k1] = nlayers
    arrays = {}
    for name, (i, j) in zip(VOIGT_NAMES, VOIGT_IJ):
        arrays[name] = np.asfortranarray(np.array([C[i, j] for C in Cs])[mid])
    arrays["rho"] = np.asfortranarray(np.array(rhos)[mid])
    for q, gname in enumerate(["gamma_x", "gamma_y", "gamma_z", "gamma_xy", "gamma_xz", "gamma_yz"]):
        arrays[gname] = np.asfortranarray(np.array([g[q] for g in gams])[mid])
    vmax = max_speed(Cs, min(rhos))
    dt = cfl * min(dx, dz) / (math.sqrt(2.0) * vmax)
    arrays.update(cpml_arrays(shape, (dx, dz), cpml, vmax, dt, f0) if with_cpml else no_cpml(shape))
    for c in ("ic_x", "ic_y", "ic_z"):
        arrays[c] = np.zeros(shape, dtype=np.complex128, order="F")
    nrec = min(nrec, nx)
    recs = np.array([[1 + (nx - 1) * k // max(nrec - 1, 1), 1, src[1]] for k in range(nrec)], dtype=np.int32)
    return dict(physics="elastic25", shape=shape, spacing=(dx, dz), cpml=cpml, dt=dt, time_steps=steps,
                src=src, arrays=arrays, ky=float(ky),
                sources={"x": np.zeros(steps), "y": np.zeros(steps), "z": ricker(f0, dt, steps)},
                receivers=recs, density_method=density_method, vmax=vmax)


def elastic25_iso(nx=120, nz=100, dx=1.0, dz=1.0, cpml=10, steps=400, seed=12, f0=60.0, src=(60, 30), cfl=0.4,
                  nrec=24, ky=0.05, density_method="harmonic", with_cpml=True):
    """2.5D elastic, smoothly varying isotropic model with 0.5 % per-cell
    jitter (as elastic3_iso, on the x-z grid); the repeated stiffness entries
    share one array, the other twelve are absent."""
    rng = np.random.default_rng(seed)
    NX, NZ = nx + 2 * cpml, nz + 2 * cpml
    shape = (NX, NZ)
    x = np.linspace(0.0, 1.0, NX)[:, None]
    z = np.linspace(0.0, 1.0, NZ)[None, :]
    jit = lambda: 1.0 + 0.005 * rng.standard_normal(shape)
    vp = (3000.0 + 1200.0 * z + 150.0 * np.sin(2 * math.pi * 1.5 * x)) * jit()
    vs = vp / (1.8 + 0.1 * np.cos(2 * math.pi * z)) * jit()
    rho = (1800.0 + 500.0 * z + 50.0 * np.sin(2 * math.pi * x)) * jit()
    mu = np.asfortranarray(rho * vs * vs)
    lam = np.asfortranarray(rho * (vp * vp - 2.0 * vs * vs))
    c11 = np.asfortranarray(lam + 2.0 * mu)
    arrays = {"c11": c11, "c22": c11, "c33": c11, "c12": lam, "c13": lam, "c23": lam,
              "c44": mu, "c55": mu, "c66": mu, "rho": np.asfortranarray(rho)}
    g = np.asfortranarray(0.002 + 0.006 * z * np.ones(shape))
    for gname in ("gamma_x", "gamma_y", "gamma_z", "gamma_xy", "gamma_xz", "gamma_yz"):
        arrays[gname] = g
    vmax = float(np.sqrt(c11 / rho).max())
    dt = cfl * min(dx, dz) / (math.sqrt(2.0) * vmax)
    arrays.update(cpml_arrays(shape, (dx, dz), cpml, vmax, dt, f0) if with_cpml else no_cpml(shape))
    for c in ("ic_x", "ic_y", "ic_z"):
        arrays[c] = np.zeros(shape, dtype=np.complex128, order="F")
    nrec = min(nrec, nx)
    recs = np.array([[1 + (nx - 1) * k // max(nrec - 1, 1), 1, src[1]] for k in range(nrec)], dtype=np.int32)
    return dict(physics="elastic25", shape=shape, spacing=(dx, dz), cpml=cpml, dt=dt, time_steps=steps,
                src=src, arrays=arrays, ky=float(ky),
                sources={"x": np.zeros(steps), "y": np.zeros(steps), "z": ricker(f0, dt, steps)},
                receivers=recs, density_method=density_method, vmax=vmax)


def em3(nx=384, ny=384, nz=192, d=0.05, cpml=10, steps=2000, seed=2, nlayers=5, f0=100e6,
        src=(192, 192, 20), cfl=0.5, nrec=128, with_cpml=True):
    """3D EM, five seeded SPD permittivity/conductivity layers (the
    384x384x192, 2000-step case by default)."""
    rng = np.random.default_rng(seed)
    NX, NY, NZ = nx + 2 * cpml, ny + 2 * cpml, nz + 2 * cpml
    shape = (NX, NY, NZ)
    lid = layer_index(NZ, cpml, nlayers, rng, min_thick=min(4, max(1, nz // (2 * nlayers))))
    eps = [spd_tensor(rng, 3.0, 12.0, 0.3) for _ in range(nlayers)]
    sig = [spd_tensor(rng, 1e-5, 1e-2, 0.1, log=True) for _ in range(nlayers)]
    arrays = {}

    def zcol(vals):
        col = np.array([vals[l] for l in lid])
        return np.asfortranarray(np.broadcast_to(col[None, None, :], shape))

    for name, (i, j) in zip(["11", "12", "13", "22", "23", "33"],
                            [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]):
        arrays["e" + name] = zcol([e[i, j] for e in eps])
        arrays["s" + name] = zcol([s[i, j] for s in sig])
    lam_min = min(np.linalg.eigvalsh(e).min() for e in eps)
    vmax = CLIGHT / math.sqrt(lam_min)
    dt = cfl * d / (math.sqrt(3.0) * vmax)
    arrays.update(cpml_arrays(shape, (d, d, d), cpml, vmax, dt, f0) if with_cpml else no_cpml(shape))
    for c in ("ic_x", "ic_y", "ic_z"):
        arrays[c] = np.zeros(shape, order="F")
    nrec = min(nrec, max(1, (nx - 1) // 3 + 1))
    recs = np.array([[1 + 3 * k, ny // 2, 1] for k in range(nrec)], dtype=np.int32)
    return dict(physics="em3", shape=shape, spacing=(d, d, d), cpml=cpml, dt=dt, time_steps=steps,
                src=src, arrays=arrays,
                sources={"x": ricker(f0, dt, steps), "y": np.zeros(steps), "z": np.zeros(steps)},
                receivers=recs, vmax=vmax)


def em25(nx=120, nz=90, d=0.05, cpml=10, steps=400, seed=6, nlayers=4, f0=100e6, src=(60, 25), cfl=0.4,
         nrec=24, ky=5.0, with_cpml=True):
    """2.5D EM (x-z grid, y handled as i*ky): seeded layers of SPD
    permittivity and conductivity tensors, Ricker Ex/Ez dipole."""
    rng = np.random.default_rng(seed)
    NX, NZ = nx + 2 * cpml, nz + 2 * cpml
    shape = (NX, NZ)
    lid = layer_index(NZ, cpml, nlayers, rng, min_thick=min(4, max(1, nz // (2 * nlayers))))
    eps = [spd_tensor(rng, 3.0, 12.0, 0.3) for _ in range(nlayers)]
    sig = [spd_tensor(rng, 1e-5, 1e-2, 0.1, log=True) for _ in range(nlayers)]
    arrays = {}
    for name, (i, j) in zip(["11", "12", "13", "22", "23", "33"],
                            [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]):
        arrays["e" + name] = np.asfortranarray(np.broadcast_to(np.array([eps[l][i, j] for l in lid])[None, :], shape))
        arrays["s" + name] = np.asfortranarray(np.broadcast_to(np.array([sig[l][i, j] for l in lid])[None, :], shape))
    lam_min = min(np.linalg.eigvalsh(e).min() for e in eps)
    vmax = CLIGHT / math.sqrt(lam_min)
    dt = cfl * d / (math.sqrt(2.0) * vmax)
    arrays.update(cpml_arrays(shape, (d, d), cpml, vmax, dt, f0) if with_cpml else no_cpml(shape))
    for c in ("ic_x", "ic_y", "ic_z"):
        arrays[c] = np.zeros(shape, dtype=np.complex128, order="F")
    nrec = min(nrec, nx)
    recs = np.array([[1 + (nx - 1) * k // max(nrec - 1, 1), 1, src[1]] for k in range(nrec)], dtype=np.int32)
    w = ricker(f0, dt, steps)
    return dict(physics="em25", shape=shape, spacing=(d, d), cpml=cpml, dt=dt, time_steps=steps, src=src,
                arrays=arrays, ky=float(ky), sources={"x": w, "y": 0.3 * w, "z": -0.5 * w},
                receivers=recs, vmax=vmax)


def em2(nx=600, nz=300, d=0.05, cpml=10, steps=3000, seed=4, nlayers=6, f0=100e6,
        src=(300, 20), cfl=0.5, nrec=64, with_cpml=True):
    """2D EM, seeded layered anisotropic permittivity and conductivity."""
    rng = np.random.default_rng(seed)
    NX, NZ = nx + 2 * cpml, nz + 2 * cpml
    shape = (NX, NZ)
    lid = layer_index(NZ, cpml, nlayers, rng, min_thick=min(4, max(1, nz // (2 * nlayers))))
    eps = [spd_tensor(rng, 3.0, 12.0, 0.3) for _ in range(nlayers)]
    sig = [spd_tensor(rng, 1e-5, 1e-2, 0.1, log=True) for _ in range(nlayers)]
    arrays = {}
    for name, (i, j) in zip(["11", "13", "33"], [(0, 0), (0, 2), (2, 2)]):
        arrays["e" + name] = np.asfortranarray(np.broadcast_to(np.array([eps[l][i, j] for l in lid])[None, :], shape))
        arrays["s" + name] = np.asfortranarray(np.broadcast_to(np.array([sig[l][i, j] for l in lid])[None, :], shape))
    lam_min = min(np.linalg.eigvalsh(e[np.ix_([0, 2], [0, 2])]).min() for e in eps)
    vmax = CLIGHT / math.sqrt(lam_min)
    dt = cfl * d / (math.sqrt(2.0) * vmax)
    arrays.update(cpml_arrays(shape, (d, d), cpml, vmax, dt, f0) if with_cpml else no_cpml(shape))
    arrays["ic_x"] = np.zeros(shape, order="F")
    arrays["ic_z"] = np.zeros(shape, order="F")
    nrec = min(nrec, nx)
    recs = np.array([[1 + (nx - 1) * k // max(nrec - 1, 1), 1, src[1]] for k in range(nrec)], dtype=np.int32)
    return dict(physics="em2", shape=shape, spacing=(d, d), cpml=cpml, dt=dt, time_steps=steps,
                src=src, arrays=arrays, sources={"x": ricker(f0, dt, steps), "z": np.zeros(steps)},
                receivers=recs, vmax=vmax)


def batch2d(physics="elastic2", nshots=64, seed=3, nx=(300, 800), nz=(150, 400), steps=(1500, 3000),
            nlayers=4, stencil="reference", tilt=(0.0, 30.0)):
    """Uneven multi-shot 2D set: per shot a seeded grid size (nx, nz uniform in
    the given ranges), its own `nlayers`-layer anisotropic model (as
    `elastic2` / `em2`), its own source position and step count.

    Elastic only: `tilt` is the range of symmetry-axis tilts in degrees and
    `stencil` the velocity differences.  The reference's stress update adds
    c15 * (dvz_dx + dvx_dz) to sigma_xx with both shear derivatives taken at
    the sigma_xx point without averaging (cpmlfdtd.f08:325-345); for tilted
    axes (c15, c35 != 0) that coupling grows exponentially over a few
    thousand steps, so long runs need tilt (0, 0).  The reference's 3-term
    velocity differences (cpmlfdtd.f08:405-446) add a slower zero-order
    growth; "corrected" selects the 4-term differences."""
    rng = np.random.default_rng(seed)
    shots = []
    for q in range(nshots):
        sx = int(rng.integers(nx[0], nx[1] + 1))
        sz = int(rng.integers(nz[0], nz[1] + 1))
        n = int(rng.integers(steps[0], steps[1] + 1))
        src = (int(rng.integers(sx // 4, 3 * sx // 4)), int(rng.integers(10, sz // 3)))
        sd = int(rng.integers(0, 2**31 - 1))
        if physics == "elastic2":
            shots.append(elastic2(nx=sx, nz=sz, steps=n, seed=sd, nlayers=nlayers, src=src,
                                  rec_z=src[1], nrec=20, tilt=tilt))
            shots[-1]["stencil"] = stencil
        else:
            shots.append(em2(nx=sx, nz=sz, steps=n, seed=sd, nlayers=nlayers, src=src, nrec=20))
    return shots


def _gaussians(shape, cpml, n, rng, amp=0.15, sigma=(8.0, 25.0)):
    """Sum of n seeded Gaussian bumps of relative amplitude +-amp inside the
    unpadded grid (one multiplicative factor field, 1 + sum)."""
    NX, NZ = shape
    ii, jj = np.meshgrid(np.arange(NX), np.arange(NZ), indexing="ij")
    f = np.zeros(shape)
    for _ in range(n):
        ci = rng.uniform(cpml + 20, NX - cpml - 20)
        cj = rng.uniform(cpml + 20, NZ - cpml - 20)
        s = rng.uniform(*sigma)
        f += rng.choice([-1.0, 1.0]) * amp * np.exp(-((ii - ci) ** 2 + (jj - cj) ** 2) / (2 * s * s))
    return 1.0 + f


def fwi2d(physics="em2", nx=600, nz=300, steps=3000, nshots=16, nrec=100, seed=8, nlayers=6, anomalies=10,
          smooth_passes=30):
    """Full-waveform-inversion problem: a seeded nlayers-layer anisotropic
    model plus `anomalies` Gaussian anomalies of +-15% in permittivity (EM)
    or stiffness (elastic) as the true model, the smoothed true model as the
    starting model, nshots sources along a line near the top and nrec
    receivers on that line.  Returns a dict: cfg (the true model's
    configuration dict), true / start (parameter arrays by name), shots (one
    dict per shot for seidart_hip.fwi), wavelet (steps,)."""
    rng = np.random.default_rng(seed)
    if physics == "em2":
        cfg = em2(nx=nx, nz=nz, steps=steps, seed=seed, nlayers=nlayers, src=(nx // 2, 10), nrec=nrec)
        names, scaled = ("e11", "e13", "e33", "s11", "s13", "s33"), ("e11", "e13", "e33")
        f0 = 100e6
    else:
        cfg = elastic2(nx=nx, nz=nz, steps=steps, seed=seed, nlayers=nlayers, src=(nx // 2, 10), rec_z=10,
                       nrec=nrec, tilt=(0.0, 0.0))
        names = ("c11", "c13", "c15", "c33", "c35", "c55", "rho", "gamma_x", "gamma_z", "gamma_xz")
        scaled = ("c11", "c13", "c15", "c33", "c35", "c55")
        f0 = 50.0
    A = cfg["arrays"]
    bump = _gaussians(cfg["shape"], cfg["cpml"], anomalies, rng)
    true = {n: np.asfortranarray(np.asarray(A[n], dtype=np.float64) * (bump if n in scaled else 1.0))
            for n in names}
    import torch  # the smoothing uses torch's pooling; models is imported without torch elsewhere

    def smooth(a):
        x = torch.as_tensor(a, dtype=torch.float64)[None, None]
        for _ in range(smooth_passes):
            x = torch.nn.functional.avg_pool2d(torch.nn.functional.pad(x, (1, 1, 1, 1), mode="replicate"), 3, 1)
        return np.asfortranarray(x[0, 0].numpy())

    start = {n: smooth(v) for n, v in true.items()}
    cp = cfg["cpml"]
    xs = np.linspace(40, nx - 40, nshots).round().astype(int)
    cpml_arrays = {k: v for k, v in A.items() if k.endswith("_cpml")}
    shots = [dict(shape=cfg["shape"], spacing=cfg["spacing"], cpml=cp, dt=cfg["dt"], src=(int(x), 10),
                  receivers=cfg["receivers"], cpml_arrays=cpml_arrays,
                  density_method=cfg.get("density_method", "arithmetic"))
             for x in xs]
    return dict(physics=physics, cfg=cfg, true=true, start=start, shots=shots,
                wavelet=ricker(f0, cfg["dt"], steps))


# ------------------------------------------------------ reference file set ----
_IC_PREFIX = {"elastic2": "initialconditionV", "elastic25": "initialconditionV", "elastic3": "initialconditionV",
              "em25": "initialconditionE",
              "em2": "initialconditionE", "em3": "initialconditionE"}


def write_inputs(cfg, workdir, source_type="ac"):
    """Write a configuration as the reference's input files (one Fortran
    record each, padded, column-major) plus input.json, as the SeidarT
    frontend does before running seidartfdtd."""
    import json
    import os

    from . import fortran_io as fio

    phys = cfg["physics"]
    is3d = phys.endswith("3")
    is25 = phys in ("elastic25", "em25")
    os.makedirs(workdir, exist_ok=True)
    shape = cfg["shape"]
    defaults = no_cpml(shape)
    ic_dtype = np.complex128 if is25 else np.float64   # material_rw2c for 2.5D
    for name, a in {**defaults, **cfg["arrays"]}.items():
        if name.startswith("ic_"):
            fname = _IC_PREFIX[phys] + name[-1]
            dt_ = ic_dtype
        else:
            fname = name
            dt_ = np.float64
        fio.write_record(os.path.join(workdir, fname + ".dat"), np.asarray(a, dtype=dt_).reshape(shape))
    comps = "xyz" if (is3d or is25) else "xz"
    for c in comps:
        if "ic_" + c not in cfg["arrays"]:
            fio.write_record(os.path.join(workdir, _IC_PREFIX[phys] + c + ".dat"), np.zeros(shape, dtype=ic_dtype))
    n = cfg["time_steps"]
    stem = "seismicsource" if phys.startswith("elastic") else "electromagneticsource"
    names = (list(comps) if (source_type == "ac" or not phys.startswith("elastic"))
             else (["xx", "xy", "xz", "yy", "yz", "zz"] if (is3d or is25) else ["xx", "xz", "zz"]))
    for c in names:
        s = np.asarray(cfg["sources"].get(c, np.zeros(n)), dtype=np.float64)
        fio.write_record(os.path.join(workdir, stem + c + ".dat"), s)
    cp = cfg["cpml"]
    sp = cfg["spacing"]
    if is3d:
        nx, ny, nz = (s_ - 2 * cp for s_ in shape)
        dx, dy, dz = sp
        xind, yind, zind = cfg["src"]
    else:
        nx, nz = (s_ - 2 * cp for s_ in shape)
        ny, dy = 1, 1.0
        dx, dz = sp
        xind, zind = cfg["src"]
        yind = 1
    src = {"x": 0.0, "y": 0.0, "z": 0.0, "xind": int(xind), "yind": int(yind), "zind": int(zind),
           "source_frequency": 0.0, "x-z_rotation": 0.0, "x-y_rotation": 0.0, "y-z_rotation": 0.0,
           "amplitude": 1.0, "source_type": source_type, "source_wavelet": "gaus2",
           "dt": float(cfg["dt"]), "time_steps": int(n)}
    pw = cfg.get("plane_wave")
    if pw:   # the JSON fields of a plane-wave source (seidartio.f08:82-98)
        src.update({"x-z_rotation": float(pw["angles"][0]), "x-y_rotation": float(pw["angles"][1]),
                    "y-z_rotation": float(pw["angles"][2]), "source_frequency": float(pw["f0"]),
                    "amplitude": float(pw["amplitude"]),
                    "source_wavelet": ("gaus0", "gaus1", "gaus2")[int(pw.get("wavelet", 0))]})
        if source_type == "ac":
            src["source_type"] = "pwp" if pw["kind"] == "p" else "pw"
    doc = {"Domain": {"dim": 3.0 if is3d else (2.5 if is25 else 2.0), "nx": int(nx), "ny": int(ny), "nz": int(nz),
                      "dx": float(dx), "dy": float(dy), "dz": float(dz), "cpml": int(cp), "nmats": 1,
                      "image_file": "model.png"},
           "Seismic": {"Source": dict(src)}, "Electromagnetic": {"Source": dict(src)}}
    with open(os.path.join(workdir, "input.json"), "w") as f:
        json.dump(doc, f, indent=2)
    return os.path.join(workdir, "input.json")
