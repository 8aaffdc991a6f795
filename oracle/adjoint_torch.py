"""Automatic-differentiation oracle for the adjoint-state gradients.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the GPU adjoint
(sdt_adjoint); nothing in the package imports it.

The 2D time stepping is restated here in float64 torch on the CPU, with the
per-cell update coefficients and the source series as differentiable inputs;
torch.autograd then gives the exact gradient of the discrete forward map,
which the hand-written GPU adjoint must reproduce.  The restatement follows
the reference loops:

* 2D EM (electromag2): H update cpmlfdtd.f08:2576-2593, E update 2605-2620,
  Dirichlet faces 2692-2721, point source 2678-2681 (here the series is
  injected as given: the raw-source mode of the library), coefficients
  2519-2540 (`em2_coefficients_np`).
* 2D elastic (seismic2): sigma_xx / sigma_zz cpmlfdtd.f08:324-353, sigma_xz
  362-386, vx 399-416, vz 428-446, point sources 519-523 (raw series),
  Dirichlet edges 534-549; the attenuation divisor 1 + gamma dt enters as
  its reciprocal (a coefficient), density averages as averaging.f08:5-25.

Its forward traces are pinned against the C oracle (oracle.c), which is
pinned against the reference's own cases.
"""
from __future__ import annotations

import math

import numpy as np
import torch
import torch.nn.functional as F

EPS0 = 8.85418782e-12                        # constants.f08:10
MU0 = 4.0 * math.pi * 1.0e-7                 # constants.f08:9


def em2_coefficients_np(e11, e13, e33, s11, s13, s33):
    """The seven E-update coefficients of cpmlfdtd.f08:2519-2540 (relative
    permittivity scaled by eps0; determinant clamped at 1e-18 eps0^2)."""
    e11, e13, e33 = e11 * EPS0, e13 * EPS0, e33 * EPS0
    det = e11 * e33 - e13 * e13
    det = np.where(det <= 1.0e-18 * EPS0 * EPS0, 1.0e-18 * EPS0 * EPS0, det)
    return np.stack([e33 / det, -e13 / det, e11 / det, (e33 * s11 - e13 * s13) / det,
                     (e33 * s13 - e13 * s33) / det, (e11 * s13 - e13 * s11) / det,
                     (e11 * s33 - e13 * s13) / det])


def em2_traces(coef, src_x, src_z, cpml, dt, dx, dz, src_ij, rec_ij):
    """Receiver traces (steps, nrec, 2) = (Ex, Ez) after every step.

    coef: (7, nx, nz) tensor (aEx, bEx, bEz, dEx, eEx, dEz, eEz);
    src_x, src_z: (steps,) tensors added to Ex, Ez at src_ij (0-based) after
    the E update (dropped on a face, as the reference's Dirichlet step does);
    cpml: dict of the kappa/acoef/bcoef (and _half) arrays; rec_ij: (nrec, 2)
    0-based indices."""
    dev = dict(dtype=torch.float64)
    aEx, bEx, bEz, dEx, eEx, dEz, eEz = coef
    nx, nz = aEx.shape
    k, a, b = (torch.as_tensor(np.asarray(cpml[n]), **dev) for n in ("kappa_cpml", "acoef_cpml", "bcoef_cpml"))
    kh, ah, bh = (torch.as_tensor(np.asarray(cpml[n]), **dev)
                  for n in ("kappa_half_cpml", "acoef_half_cpml", "bcoef_half_cpml"))
    nsteps = src_x.shape[0]
    Ex = torch.zeros((nx, nz), **dev)
    Ez = torch.zeros((nx, nz), **dev)
    Hy = torch.zeros((nx, nz), **dev)
    R = (slice(0, nx - 1), slice(0, nz - 1))      # cells updated by the H step (i < nx, j < nz)
    I = (slice(1, nx - 1), slice(1, nz - 1))      # cells updated by the E step
    m_dex_dz = torch.zeros((nx - 1, nz - 1), **dev)
    m_dez_dx = torch.zeros((nx - 1, nz - 1), **dev)
    m_dhy_dx = torch.zeros((nx - 2, nz - 2), **dev)
    m_dhy_dz = torch.zeros((nx - 2, nz - 2), **dev)
    hold = torch.ones((nx - 1, nz - 1), **dev)
    hold[0, :] = 0.0     # Hy on the i = 1 and j = 1 faces is zeroed after every step
    hold[:, 0] = 0.0
    dbhy = dt / MU0
    si, sj = src_ij
    face = si in (0, nx - 1) or sj in (0, nz - 1)
    onehot = torch.zeros((nx, nz), **dev)
    onehot[si, sj] = 0.0 if face else 1.0
    ri = torch.as_tensor(np.asarray(rec_ij)[:, 0], dtype=torch.long)
    rj = torch.as_tensor(np.asarray(rec_ij)[:, 1], dtype=torch.long)
    out = []
    for t in range(nsteps):
        gz = (Ex[:-1, 1:] - Ex[:-1, :-1]) / dz
        gx = (Ez[1:, :-1] - Ez[:-1, :-1]) / dx
        m_dex_dz = b[R] * m_dex_dz + a[R] * gz
        m_dez_dx = bh[R] * m_dez_dx + ah[R] * gx
        gx = gx / kh[R] + m_dez_dx
        gz = gz / k[R] + m_dex_dz
        Hy = F.pad(hold * Hy[R] + dbhy * (gx - gz), (0, 1, 0, 1))
        fz = (Hy[1:-1, 1:-1] - Hy[1:-1, :-2]) / dz
        fx = (Hy[1:-1, 1:-1] - Hy[:-2, 1:-1]) / dx
        m_dhy_dx = bh[I] * m_dhy_dx + ah[I] * fx
        m_dhy_dz = b[I] * m_dhy_dz + a[I] * fz
        fz = fz / k[I] + m_dhy_dz
        fx = fx / kh[I] + m_dhy_dx
        exo, ezo = Ex[I], Ez[I]
        ez = ezo + (-bEx[I] * fz + bEz[I] * fx - dEz[I] * exo - eEz[I] * ezo) * dt
        ex = exo + (-aEx[I] * fz + bEx[I] * fx - dEx[I] * exo - eEx[I] * ezo) * dt
        Ex = F.pad(ex, (1, 1, 1, 1)) + onehot * src_x[t]
        Ez = F.pad(ez, (1, 1, 1, 1)) + onehot * src_z[t]
        out.append(torch.stack([Ex[ri, rj], Ez[ri, rj]], dim=1))
    return torch.stack(out)


def _mean(a, b, code):
    """scalar_mean of averaging.f08:5-25 (1 harmonic, 2 geometric, 3 arithmetic, 4 none)."""
    if code == 1:
        return 2.0 * a * b / (a + b)
    if code == 2:
        return torch.sqrt(a * b)
    if code == 3:
        return 0.5 * (a + b)
    return a


def el2_traces(coef, src, cpml, dt, dx, dz, src_ij, rec_ij, density_code=3, corrected=False):
    """Receiver traces (steps, nrec, 2) = (vx, vz) after every step.

    coef: (10, nx, nz) tensor (c11, c13, c15, c33, c35, c55, rho, 1/(1 +
    gamma_x dt), 1/(1 + gamma_z dt), 1/(1 + gamma_xz dt)); src: (5, steps)
    raw series added to sigma_xx(is, js), sigma_xz(is+1, js+1),
    sigma_zz(is, js), vx(is, js), vz(is, js) (0-based src_ij = (is, js))."""
    dev = dict(dtype=torch.float64)
    c11, c13, c15, c33, c35, c55, rho, igx, igz, igxz = coef
    nx, nz = c11.shape
    k, a, b = (torch.as_tensor(np.asarray(cpml[n]), **dev) for n in ("kappa_cpml", "acoef_cpml", "bcoef_cpml"))
    kh, ah, bh = (torch.as_tensor(np.asarray(cpml[n]), **dev)
                  for n in ("kappa_half_cpml", "acoef_half_cpml", "bcoef_half_cpml"))
    nsteps = src.shape[1]
    r1 = (1, nx - 2, 2, nz - 1)     # sigma_xx, sigma_zz: i = 2..nx-2, j = 3..nz-1 (1-based)
    r2 = (2, nx - 1, 1, nz - 2)     # sigma_xz: i = 3..nx-1, j = 2..nz-2
    rx = (2, nx - 1, 2, nz - 1)     # vx: i = 3..nx-1, j = 3..nz-1
    rz = (1, nx - 2, 1, nz - 2)     # vz: i = 2..nx-2, j = 2..nz-2

    def at(f, r, di=0, dj=0):
        return f[r[0] + di:r[1] + di, r[2] + dj:r[3] + dj]

    def pad(v, r):
        return F.pad(v, (r[2], nz - r[3], r[0], nx - r[1]))

    def zeros(r):
        return torch.zeros((r[1] - r[0], r[3] - r[2]), **dev)

    vx = torch.zeros((nx, nz), **dev)
    vz = torch.zeros((nx, nz), **dev)
    sxx = torch.zeros((nx, nz), **dev)
    szz = torch.zeros((nx, nz), **dev)
    sxz = torch.zeros((nx, nz), **dev)
    m1 = [zeros(r1) for _ in range(4)]
    m2 = [zeros(r2) for _ in range(4)]
    mx = [zeros(rx) for _ in range(2)]
    mz = [zeros(rz) for _ in range(2)]
    inner = torch.zeros((nx, nz), **dev)
    inner[1:-1, 1:-1] = 1.0
    si, sj = src_ij
    ones = []
    for (i, j) in ((si, sj), (si + 1, sj + 1), (si, sj), (si, sj), (si, sj)):
        o = torch.zeros((nx, nz), **dev)
        o[i, j] = 1.0
        ones.append(o)
    ri = torch.as_tensor(np.asarray(rec_ij)[:, 0], dtype=torch.long)
    rj = torch.as_tensor(np.asarray(rec_ij)[:, 1], dtype=torch.long)
    X, Z = 24.0 * dx, 24.0 * dz

    def dplus_x(f, r):
        return (27.0 * (at(f, r, 1) - at(f, r)) - (at(f, r, 2) - at(f, r, -1))) / X

    def dminus_x(f, r):
        return (27.0 * (at(f, r) - at(f, r, -1)) - (at(f, r, 1) - at(f, r, -2))) / X

    def dplus_z(f, r):
        return (27.0 * (at(f, r, 0, 1) - at(f, r)) - (at(f, r, 0, 2) - at(f, r, 0, -1))) / Z

    def dminus_z(f, r):
        return (27.0 * (at(f, r) - at(f, r, 0, -1)) - (at(f, r, 0, 1) - at(f, r, 0, -2))) / Z

    def tminus(s0, sm1, sm2, sp1):
        return 27.0 * (s0 - sm1) - (sp1 - sm2) if corrected else 27.0 * s0 - 27.0 * sm1 + sm2

    def tplus(s0, sp1, sp2, sm1):
        return 27.0 * (sp1 - s0) - (sp2 - sm1) if corrected else -27.0 * s0 + 27.0 * sp1 - sp2

    out = []
    for t in range(nsteps):
        ux, wz, wx, uz = dplus_x(vx, r1), dminus_z(vz, r1), dplus_x(vz, r1), dminus_z(vx, r1)
        m1[0] = at(bh, r1) * m1[0] + at(ah, r1) * ux
        m1[1] = at(bh, r1) * m1[1] + at(ah, r1) * uz
        m1[2] = at(b, r1) * m1[2] + at(a, r1) * wz
        m1[3] = at(b, r1) * m1[3] + at(a, r1) * wx
        ux = ux / at(kh, r1) + m1[0]
        uz = uz / at(kh, r1) + m1[1]
        wz = wz / at(k, r1) + m1[2]
        wx = wx / at(k, r1) + m1[3]
        sh = wx + uz
        nxx = (at(sxx, r1) + (at(c11, r1) * ux + at(c13, r1) * wz + at(c15, r1) * sh) * dt) * at(igx, r1)
        nzz = (at(szz, r1) + (at(c13, r1) * ux + at(c33, r1) * wz + at(c35, r1) * sh) * dt) * at(igz, r1)
        ux, wz, wx, uz = dminus_x(vx, r2), dplus_z(vz, r2), dminus_x(vz, r2), dplus_z(vx, r2)
        m2[0] = at(bh, r2) * m2[0] + at(ah, r2) * ux
        m2[1] = at(bh, r2) * m2[1] + at(ah, r2) * uz
        m2[2] = at(bh, r2) * m2[2] + at(ah, r2) * wz
        m2[3] = at(bh, r2) * m2[3] + at(ah, r2) * wx
        ux = ux / at(kh, r2) + m2[0]
        uz = uz / at(kh, r2) + m2[1]
        wz = wz / at(kh, r2) + m2[2]
        wx = wx / at(kh, r2) + m2[3]
        nxz = (at(sxz, r2) + (at(c15, r2) * ux + at(c35, r2) * wz + at(c55, r2) * (wx + uz)) * dt) * at(igxz, r2)
        sxx = sxx * (1 - pad(torch.ones_like(nxx), r1)) + pad(nxx, r1)
        szz = szz * (1 - pad(torch.ones_like(nzz), r1)) + pad(nzz, r1)
        sxz = sxz * (1 - pad(torch.ones_like(nxz), r2)) + pad(nxz, r2)
        rhoxx = _mean(at(rho, rx), at(rho, rx, -1), density_code)
        rhozx = _mean(at(rho, rx), at(rho, rx, 0, -1), density_code)
        d1 = tminus(at(sxx, rx), at(sxx, rx, -1), at(sxx, rx, -2), at(sxx, rx, 1)) / X
        d2 = tminus(at(sxz, rx), at(sxz, rx, 0, -1), at(sxz, rx, 0, -2), at(sxz, rx, 0, 1)) / Z
        mx[0] = at(b, rx) * mx[0] + at(a, rx) * d1
        mx[1] = at(b, rx) * mx[1] + at(a, rx) * d2
        d1 = d1 / at(k, rx) + mx[0]
        d2 = d2 / at(k, rx) + mx[1]
        nvx = at(vx, rx) + dt * (d1 / rhoxx + d2 / rhozx)
        rhoxz = _mean(at(rho, rz), at(rho, rz, 1), density_code)
        rhozz = _mean(at(rho, rz), at(rho, rz, 0, 1), density_code)
        d1 = tplus(at(sxz, rz), at(sxz, rz, 1), at(sxz, rz, 2), at(sxz, rz, -1)) / X
        d2 = tplus(at(szz, rz), at(szz, rz, 0, 1), at(szz, rz, 0, 2), at(szz, rz, 0, -1)) / Z
        mz[0] = at(bh, rz) * mz[0] + at(ah, rz) * d1
        mz[1] = at(bh, rz) * mz[1] + at(ah, rz) * d2
        d1 = d1 / at(kh, rz) + mz[0]
        d2 = d2 / at(kh, rz) + mz[1]
        nvz = at(vz, rz) + dt * (d1 / rhoxz + d2 / rhozz)
        vx = vx * (1 - pad(torch.ones_like(nvx), rx)) + pad(nvx, rx)
        vz = vz * (1 - pad(torch.ones_like(nvz), rz)) + pad(nvz, rz)
        sxx = sxx + ones[0] * src[0, t]
        sxz = sxz + ones[1] * src[1, t]
        szz = szz + ones[2] * src[2, t]
        vx = (vx + ones[3] * src[3, t]) * inner
        vz = (vz + ones[4] * src[4, t]) * inner
        out.append(torch.stack([vx[ri, rj], vz[ri, rj]], dim=1))
    return torch.stack(out)


def el2_cfg_inputs(cfg):
    """(coef (10, nx, nz), raw source series (5, steps) with the default
    scaling 1/rho at the source cells (cpmlfdtd.f08:519-523), cpml dict, dt,
    dx, dz, 0-based source and receiver indices) of an `elastic2`
    configuration dict."""
    A = cfg["arrays"]
    dt = cfg["dt"]
    g = [np.asarray(A.get(n, np.zeros(A["rho"].shape)), dtype=np.float64) for n in ("gamma_x", "gamma_z", "gamma_xz")]
    coef = np.stack([np.asarray(A[n], dtype=np.float64) for n in ("c11", "c13", "c15", "c33", "c35", "c55", "rho")]
                    + [1.0 / (1.0 + x * dt) for x in g])
    cp = cfg["cpml"]
    si, sj = cfg["src"][0] + cp - 1, cfg["src"][1] + cp - 1
    rho = A["rho"]
    S = cfg["sources"]
    steps = cfg["time_steps"]

    def series(name):
        return np.asarray(S.get(name, np.zeros(steps)), dtype=np.float64)

    src = np.stack([series("xx") / rho[si, sj], series("xz") / rho[si + 1, sj + 1], series("zz") / rho[si, sj],
                    series("x") / rho[si, sj], series("z") / rho[si, sj]])
    rec = np.asarray(cfg["receivers"])
    rec_ij = np.stack([rec[:, 0] + cp - 1, rec[:, 2] + cp - 1], axis=1)
    return coef, src, A, dt, cfg["spacing"][0], cfg["spacing"][1], (si, sj), rec_ij


def em2_cfg_inputs(cfg):
    """(coef (7, nx, nz), scaled source series (2, steps), cpml dict, dt, dx,
    dz, 0-based source and receiver indices) of an `em2` configuration dict,
    with the library's default source scaling dt / (eps11 eps0) at the source
    cell (cpmlfdtd.f08:2678-2681)."""
    A = cfg["arrays"]
    coef = em2_coefficients_np(*(np.asarray(A[n], dtype=np.float64) for n in ("e11", "e13", "e33", "s11", "s13", "s33")))
    cp = cfg["cpml"]
    si, sj = cfg["src"][0] + cp - 1, cfg["src"][1] + cp - 1
    div = A["e11"][si, sj] * EPS0
    sx = np.asarray(cfg["sources"]["x"], dtype=np.float64) * cfg["dt"] / div
    sz = np.asarray(cfg["sources"]["z"], dtype=np.float64) * cfg["dt"] / div
    rec = np.asarray(cfg["receivers"])
    rec_ij = np.stack([rec[:, 0] + cp - 1, rec[:, 2] + cp - 1], axis=1)
    return coef, np.stack([sx, sz]), A, cfg["dt"], cfg["spacing"][0], cfg["spacing"][1], (si, sj), rec_ij
