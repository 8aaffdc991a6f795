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

Its forward traces are pinned against the C oracle (oracle.c), which is
pinned against the reference's own cases.
"""
from __future__ import annotations

import math

import numpy as np
import torch
import torch.nn.functional as F

EPS0 = 8.85418782e-12                        # constants.f08:11
MU0 = 4.0 * math.pi * 1.0e-7                 # constants.f08:10


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
