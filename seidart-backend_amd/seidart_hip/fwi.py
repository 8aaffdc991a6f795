"""Full-waveform-inversion gradients as torch.autograd Functions.

The forward pass runs the 2D solvers on the GPU with the per-cell update
coefficients and the (already scaled) source series as inputs, saving
checkpoints; the backward pass runs the hand-written discrete adjoint
(sdt_adjoint / sdt_batch_adjoint), which recomputes each checkpoint segment
and correlates forward and adjoint fields into FP64 gradient accumulators.
Gradients with respect to physical parameters follow from torch autograd
through the coefficient formulas below (FP64, differentiable).

Coefficient order: see `EM2_COEFFS` / `EL2_COEFFS`.  A shot is a dict with
keys shape, spacing, cpml (the padding width), dt, src (1-based unpadded
(x, z)), receivers ((n, 3) 1-based unpadded), cpml_arrays (the kappa / alpha
/ acoef / bcoef arrays and their _half forms) and, for the elastic solver,
optionally density_method and stencil.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .solver import Batch, Solver, flatten_coefficients

EPS0 = 8.85418782e-12                        # constants.f08:10
EM2_COEFFS = ("aEx", "bEx", "bEz", "dEx", "eEx", "dEz", "eEz")
EL2_COEFFS = ("c11", "c13", "c15", "c33", "c35", "c55", "rho", "igx", "igz", "igxz")
# source components the raw series feed, in the order of the src rows
_SRC = {"em2": ("x", "z"), "elastic2": ("xx", "xz", "zz", "x", "z")}
_TRACE = {"em2": (0, 2), "elastic2": (0, 2)}   # recorded components kept: (Ex, Ez) / (vx, vz)


def em2_coefficients(eps, sig):
    """(7, nx, nz) E-update coefficients from relative permittivity
    eps = (e11, e13, e33) and conductivity sig = (s11, s13, s33), as
    cpmlfdtd.f08:2519-2540 computes them (permittivity scaled by eps0,
    determinant clamped at 1e-18 eps0^2)."""
    e11, e13, e33 = eps[0] * EPS0, eps[1] * EPS0, eps[2] * EPS0
    s11, s13, s33 = sig[0], sig[1], sig[2]
    det = e11 * e33 - e13 * e13
    det = torch.clamp(det, min=1.0e-18 * EPS0 * EPS0)
    return torch.stack([e33 / det, -e13 / det, e11 / det, (e33 * s11 - e13 * s13) / det,
                        (e33 * s13 - e13 * s33) / det, (e11 * s13 - e13 * s11) / det,
                        (e11 * s33 - e13 * s13) / det])


def el2_coefficients(stiffness, rho, gamma, dt):
    """(10, nx, nz) update coefficients of the 2D elastic solver from the
    stiffness set (c11, c13, c15, c33, c35, c55), density and attenuation
    (gamma_x, gamma_z, gamma_xz): the first seven as given, then
    1/(1 + gamma dt) (cpmlfdtd.f08:344-353, 382-386)."""
    return torch.stack(list(stiffness) + [rho] + [1.0 / (1.0 + g * dt) for g in gamma])


def em2_source_scale(eps11_at_src, dt):
    """The library's default EM source scaling dt / (eps11 eps0) at the
    source cell (cpmlfdtd.f08:2678-2681); multiply a wavelet by it to get the
    raw series the Functions take."""
    return dt / (eps11_at_src * EPS0)


def _make_solver(physics, shot, precision, checkpoint, coef, src):
    s = Solver(physics, shot["shape"], shot["spacing"], shot["cpml"], shot["dt"], src.shape[-1], shot["src"],
               precision=precision, density_method=shot.get("density_method", "arithmetic"),
               stencil=shot.get("stencil", "reference"))
    try:
        s.set_option("checkpoint", checkpoint)
        s.set_option("raw_source", 1)
        s.set_arrays(shot["cpml_arrays"])
        s.set_coefficients(coef)
        for k, comp in enumerate(_SRC[physics]):
            s.set_source(comp, src[k])
        s.set_receivers(shot["receivers"])
    except Exception:
        s.close()
        raise
    return s


def _trace_grad(physics, g, nt, nrec):
    tg = np.zeros((nt, nrec, 3))
    for k, m in enumerate(_TRACE[physics]):
        tg[:, :, m] = g[:, :, k]
    return tg


class _Traces(torch.autograd.Function):
    """One shot: (coef, src) -> traces (steps, nrec, 2)."""

    @staticmethod
    def forward(ctx, coef, src, physics, shot, precision, checkpoint):
        s = _make_solver(physics, shot, precision, checkpoint, coef.detach().cpu().numpy(),
                         src.detach().cpu().numpy())
        try:
            s.run()
            rec = s.receivers()
        except Exception:
            s.close()
            raise
        ctx.solver = s
        ctx.physics = physics
        return torch.from_numpy(rec[:, :, list(_TRACE[physics])].copy())

    @staticmethod
    def backward(ctx, g):
        s = ctx.solver
        try:
            g = g.detach().cpu().numpy()
            cg, sg = s.adjoint(_trace_grad(ctx.physics, g, s.time_steps, s.nrec))
        finally:
            s.close()
            ctx.solver = None
        gsrc = np.stack([sg[L.SOURCE_IDS[c]] for c in _SRC[ctx.physics]])
        return torch.from_numpy(cg), torch.from_numpy(gsrc), None, None, None, None


def traces(physics, coef, src, shot, *, precision="fp32", checkpoint=100):
    """Differentiable receiver traces (steps, nrec, 2) of one shot.

    coef: (ncoef, nx, nz) float64 tensor; src: float64 tensor of raw series,
    (2, steps) for "em2" (added to Ex, Ez at the source cell) or (5, steps)
    for "elastic2" (added to sigma_xx, sigma_xz at (+1, +1), sigma_zz, vx,
    vz; the library's default scaling divides each by the density there)."""
    if not (torch.is_grad_enabled() and (coef.requires_grad or src.requires_grad)):
        checkpoint = 0   # no backward pass can follow: no checkpoints
    return _Traces.apply(coef, src, physics, shot, precision, int(checkpoint))


class _BatchTraces(torch.autograd.Function):
    """Many shots sharing one model: (coef, src[nshot]) -> traces[nshot]."""

    @staticmethod
    def forward(ctx, coef, src, physics, shots, precision, checkpoint):
        c = flatten_coefficients(coef.detach().cpu().numpy())   # once for every shot
        sv = src.detach().cpu().numpy()
        solvers = []
        b = None
        try:
            for q, shot in enumerate(shots):
                solvers.append(_make_solver(physics, shot, precision, checkpoint, c, sv[q]))
            b = Batch(solvers)
            b.run()
            out = np.stack([s.receivers()[:, :, list(_TRACE[physics])] for s in solvers])
        except Exception:
            if b is not None:
                b.close()   # before the handles it points to
            for s in solvers:
                s.close()
            raise
        ctx.batch, ctx.solvers, ctx.physics = b, solvers, physics
        return torch.from_numpy(out)

    @staticmethod
    def backward(ctx, g):
        b, solvers, physics = ctx.batch, ctx.solvers, ctx.physics
        try:
            g = g.detach().cpu().numpy()
            tgs = [np.ascontiguousarray(_trace_grad(physics, g[q], s.time_steps, s.nrec))
                   for q, s in enumerate(solvers)]
            nc = solvers[0].info("coefficients")
            cgs = [np.empty((nc, s.ncell)) for s in solvers]
            sgs = [np.empty((L.NUM_SRC, s.time_steps)) for s in solvers]
            D = C.POINTER(C.c_double)
            tp = (D * len(solvers))(*[t.ctypes.data_as(D) for t in tgs])
            cp = (D * len(solvers))(*[x.ctypes.data_as(D) for x in cgs])
            sp = (D * len(solvers))(*[x.ctypes.data_as(D) for x in sgs])
            L.check(L.lib.sdt_batch_adjoint(b._b, tp, cp, sp))
        finally:
            b.close()
            for s in solvers:
                s.close()
            ctx.batch = ctx.solvers = None
        shape = solvers[0].shape
        acc = np.zeros_like(cgs[0])
        for x in cgs:   # shot order, as a running sum
            acc += x
        gc = np.stack([row.reshape(shape, order="F") for row in acc])
        gsrc = np.stack([np.stack([sg[L.SOURCE_IDS[c]] for c in _SRC[physics]]) for sg in sgs])
        return torch.from_numpy(gc), torch.from_numpy(gsrc), None, None, None, None


def batch_traces(physics, coef, src, shots, *, precision="fp32", checkpoint=100):
    """Differentiable traces (nshot, steps, nrec, 2) of shots that share one
    model (same grid), run as one batch; the backward pass sums the
    per-shot coefficient gradients."""
    if not (torch.is_grad_enabled() and (coef.requires_grad or src.requires_grad)):
        checkpoint = 0
    return _BatchTraces.apply(coef, src, physics, list(shots), precision, int(checkpoint))


def l2_misfit(pred, obs, weight=None):
    """0.5 * sum(weight * (pred - obs)^2), the L2 trace misfit.  `weight`
    (optional) holds one factor per receiver: shape (nrec,) for the traces
    of one shot (steps, nrec, 2), (nshot, nrec) for a batch (nshot, steps,
    nrec, 2); see near_offset_weights."""
    r2 = (pred - obs) ** 2
    if weight is not None:
        w = torch.as_tensor(weight, dtype=r2.dtype)
        r2 = r2 * (w[:, None, :, None] if w.dim() == 2 else w[None, :, None])
    return 0.5 * torch.sum(r2)


def near_offset_weights(shots, radius=5):
    """Per-shot receiver weights of a near-offset mute: 0 for receivers within
    `radius` cells of the shot's source cell (in x and in z), 1 otherwise, a
    standard FWI data mute (the `weight` of l2_misfit).

    FP32 accuracy does not need it: a receiver on or next to the source
    records the injected pulse itself, and its residual is a small
    difference of large traces; 2D EM shots run in FP32 compute a window
    around the source in FP64 (solver option 'source_window', default on),
    so those traces carry no FP32 rounding of the pulse.  On a 300 x 150,
    1500-step EM problem with a receiver on a source cell the FP32 and FP64
    gradients agree to 5e-6 with or without this mute (5e-3 without the
    window and without a mute)."""
    out = []
    for sh in shots:
        rec = np.asarray(sh["receivers"]).reshape(-1, 3)
        sx, sz = sh["src"][0], sh["src"][-1]
        near = (np.abs(rec[:, 0] - sx) <= radius) & (np.abs(rec[:, 2] - sz) <= radius)
        out.append(np.where(near, 0.0, 1.0))
    return np.stack(out)


def smooth(field, passes=10):
    """Starting model: repeated 3x3 box smoothing of a (nx, nz) array."""
    x = torch.as_tensor(field, dtype=torch.float64)[None, None]
    for _ in range(passes):
        x = torch.nn.functional.avg_pool2d(torch.nn.functional.pad(x, (1, 1, 1, 1), mode="replicate"), 3, 1)
    return x[0, 0]


__all__ = ["EM2_COEFFS", "EL2_COEFFS", "em2_coefficients", "el2_coefficients", "em2_source_scale", "traces",
           "batch_traces", "l2_misfit", "near_offset_weights", "smooth"]
