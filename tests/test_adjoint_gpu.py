"""Adjoint-state gradients (sdt_adjoint through seidart_hip.fwi) against the
autodiff oracle (torch autograd of the float64 CPU restatement of the time
stepping), against central finite differences, and FP32 against FP64."""
import numpy as np
import pytest

from conftest import rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _shot(cfg):
    return dict(shape=cfg["shape"], spacing=cfg["spacing"], cpml=cfg["cpml"], dt=cfg["dt"], src=cfg["src"],
                receivers=cfg["receivers"],
                cpml_arrays={k: v for k, v in cfg["arrays"].items() if k.endswith("_cpml")})


def _em2_case(steps=150, seed=4):
    from oracle.adjoint_torch import em2_cfg_inputs
    from seidart_hip import models
    cfg = models.em2(nx=50, nz=40, steps=steps, src=(25, 12), nrec=8, seed=seed)
    coef, src, A, dt, dx, dz, sij, rij = em2_cfg_inputs(cfg)
    return cfg, coef, src, (A, dt, dx, dz, sij, rij)


def _gpu_grad(cfg, coef, src, w, precision, checkpoint):
    from seidart_hip import fwi
    c = torch.tensor(coef, requires_grad=True)
    s = torch.tensor(src, requires_grad=True)
    tr = fwi.traces("em2", c, s, _shot(cfg), precision=precision, checkpoint=checkpoint)
    (tr * torch.tensor(w)).sum().backward()
    return tr.detach().numpy(), c.grad.numpy(), s.grad.numpy()


@pytest.mark.parametrize("checkpoint", [40, 150])
def test_em2_adjoint_matches_autograd(checkpoint):
    from oracle.adjoint_torch import em2_traces
    cfg, coef, src, o = _em2_case()
    print(f"[adjoint] em2 shape={cfg['shape']} steps={cfg['time_steps']} checkpoint={checkpoint}", flush=True)
    w = np.random.default_rng(0).standard_normal((cfg["time_steps"], len(cfg["receivers"]), 2))
    tr, gc, gs = _gpu_grad(cfg, coef, src, w, "fp64", checkpoint)
    c2 = torch.tensor(coef, requires_grad=True)
    s2 = torch.tensor(src, requires_grad=True)
    tr2 = em2_traces(c2, s2[0], s2[1], *o)
    (tr2 * torch.tensor(w)).sum().backward()
    assert rel_l2(tr, tr2.detach().numpy()) <= 1e-10
    errs = [rel_l2(gc[q], c2.grad.numpy()[q]) for q in range(7)]
    print("coefficient gradients rel L2:", [f"{e:.2e}" for e in errs], "source:",
          f"{rel_l2(gs, s2.grad.numpy()):.2e}", flush=True)
    assert max(errs) <= 1e-9
    assert rel_l2(gs, s2.grad.numpy()) <= 1e-9


def test_em2_adjoint_checkpoint_interval_invariant():
    cfg, coef, src, _ = _em2_case(steps=120)
    w = np.random.default_rng(1).standard_normal((120, len(cfg["receivers"]), 2))
    ref = _gpu_grad(cfg, coef, src, w, "fp64", 120)
    for every in (1, 7, 32):
        got = _gpu_grad(cfg, coef, src, w, "fp64", every)
        assert rel_l2(got[1], ref[1]) <= 1e-12, every
        assert rel_l2(got[2], ref[2]) <= 1e-12, every


def test_em2_adjoint_fp32_close_to_fp64():
    cfg, coef, src, _ = _em2_case(steps=200)
    w = np.random.default_rng(2).standard_normal((200, len(cfg["receivers"]), 2))
    _, g64, s64 = _gpu_grad(cfg, coef, src, w, "fp64", 50)
    _, g32, s32 = _gpu_grad(cfg, coef, src, w, "fp32", 50)
    errs = [rel_l2(g32[q], g64[q]) for q in range(7)]
    print("fp32 vs fp64 coefficient gradients rel L2:", [f"{e:.2e}" for e in errs], flush=True)
    assert rel_l2(g32, g64) <= 1e-4
    assert max(errs) <= 1e-3
    assert rel_l2(s32, s64) <= 1e-4


def test_em2_gradient_matches_finite_differences():
    """dJ/d eps11 and dJ/d sigma11 at interior cells, J the L2 misfit of
    traces against those of a perturbed model; central differences with a
    1e-4 relative step in FP64."""
    from seidart_hip import fwi, models
    cfg = models.em2(nx=40, nz=30, steps=160, src=(20, 8), nrec=6, seed=9)
    A = cfg["arrays"]
    shot = _shot(cfg)
    names = ("e11", "e13", "e33", "s11", "s13", "s33")
    eps0 = [torch.tensor(A[n]) for n in names[:3]]
    sig0 = [torch.tensor(A[n]) for n in names[3:]]
    cp = cfg["cpml"]
    si, sj = cfg["src"][0] + cp - 1, cfg["src"][1] + cp - 1
    scale = fwi.em2_source_scale(A["e11"][si, sj], cfg["dt"])
    src = torch.tensor(np.stack([cfg["sources"]["x"], cfg["sources"]["z"]]) * scale)
    # observed data from a model with a bump in eps11 and sigma11
    eps_t = [e.clone() for e in eps0]
    sig_t = [s.clone() for s in sig0]
    eps_t[0][cp + 15:cp + 25, cp + 14:cp + 20] *= 1.1
    sig_t[0][cp + 15:cp + 25, cp + 14:cp + 20] *= 1.5
    with torch.no_grad():
        obs = fwi.traces("em2", fwi.em2_coefficients(eps_t, sig_t), src, shot, precision="fp64", checkpoint=40)

    def J(eps, sig):
        return fwi.l2_misfit(fwi.traces("em2", fwi.em2_coefficients(eps, sig), src, shot, precision="fp64",
                                        checkpoint=40), obs)

    e = [x.clone().requires_grad_(True) for x in eps0]
    s = [x.clone().requires_grad_(True) for x in sig0]
    J(e, s).backward()
    rng = np.random.default_rng(3)
    cells = [(int(rng.integers(cp + 2, cp + 38)), int(rng.integers(cp + 2, cp + 28))) for _ in range(8)]
    fd, ad = [], []
    for which, base, grad in (("eps11", eps0, e[0].grad), ("sig11", sig0, s[0].grad)):
        for (i, j) in cells:
            h = 1e-4 * float(base[0][i, j])
            vals = []
            for sign in (1.0, -1.0):
                p = [x.clone() for x in base]
                p[0][i, j] += sign * h
                with torch.no_grad():
                    vals.append(float(J(p, sig0) if which == "eps11" else J(eps0, p)))
            fd.append((vals[0] - vals[1]) / (2 * h))
            ad.append(float(grad[i, j]))
    fd, ad = np.array(fd), np.array(ad)
    err = np.linalg.norm(fd - ad) / np.linalg.norm(fd)
    print("finite differences vs adjoint rel L2:", f"{err:.2e}", flush=True)
    assert err <= 1e-3


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_em2_batch_adjoint_equals_single_shots(precision):
    """Shots sharing one model, run as one batch: traces and the summed
    gradient equal those of the shots run one at a time."""
    from seidart_hip import fwi, models
    cfg = models.em2(nx=56, nz=44, steps=130, src=(28, 12), nrec=8, seed=5)
    base = _shot(cfg)
    shots = []
    for q, sx in enumerate((10, 28, 45)):
        sh = dict(base)
        sh["src"] = (sx, 10 + 4 * q)
        shots.append(sh)
    rng = np.random.default_rng(4)
    nt = cfg["time_steps"]
    src = torch.tensor(rng.standard_normal((3, 2, nt)) * 1e3)
    from oracle.adjoint_torch import em2_cfg_inputs
    coef_np = em2_cfg_inputs(cfg)[0]
    w = torch.tensor(rng.standard_normal((3, nt, len(cfg["receivers"]), 2)))
    print(f"[adjoint] batch of 3 em2 shots {precision}", flush=True)
    c = torch.tensor(coef_np, requires_grad=True)
    s = src.clone().requires_grad_(True)
    tb = fwi.batch_traces("em2", c, s, shots, precision=precision, checkpoint=30)
    (tb * w).sum().backward()
    gsum = torch.zeros_like(c)
    for q in range(3):
        c1 = torch.tensor(coef_np, requires_grad=True)
        s1 = src[q].clone().requires_grad_(True)
        t1 = fwi.traces("em2", c1, s1, shots[q], precision=precision, checkpoint=30)
        (t1 * w[q]).sum().backward()
        assert rel_l2(tb[q].detach().numpy(), t1.detach().numpy()) <= 1e-13
        assert rel_l2(s.grad[q].numpy(), s1.grad.numpy()) <= 1e-12
        gsum += c1.grad
    err = rel_l2(c.grad.numpy(), gsum.numpy())
    print("batch vs single-shot gradient rel L2:", f"{err:.2e}", flush=True)
    assert err <= 1e-12


# ---- 2D elastic --------------------------------------------------------------
def _el2_case(steps=140, density="arithmetic", stencil="reference", gamma=0.0, seed=3):
    from oracle.adjoint_torch import el2_cfg_inputs
    from seidart_hip import models
    cfg = models.elastic2(nx=44, nz=38, steps=steps, src=(22, 12), nrec=7, seed=seed, nlayers=3,
                          density_method=density, gamma=gamma, tilt=(0.0, 25.0))
    cfg["stencil"] = stencil
    coef, src, A, dt, dx, dz, sij, rij = el2_cfg_inputs(cfg)
    src = np.random.default_rng(seed).standard_normal(src.shape) * np.abs(src).max()
    return cfg, coef, src, (A, dt, dx, dz, sij, rij)


def _gpu_grad_el2(cfg, coef, src, w, precision, checkpoint):
    from seidart_hip import fwi
    c = torch.tensor(coef, requires_grad=True)
    s = torch.tensor(src, requires_grad=True)
    shot = _shot(cfg)
    shot["density_method"] = cfg.get("density_method", "arithmetic")
    shot["stencil"] = cfg.get("stencil", "reference")
    tr = fwi.traces("elastic2", c, s, shot, precision=precision, checkpoint=checkpoint)
    (tr * torch.tensor(w)).sum().backward()
    return tr.detach().numpy(), c.grad.numpy(), s.grad.numpy()


@pytest.mark.parametrize("density,stencil,gamma", [("arithmetic", "reference", 0.0),
                                                   ("harmonic", "corrected", 3.0),
                                                   ("geometric", "reference", 1.0)])
def test_el2_adjoint_matches_autograd(density, stencil, gamma):
    from oracle.adjoint_torch import el2_traces
    cfg, coef, src, o = _el2_case(density=density, stencil=stencil, gamma=gamma)
    print(f"[adjoint] elastic2 {density} {stencil} gamma={gamma} shape={cfg['shape']}", flush=True)
    w = np.random.default_rng(0).standard_normal((cfg["time_steps"], len(cfg["receivers"]), 2))
    tr, gc, gs = _gpu_grad_el2(cfg, coef, src, w, "fp64", 45)
    c2 = torch.tensor(coef, requires_grad=True)
    s2 = torch.tensor(src, requires_grad=True)
    code = {"harmonic": 1, "geometric": 2, "arithmetic": 3}[density]
    tr2 = el2_traces(c2, s2, *o, density_code=code, corrected=stencil == "corrected")
    (tr2 * torch.tensor(w)).sum().backward()
    assert rel_l2(tr, tr2.detach().numpy()) <= 1e-10
    errs = [rel_l2(gc[q], c2.grad.numpy()[q]) for q in range(10)]
    print("coefficient gradients rel L2:", [f"{e:.2e}" for e in errs], "source:",
          f"{rel_l2(gs, s2.grad.numpy()):.2e}", flush=True)
    assert max(errs) <= 1e-9
    assert rel_l2(gs, s2.grad.numpy()) <= 1e-9


def test_el2_adjoint_fp32_close_to_fp64():
    cfg, coef, src, _ = _el2_case(steps=200, gamma=1.0)
    w = np.random.default_rng(2).standard_normal((200, len(cfg["receivers"]), 2))
    _, g64, s64 = _gpu_grad_el2(cfg, coef, src, w, "fp64", 50)
    _, g32, s32 = _gpu_grad_el2(cfg, coef, src, w, "fp32", 50)
    errs = [rel_l2(g32[q], g64[q]) for q in range(10)]
    print("fp32 vs fp64 coefficient gradients rel L2:", [f"{e:.2e}" for e in errs], flush=True)
    assert max(errs) <= 1e-3
    assert rel_l2(s32, s64) <= 1e-4


def test_el2_gradient_matches_finite_differences():
    """dJ/dc11 and dJ/drho at interior cells, J the L2 misfit against traces
    of a model with a stiffness and density bump; central differences with a
    1e-4 relative step in FP64."""
    from seidart_hip import fwi
    cfg, coef, src, _ = _el2_case(steps=160, seed=6)
    shot = _shot(cfg)
    src_t = torch.tensor(src)
    cp = cfg["cpml"]
    true = coef.copy()
    true[0, cp + 14:cp + 24, cp + 10:cp + 18] *= 1.1
    true[6, cp + 14:cp + 24, cp + 10:cp + 18] *= 0.9
    with torch.no_grad():
        obs = fwi.traces("elastic2", torch.tensor(true), src_t, shot, precision="fp64", checkpoint=40)

    def J(c):
        return fwi.l2_misfit(fwi.traces("elastic2", c, src_t, shot, precision="fp64", checkpoint=40), obs)

    c = torch.tensor(coef, requires_grad=True)
    J(c).backward()
    rng = np.random.default_rng(5)
    cells = [(int(rng.integers(cp + 3, cp + 40)), int(rng.integers(cp + 3, cp + 34))) for _ in range(8)]
    fd, ad = [], []
    for q in (0, 6):
        for (i, j) in cells:
            h = 1e-4 * abs(coef[q, i, j])
            vals = []
            for sign in (1.0, -1.0):
                p = coef.copy()
                p[q, i, j] += sign * h
                with torch.no_grad():
                    vals.append(float(J(torch.tensor(p))))
            fd.append((vals[0] - vals[1]) / (2 * h))
            ad.append(float(c.grad[q, i, j]))
    fd, ad = np.array(fd), np.array(ad)
    err = np.linalg.norm(fd - ad) / np.linalg.norm(fd)
    print("finite differences vs adjoint rel L2:", f"{err:.2e}", flush=True)
    assert err <= 1e-3


def test_el2_batch_adjoint_equals_single_shots():
    from seidart_hip import fwi
    cfg, coef, _, _ = _el2_case(steps=120)
    base = _shot(cfg)
    shots = []
    for q, sx in enumerate((8, 22, 36)):
        sh = dict(base)
        sh["src"] = (sx, 8 + 5 * q)
        shots.append(sh)
    rng = np.random.default_rng(7)
    nt = cfg["time_steps"]
    src = torch.tensor(rng.standard_normal((3, 5, nt)) * 1e-3)
    w = torch.tensor(rng.standard_normal((3, nt, len(cfg["receivers"]), 2)))
    c = torch.tensor(coef, requires_grad=True)
    s = src.clone().requires_grad_(True)
    tb = fwi.batch_traces("elastic2", c, s, shots, precision="fp64", checkpoint=30)
    (tb * w).sum().backward()
    gsum = torch.zeros_like(c)
    for q in range(3):
        c1 = torch.tensor(coef, requires_grad=True)
        s1 = src[q].clone().requires_grad_(True)
        t1 = fwi.traces("elastic2", c1, s1, shots[q], precision="fp64", checkpoint=30)
        (t1 * w[q]).sum().backward()
        assert rel_l2(tb[q].detach().numpy(), t1.detach().numpy()) <= 1e-13
        assert rel_l2(s.grad[q].numpy(), s1.grad.numpy()) <= 1e-12
        gsum += c1.grad
    err = rel_l2(c.grad.numpy(), gsum.numpy())
    print("batch vs single-shot gradient rel L2:", f"{err:.2e}", flush=True)
    assert err <= 1e-12


def _em2_fwi_case(nshots=3, steps=1500):
    """A 300 x 150 FWI problem whose first shot's source cell carries a
    receiver (the geometry that exposes FP32 rounding of the recorded source
    pulse), the others between receivers."""
    from seidart_hip import fwi, models
    prob = models.fwi2d("em2", nx=300, nz=150, steps=steps, nshots=nshots, nrec=50)
    rx = np.asarray(prob["shots"][0]["receivers"])[:, 0]
    prob["shots"][0]["src"] = (int(rx[7]), prob["shots"][0]["src"][1])
    cp, dt = prob["cfg"]["cpml"], prob["cfg"]["dt"]
    names = ("e11", "e13", "e33", "s11", "s13", "s33")
    src = torch.tensor(np.stack([np.stack([prob["wavelet"], np.zeros(steps)]) *
                                 fwi.em2_source_scale(prob["start"]["e11"][s["src"][0] + cp - 1,
                                                                           s["src"][1] + cp - 1], dt)
                                 for s in prob["shots"]]))
    true = [torch.tensor(prob["true"][n]) for n in names]
    start = [torch.tensor(prob["start"][n]) for n in names]
    with torch.no_grad():
        obs = fwi.batch_traces("em2", fwi.em2_coefficients(true[:3], true[3:]), src, prob["shots"],
                               precision="fp64")
    return prob, src, start, obs


def test_em2_fp32_batch_gradient_within_1e4_of_fp64():
    """Batched FP32 forward + adjoint against serial FP64 forward + adjoint,
    L2 misfit over every trace (no mute), a receiver on shot 0's source cell:
    every permittivity and conductivity gradient within 1e-4 relative L2
    (observed <= 5.1e-6), and likewise with 2- and 5-cell near-offset mutes.
    The FP64 window around each FP32 shot's source (SDT_OPT_SOURCE_WINDOW)
    keeps that receiver's trace free of FP32 rounding of the source pulse,
    which otherwise enters the adjoint source (up to 5e-3 here without the
    window and without a mute)."""
    from seidart_hip import fwi
    prob, src, start, obs = _em2_fwi_case()
    shots = prob["shots"]
    print(f"[adjoint] em2 fwi {prob['cfg']['shape']} {len(shots)} shots {src.shape[-1]} steps", flush=True)
    out = {}
    for mute in (5, 2, -1):
        w = fwi.near_offset_weights(shots, mute) if mute >= 0 else None
        g = {}
        for prec in ("fp32", "fp64"):
            p = [x.clone().requires_grad_(True) for x in start]
            c = fwi.em2_coefficients(p[:3], p[3:])
            if prec == "fp32":
                fwi.l2_misfit(fwi.batch_traces("em2", c, src, shots, precision="fp32", checkpoint=100),
                              obs, w).backward()
            else:
                J = 0.0
                for q in range(len(shots)):
                    J = J + fwi.l2_misfit(fwi.traces("em2", c, src[q], shots[q], precision="fp64",
                                                     checkpoint=100), obs[q], None if w is None else w[q])
                J.backward()
            g[prec] = np.stack([x.grad.numpy() for x in p])
        out[mute] = [rel_l2(g["fp32"][k], g["fp64"][k]) for k in range(6)]
        print(f"near-offset mute {mute}: fp32 vs fp64 gradient rel L2", [f"{e:.2e}" for e in out[mute]], flush=True)
    assert max(max(e) for e in out.values()) <= 1e-4
