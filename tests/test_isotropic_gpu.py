"""The 3D elastic isotropic fast path (only c11, c12, c44 held and read when
every cell's stiffness has the isotropic form and no material table applies)
against the oracle, against the general 21-array path on the same model, and
its selection rules."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
FIELDS = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")


def _cfg(**kw):
    from seidart_hip import models
    return models.elastic3_iso(nx=28, ny=26, nz=30, cpml=6, steps=160, src=(14, 13, 5), nrec=6, **kw)


def _run(cfg, precision="fp64", **options):
    from seidart_hip.solver import make_solver
    s = make_solver(cfg, precision=precision, options=options or None)
    try:
        s.run(cfg["time_steps"])
        return s.receivers(), s.norms(), {f: s.field(f) for f in FIELDS}, s.info("isotropic"), s.materials()
    finally:
        s.close()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 1e-3)])
def test_isotropic_path_matches_oracle(precision, tol):
    from oracle.oracle_ctypes import run_oracle
    cfg = _cfg()
    print(f"[isotropic] oracle parity {precision}", flush=True)
    g_rec, (g_vn, g_sn), g_f, iso, nmat = _run(cfg, precision)
    assert iso == 1 and nmat == 0
    o_rec, (o_vn, o_sn), o_f, us = run_oracle(cfg, fields=FIELDS)
    assert us == 0 and np.abs(o_rec).max() > 0
    errs = {"receivers": rel_l2(g_rec, o_rec), "velnorm": rel_l2(g_vn, o_vn), "stressnorm": rel_l2(g_sn, o_sn)}
    errs.update({k: rel_l2(g_f[k], o_f[k]) for k in FIELDS})
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= tol, (k, v)


def test_isotropic_path_equals_general_path():
    """Same arithmetic as the 21-array path (the zero entries add exact
    zeros): the two agree to rounding, with a plane-wave source too."""
    cfg = _cfg()
    cfg["plane_wave"] = dict(kind="p", angles=(20.0, 30.0, 10.0), f0=2500.0, amplitude=1e-3, wavelet=2)
    print("[isotropic] fast vs general path", flush=True)
    a_rec, (a_vn, a_sn), a_f, a_iso, _ = _run(cfg)
    b_rec, (b_vn, b_sn), b_f, b_iso, _ = _run(cfg, isotropic=0)
    assert a_iso == 1 and b_iso == 0
    errs = {"receivers": rel_l2(a_rec, b_rec), "velnorm": rel_l2(a_vn, b_vn), "stressnorm": rel_l2(a_sn, b_sn)}
    errs.update({k: rel_l2(a_f[k], b_f[k]) for k in FIELDS})
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= 1e-13, (k, v)


def test_isotropic_selection_rules():
    """Not selected when one cell breaks the isotropic form, nor when the
    model fits the material table (which reads less)."""
    cfg = _cfg(attenuation=False)
    c14 = np.zeros(cfg["shape"], order="F")
    c14[10, 11, 12] = 1.0e6
    cfg["arrays"]["c14"] = c14
    print("[isotropic] selection rules", flush=True)
    _, _, _, iso, nmat = _run(cfg)
    assert iso == 0 and nmat == 0
    from seidart_hip import models
    layered = models.elastic3(nx=20, ny=18, nz=22, cpml=4, steps=10, src=(10, 9, 6), nrec=2)
    lam, mu = 7.0e9, 3.0e9
    for n in ("c11", "c22", "c33"):
        layered["arrays"][n] = np.full(layered["shape"], lam + 2 * mu, order="F")
    for n in ("c12", "c13", "c23"):
        layered["arrays"][n] = np.full(layered["shape"], lam, order="F")
    for n in ("c44", "c55", "c66"):
        layered["arrays"][n] = np.full(layered["shape"], mu, order="F")
    for n in ("c14", "c15", "c16", "c24", "c25", "c26", "c34", "c35", "c36", "c45", "c46", "c56"):
        layered["arrays"][n] = np.zeros(layered["shape"], order="F")
    _, _, _, iso, nmat = _run(layered)
    assert iso == 0 and nmat > 0
    _, _, _, iso, nmat = _run(layered, material_table=0)
    assert iso == 1 and nmat == 0


@pytest.mark.parametrize("isotropic", [1, 0])
def test_partly_equal_attenuation_arrays(isotropic):
    """Attenuation divisors: equal directions share one device array; with
    two distinct values among the six the per-direction loads stay."""
    from oracle.oracle_ctypes import run_oracle
    cfg = _cfg()
    g = cfg["arrays"]["gamma_x"]
    cfg["arrays"]["gamma_xy"] = np.asfortranarray(g * 1.7)
    cfg["arrays"]["gamma_yz"] = cfg["arrays"]["gamma_xy"]
    print(f"[isotropic] partly equal attenuation, isotropic={isotropic}", flush=True)
    g_rec, (g_vn, g_sn), g_f, iso, _ = _run(cfg, isotropic=isotropic)
    assert iso == isotropic
    o_rec, (o_vn, o_sn), o_f, us = run_oracle(cfg, fields=FIELDS)
    assert us == 0
    errs = {"receivers": rel_l2(g_rec, o_rec), "velnorm": rel_l2(g_vn, o_vn), "stressnorm": rel_l2(g_sn, o_sn)}
    errs.update({k: rel_l2(g_f[k], o_f[k]) for k in FIELDS})
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= 1e-10, (k, v)


def _cfg25(**kw):
    from seidart_hip import models
    return models.elastic25_iso(nx=70, nz=56, steps=200, src=(35, 12), nrec=10, **kw)


FIELDS25 = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")


def _run25(cfg, **options):
    from seidart_hip.solver import make_solver
    s = make_solver(cfg, options=options or None)
    try:
        s.run(cfg["time_steps"])
        return s.receivers(), s.norms()[0], {f: s.field(f) for f in FIELDS25}, s.info("isotropic")
    finally:
        s.close()


def test_isotropic_path_25d_matches_oracle_and_general_path():
    """2.5D elastic: the isotropic arrays against the oracle and against the
    21-array path on the same model."""
    from oracle.oracle_ctypes import run_oracle
    cfg = _cfg25()
    print("[isotropic] 2.5D oracle parity and general path", flush=True)
    a_rec, a_vn, a_f, a_iso = _run25(cfg)
    b_rec, b_vn, b_f, b_iso = _run25(cfg, isotropic=0)
    assert a_iso == 1 and b_iso == 0
    o_rec, (o_vn, _), o_f, us = run_oracle(cfg, fields=FIELDS25)
    assert us == 0 and np.abs(o_rec).max() > 0
    errs = {"receivers": rel_l2(a_rec, o_rec), "velnorm": rel_l2(a_vn, o_vn)}
    errs.update({k: rel_l2(a_f[k], o_f[k]) for k in FIELDS25})
    gen = {"receivers": rel_l2(a_rec, b_rec), "velnorm": rel_l2(a_vn, b_vn)}
    gen.update({k: rel_l2(a_f[k], b_f[k]) for k in FIELDS25})
    print({k: f"{v:.2e}" for k, v in errs.items()}, {k: f"{v:.2e}" for k, v in gen.items()}, flush=True)
    for k, v in errs.items():
        assert v <= 1e-10, (k, v)
    for k, v in gen.items():
        assert v <= 1e-13, (k, v)


def test_isotropic_path_25d_plane_wave_matches_oracle():
    """2.5D isotropic arrays with a plane-wave source (the injection reads the
    stiffness through the isotropic slots)."""
    from oracle.oracle_ctypes import run_oracle
    cfg = _cfg25()
    cfg["plane_wave"] = dict(kind="p", angles=(25.0, 0.0, 10.0), f0=300.0, amplitude=1e-3, wavelet=2)
    print("[isotropic] 2.5D plane wave", flush=True)
    g_rec, g_vn, g_f, iso = _run25(cfg)
    assert iso == 1
    o_rec, (o_vn, _), o_f, us = run_oracle(cfg, fields=FIELDS25)
    assert us == 0 and np.abs(o_rec).max() > 0
    errs = {"receivers": rel_l2(g_rec, o_rec), "velnorm": rel_l2(g_vn, o_vn)}
    errs.update({k: rel_l2(g_f[k], o_f[k]) for k in FIELDS25})
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= 1e-10, (k, v)
