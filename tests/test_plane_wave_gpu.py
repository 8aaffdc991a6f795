"""Plane-wave sources (sdt_set_plane_wave) against the oracle: 2D, 2.5D and
3D elastic (S and P polarisation) and 2D, 3D EM, oblique directions entering
through one, two or three faces, and faces on the grid edge (no absorbing
layer); FP64 bar 1e-10 on receivers, norms and fields."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL64 = 1e-10


def _check(cfg, fields, options=None, stress_norm=False):
    from oracle.oracle_ctypes import run_oracle
    from seidart_hip.solver import run_config
    print(f"[plane wave] {cfg['physics']} {cfg['plane_wave']} {options or ''}", flush=True)
    g_rec, (g_vn, g_sn), g_f = run_config(cfg, fields=fields, options=options)
    o_rec, (o_vn, o_sn), o_f, us = run_oracle(cfg, fields=fields)
    assert us == 0
    assert np.abs(o_rec).max() > 0 and np.abs(o_vn).max() > 0
    errs = {"receivers": rel_l2(g_rec, o_rec), "velnorm": rel_l2(g_vn, o_vn)}
    if stress_norm:
        errs["stressnorm"] = rel_l2(g_sn, o_sn)
    errs.update({k: rel_l2(g_f[k], o_f[k]) for k in fields})
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= TOL64, (k, v)


@pytest.mark.parametrize("kind,angles,wavelet", [("s", (0.0, 0.0, 0.0), 2), ("p", (35.0, 0.0, 0.0), 1),
                                                 ("s", (-120.0, 10.0, 0.0), 0)])
def test_elastic2_plane_wave(kind, angles, wavelet):
    from seidart_hip import models
    cfg = models.elastic2(nx=70, nz=56, steps=220, src=(35, 20), nrec=10, nlayers=2, tilt=(0.0, 20.0))
    cfg["plane_wave"] = dict(kind=kind, angles=angles, f0=300.0, amplitude=1e-3, wavelet=wavelet)
    _check(cfg, ("vx", "vz", "sxx", "szz", "sxz"))


@pytest.mark.parametrize("kind,angles", [("p", (25.0, 0.0, 0.0)), ("s", (200.0, 0.0, 15.0))])
def test_elastic25_plane_wave(kind, angles):
    from seidart_hip import models
    cfg = models.elastic25(nx=60, nz=50, steps=180, src=(30, 20), nrec=8)
    cfg["plane_wave"] = dict(kind=kind, angles=angles, f0=300.0, amplitude=1e-3, wavelet=2)
    _check(cfg, ("vx", "vy", "vz", "sxz"))


@pytest.mark.parametrize("kind,angles,options", [("s", (30.0, 20.0, 10.0), None),
                                                 ("p", (-20.0, 40.0, 25.0), {"material_table": 0}),
                                                 ("p", (0.0, 0.0, 0.0), None)])
def test_elastic3_plane_wave(kind, angles, options):
    from seidart_hip import models
    cfg = models.elastic3(nx=28, ny=26, nz=30, cpml=6, steps=120, src=(14, 13, 10), nrec=6)
    cfg["plane_wave"] = dict(kind=kind, angles=angles, f0=1500.0, amplitude=1e-3, wavelet=2)
    _check(cfg, ("vx", "vy", "vz", "sxx", "syz"), options=options, stress_norm=True)


@pytest.mark.parametrize("angles,wavelet", [((0.0, 0.0, 0.0), 2), ((35.0, 0.0, 0.0), 1), ((-120.0, 10.0, 0.0), 0)])
def test_em2_plane_wave(angles, wavelet):
    from seidart_hip import models
    cfg = models.em2(nx=70, nz=56, steps=200, src=(35, 20), nrec=10, nlayers=3)
    cfg["plane_wave"] = dict(kind="s", angles=angles, f0=200e6, amplitude=1.0, wavelet=wavelet)
    _check(cfg, ("ex", "ez", "hy"))


@pytest.mark.parametrize("angles,options", [((30.0, 20.0, 10.0), None), ((-20.0, 40.0, 25.0), {"material_table": 0}),
                                            ((180.0, 0.0, 0.0), None)])
def test_em3_plane_wave(angles, options):
    from seidart_hip import models
    cfg = models.em3(nx=28, ny=26, nz=30, cpml=6, steps=150, src=(14, 13, 10), nrec=6, nlayers=3)
    cfg["plane_wave"] = dict(kind="s", angles=angles, f0=300e6, amplitude=1.0, wavelet=2)
    _check(cfg, ("ex", "ey", "ez", "hx", "hz"), options=options)


def test_elastic2_plane_wave_face_on_grid_edge():
    """No absorbing layer: the faces are the grid's edge rows, whose
    velocities the Dirichlet step zeroes after the injection; the injected
    stresses stay and drive the wave."""
    from seidart_hip import models
    cfg = models.elastic2(nx=48, nz=40, cpml=0, steps=120, src=(24, 20), rec_z=4, nrec=8, nlayers=2)
    cfg["plane_wave"] = dict(kind="p", angles=(30.0, 0.0, 0.0), f0=300.0, amplitude=1e-3, wavelet=2)
    _check(cfg, ("vx", "vz", "sxx", "szz", "sxz"))


def test_em2_plane_wave_face_on_grid_edge():
    """No absorbing layer: both faces are grid edges, where the Dirichlet
    step zeroes E and H after the injection, so nothing is ever emitted (as
    in the oracle)."""
    from oracle.oracle_ctypes import run_oracle
    from seidart_hip import models
    from seidart_hip.solver import run_config
    cfg = models.em2(nx=48, nz=40, cpml=0, steps=40, src=(24, 20), nrec=8, nlayers=2)
    cfg["plane_wave"] = dict(kind="s", angles=(30.0, 0.0, 0.0), f0=200e6, amplitude=1.0, wavelet=2)
    fields = ("ex", "ez", "hy")
    g_rec, (g_vn, _), g_f = run_config(cfg, fields=fields)
    o_rec, (o_vn, _), o_f, us = run_oracle(cfg, fields=fields)
    assert us == 0
    for a in (o_rec, o_vn, *o_f.values()):
        assert not np.any(a)
    for a in (g_rec, g_vn, *g_f.values()):
        assert not np.any(a)
