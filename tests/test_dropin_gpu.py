"""The file-based drop-in (seidart_hip.cpmlfdtd) and the seidartfdtd CLI on
the GPU: the reference's input files in, the reference's snapshot and norm
files out, checked against the CPU oracle."""
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _oracle(cfg, fields):
    from oracle.oracle_ctypes import run_oracle
    rec, norms, f, us = run_oracle(cfg, fields=fields)
    assert us == 0
    return rec, norms, f


def test_cli_seismic2_snapshots(tmp_path, monkeypatch):
    from seidart_hip import cli, fortran_io as fio, models
    cfg = models.elastic2(nx=60, nz=50, cpml=8, steps=40, src=(30, 15), rec_z=15, nrec=4)
    models.write_inputs(cfg, str(tmp_path))
    monkeypatch.chdir(tmp_path)
    print("[dropin] cli seismic2", flush=True)
    assert cli.main(["input.json", "seismic=true", "density_method=arithmetic"]) == 0
    _, _, f = _oracle(cfg, ("vx", "vz"))
    for ch, key in (("Vx", "vx"), ("Vz", "vz")):
        p = tmp_path / fio.snapshot_name(ch, 40, 30, 15)
        assert p.exists()
        got = fio.read_array(str(p), cfg["shape"], np.float32)
        assert rel_l2(got, f[key].astype(np.float32)) < 1e-6
    assert (tmp_path / fio.snapshot_name("Vx", 1, 30, 15)).exists()


def test_dropin_seismic3_double_snapshots_and_norms(tmp_path):
    from seidart_hip import cpmlfdtd, fortran_io as fio, models
    cfg = models.elastic3(nx=24, ny=22, nz=20, cpml=4, steps=30, src=(12, 11, 6), nrec=4)
    models.write_inputs(cfg, str(tmp_path))
    sh = cfg["shape"]
    dom = cpmlfdtd.Domain(dim=3.0, nx=sh[0], ny=sh[1], nz=sh[2], dx=0.5, dy=0.5, dz=0.5, cpml=4)
    src = cpmlfdtd.Source(dt=cfg["dt"], time_steps=30, xind=12, yind=11, zind=6, source_type="ac")
    print("[dropin] seismic3 fp64 snapshots", flush=True)
    cpmlfdtd.seismic3(dom, src, "arithmetic", False, workdir=str(tmp_path))
    _, (vn, sn), f = _oracle(cfg, ("vx", "vy", "vz"))
    for ch, key in (("Vx", "vx"), ("Vy", "vy"), ("Vz", "vz")):
        got = fio.read_array(str(tmp_path / fio.snapshot_name(ch, 30, 12, 6, yind=11)), sh)
        assert rel_l2(got, f[key]) < 1e-10
    np.testing.assert_allclose(fio.read_record(str(tmp_path / "velocity_norm.dat"), np.float64, 30), vn,
                               rtol=1e-10, atol=0)
    np.testing.assert_allclose(fio.read_record(str(tmp_path / "stress_norm.dat"), np.float64, 30), sn,
                               rtol=1e-10, atol=0)


def test_cli_electromag2(tmp_path, monkeypatch):
    from seidart_hip import cli, fortran_io as fio, models
    cfg = models.em2(nx=60, nz=40, steps=50, src=(30, 12), nrec=4)
    models.write_inputs(cfg, str(tmp_path))
    monkeypatch.chdir(tmp_path)
    print("[dropin] cli electromag2", flush=True)
    assert cli.main(["input.json", "seismic=false"]) == 0
    _, _, f = _oracle(cfg, ("ex", "ez"))
    for ch, key in (("Ex", "ex"), ("Ez", "ez")):
        got = fio.read_array(str(tmp_path / fio.snapshot_name(ch, 50, 30, 12)), cfg["shape"], np.float32)
        assert rel_l2(got, f[key].astype(np.float32)) < 1e-6


def test_cli_reports_blowup(tmp_path, monkeypatch):
    from seidart_hip import cli, fortran_io as fio, models
    cfg = models.elastic2(nx=40, nz=40, cpml=6, steps=10, src=(20, 10), rec_z=10, nrec=2)
    ic = np.zeros(cfg["shape"], order="F")
    ic[20, 20] = 1e26
    cfg["arrays"]["ic_x"] = ic
    models.write_inputs(cfg, str(tmp_path))
    monkeypatch.chdir(tmp_path)
    print("[dropin] forced blow-up", flush=True)
    assert cli.main(["input.json"]) == 1
    assert not os.path.exists(tmp_path / fio.snapshot_name("Vx", 1, 20, 10))
