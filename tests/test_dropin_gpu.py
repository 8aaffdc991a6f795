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


def test_cli_seismic25_zstd_blocks(tmp_path, monkeypatch):
    """2.5D through the CLI with the reference's environment (KY,
    DENSITY_METHOD, FDTD_STEPS_PER_BLOCK): zstd blocks of the CPML-stripped
    real parts, named S25.<first>-<last>.<xind>.<yind>.<zind>.blk.zst."""
    from seidart_hip import cli, fortran_io as fio, models
    cfg = models.elastic25(nx=48, nz=40, cpml=6, steps=40, src=(24, 12), nrec=4, ky=0.15,
                           density_method="geometric")
    models.write_inputs(cfg, str(tmp_path))
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("KY", "0.15")
    monkeypatch.setenv("DENSITY_METHOD", "geometric")
    monkeypatch.setenv("FDTD_STEPS_PER_BLOCK", "16")
    monkeypatch.delenv("FDTD_LEGACY_OUTPUT", raising=False)
    print("[dropin] cli seismic25 blocks", flush=True)
    assert cli.main(["input.json", "seismic=true"]) == 0
    names = [fio.block_name("S25", a, b, 24, 1, 12) for a, b in ((1, 16), (17, 32), (33, 40))]
    for n in names:
        assert (tmp_path / n).exists(), n
    _, _, f = _oracle(cfg, ("vx", "vy", "vz"))
    onx, onz = 48, 40
    raw = fio.zstd_decompress((tmp_path / names[2]).read_bytes(), onx * onz * 3 * 8 * 4)
    blk = np.frombuffer(raw, dtype="<f4").reshape((onx, 1, onz, 3, 8), order="F")
    for q, key in enumerate(("vx", "vy", "vz")):
        want = f[key].real[6:-6, 6:-6].astype(np.float32)
        assert rel_l2(blk[:, 0, :, q, 7], want) < 1e-6, key


def test_dropin_seismic25_legacy_images(tmp_path, monkeypatch):
    from seidart_hip import cpmlfdtd, fortran_io as fio, models
    cfg = models.elastic25(nx=40, nz=36, cpml=6, steps=20, src=(20, 10), nrec=4, ky=0.0,
                           density_method="harmonic")
    s = cfg["sources"]["z"]
    cfg["sources"] = {"xx": s, "xy": 0.4 * s, "xz": -0.3 * s, "yy": 0.8 * s, "yz": 0.2 * s, "zz": -s}
    models.write_inputs(cfg, str(tmp_path), source_type="tensor")
    sh = cfg["shape"]
    dom = cpmlfdtd.Domain(dim=2.5, nx=sh[0], ny=1, nz=sh[1], dx=1.0, dy=1.0, dz=1.0, cpml=6)
    src = cpmlfdtd.Source(dt=cfg["dt"], time_steps=20, xind=20, yind=1, zind=10, source_type="tensor")
    monkeypatch.delenv("KY", raising=False)
    monkeypatch.delenv("DENSITY_METHOD", raising=False)
    monkeypatch.setenv("FDTD_LEGACY_OUTPUT", "TRUE")
    monkeypatch.setenv("SINGLE_OUTPUT", ".FALSE.")
    print("[dropin] seismic25 legacy images", flush=True)
    cpmlfdtd.seismic25(dom, src, workdir=str(tmp_path))
    _, _, f = _oracle(cfg, ("vx", "vz"))
    assert np.abs(f["vz"]).max() > 0
    for ch, key in (("Vx", "vx"), ("Vz", "vz")):
        got = fio.read_array(str(tmp_path / fio.snapshot_name(ch, 20, 20, 10)), sh)
        assert rel_l2(got, f[key].real) < 1e-10


def test_cli_electromag25_blocks(tmp_path, monkeypatch):
    """2.5D EM through the CLI: EM25.<first>-<last>.<x>.<y>.<z>.blk.zst blocks."""
    from seidart_hip import cli, fortran_io as fio, models
    cfg = models.em25(nx=40, nz=36, cpml=6, steps=24, src=(20, 12), nrec=4, ky=8.0)
    models.write_inputs(cfg, str(tmp_path))
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("KY", "8.0")
    monkeypatch.setenv("FDTD_STEPS_PER_BLOCK", "10")
    monkeypatch.delenv("FDTD_LEGACY_OUTPUT", raising=False)
    print("[dropin] cli electromag25 blocks", flush=True)
    assert cli.main(["input.json", "seismic=false"]) == 0
    names = [fio.block_name("EM25", a, b, 20, 1, 12) for a, b in ((1, 10), (11, 20), (21, 24))]
    for n in names:
        assert (tmp_path / n).exists(), n
    _, _, f = _oracle(cfg, ("ex", "ey", "ez"))
    raw = fio.zstd_decompress((tmp_path / names[2]).read_bytes(), 40 * 36 * 3 * 4 * 4)
    blk = np.frombuffer(raw, dtype="<f4").reshape((40, 1, 36, 3, 4), order="F")
    for q, key in enumerate(("ex", "ey", "ez")):
        want = f[key].real[6:-6, 6:-6].astype(np.float32)
        assert rel_l2(blk[:, 0, :, q, 3], want) < 1e-6, key


@pytest.mark.parametrize("physics,source_type", [("elastic2", "pw_p"), ("em2", "pw"), ("elastic3", " pws "),
                                                 ("em3", "pw")])
def test_cli_plane_wave(tmp_path, monkeypatch, physics, source_type):
    """Plane-wave sources from input.json (rotations, source_frequency,
    amplitude, source_wavelet) through the CLI, snapshots against the
    oracle's plane wave."""
    from seidart_hip import cli, fortran_io as fio, models
    if physics == "elastic2":
        cfg = models.elastic2(nx=50, nz=40, cpml=6, steps=60, src=(25, 10), rec_z=10, nrec=4)
        pw = dict(kind="p", angles=(30.0, 0.0, 0.0), f0=400.0, amplitude=1e-3, wavelet=1)
        chans = (("Vx", "vx"), ("Vz", "vz"))
    elif physics == "elastic3":
        cfg = models.elastic3(nx=20, ny=18, nz=22, cpml=4, steps=40, src=(10, 9, 6), nrec=4)
        pw = dict(kind="s", angles=(20.0, 30.0, 10.0), f0=2500.0, amplitude=1e-3, wavelet=2)
        chans = (("Vx", "vx"), ("Vy", "vy"), ("Vz", "vz"))
    elif physics == "em2":
        cfg = models.em2(nx=50, nz=40, steps=60, src=(25, 12), nrec=4)
        pw = dict(kind="s", angles=(-40.0, 0.0, 0.0), f0=300e6, amplitude=1.0, wavelet=0)
        chans = (("Ex", "ex"), ("Ez", "ez"))
    else:
        cfg = models.em3(nx=20, ny=18, nz=22, cpml=4, steps=50, src=(10, 9, 6), nrec=4)
        pw = dict(kind="s", angles=(30.0, 20.0, 10.0), f0=400e6, amplitude=1.0, wavelet=2)
        chans = (("Ex", "ex"), ("Ey", "ey"), ("Ez", "ez"))
    cfg["plane_wave"] = pw
    models.write_inputs(cfg, str(tmp_path), source_type=source_type)
    monkeypatch.chdir(tmp_path)
    print(f"[dropin] cli plane wave {physics} {source_type!r}", flush=True)
    seismic = "true" if physics.startswith("elastic") else "false"
    assert cli.main(["input.json", f"seismic={seismic}", "density_method=arithmetic"]) == 0
    _, _, f = _oracle(cfg, tuple(k for _, k in chans))
    src = cfg["src"]
    three = physics.endswith("3")
    for ch, key in chans:
        name = fio.snapshot_name(ch, cfg["time_steps"], src[0], src[-1], src[1] if three else None)
        got = fio.read_array(str(tmp_path / name), cfg["shape"], np.float32)
        assert np.abs(f[key]).max() > 0
        assert rel_l2(got, f[key].astype(np.float32)) < 1e-6, key
