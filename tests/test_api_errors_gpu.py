"""Error paths of the C-ABI through the Python wrapper: plane-wave kinds the
solver does not have, option values out of range, and inputs changed after
the first step (fixed from then on)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small(physics):
    from seidart_hip import models
    from seidart_hip.solver import make_solver
    cfg = {"elastic2": lambda: models.elastic2(nx=30, nz=24, cpml=4, steps=4, src=(15, 10), rec_z=10, nrec=2),
           "elastic3": lambda: models.elastic3(nx=16, ny=14, nz=18, cpml=3, steps=4, src=(8, 7, 6), nrec=2),
           "em2": lambda: models.em2(nx=30, nz=24, cpml=4, steps=4, src=(15, 10), nrec=2),
           "em25": lambda: models.em25(nx=30, nz=24, cpml=4, steps=4, src=(15, 10), nrec=2)}[physics]()
    return cfg, make_solver(cfg)


def test_plane_wave_kinds():
    from seidart_hip._lib import SdtError
    print("[api errors] plane-wave kinds", flush=True)
    _, s = _small("em2")
    try:
        with pytest.raises(SdtError, match="transverse"):
            s.set_plane_wave("p", (0.0, 0.0, 0.0), 1e8, 1.0)
        s.set_plane_wave("s", (0.0, 0.0, 0.0), 1e8, 1.0)       # the EM 'pw' source
    finally:
        s.close()
    _, s = _small("em25")
    try:
        with pytest.raises(SdtError, match="no plane-wave source"):
            s.set_plane_wave("s", (0.0, 0.0, 0.0), 1e8, 1.0)
    finally:
        s.close()
    _, s = _small("elastic2")
    try:
        with pytest.raises(SdtError, match="f0 > 0"):
            s.set_plane_wave("p", (0.0, 0.0, 0.0), 0.0, 1.0)
        with pytest.raises(SdtError, match="finite"):
            s.set_plane_wave("p", (float("nan"), 0.0, 0.0), 50.0, 1.0)
    finally:
        s.close()


def test_isotropic_option_range_and_state():
    from seidart_hip._lib import SdtError
    print("[api errors] isotropic option", flush=True)
    cfg, s = _small("elastic3")
    try:
        with pytest.raises(SdtError, match="0 or 1"):
            s.set_option("isotropic", 2)
        s.set_option("isotropic", 0)
        s.run(1)
        assert s.info("isotropic") == 0
        with pytest.raises(SdtError, match="fixed"):
            s.set_option("isotropic", 1)
        with pytest.raises(SdtError, match="fixed"):
            s.set_plane_wave("s", (0.0, 0.0, 0.0), 50.0, 1.0)
        s.run(cfg["time_steps"] - 1)
        assert np.all(np.isfinite(s.receivers()))
    finally:
        s.close()


def test_source_window_option_range_and_state():
    """SDT_OPT_SOURCE_WINDOW takes 0 or 1 before the first step and is fixed
    after it; a 2D EM FP32 run with a plane-wave source (no point source, so
    no window) still runs."""
    from seidart_hip import models
    from seidart_hip._lib import SdtError
    from seidart_hip.solver import make_solver
    print("[api errors] source window option", flush=True)
    cfg = models.em2(nx=30, nz=24, cpml=4, steps=4, src=(15, 10), nrec=2)
    s = make_solver(cfg, precision="fp32")
    try:
        with pytest.raises(SdtError, match="0 or 1"):
            s.set_option("source_window", 2)
        s.set_option("source_window", 0)
        s.run(1)
        with pytest.raises(SdtError, match="fixed"):
            s.set_option("source_window", 1)
    finally:
        s.close()
    s = make_solver(cfg, precision="fp32")
    try:
        s.set_plane_wave("s", (0.0, 0.0, 0.0), 1e8, 1.0)
        s.run()
        assert np.all(np.isfinite(s.receivers()))
    finally:
        s.close()


def test_adjoint_takes_point_sources():
    from seidart_hip._lib import SdtError
    print("[api errors] adjoint with a plane wave", flush=True)
    _, s = _small("elastic2")
    try:
        s.set_option("checkpoint", 2)
        s.set_plane_wave("p", (0.0, 0.0, 0.0), 300.0, 1e-3)
        with pytest.raises(SdtError, match="point sources"):
            s.run(1)
    finally:
        s.close()
