"""Device-packed snapshots (sdt_snapshot) against the blocking FP64 field
read-back (sdt_get_field): FP64 snapshots equal the field, FP32 snapshots
equal its FP32 rounding, for every real-valued solver, both precisions, the
lazily zeroed H faces of the EM solvers and a snapshot before any step."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cfgs():
    from seidart_hip import models
    return {
        "elastic2": (models.elastic2(nx=70, nz=50, cpml=6, steps=40, src=(30, 20), nrec=4), ("vx", "vz", "sxz")),
        "elastic3": (models.elastic3(nx=30, ny=26, nz=22, cpml=5, steps=30, src=(15, 13, 9), nrec=4),
                     ("vx", "vy", "vz", "sxx", "syz")),
        "em2": (models.em2(nx=60, nz=44, steps=40, src=(30, 15), nrec=4), ("ex", "ez", "hy")),
        "em3": (models.em3(nx=30, ny=28, nz=24, cpml=5, steps=30, src=(15, 14, 10), nrec=4),
                ("ex", "ey", "ez", "hx", "hy", "hz")),
    }


@pytest.mark.parametrize("physics", ["elastic2", "elastic3", "em2", "em3"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_snapshot_equals_field(physics, precision):
    from seidart_hip.solver import PinnedArray, make_solver
    cfg, names = _cfgs()[physics]
    print(f"[snapshot] {physics} {precision} {cfg['shape']}", flush=True)
    s = make_solver(cfg, precision=precision)
    try:
        b32 = PinnedArray(s.ncell, np.float32)
        b64 = PinnedArray(s.ncell, np.float64)
        for steps in (0, cfg["time_steps"]):
            s.run(steps)
            for name in names:
                ref = s.field(name).ravel(order="F")
                s.snapshot(name, b64.array, single=False)
                s.snapshot(name, b32.array, single=True)
                s.synchronize()
                assert np.array_equal(b64.array, ref), (name, steps)
                assert np.array_equal(b32.array, ref.astype(np.float32)), (name, steps)
        assert np.any(b64.array != 0)
    finally:
        s.close()


def test_snapshot_rejects_bad_buffers():
    from seidart_hip import SdtError, models
    from seidart_hip.solver import make_solver
    cfg = models.elastic25(nx=30, nz=24, cpml=4, steps=5, src=(15, 8), nrec=2)
    s = make_solver(cfg)
    try:
        with pytest.raises(SdtError):
            s.snapshot("vx", np.zeros(s.ncell, np.float32))   # complex fields: no snapshots
    finally:
        s.close()
    cfg = models.elastic2(nx=30, nz=24, cpml=4, steps=5, src=(15, 8), rec_z=8, nrec=2)
    s = make_solver(cfg)
    try:
        with pytest.raises(ValueError):
            s.snapshot("vx", np.zeros(s.ncell - 1, np.float32))
        with pytest.raises(ValueError):
            s.snapshot("vx", np.zeros(s.ncell, np.float64), single=True)
    finally:
        s.close()
