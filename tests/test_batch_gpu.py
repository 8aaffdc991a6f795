"""Batched shots (2D elastic, 2D EM) through sdt_batch_*: every shot of a
batch must give what the same handle gives when run alone, and one shot is
also held against the oracle."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _shots(physics):
    from seidart_hip import models
    if physics == "elastic2":
        return [models.elastic2(nx=90, nz=70, steps=260, seed=11, src=(45, 20), nrec=8),
                models.elastic2(nx=131, nz=57, steps=300, seed=12, src=(60, 15), nrec=5, gamma=2.0),
                models.elastic2(nx=64, nz=64, steps=90, seed=13, src=(30, 30), nrec=4),
                models.elastic2(nx=47, nz=153, steps=300, seed=14, src=(20, 70), nrec=6, nlayers=3),
                models.elastic2(nx=200, nz=33, steps=233, seed=15, src=(100, 10), nrec=9,
                                density_method="harmonic")]
    return [models.em2(nx=120, nz=80, steps=300, seed=21, src=(60, 20), nrec=8),
            models.em2(nx=75, nz=101, steps=200, seed=22, src=(30, 40), nrec=5),
            models.em2(nx=66, nz=40, steps=77, seed=23, src=(33, 10), nrec=4, nlayers=3),
            models.em2(nx=140, nz=61, steps=300, seed=24, src=(70, 25), nrec=7)]


def _outputs(s):
    names = ("vx", "vz", "sxz") if s.physics == "elastic2" else ("ex", "ez", "hy")
    return s.receivers(), s.norms()[0], {f: s.field(f) for f in names}


@pytest.mark.parametrize("physics", ["elastic2", "em2"])
def test_batch_mixes_material_layouts(physics):
    """Shots with the material table and with per-cell arrays in one batch:
    each equals its own run alone to FP32 rounding (a batch holding a table
    shot runs the table-capable kernel for every shot, whose per-cell-array
    branch the compiler may contract differently from the table-free kernel
    a lone per-cell-array shot runs; a wrong layout would differ by O(1))."""
    from seidart_hip.solver import Batch, make_solver
    cfgs = _shots(physics)
    opts = [{"material_table": q % 2} for q in range(len(cfgs))]
    solo = []
    for c, o in zip(cfgs, opts):
        s = make_solver(c, precision="fp32", options=o)
        try:
            s.run()
            solo.append(_outputs(s) + (s.materials(),))
        finally:
            s.close()
    assert [m > 0 for *_, m in solo] == [bool(o["material_table"]) for o in opts]
    shots = [make_solver(c, precision="fp32", options=o) for c, o in zip(cfgs, opts)]
    try:
        with Batch(shots) as b:
            b.run()
        for q, s in enumerate(shots):
            rec, vn, f = _outputs(s)
            r0, v0, f0, _ = solo[q]
            assert rel_l2(rec, r0) <= 1e-5 and rel_l2(vn, v0) <= 1e-5, q
            for k in f:
                assert rel_l2(f[k], f0[k]) <= 1e-5, (q, k)
    finally:
        for s in shots:
            s.close()


@pytest.mark.parametrize("physics,precision", [("elastic2", "fp64"), ("elastic2", "fp32"),
                                                ("em2", "fp64"), ("em2", "fp32")])
def test_batch_equals_single_runs(physics, precision):
    from seidart_hip.solver import Batch, make_solver
    cfgs = _shots(physics)
    print(f"[batch] {physics} {precision} shapes={[c['shape'] for c in cfgs]}", flush=True)
    solo = []
    for c in cfgs:
        s = make_solver(c, precision=precision)
        try:
            s.run()
            solo.append(_outputs(s))
        finally:
            s.close()
    shots = [make_solver(c, precision=precision) for c in cfgs]
    try:
        with Batch(shots) as b:
            b.run(50)       # tuning steps, then a partial chunk
            assert all(s.steps_done == min(50, c["time_steps"]) for s, c in zip(shots, cfgs))
            b.run()         # graph replays, then the remainder; short shots stop early
            rows = b.info("tile_rows")
            assert rows in (4, 16, 32)
            assert b.info("xcd_order") in (0, 1)
            print(f"[batch] tile rows {rows}, tiles {b.info('tiles')}", flush=True)
        for q, (s, c) in enumerate(zip(shots, cfgs)):
            assert s.steps_done == c["time_steps"]
            rec, vn, f = _outputs(s)
            r0, v0, f0 = solo[q]
            errs = {"receivers": rel_l2(rec, r0), "velnorm": rel_l2(vn, v0)}
            errs.update({k: rel_l2(f[k], f0[k]) for k in f0})
            print(q, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
            # same per-cell arithmetic as the single-shot kernels
            for k, v in errs.items():
                assert v <= 1e-13, (q, k, v)
    finally:
        for s in shots:
            s.close()


def test_batch_shot_matches_oracle():
    from oracle.oracle_ctypes import run_oracle
    from seidart_hip.solver import Batch, make_solver
    cfgs = _shots("elastic2")[:3]
    shots = [make_solver(c) for c in cfgs]
    try:
        with Batch(shots) as b:
            b.run()
        rec = shots[1].receivers()
        f = shots[1].field("vz")
    finally:
        for s in shots:
            s.close()
    o_rec, _, o_f, us = run_oracle(cfgs[1], fields=("vz",))
    assert us == 0
    assert rel_l2(rec, o_rec) <= 1e-10
    assert rel_l2(f, o_f["vz"]) <= 1e-10


def test_batch_reports_unstable_shot():
    from seidart_hip import UnstableError
    from seidart_hip.solver import Batch, make_solver
    cfgs = _shots("em2")[:3]
    ic = np.zeros(cfgs[1]["shape"], order="F")
    ic[30, 30] = 1e26
    cfgs[1]["arrays"]["ic_x"] = ic
    shots = [make_solver(c) for c in cfgs]
    try:
        with Batch(shots) as b:
            with pytest.raises(UnstableError) as e:
                b.run(40)
            assert e.value.shot == 1 and e.value.step == 1
        # the other shots advanced normally
        assert shots[0].steps_done == 40 and shots[2].steps_done == 40
    finally:
        for s in shots:
            s.close()


def test_batch_rejects_mixed_kinds():
    from seidart_hip import SdtError, models
    from seidart_hip.solver import Batch, make_solver
    a = make_solver(models.elastic2(nx=40, nz=40, steps=20, src=(20, 20), nrec=2))
    c = make_solver(models.em2(nx=40, nz=40, steps=20, src=(20, 20), nrec=2))
    try:
        with pytest.raises(SdtError):
            Batch([a, c])
    finally:
        a.close()
        c.close()


def test_batch_owns_its_handles_until_destroyed():
    """While a batch lives its handles cannot be run, adjointed or destroyed
    on their own, nor join a second batch; destroying the batch releases
    them."""
    import ctypes as C
    from seidart_hip import SdtError, models
    from seidart_hip import _lib as L
    from seidart_hip.solver import Batch, make_solver
    a = make_solver(models.elastic2(nx=40, nz=40, steps=20, src=(20, 20), nrec=2))
    c = make_solver(models.elastic2(nx=44, nz=40, steps=20, src=(20, 20), nrec=2))
    b = Batch([a, c])
    try:
        assert L.lib.sdt_run(a._h, 1, None) == L.ERR_STATE
        assert L.lib.sdt_destroy(c._h) == L.ERR_STATE
        assert "belongs to a batch" in L.lib.sdt_last_error(c._h).decode()
        with pytest.raises(SdtError):
            a.close()
        assert a._h   # still alive
        with pytest.raises(SdtError):
            Batch([c])
        b.run(5)
    finally:
        b.close()
    a.run(3)   # released: runs on its own again
    assert a.steps_done == 8
    a.close()
    c.close()
    assert not a._h and not c._h
