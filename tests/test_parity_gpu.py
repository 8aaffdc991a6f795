"""GPU parity: libseidart_hip.so (through its C-ABI) against the CPU oracle on
the same seeded inputs.  FP64 bar: relative L2 <= 1e-10 on receiver traces,
per-step norms and final fields."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def _run_both(cfg, fields, steps=None, precision="fp64", options=None):
    from seidart_hip.solver import run_config
    from oracle.oracle_ctypes import run_oracle
    print(f"[parity] {cfg['physics']} shape={cfg['shape']} steps={steps or cfg['time_steps']} "
          f"{precision} {options or ''}", flush=True)
    g_rec, (g_vn, g_sn), g_f = run_config(cfg, precision=precision, steps=steps, fields=fields,
                                          options=options)
    o_rec, (o_vn, o_sn), o_f, us = run_oracle(cfg, steps=steps, fields=fields)
    assert us == 0
    return (g_rec, g_vn, g_sn, g_f), (o_rec, o_vn, o_sn, o_f)


def _check(g, o, tol, stress_norm=False):
    g_rec, g_vn, g_sn, g_f = g
    o_rec, o_vn, o_sn, o_f = o
    errs = {"receivers": rel_l2(g_rec, o_rec), "velnorm": rel_l2(g_vn, o_vn)}
    if stress_norm:
        errs["stressnorm"] = rel_l2(g_sn, o_sn)
    for k in o_f:
        errs[k] = rel_l2(g_f[k], o_f[k])
    print({k: f"{v:.3e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= tol, f"{k}: rel L2 {v:.3e} > {tol:.0e}"
    return errs


@pytest.mark.parametrize("options", [{}, {"material_table": 0}])
def test_elastic2_config0_fp64(options):
    """2D elastic, 200x200 (+10 CPML), 1000 steps, two tilted ice layers;
    stiffness from the material table (default) and from per-cell arrays."""
    from seidart_hip import models
    cfg = models.elastic2()
    g, o = _run_both(cfg, ("vx", "vz", "sxx", "szz", "sxz"), options=options)
    _check(g, o, TOL64)


def test_elastic2_material_table_equals_per_cell_arrays():
    """Table and per-cell arrays hold the same values: the runs agree to
    rounding; the table is used without an adjoint and dropped with one."""
    from seidart_hip import models
    from seidart_hip.solver import make_solver
    cfg = models.elastic2(nx=90, nz=70, cpml=8, steps=300, src=(45, 20), rec_z=20, nrec=8, gamma=2.0,
                          density_method="geometric")
    out = []
    for opts in ({"material_table": 1}, {"material_table": 0}, {"material_table": 1, "checkpoint": 50}):
        s = make_solver(cfg, options=opts)
        try:
            s.run(cfg["time_steps"])
            out.append((s.receivers(), s.norms()[0], {f: s.field(f) for f in ("vx", "vz", "sxx", "sxz")},
                        s.materials()))
        finally:
            s.close()
    (r1, n1, f1, m1), (r0, n0, f0, m0), (r2, n2, f2, m2) = out
    assert m1 > 1 and m0 == 0 and m2 == 0
    assert rel_l2(r1, r0) < 1e-13 and rel_l2(n1, n0) < 1e-13
    assert np.array_equal(r2, r0) and np.array_equal(n2, n0)
    for k in f0:
        assert rel_l2(f1[k], f0[k]) < 1e-13, k


@pytest.mark.parametrize("density", ["harmonic", "geometric", "none"])
def test_elastic2_density_methods(density):
    from seidart_hip import models
    cfg = models.elastic2(nx=80, nz=70, cpml=8, steps=300, src=(40, 20), rec_z=20, nrec=8,
                          density_method=density)
    g, o = _run_both(cfg, ("vx", "vz", "sxz"))
    _check(g, o, TOL64)


def test_elastic2_stress_source_and_attenuation():
    from seidart_hip import models
    cfg = models.elastic2(nx=80, nz=70, cpml=8, steps=300, src=(40, 20), rec_z=20, nrec=8,
                          gamma=50.0)
    s = cfg["sources"]["z"]
    cfg["sources"] = {"xx": s, "xz": 0.5 * s, "zz": -0.7 * s}
    g, o = _run_both(cfg, ("vx", "vz", "sxx", "szz", "sxz"))
    _check(g, o, TOL64)


@pytest.mark.parametrize("options", [{}, {"zchunk": 5}, {"zchunk": 1}, {"zchunk": 7, "stagger": 0},
                                     {"material_table": 0}, {"material_table": 0, "zchunk": 3},
                                     {"concurrent": 1}, {"concurrent": 1, "graph": 0}, {"graph": 0}])
def test_elastic3_small_fp64(options):
    """z-marching chunk heights that put chunk seams inside and outside the
    CPML slabs and the interior box; the shell kernels on a second stream
    beside the interior kernels (concurrent: the shell then stores the column
    next to the box's last one itself)."""
    from seidart_hip import models
    cfg = models.elastic3(nx=40, ny=36, nz=32, cpml=6, steps=200, src=(20, 18, 10), nrec=16)
    fields = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")
    g, o = _run_both(cfg, fields, options=options)
    _check(g, o, TOL64, stress_norm=True)


@pytest.mark.parametrize("ny,cpml", [(2, 5), (3, 5), (1, 5)])
def test_elastic3_thin_interior_box(ny, cpml):
    """Interior boxes one row (ny 2) and two rows (ny 3) tall: the last 4-row
    y-strip tile below the box reaches past it into rows the strip above owns,
    and every CPML-shell cell must still be updated by exactly one tile (a
    cell updated by two tiles gets its stress increment twice, depending on
    their timing).  ny 1 leaves no CPML-free centre: every cell takes the
    shell kernels."""
    import numpy as np
    from seidart_hip import models
    from seidart_hip.solver import make_solver
    from oracle.oracle_ctypes import run_oracle
    cfg = models.elastic3(nx=20, ny=ny, nz=18, cpml=cpml, steps=30, src=(10, 1, 6), nrec=4)
    fields = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")
    rng = np.random.default_rng(ny * 10 + cpml)
    # every cell active from the first step: velocities O(1), stresses O(rho c v)
    ic = {f: rng.uniform(-1.0, 1.0, cfg["shape"]) * (1.0 if f.startswith("v") else 4e6) for f in fields}
    print(f"[parity] elastic3 thin box shape={cfg['shape']}", flush=True)
    s = make_solver(cfg)
    try:
        for k, a in ic.items():
            s.set_field(k, a)
        s.run()
        g = (s.receivers(), *s.norms(), {f: s.field(f) for f in fields})
    finally:
        s.close()
    o_rec, (o_vn, o_sn), o_f, us = run_oracle(cfg, fields=fields, field_in=ic)
    assert us == 0
    _check(g, (o_rec, o_vn, o_sn, o_f), TOL64, stress_norm=True)


def test_elastic3_cpml_wider_than_a_wavefront():
    """A 66-cell CPML on every side (144^3 with a 12^3 interior box): the
    x-strip tiles number a 66-cell strip's cells row by row, so a strip row
    spans two wavefronts and a wavefront holds the end of one row and the
    start of the next; the z slabs and y strips take linear tiles of whole
    rows.  Seeded random fields on every cell, 4 steps, every field against
    the oracle."""
    import numpy as np
    from seidart_hip import models
    from seidart_hip.solver import make_solver
    from oracle.oracle_ctypes import run_oracle
    cfg = models.elastic3(nx=12, ny=12, nz=12, cpml=66, steps=4, src=(6, 6, 6), nrec=4, nlayers=2)
    fields = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")
    rng = np.random.default_rng(66)
    ic = {f: rng.uniform(-1.0, 1.0, cfg["shape"]) * (1.0 if f.startswith("v") else 4e6) for f in fields}
    print(f"[parity] elastic3 wide CPML shape={cfg['shape']}", flush=True)
    s = make_solver(cfg)
    try:
        for k, a in ic.items():
            s.set_field(k, a)
        s.run()
        g = (s.receivers(), *s.norms(), {f: s.field(f) for f in fields})
    finally:
        s.close()
    o_rec, (o_vn, o_sn), o_f, us = run_oracle(cfg, fields=fields, field_in=ic)
    assert us == 0
    _check(g, (o_rec, o_vn, o_sn, o_f), TOL64, stress_norm=True)


def test_elastic3_corrected_stencil():
    from seidart_hip import models
    cfg = models.elastic3(nx=30, ny=26, nz=28, cpml=5, steps=120, src=(15, 13, 9), nrec=8)
    cfg["stencil"] = "corrected"
    fields = ("vx", "vy", "vz", "sxx", "syz")
    g, o = _run_both(cfg, fields)
    _check(g, o, TOL64, stress_norm=True)


@pytest.mark.parametrize("remap", [1, 2, 3])
def test_elastic3_block_orders_agree(remap):
    """The XCD-contiguous block orders are permutations of the workgroups:
    the fields equal those of the hardware order bit for bit (workgroup
    counts not divisible by 8 in both kernel kinds)."""
    from seidart_hip import models
    from seidart_hip.solver import run_config
    cfg = models.elastic3(nx=96, ny=88, nz=72, cpml=6, steps=40, src=(48, 44, 20), nrec=8)
    fields = ("vx", "vy", "vz", "sxx", "syz")
    r0, n0, f0 = run_config(cfg, fields=fields, options={"xcd_remap": 0})
    r1, n1, f1 = run_config(cfg, fields=fields, options={"xcd_remap": remap})
    assert np.array_equal(r0, r1) and np.array_equal(n0[0], n1[0]) and np.array_equal(n0[1], n1[1])
    for k in fields:
        assert np.array_equal(f0[k], f1[k]), k


@pytest.mark.parametrize("perturb", [False, True])
def test_elastic3_cpml_coefficient_forms(perturb):
    """The CPML shell's coefficients as a dictionary (a one-byte id per slot)
    and, once they differ in every cell (more than 256 tuples), as six
    streams inside the tile-plane blocks of the memory variables."""
    from seidart_hip import models
    from seidart_hip.solver import make_solver
    from oracle.oracle_ctypes import run_oracle
    cfg = models.elastic3(nx=30, ny=26, nz=28, cpml=5, steps=100, src=(15, 13, 9), nrec=8)
    if perturb:
        rng = np.random.default_rng(13)
        for k in ("acoef_cpml", "bcoef_cpml", "acoef_half_cpml", "bcoef_half_cpml"):
            a = cfg["arrays"][k]
            cfg["arrays"][k] = np.asfortranarray(a * (1.0 + 1e-3 * rng.random(a.shape)))
    fields = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")
    print(f"[parity] elastic3 CPML coefficients perturbed={perturb}", flush=True)
    s = make_solver(cfg)
    try:
        s.run()
        g = (s.receivers(), *s.norms(), {f: s.field(f) for f in fields})
        tuples = s.info("cpml_tuples")
    finally:
        s.close()
    assert (tuples == 0) == perturb and tuples <= 256
    o_rec, (o_vn, o_sn), o_f, us = run_oracle(cfg, fields=fields)
    assert us == 0
    _check(g, (o_rec, o_vn, o_sn, o_f), TOL64, stress_norm=True)


def test_elastic3_interior_and_cpml_paths_agree():
    """A vanishing but non-zero CPML coefficient everywhere forces every tile
    through the CPML body; results must equal the face-specialised run."""
    from seidart_hip import models
    from seidart_hip.solver import run_config
    cfg = models.elastic3(nx=30, ny=28, nz=26, cpml=5, steps=80, src=(15, 14, 9), nrec=8)
    r1, _, f1 = run_config(cfg, fields=("vx", "szz"))
    a = cfg["arrays"]["acoef_cpml"].copy()
    a[a == 0] = 1e-300
    cfg["arrays"]["acoef_cpml"] = a
    r2, _, f2 = run_config(cfg, fields=("vx", "szz"))
    assert rel_l2(r2, r1) < 1e-14 and rel_l2(f2["vx"], f1["vx"]) < 1e-14


def test_elastic3_material_table_equals_per_cell_arrays():
    """The material-table layout reads the same FP64 values as the per-cell
    arrays; the two runs agree to rounding (the compiler may contract the
    stiffness products into FMAs differently when the row comes from scalar
    registers)."""
    from seidart_hip import models
    from seidart_hip.solver import run_config
    cfg = models.elastic3(nx=36, ny=30, nz=34, cpml=6, steps=120, src=(18, 15, 10), nrec=8,
                          density_method="geometric")
    fields = ("vx", "vy", "vz", "sxx", "syz")
    r1, n1, f1 = run_config(cfg, fields=fields, options={"material_table": 1})
    r0, n0, f0 = run_config(cfg, fields=fields, options={"material_table": 0})
    assert rel_l2(r1, r0) < 1e-13 and rel_l2(n1[0], n0[0]) < 1e-13 and rel_l2(n1[1], n0[1]) < 1e-13
    for k in fields:
        assert rel_l2(f1[k], f0[k]) < 1e-13, k


def test_elastic3_material_count_reported():
    from seidart_hip import models
    from seidart_hip.solver import Solver
    cfg = models.elastic3(nx=24, ny=22, nz=20, cpml=4, steps=4, src=(12, 11, 6), nrec=2, nlayers=5)
    counts = []
    for opt in (1, 0):
        with Solver("elastic3", cfg["shape"], cfg["spacing"], cfg["cpml"], cfg["dt"], 4, cfg["src"]) as s:
            s.set_option("material_table", opt)
            s.set_arrays(cfg["arrays"])
            s.run(1)
            counts.append(s.materials())
    assert counts == [5, 0]


def test_elastic3_heterogeneous_model_uses_per_cell_arrays():
    """More than 256 distinct materials (per-cell density jitter): the
    per-cell layout is chosen and parity holds."""
    from seidart_hip import models
    cfg = models.elastic3(nx=30, ny=28, nz=26, cpml=5, steps=100, src=(15, 14, 9), nrec=8)
    rng = np.random.default_rng(7)
    cfg["arrays"]["rho"] = cfg["arrays"]["rho"] * (1.0 + 0.05 * rng.random(cfg["shape"]))
    g, o = _run_both(cfg, ("vx", "vz", "sxy"))
    _check(g, o, TOL64, stress_norm=True)


def test_elastic3_stress_source_harmonic():
    from seidart_hip import models
    cfg = models.elastic3(nx=30, ny=28, nz=26, cpml=5, steps=150, src=(15, 14, 9), nrec=8,
                          density_method="harmonic")
    s = cfg["sources"]["z"]
    cfg["sources"] = {"xx": s, "xy": 0.3 * s, "xz": -0.2 * s, "yy": 0.9 * s, "yz": 0.4 * s,
                      "zz": -s, "x": 0.1 * s, "y": -0.2 * s, "z": 0.3 * s}
    fields = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")
    g, o = _run_both(cfg, fields)
    _check(g, o, TOL64, stress_norm=True)


def test_elastic2_corrected_stencil():
    from seidart_hip import models
    cfg = models.elastic2(nx=80, nz=70, cpml=8, steps=300, src=(40, 20), rec_z=20, nrec=8)
    cfg["stencil"] = "corrected"
    g, o = _run_both(cfg, ("vx", "vz", "sxz"))
    _check(g, o, TOL64)


def test_unstable_run_reports_step():
    from seidart_hip import models, UnstableError
    from seidart_hip.solver import run_config
    cfg = models.elastic3(nx=20, ny=20, nz=20, cpml=4, steps=10, src=(10, 10, 8), nrec=2)
    ic = np.zeros(cfg["shape"], order="F")
    ic[12, 12, 12] = 1e26
    cfg["arrays"]["ic_y"] = ic
    with pytest.raises(UnstableError) as e:
        run_config(cfg)
    assert e.value.step == 1


@pytest.mark.parametrize("options", [{}, {"material_table": 0}])
def test_em2_fp64(options):
    from seidart_hip import models
    cfg = models.em2(nx=120, nz=80, steps=400, src=(60, 20), nrec=16)
    g, o = _run_both(cfg, ("ex", "ez", "hy"), options=options)
    _check(g, o, TOL64)


def test_em2_material_table_equals_per_cell_arrays():
    """Table and per-cell coefficient arrays hold the same values: the runs
    agree to rounding; with an adjoint planned no table is built."""
    from seidart_hip import models
    from seidart_hip.solver import make_solver
    cfg = models.em2(nx=110, nz=84, steps=300, src=(55, 20), nrec=8)
    out = []
    for opts in ({"material_table": 1}, {"material_table": 0}, {"material_table": 1, "checkpoint": 60}):
        s = make_solver(cfg, options=opts)
        try:
            s.run(cfg["time_steps"])
            out.append((s.receivers(), s.norms()[0], {f: s.field(f) for f in ("ex", "ez", "hy")},
                        s.materials()))
        finally:
            s.close()
    (r1, n1, f1, m1), (r0, n0, f0, m0), (r2, n2, f2, m2) = out
    assert m1 > 1 and m0 == 0 and m2 == 0
    assert rel_l2(r1, r0) < 1e-13 and rel_l2(n1, n0) < 1e-13
    assert np.array_equal(r2, r0) and np.array_equal(n2, n0)
    for k in f0:
        assert rel_l2(f1[k], f0[k]) < 1e-13, k


@pytest.mark.parametrize("options", [{}, {"material_table": 0}, {"xcd_remap": 1}, {"xcd_remap": 1, "material_table": 0},
                                     {"fused": 1}, {"fused": 1, "material_table": 0}, {"fused": 1, "zchunk": 5},
                                     {"fused": 1, "tile_rows": 4, "zchunk": 7},
                                     {"fused": 1, "tile_rows": 16, "zchunk": 3}])
def test_em3_fp64(options):
    """Half-step kernels (default; grid order and XCD column order) and the
    fused step (chunks of planes, 4-, 8- and 16-row tiles)."""
    from seidart_hip import models
    cfg = models.em3(nx=40, ny=36, nz=32, cpml=6, steps=200, src=(20, 18, 10), nrec=8)
    cfg["sources"]["y"] = 0.5 * cfg["sources"]["x"]
    cfg["sources"]["z"] = -0.25 * cfg["sources"]["x"]
    g, o = _run_both(cfg, ("ex", "ey", "ez", "hx", "hy", "hz"), options=options)
    _check(g, o, TOL64)


@pytest.mark.parametrize("options", [{}, {"material_table": 0}, {"xcd_remap": 1}, {"fused": 1}])
def test_em3_odd_plane_count(options):
    """An odd padded grid (46 x 42 x 39) with its point source near the
    bottom CPML slab."""
    from seidart_hip import models
    cfg = models.em3(nx=34, ny=30, nz=27, cpml=6, steps=150, src=(17, 15, 9), nrec=8)
    g, o = _run_both(cfg, ("ex", "ey", "ez", "hx", "hy", "hz"), options=options)
    _check(g, o, TOL64)


@pytest.mark.parametrize("fused", [1, 0])
def test_em3_runs_split_across_graph_replays(fused):
    """The fused step swaps E / H buffers every step and replays captured
    32-step graphs per buffer parity: a run split into pieces of odd and
    even lengths (direct steps and replays at both parities) must equal the
    oracle's uninterrupted run, in FP64 and in FP32 (against FP64 to 1e-3)."""
    from seidart_hip import models
    from seidart_hip.solver import make_solver
    from oracle.oracle_ctypes import run_oracle
    cfg = models.em3(nx=30, ny=28, nz=24, cpml=5, steps=150, src=(15, 14, 9), nrec=6)
    o_rec, (o_vn, _), o_f, _ = run_oracle(cfg, fields=("ex", "hz"))
    for precision, tol in (("fp64", TOL64), ("fp32", 1e-3)):
        s = make_solver(cfg, precision=precision, options={"fused": fused})
        try:
            for n in (5, 40, 3, 37, 33, 32):
                s.run(n)
            rec, (vn, _) = s.receivers(), s.norms()
            errs = [rel_l2(rec, o_rec), rel_l2(vn, o_vn), rel_l2(s.field("ex"), o_f["ex"]),
                    rel_l2(s.field("hz"), o_f["hz"])]
            print(precision, [f"{e:.2e}" for e in errs], flush=True)
            assert max(errs) <= tol
        finally:
            s.close()


def test_em3_heterogeneous_permittivity():
    """More than 256 distinct materials: the per-cell arrays are used."""
    from seidart_hip import models
    cfg = models.em3(nx=30, ny=28, nz=26, cpml=5, steps=120, src=(15, 14, 9), nrec=8)
    rng = np.random.default_rng(11)
    cfg["arrays"]["e11"] = cfg["arrays"]["e11"] * (1.0 + 0.05 * rng.random(cfg["shape"]))
    g, o = _run_both(cfg, ("ex", "hz"))
    _check(g, o, TOL64)


@pytest.mark.parametrize("fused", [1, 0])
@pytest.mark.parametrize("perturb", [False, True])
def test_em3_cpml_coefficient_forms(fused, perturb):
    """The CPML coefficients as a dictionary (a one-byte id per cell: the
    layered profiles give few distinct tuples) and, once they differ in every
    cell (more than 256 tuples), as the six per-cell arrays."""
    from seidart_hip import models
    from seidart_hip.solver import make_solver
    from oracle.oracle_ctypes import run_oracle
    cfg = models.em3(nx=30, ny=28, nz=26, cpml=5, steps=120, src=(15, 14, 9), nrec=8)
    if perturb:
        rng = np.random.default_rng(12)
        for k in ("acoef_cpml", "bcoef_cpml", "acoef_half_cpml", "bcoef_half_cpml"):
            a = cfg["arrays"][k]
            cfg["arrays"][k] = np.asfortranarray(a * (1.0 + 1e-3 * rng.random(a.shape)))
    fields = ("ex", "ey", "ez", "hx", "hy", "hz")
    s = make_solver(cfg, options={"fused": fused})
    try:
        s.run()
        g = (s.receivers(), *s.norms(), {f: s.field(f) for f in fields})
        tuples = s.info("cpml_tuples")
    finally:
        s.close()
    assert (tuples == 0) == perturb and tuples <= 256
    o_rec, (o_vn, o_sn), o_f, us = run_oracle(cfg, fields=fields)
    assert us == 0
    _check(g, (o_rec, o_vn, o_sn, o_f), TOL64)


@pytest.mark.parametrize("physics", ["elastic2", "elastic25", "elastic3", "em2", "em25", "em3"])
def test_fp32_close_to_fp64(physics):
    """FP32 wavefields with FP64-precomputed coefficients: trace error vs the
    FP64 oracle stated per solver (bar 1e-3 relative L2 on these short runs)."""
    from seidart_hip import models
    cfg = {"elastic2": lambda: models.elastic2(nx=80, nz=70, cpml=8, steps=300, src=(40, 20), rec_z=20, nrec=8),
           "elastic25": lambda: models.elastic25(nx=70, nz=60, cpml=8, steps=250, src=(35, 20), nrec=8),
           "em25": lambda: models.em25(nx=70, nz=60, cpml=8, steps=250, src=(35, 20), nrec=8),
           "elastic3": lambda: models.elastic3(nx=40, ny=36, nz=32, cpml=6, steps=200, src=(20, 18, 10), nrec=16),
           "em2": lambda: models.em2(nx=120, nz=80, steps=400, src=(60, 20), nrec=16),
           "em3": lambda: models.em3(nx=40, ny=36, nz=32, cpml=6, steps=200, src=(20, 18, 10), nrec=8)}[physics]()
    g, o = _run_both(cfg, (), precision="fp32")
    err = rel_l2(g[0], o[0])
    print(f"[fp32] {physics} receiver rel L2 vs FP64 oracle: {err:.3e}", flush=True)
    assert err <= 1e-3


@pytest.mark.parametrize("ic,src", [(False, (60, 20)), (True, (60, 20)), (False, (2, 3))])
def test_em2_fp32_source_window(ic, src):
    """2D EM in FP32 with the FP64 window around the point source (the
    default) and without it, both against the FP64 oracle, receiver 0 on the
    source cell.  Without initial fields the source pulse dominates that
    receiver and the window's FP64 values cut its error at least tenfold
    (FP32 rounding of the large pulse otherwise sets it).  With random
    initial E on every cell, set after setup, the run stays FP32-close to
    the oracle: the window takes the set values, not stale ones.  A source
    next to the CPML corner gets a window clipped to the CPML-free box, the
    source on its low edges: still correct, and no worse than FP32."""
    import numpy as np
    from seidart_hip import models
    from seidart_hip.solver import make_solver
    from oracle.oracle_ctypes import run_oracle
    cfg = models.em2(nx=120, nz=80, steps=600, src=src, nrec=16)
    cfg["receivers"] = np.vstack([[cfg["src"][0], 1, cfg["src"][-1]], np.asarray(cfg["receivers"])[1:]])
    fields = ("ex", "ez", "hy")
    icd = {}
    if ic:
        rng = np.random.default_rng(12)
        icd = {f: rng.uniform(-1.0, 1.0, cfg["shape"]) for f in ("ex", "ez")}
    print(f"[parity] em2 fp32 source window, initial fields {ic}, source {src}", flush=True)
    o_rec, _, o_f, us = run_oracle(cfg, fields=fields, field_in=icd or None)
    assert us == 0
    errs = {}
    for win in (1, 0):
        s = make_solver(cfg, precision="fp32", options={"source_window": win})
        try:
            for k, a in icd.items():
                s.set_field(k, a)
            s.run()
            rec = s.receivers()
            f = {k: s.field(k) for k in fields}
        finally:
            s.close()
        near = np.abs(rec[:, 0] - o_rec[:, 0]).max() / np.abs(o_rec[:, 0]).max()
        errs[win] = (near, rel_l2(rec, o_rec), max(rel_l2(f[k], o_f[k]) for k in fields))
        print(f"window {win}: source-cell receiver max err / max {near:.2e}, receivers rel L2 "
              f"{errs[win][1]:.2e}, fields {errs[win][2]:.2e}", flush=True)
    assert errs[1][1] <= 1e-4 and errs[1][2] <= 1e-4
    if src != (60, 20):
        assert errs[1][0] <= errs[0][0]
    elif not ic:
        assert errs[1][0] <= 0.1 * errs[0][0]


EL25_FIELDS = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")


@pytest.mark.parametrize("options", [{}, {"material_table": 0}])
def test_elastic25_fp64(options):
    """2.5D elastic: layered rotated-ice model with a lateral anomaly,
    attenuation, harmonic density smear, ky = 0.05, 400 steps; stiffness
    from the material table (default) and from the 27 per-cell arrays."""
    from seidart_hip import models
    cfg = models.elastic25()
    g, o = _run_both(cfg, EL25_FIELDS, options=options)
    _check(g, o, TOL64)


def test_elastic25_material_table_equals_per_cell_arrays():
    """The table holds the same FP64 stiffness and divisors as the per-cell
    arrays; the runs agree to rounding (FMA contraction may differ when the
    row sits in scalar registers).  The anomaly's edge puts wavefronts with
    two materials on the per-lane row path."""
    from seidart_hip import models
    from seidart_hip.solver import make_solver
    cfg = models.elastic25(nx=70, nz=60, cpml=8, steps=200, src=(35, 20), nrec=8, ky=0.1)
    out = []
    for opt in (1, 0):
        s = make_solver(cfg, options={"material_table": opt})
        try:
            s.run(cfg["time_steps"])
            out.append((s.receivers(), s.norms()[0], {f: s.field(f) for f in EL25_FIELDS}, s.materials()))
        finally:
            s.close()
    (r1, n1, f1, m1), (r0, n0, f0, m0) = out
    assert m1 > 1 and m0 == 0
    assert rel_l2(r1, r0) < 1e-13 and rel_l2(n1, n0) < 1e-13
    for k in EL25_FIELDS:
        assert rel_l2(f1[k], f0[k]) < 1e-13, k


@pytest.mark.parametrize("density,ky", [("arithmetic", 0.0), ("geometric", 0.3), ("none", 0.12)])
def test_elastic25_sources_and_density(density, ky):
    from seidart_hip import models
    cfg = models.elastic25(nx=70, nz=60, cpml=8, steps=250, src=(35, 20), nrec=8, ky=ky,
                           density_method=density)
    s = cfg["sources"]["z"]
    cfg["sources"] = {"xx": s, "xy": 0.3 * s, "xz": -0.2 * s, "yy": 0.9 * s, "yz": 0.4 * s,
                      "zz": -s, "x": 0.1 * s, "y": -0.2 * s, "z": 0.3 * s}
    g, o = _run_both(cfg, EL25_FIELDS)
    _check(g, o, TOL64)


def test_elastic25_complex_ic_corrected_stencil():
    from seidart_hip import models
    cfg = models.elastic25(nx=64, nz=56, cpml=8, steps=200, src=(32, 20), nrec=8, ky=0.2)
    cfg["stencil"] = "corrected"
    shape = cfg["shape"]
    x, z = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    g0 = np.exp(-((x - 40) ** 2 + (z - 30) ** 2) / 20.0)
    cfg["arrays"]["ic_x"] = (0.3 + 0.7j) * g0
    cfg["arrays"]["ic_z"] = (-0.5 + 0.2j) * np.roll(g0, 5, axis=1)
    g, o = _run_both(cfg, EL25_FIELDS)
    _check(g, o, TOL64)


EM25_FIELDS = ("ex", "ey", "ez", "hx", "hy", "hz")


@pytest.mark.parametrize("ky,options", [(0.0, {}), (5.0, {}), (30.0, {}), (5.0, {"material_table": 0})])
def test_em25_fp64(ky, options):
    """2.5D EM: layered SPD permittivity / conductivity, 4th-order
    differences, i*ky y derivatives, three-component dipole; coefficients
    from the material table (default) and from per-cell arrays."""
    from seidart_hip import models
    cfg = models.em25(ky=ky)
    g, o = _run_both(cfg, EM25_FIELDS, options=options)
    _check(g, o, TOL64)


def test_em25_material_table_equals_per_cell_arrays():
    """Table and per-cell arrays hold the same values: the runs agree to
    rounding; a per-cell jitter of one conductivity overflows the table and
    falls back to the arrays."""
    from seidart_hip import models
    from seidart_hip.solver import make_solver
    cfg = models.em25(nx=70, nz=60, cpml=8, steps=200, src=(35, 20), nrec=8, ky=12.0)
    jit = dict(cfg)
    jit["arrays"] = dict(cfg["arrays"])
    rng = np.random.default_rng(5)
    jit["arrays"]["s11"] = cfg["arrays"]["s11"] * (1.0 + 0.01 * rng.random(cfg["shape"]))
    out = []
    for c, opts in ((cfg, {"material_table": 1}), (cfg, {"material_table": 0}), (jit, {})):
        s = make_solver(c, options=opts)
        try:
            s.run(c["time_steps"])
            out.append((s.receivers(), s.norms()[0], {f: s.field(f) for f in EM25_FIELDS}, s.materials()))
        finally:
            s.close()
    (r1, n1, f1, m1), (r0, n0, f0, m0), (_, _, _, mj) = out
    assert m1 > 1 and m0 == 0 and mj == 0
    assert rel_l2(r1, r0) < 1e-13 and rel_l2(n1, n0) < 1e-13
    for k in EM25_FIELDS:
        assert rel_l2(f1[k], f0[k]) < 1e-13, k


def test_em25_complex_ic():
    from seidart_hip import models
    cfg = models.em25(nx=70, nz=60, cpml=8, steps=200, src=(35, 20), nrec=8, ky=12.0)
    shape = cfg["shape"]
    x, z = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    g0 = np.exp(-((x - 40) ** 2 + (z - 30) ** 2) / 20.0)
    cfg["arrays"]["ic_y"] = (0.2 - 0.9j) * g0
    cfg["arrays"]["ic_z"] = (0.6 + 0.1j) * np.roll(g0, 4, axis=0)
    g, o = _run_both(cfg, EM25_FIELDS)
    _check(g, o, TOL64)
