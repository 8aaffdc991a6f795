"""Throughput of the CPML-FDTD solvers on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|fp32]
                    [--workload seismic3|electromag3|seismic2|batch2d|fwi2d|...]

Default workload: the 3D anisotropic elastic solver (seismic3) on 256^3 cells
+ 12-cell CPML (280^3 padded), dx = 0.5 m, eight seeded rotated-ice layers
with attenuation, Ricker 100 Hz vertical force, 256 receivers, FP64 (the
reference's arithmetic).  One step = one time step over the whole padded
grid (stress pass, velocity pass, source/receiver tail).  It prints one JSON
line: `value` is GCell-updates/s over the K timed steps with every input
resident in HBM.

`roofline` prices the stress pass (the interior and CPML-shell stress
kernels, one launch each per step) by the bytes its kernels must move -- each
operand array element read once and each result written once, for the
material layout that ran -- over the pass's mean duration, timed with HIP
events on the solver's stream, against the 8 TB/s HBM peak.  The byte model
that counts every material parameter as a per-cell array (the contract
figure) is reported beside it under its own keys, and `traffic` is the HBM
traffic measured with rocprofv3 counters, with its ratio to the bytes the
kernels must move.

The default run also times the other benchmark configurations and adds them
to the same line under "configs": the 2D elastic 200x200 case beside the CPU
oracle, the 3D EM 384x384x192 case, the FP32 forms, the batched uneven
multi-shot 2D case and the FWI gradient case (see --subrecords).

`cpu_baseline` times the oracle (a C restatement of cpmlfdtd.f08's loops,
built with the reference's production compiler flags, OpenMP) on a bounded
sample of the same workload on the host cores.
"""
from __future__ import annotations

import argparse
import copy
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "seidart-backend_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0


def log(msg):
    """Progress lines go to stderr; stdout carries the one JSON line."""
    print(msg, file=sys.stderr, flush=True)


# ------------------------------------------------------------------ byte models
# Contract model (per step): B = w [N (2F + M) + N_pml (2 Psi + C)], every
# material parameter a per-cell array, each field read and written once per
# step, live unique CPML recursions read and written once, six CPML
# coefficient arrays read once.
EL3 = dict(F=9, M=28, PSI=29, C=6)
EM3 = dict(F=6, M=12, PSI=12, C=6)
EL2 = dict(F=5, M=10, PSI=12, C=6)
EM2 = dict(F=3, M=7, PSI=4, C=6)


def _cells3(shape, cpml):
    nx, ny, nz = shape
    n = nx * ny * nz
    n_int = max(nx - 2 * cpml, 0) * max(ny - 2 * cpml, 0) * max(nz - 2 * cpml, 0)
    return n, n_int, n - n_int


def contract_bytes(P, n, n_pml, w):
    return w * (n * (2 * P["F"] + P["M"]) + n_pml * (2 * P["PSI"] + P["C"]))


def el3_contract_stress_pass(shape, cpml, w):
    """The contract model split to the stress pass: velocities read, six
    stresses read and written, the 27 stiffness and attenuation arrays, and in
    CPML cells the 20 stress-side recursions and the six coefficients."""
    n, _, n_pml = _cells3(shape, cpml)
    return w * (n * (3 + 12 + 27) + n_pml * (40 + 6))


def el3_path_bytes(shape, cpml, w, layout, n_ig=6, ctab=True):
    """Bytes each 3D elastic kernel must move per launch for the material
    layout that ran ("table": one-byte material id per cell and a cached
    table; "dense": 21 stiffness arrays + n_ig attenuation-divisor arrays +
    density; "iso": c11, c12, c44 + n_ig divisors + density).

    interior stress:   3 velocities read, 6 stresses read + written, material
    interior velocity: 6 stresses read, 3 velocities read + written, density
    CPML shell: the same plus the live memory recursions read + written (20
    stress side, 9 velocity side) and the CPML coefficients (a one-byte
    dictionary id per cell, or six values)."""
    n, n_int, n_pml = _cells3(shape, cpml)
    idb = 1 if layout == "table" else 0
    mat = {"table": 0, "dense": 21 + n_ig, "iso": 3 + n_ig}[layout]
    rho = 0 if layout == "table" else 1
    cf = 1 if ctab else 6 * w
    s_int = w * (3 + 12 + mat) + idb
    v_int = w * (6 + 6 + rho) + idb
    out = {"int_stress": n_int * s_int, "shell_stress": n_pml * (s_int + 40 * w + cf),
           "int_velocity": n_int * v_int, "shell_velocity": n_pml * (v_int + 18 * w + cf)}
    out["stress"] = out["int_stress"] + out["shell_stress"]
    out["velocity"] = out["int_velocity"] + out["shell_velocity"]
    out["step"] = out["stress"] + out["velocity"]
    return out


def em3_path_bytes(shape, cpml, w, table):
    """H update: E read, H read + written; E update: H read, E read + written,
    the material (one-byte id, or the 12 inverse-permittivity and
    conductivity arrays); CPML cells of either: six memory variables read +
    written and the six coefficient arrays.  The fused step (one kernel):
    E and H read once and written once, the material, and in CPML cells the
    twelve memory variables read + written and the six coefficients once."""
    n, _, n_pml = _cells3(shape, cpml)
    mat = 1 if table else 12 * w
    h = n * 9 * w + n_pml * 18 * w
    e = n * (9 * w + mat) + n_pml * 18 * w
    fused = n * (12 * w + mat) + n_pml * 30 * w
    return {"h": h, "e": e, "step": h + e, "fused": fused}


def path_bytes_2d(physics, n, n_pml, w, table):
    """Per step of one 2D grid.  elastic2: stress pass 2 velocities read, 3
    stresses read + written, material (id byte, or 6 stiffness + 3 divisor
    arrays); velocity pass 3 stresses read, 2 velocities read + written,
    density; CPML cells 12 recursions read + written and the 6 coefficient
    arrays in each pass.  em2: H pass 2 E read, Hy read + written; E pass Hy
    read, 2 E read + written, material (id byte or 7 coefficient arrays);
    CPML cells 4 recursions and the 6 coefficient arrays in each pass."""
    if physics == "elastic2":
        per = w * 16 + (1 if table else 9 * w)
        extra = w * (24 + 12)
    else:
        per = w * 9 + (1 if table else 7 * w)
        extra = w * (8 + 12)
    return n * per + n_pml * extra


# --------------------------------------------------------------- host / CPU leg
def cpu_threads():
    """Host threads for the CPU baseline: OMP_NUM_THREADS when set, else the
    host's cores capped at 16 (one GPU's share of a shared host)."""
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)


def host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


CPU_BUILD = ("C restatement of cpmlfdtd.f08 (oracle/oracle.c) built with the reference's production flags "
             "(build.sh:123: -O3 -funroll-loops -ffast-math -ffp-contract=fast, -march=x86-64-v3 for "
             "-march=native), OpenMP")


def cpu_baseline(cfg, steps, label):
    from oracle.oracle_ctypes import run_oracle
    import numpy as np
    threads = cpu_threads()
    run_oracle(cfg, steps=1, nthreads=threads, fast=True)   # first touch of the oracle's arrays
    t1 = time.perf_counter()
    run_oracle(cfg, steps=steps, nthreads=threads, fast=True)
    cpu_s = time.perf_counter() - t1
    n = int(np.prod(cfg["shape"]))
    return {"value": round(n * steps / cpu_s / 1e9, 5), "unit": "GCell-updates/s", "cores": threads,
            "kind": "port", "sample": f"{steps} time steps of {label} ({cpu_s:.1f} s after a 1-step warm-up), "
                                      f"{threads} OpenMP threads",
            "build": CPU_BUILD, "host": host_info()}


# ------------------------------------------------------------------ GPU timing
def timed_run(cfg, args, material_table, isotropic=1, precision=None, extra=None):
    """W untimed steps (the first uploads every input), then K steps timed by
    the wall clock between synchronisations (captured step graphs, no per-
    kernel events), then KT more steps with HIP events around each kernel pass
    on the solver's stream for the per-pass durations.  Returns the timings,
    the material count, whether the isotropic arrays were used and the
    receiver traces of all W + K + KT steps."""
    import numpy as np
    from seidart_hip.solver import Solver
    K, W = args.steps, args.warmup
    KT = max(1, min(K, 100))
    s = Solver(cfg["physics"], cfg["shape"], cfg["spacing"], cfg["cpml"], cfg["dt"], W + K + KT, cfg["src"],
               precision=precision or args.precision, density_method=cfg.get("density_method", "arithmetic"),
               ky=cfg.get("ky", 0.0))
    s.set_option("zchunk", args.zchunk)
    s.set_option("stagger", args.stagger)
    s.set_option("tile_rows", args.tile_rows)
    s.set_option("material_table", material_table)
    s.set_option("concurrent", args.concurrent)
    s.set_option("xcd_remap", args.xcd_remap)
    s.set_option("isotropic", isotropic)
    if cfg["physics"] == "em3":
        s.set_option("fused", args.fused)
    s.set_arrays(cfg["arrays"])
    for comp, series in cfg["sources"].items():
        series = np.asarray(series, dtype=np.float64)
        if series.size < W + K + KT:
            series = np.concatenate([series, np.zeros(W + K + KT - series.size)])
        s.set_source(comp, series[:W + K + KT])
    s.set_receivers(cfg["receivers"])
    s.run(W)
    s.synchronize()
    t0 = time.perf_counter()
    s.run(K)
    s.synchronize()
    elapsed = time.perf_counter() - t0
    s.set_timing(True)
    s.run(KT)
    s.synchronize()
    ms_stress, ms_vel, launches = s.kernel_times()
    nmat = s.materials()
    iso = s.info("isotropic")
    more = s.info(extra) if extra else None
    rec = s.receivers()
    s.close()
    if not np.all(np.isfinite(rec)):
        raise RuntimeError("non-finite receiver traces")
    if extra:
        return elapsed, ms_stress, ms_vel, launches, nmat, iso, rec, more
    return elapsed, ms_stress, ms_vel, launches, nmat, iso, rec


def rel_l2(a, b):
    import numpy as np
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def measured_traffic(name, key="traffic_bytes_per_launch"):
    """HBM bytes per launch of a pass from a committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel calibration),
    or None when the summary or the key is absent."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    m = re.search(r"^" + key + r"\s*=\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


CONTRACT_NOTE = ("contract_bytes counts every material parameter as a per-cell array; the layout that ran "
                 "reads a one-byte id and a cached table row instead, so contract_frac is not a bandwidth "
                 "(it can exceed 1)")


def calibrate_copy(nbytes=2 << 30, reps=20):
    """Plain element-wise copy of a known size (one dispatch per copy) whose
    rocprofv3 FETCH_SIZE / WRITE_SIZE readings scale the counters of the
    solver kernels to bytes."""
    import torch
    n = nbytes // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        torch.neg(x, out=y)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(json.dumps({"calibration": "copy", "bytes_read": n * 8, "bytes_written": n * 8,
                      "ms": round(ms, 4), "GB/s": round(2 * n * 8 / ms / 1e6, 1)}), flush=True)


# ------------------------------------------------------------------ 3D elastic
def bench_el3(args, rec64=None):
    """seismic3 256^3 + 12-cell CPML in args.precision: the main line, the
    per-cell-array layout beside it (FP64), the CPU baseline."""
    import numpy as np
    from seidart_hip import models

    K, W = args.steps, args.warmup
    n = args.n
    t_gen = time.perf_counter()
    cfg = models.elastic3(nx=n, ny=n, nz=n, steps=W + K, src=(n // 2, n // 2, 20), nrec=min(256, n))
    t_gen = time.perf_counter() - t_gen
    shape = cfg["shape"]
    ncell = int(np.prod(shape))
    w = 8 if args.precision == "fp64" else 4
    fp64 = args.precision == "fp64"
    log(f"[seismic3] {shape} {args.precision} K={K} W={W}")

    main_run = timed_run(cfg, args, args.material_table)
    dense_run = timed_run(cfg, args, 0) if (args.material_table and args.compare_layouts and fp64) else None
    elapsed, ms_stress, ms_vel, launches, nmat, _, rec = main_run
    table = bool(args.material_table) and nmat > 0
    layout = "table" if table else "dense"

    n_all, _, n_pml = _cells3(shape, cfg["cpml"])
    contract_step = contract_bytes(EL3, n_all, n_pml, w)
    contract_stress = el3_contract_stress_pass(shape, cfg["cpml"], w)
    path = el3_path_bytes(shape, cfg["cpml"], w, layout)
    avg_stress_s = ms_stress / max(launches, 1) / 1e3
    avg_vel_s = ms_vel / max(launches, 1) / 1e3
    achieved = path["stress"] / avg_stress_s / 1e9
    prof = "el3_seismic3_280cube_fp64_hbm_traffic.txt"
    traffic = measured_traffic(prof) if fp64 and n == 256 and table else None
    traffic_v = measured_traffic(prof, "velocity_traffic_bytes_per_launch") if traffic else None

    line = {
        "metric": f"GCell-updates/s (3D elastic CPML-FDTD, {args.precision.upper()})",
        "value": round(ncell * K / elapsed / 1e9, 4),
        "unit": "GCell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(elapsed / K * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if fp64 else "f32",
        "data": "synthetic (seeded 8-layer rotated-ice model, Ricker source)",
        "config": {"workload": f"seismic3 {n}^3 + {cfg['cpml']}-cell CPML ({shape[0]}^3 padded), "
                               f"dx=0.5 m, 8 layers, attenuation, 256 receivers",
                   "cells": ncell, "precision": args.precision,
                   "material_layout": f"table ({nmat} materials, 1-byte id per cell)" if table
                                      else "per-cell arrays",
                   "zchunk": args.zchunk, "stagger": args.stagger, "tile_rows": args.tile_rows,
                   "concurrent_shell": bool(args.concurrent)},
        # bytes the four kernels of a step must move for this layout, per step
        "hbm_fraction_step": round(path["step"] * K / elapsed / 1e9 / HBM_PEAK_GBS, 4),
        "step_bytes": path["step"],
        # the contract byte model (every material parameter a per-cell array)
        "hbm_fraction_step_contract": round(contract_step * K / elapsed / 1e9 / HBM_PEAK_GBS, 4),
        "step_bytes_contract": contract_step,
        "roofline": {"bound": "hbm",
                     "kernel": "el3 stress pass (el3_int_stress_kernel + el3_shell_stress_kernel)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes": path["stress"],
                     "waste_ratio": round(traffic / path["stress"], 3) if traffic else None,
                     "contract_bytes": contract_stress,
                     "contract_frac": round(contract_stress / avg_stress_s / 1e9 / HBM_PEAK_GBS, 4),
                     "contract_note": CONTRACT_NOTE if table else None,
                     "avg_launch_ms": round(avg_stress_s * 1e3, 4),
                     "velocity_pass": {"avg_launch_ms": round(avg_vel_s * 1e3, 4),
                                       "algorithmic_bytes": path["velocity"],
                                       "frac": round(path["velocity"] / avg_vel_s / 1e9 / HBM_PEAK_GBS, 4),
                                       "traffic": traffic_v,
                                       "waste_ratio": round(traffic_v / path["velocity"], 3) if traffic_v
                                       else None}},
        "cpu_baseline": None,
        "setup_s": round(t_gen, 2),
    }
    if dense_run is not None:
        e2, s2, v2, l2, _, _, _ = dense_run
        pd = el3_path_bytes(shape, cfg["cpml"], w, "dense")
        line["per_cell_arrays_layout"] = {
            "value": round(ncell * K / e2 / 1e9, 4), "ms_per_step": round(e2 / K * 1e3, 4),
            "hbm_fraction_step": round(pd["step"] * K / e2 / 1e9 / HBM_PEAK_GBS, 4),
            "hbm_fraction_step_contract": round(contract_step * K / e2 / 1e9 / HBM_PEAK_GBS, 4),
            "stress_pass_ms": round(s2 / max(l2, 1), 4), "velocity_pass_ms": round(v2 / max(l2, 1), 4),
            "stress_pass_frac": round(pd["stress"] / (s2 / max(l2, 1) / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
    if rec64 is not None:
        line["fp32_vs_fp64_trace_rel_l2"] = float(f"{rel_l2(rec, rec64):.3e}")
    if args.cpu_steps > 0 and fp64:
        log(f"[seismic3] CPU baseline, {args.cpu_steps} steps")
        line["cpu_baseline"] = cpu_baseline(cfg, args.cpu_steps, f"the same {shape[0]}^3 model")
    return line, rec


# ------------------------------------------------------------------ 3D EM
def bench_em3(args):
    """3D anisotropic EM (electromag3) on 384 x 384 x 192 + 10-cell CPML
    (404 x 404 x 212 padded), dx = 0.05 m, five seeded SPD permittivity /
    conductivity layers, 100 MHz Ricker Ex dipole, 128 receivers.  FP64
    unless --precision fp32 (which also reports the FP32-vs-FP64 trace
    error)."""
    import numpy as np
    from seidart_hip import models
    K, W = args.steps, args.warmup
    cfg = models.em3(steps=W + K)
    shape = cfg["shape"]
    ncell = int(np.prod(shape))
    fp64 = args.precision == "fp64"
    w = 8 if fp64 else 4
    log(f"[electromag3] {shape} {args.precision} K={K} W={W}")
    elapsed, ms_h, ms_e, launches, nmat, _, rec, fused = timed_run(cfg, args, args.material_table, extra="fused")
    n_all, _, n_pml = _cells3(shape, cfg["cpml"])
    table = nmat > 0
    path = em3_path_bytes(shape, cfg["cpml"], w, table)
    if fused:   # one kernel per step: the pass timed as "H" is the whole step
        path = {"step": path["fused"], "fused": path["fused"]}
    contract_step = contract_bytes(EM3, n_all, n_pml, w)
    # contract model split to the E update: E read + written, H read, the 12
    # material arrays; CPML cells 6 recursions and 6 coefficients
    contract_e = w * (n_all * (6 + 3 + 12) + n_pml * (12 + 6))
    avg_e = ms_e / max(launches, 1) / 1e3
    avg_h = ms_h / max(launches, 1) / 1e3
    prof = "em3_electromag3_404x404x212_fp64_hbm_traffic.txt"
    # the committed counter summary is of the default step form (the half-step kernels)
    traffic = measured_traffic(prof) if fp64 and table and not fused else None
    traffic_h = measured_traffic(prof, "h_traffic_bytes_per_launch") if traffic else None
    line = {
        "metric": f"GCell-updates/s (3D EM CPML-FDTD, {args.precision.upper()})",
        "value": round(ncell * K / elapsed / 1e9, 4), "unit": "GCell-updates/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": round(elapsed / K * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if fp64 else "f32",
        "data": "synthetic (seeded 5-layer SPD permittivity/conductivity, Ricker Ex dipole)",
        "config": {"workload": f"electromag3 384x384x192 + {cfg['cpml']}-cell CPML "
                               f"({shape[0]}x{shape[1]}x{shape[2]} padded), dx=0.05 m",
                   "cells": ncell, "precision": args.precision,
                   "material_layout": f"table ({nmat} materials, 1-byte id per cell)" if table
                                      else "per-cell arrays"},
        "hbm_fraction_step": round(path["step"] * K / elapsed / 1e9 / HBM_PEAK_GBS, 4),
        "step_bytes": path["step"],
        "hbm_fraction_step_contract": round(contract_step * K / elapsed / 1e9 / HBM_PEAK_GBS, 4),
        "cpu_baseline": None,
    }
    if fused:
        line["config"]["step"] = "fused (one kernel per step, E / H double-buffered)"
        line["roofline"] = {"bound": "hbm", "kernel": "em3_fused_kernel (H and E updates of a step)",
                            "achieved": round(path["fused"] / avg_h / 1e9, 1), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(path["fused"] / avg_h / 1e9 / HBM_PEAK_GBS, 4),
                            "traffic": traffic, "algorithmic_bytes": path["fused"],
                            "waste_ratio": round(traffic / path["fused"], 3) if traffic else None,
                            "contract_bytes": contract_step,
                            "contract_frac": round(contract_step / avg_h / 1e9 / HBM_PEAK_GBS, 4),
                            "contract_note": CONTRACT_NOTE if table else None,
                            "avg_launch_ms": round(avg_h * 1e3, 4)}
    else:
        line["config"]["step"] = "H and E updates, one kernel each"
        line["roofline"] = {"bound": "hbm", "kernel": "em3_e_kernel (E update)",
                            "achieved": round(path["e"] / avg_e / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(path["e"] / avg_e / 1e9 / HBM_PEAK_GBS, 4), "traffic": traffic,
                            "algorithmic_bytes": path["e"],
                            "waste_ratio": round(traffic / path["e"], 3) if traffic else None,
                            "contract_bytes": contract_e,
                            "contract_frac": round(contract_e / avg_e / 1e9 / HBM_PEAK_GBS, 4),
                            "contract_note": CONTRACT_NOTE if table else None,
                            "avg_launch_ms": round(avg_e * 1e3, 4),
                            "h_kernel": {"avg_launch_ms": round(avg_h * 1e3, 4), "algorithmic_bytes": path["h"],
                                         "frac": round(path["h"] / avg_h / 1e9 / HBM_PEAK_GBS, 4),
                                         "traffic": traffic_h,
                                         "waste_ratio": round(traffic_h / path["h"], 3) if traffic_h else None}}
    if not fp64:
        a64 = copy.copy(args)
        a64.precision = "fp64"
        rec64 = timed_run(cfg, a64, args.material_table)[6]
        line["fp32_vs_fp64_trace_rel_l2"] = float(f"{rel_l2(rec, rec64):.3e}")
    elif args.cpu_steps > 0:
        log(f"[electromag3] CPU baseline, {args.em3_cpu_steps} steps")
        line["cpu_baseline"] = cpu_baseline(cfg, args.em3_cpu_steps, f"the same {shape[0]}x{shape[1]}x"
                                                                     f"{shape[2]} model")
    return line


# ------------------------------------------------------------------ 2D elastic, config 0
def bench_seismic2(args):
    """The reference's CPU-runnable case: 2D elastic, 200 x 200 (+10-cell
    CPML), two seeded tilted ice layers, 1000 steps.  About 10 MB per step,
    so the GPU step is bound by launch latency (captured step graphs), not
    by HBM; the CPU oracle runs the whole case beside it."""
    import numpy as np
    from seidart_hip import models

    K, W = args.steps, args.warmup
    cfg = models.elastic2(steps=W + K)
    shape = cfg["shape"]
    n = int(np.prod(shape))
    log(f"[seismic2] {shape} {args.precision} K={K}")
    elapsed, ms_s, ms_v, launches, nmat, _, _ = timed_run(cfg, args, args.material_table)
    cpu = None
    if args.cpu_steps > 0:
        ref = models.elastic2(steps=K)
        cpu = cpu_baseline(ref, K, f"the whole {K}-step {shape[0]}x{shape[1]} case")
    path = path_bytes_2d("elastic2", n, n - (shape[0] - 20) * (shape[1] - 20), 8 if args.precision == "fp64" else 4,
                         nmat > 0)
    return {
        "metric": f"GCell-updates/s (2D elastic CPML-FDTD, {args.precision.upper()})",
        "value": round(n * K / elapsed / 1e9, 4), "unit": "GCell-updates/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(elapsed / K * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded 2-layer tilted ice model, Ricker source)",
        "config": {"workload": f"seismic2 200x200 + 10-cell CPML ({shape[0]}x{shape[1]} padded), {K} steps",
                   "precision": args.precision,
                   "material_layout": f"table ({nmat} materials, 1-byte id per cell)" if nmat else "per-cell arrays"},
        "roofline": {"bound": "launch latency", "achieved": round(path * K / elapsed / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(path * K / elapsed / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "algorithmic_bytes": path,
                     "note": f"{path / 1e6:.1f} MB per step: {path / HBM_PEAK_GBS / 1e3:.2f} us at 8 TB/s, "
                             "below one kernel launch"},
        "stress_pass_ms": round(ms_s / max(launches, 1), 4), "velocity_pass_ms": round(ms_v / max(launches, 1), 4),
        "cpu_baseline": cpu,
    }


# ------------------------------------------------------------------ batches, config 3
def _cells_2d(cfg):
    nx, nz = cfg["shape"]
    cp = cfg["cpml"]
    n = nx * nz
    return n, n - max(nx - 2 * cp, 0) * max(nz - 2 * cp, 0)


def bench_batch2d(args):
    """Uneven multi-shot 2D forward: `--shots` seeded shots (nx 300-800, nz
    150-400, 1500-3000 steps, 4-layer anisotropic models, own source each),
    advanced together by one batch (one launch per half step over every
    shot's tiles, autotuned tile height, captured step graphs) in FP32, timed
    against the same shots run one after another as single handles in FP32
    and in FP64.  The whole run of every shot is timed (no warmup steps: the
    batch's tuning steps are real steps inside the timed region); uploads
    happen before the clock starts."""
    import numpy as np
    from seidart_hip import models
    from seidart_hip.solver import Batch, make_solver

    out = {}
    for physics in ("elastic2", "em2") if args.batch_physics == "both" else (args.batch_physics,):
        t_gen = time.perf_counter()
        cfgs = models.batch2d(physics, nshots=args.shots, stencil=args.batch_stencil,
                              tilt=(0.0, args.batch_tilt))
        t_gen = time.perf_counter() - t_gen
        cell_steps = sum(_cells_2d(c)[0] * c["time_steps"] for c in cfgs)
        log(f"[batch2d] {physics}: {len(cfgs)} shots, {cell_steps / 1e9:.2f} G cell-steps (models {t_gen:.1f} s)")

        shots = [make_solver(c, precision="fp32") for c in cfgs]
        b = Batch(shots)
        t0 = time.perf_counter()
        b.run()
        t_batch = time.perf_counter() - t0
        rows, tiles, xcd = b.info("tile_rows"), b.info("tiles"), b.info("xcd_order")
        b.close()
        rec32 = [s.receivers() for s in shots]
        # growth of each shot's velocity norm: last step over the peak of its
        # first 600 steps (the source pulse); tilted elastic axes grow in the
        # reference scheme's own arithmetic
        growth = []
        for s in shots:
            vn = s.norms()[0]
            growth.append(float(vn[-1] / max(vn[:600].max(), 1e-300)))
        table = all(s.materials() > 0 for s in shots)
        for s in shots:
            s.close()
        bytes32 = sum(path_bytes_2d(physics, *_cells_2d(c), 4, table) * c["time_steps"] for c in cfgs)
        log(f"[batch2d] {physics}: batched fp32 {t_batch:.3f} s (tile rows {rows}, xcd order {xcd})")

        serial = {}
        rec64 = []
        for prec in ("fp32", "fp64") if args.batch_serial else ():
            total = 0.0
            for c in cfgs:
                s = make_solver(c, precision=prec)
                s.run(0)          # uploads
                s.synchronize()
                t0 = time.perf_counter()
                s.run()
                s.synchronize()
                total += time.perf_counter() - t0
                if prec == "fp64":
                    rec64.append(s.receivers())
                s.close()
            serial[prec] = total
            log(f"[batch2d] {physics}: serial {prec} {total:.3f} s")
        if not all(np.isfinite(r).all() for r in rec32):
            raise RuntimeError("non-finite receiver traces in the batch")
        rec = {"shots": len(cfgs), "cell_updates": cell_steps,
               "batched_fp32": {"s": round(t_batch, 4), "GCell_per_s": round(cell_steps / t_batch / 1e9, 3),
                                "shots_per_s": round(len(cfgs) / t_batch, 3),
                                "tile_rows": rows, "xcd_order": xcd, "tiles_per_launch": tiles},
               "roofline": {"bound": "hbm", "kernel": "whole batched step (every shot's kernels)",
                            "achieved": round(bytes32 / t_batch / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(bytes32 / t_batch / 1e9 / HBM_PEAK_GBS, 4), "algorithmic_bytes": bytes32,
                            "material_layout": "table" if table else "per-cell arrays"},
               "velnorm_growth": {"max": float(f"{max(growth):.3e}"),
                                  "median": float(f"{float(np.median(growth)):.3e}"),
                                  "shots_over_1e3": int(sum(g > 1e3 for g in growth))}}
        if args.batch_serial:
            errs = [rel_l2(a, b_) for a, b_ in zip(rec32, rec64)]
            rec.update({
                "serial_fp32": {"s": round(serial["fp32"], 4),
                                "GCell_per_s": round(cell_steps / serial["fp32"] / 1e9, 3)},
                "serial_fp64": {"s": round(serial["fp64"], 4),
                                "GCell_per_s": round(cell_steps / serial["fp64"] / 1e9, 3)},
                "speedup_vs_serial_fp32": round(serial["fp32"] / t_batch, 3),
                "speedup_vs_serial_fp64": round(serial["fp64"] / t_batch, 3),
                "fp32_vs_fp64_trace_rel_l2": {"max": float(f"{max(errs):.3e}"),
                                              "median": float(f"{float(np.median(errs)):.3e}")}})
        out[physics] = rec
    main_p = next(iter(out))
    m = out[main_p]
    return {
        "metric": "GCell-updates/s (batched uneven 2D shots, FP32)",
        "value": m["batched_fp32"]["GCell_per_s"], "unit": "GCell-updates/s", "n_gpus": 1,
        "steps": None, "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded 4-layer anisotropic models per shot, Ricker sources)",
        "config": {"workload": f"batch2d: {args.shots} shots per solver kind, nx 300-800, nz 150-400 "
                               "(+10-cell CPML), 1500-3000 steps each", "kinds": list(out),
                   "elastic_stencil": args.batch_stencil, "elastic_tilt_max_deg": args.batch_tilt,
                   "tilt_note": "symmetry-axis tilts of 0-30 degrees make the reference scheme grow "
                                "exponentially (its c15/c35 coupling); with them shot 28 of 64 blows up "
                                "(velocity norm > 1e25) at step 2685, so the timed shots use vertical axes"},
        "roofline": m["roofline"],
        "cpu_baseline": None,
        "batch": out,
    }


# ------------------------------------------------------------------ FWI, config 4
def bench_fwi2d(args):
    """FWI gradient throughput: 600 x 300 (+10-cell CPML), 16 shots, 3000
    steps, seeded 6-layer anisotropic true model with 10 Gaussian anomalies
    (+-15% permittivity / stiffness), starting model = smoothed true model,
    L2 trace misfit against FP64 data of the true model; checkpoints every
    100 steps.  Timed: the batched FP32 forward + adjoint of all shots
    (seidart_hip.fwi.batch_traces + backward) against the shots run one
    after another with FP64 forward + adjoint.  Checked: the FP32 gradient
    against the FP64 one, and the FP64 gradient against central differences
    (relative step 1e-4) at --fd-cells random interior cells."""
    import numpy as np
    import torch
    from seidart_hip import fwi, models

    out = {}
    for physics in ("em2", "elastic2") if args.fwi_physics == "both" else (args.fwi_physics,):
        t0 = time.perf_counter()
        prob = models.fwi2d(physics, steps=args.fwi_steps, nshots=args.fwi_shots)
        shots = prob["shots"]
        ns, nt = len(shots), args.fwi_steps
        dt = prob["cfg"]["dt"]
        cp = prob["cfg"]["cpml"]
        if physics == "em2":
            names = ("e11", "e13", "e33", "s11", "s13", "s33")

            def coefs(p):
                return fwi.em2_coefficients(p[:3], p[3:])
            # Ex dipole, scaled as the library does by default, from the starting model
            src1 = np.stack([prob["wavelet"], np.zeros(nt)])
            src = np.stack([src1 * fwi.em2_source_scale(prob["start"]["e11"][s["src"][0] + cp - 1,
                                                                             s["src"][1] + cp - 1], dt)
                            for s in shots])
            fd_param = 0
        else:
            names = ("c11", "c13", "c15", "c33", "c35", "c55", "rho", "gamma_x", "gamma_z", "gamma_xz")

            def coefs(p):
                return fwi.el2_coefficients(p[:6], p[6], p[7:], dt)
            # vertical force, scaled by 1/rho at the source as the library does by default
            src = np.zeros((ns, 5, nt))
            for q, s in enumerate(shots):
                src[q, 4] = prob["wavelet"] / prob["start"]["rho"][s["src"][0] + cp - 1, s["src"][1] + cp - 1]
            fd_param = 3   # c33
        src_t = torch.tensor(src)
        # near-offset mute: receivers within --near-offset-mute cells of a
        # shot's source are left out of that shot's misfit (-1: none)
        wts = fwi.near_offset_weights(shots, args.near_offset_mute) if args.near_offset_mute >= 0 else None
        true = [torch.tensor(prob["true"][n]) for n in names]
        start = [torch.tensor(prob["start"][n]) for n in names]
        with torch.no_grad():
            obs = fwi.batch_traces(physics, coefs(true), src_t, shots, precision="fp64")
        t_setup = time.perf_counter() - t0
        log(f"[fwi2d] {physics}: {ns} shots, {prob['cfg']['shape']}, {nt} steps, data {t_setup:.1f} s")

        # batched FP32 gradient
        p32 = [x.clone().requires_grad_(True) for x in start]
        t0 = time.perf_counter()
        J32 = fwi.l2_misfit(fwi.batch_traces(physics, coefs(p32), src_t, shots, precision="fp32",
                                             checkpoint=args.checkpoint), obs, wts)
        t_fwd32 = time.perf_counter() - t0
        J32.backward()
        t32 = time.perf_counter() - t0
        g32 = np.stack([x.grad.numpy() for x in p32])
        log(f"[fwi2d] {physics}: batched fp32 forward {t_fwd32:.3f} s, forward+adjoint {t32:.3f} s")
        rec = {"shots": ns, "steps": nt, "grid": list(prob["cfg"]["shape"]), "checkpoint_every": args.checkpoint,
               "near_offset_mute_cells": args.near_offset_mute,
               "muted_traces": 0 if wts is None else int((wts == 0).sum()),
               "misfit": float(f"{float(J32.detach()):.6e}"),
               "batched_fp32": {"s": round(t32, 4), "forward_s": round(t_fwd32, 4),
                                "shots_per_s": round(ns / t32, 3)}}
        if not args.fwi_serial:
            out[physics] = rec
            continue

        # serial FP64 forward + adjoint
        p64 = [x.clone().requires_grad_(True) for x in start]
        t0 = time.perf_counter()
        J64 = 0.0
        for q in range(ns):
            Jq = fwi.l2_misfit(fwi.traces(physics, coefs(p64), src_t[q], shots[q], precision="fp64",
                                          checkpoint=args.checkpoint), obs[q], None if wts is None else wts[q])
            Jq.backward()
            J64 += float(Jq.detach())
        t64 = time.perf_counter() - t0
        g64 = np.stack([x.grad.numpy() for x in p64])
        log(f"[fwi2d] {physics}: serial fp64 forward+adjoint {t64:.3f} s")
        agree = {n: rel_l2(g32[k], g64[k]) for k, n in enumerate(names) if np.any(g64[k])}

        # central differences of the FP64 misfit at random interior cells
        fd = None
        if args.fd_cells > 0:
            rng = np.random.default_rng(11)
            NX, NZ = prob["cfg"]["shape"]
            cells = [(int(rng.integers(cp + 5, NX - cp - 5)), int(rng.integers(cp + 5, NZ - cp - 5)))
                     for _ in range(args.fd_cells)]
            base = [x.detach() for x in start]
            t0 = time.perf_counter()
            fdv, adv, ad32 = [], [], []
            for (i, j) in cells:
                h = 1e-4 * abs(float(base[fd_param][i, j]))
                vals = []
                for sgn in (1.0, -1.0):
                    pp = [x.clone() for x in base]
                    pp[fd_param][i, j] += sgn * h
                    with torch.no_grad():
                        vals.append(float(fwi.l2_misfit(fwi.batch_traces(physics, coefs(pp), src_t, shots,
                                                                         precision="fp64"), obs, wts)))
                fdv.append((vals[0] - vals[1]) / (2 * h))
                adv.append(g64[fd_param, i, j])
                ad32.append(g32[fd_param, i, j])
            fdv, adv, ad32 = np.array(fdv), np.array(adv), np.array(ad32)
            fd = {"parameter": names[fd_param], "cells": len(cells), "step": "1e-4 relative, central",
                  "fp64_adjoint_vs_fd_rel_l2": float(f"{rel_l2(adv, fdv):.3e}"),
                  "fp32_adjoint_vs_fd_rel_l2": float(f"{rel_l2(ad32, fdv):.3e}"),
                  "s": round(time.perf_counter() - t0, 1)}
            log(f"[fwi2d] {physics}: finite differences {fd}")
        rec.update({
            "misfit_fp64": float(f"{J64:.6e}"),
            "serial_fp64": {"s": round(t64, 4), "shots_per_s": round(ns / t64, 3)},
            "speedup": round(t64 / t32, 3),
            "fp32_vs_fp64_gradient_rel_l2": {k: float(f"{v:.3e}") for k, v in agree.items()},
            "fp32_vs_fp64_gradient_max_rel_l2": float(f"{max(agree.values()):.3e}"),
            "finite_differences": fd,
        })
        out[physics] = rec
    total_shots = sum(v["shots"] for v in out.values())
    total_s = sum(v["batched_fp32"]["s"] for v in out.values())
    return {
        "metric": "FWI gradient shots/s (2D, FP32 batched forward+adjoint; serial FP64 beside it)",
        "value": round(total_shots / total_s, 3), "unit": "shots/s", "n_gpus": 1, "steps": args.fwi_steps,
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded 6-layer anisotropic model + 10 Gaussian anomalies)",
        "config": {"workload": f"fwi2d: 600x300 + 10-cell CPML, {args.fwi_steps} steps, {args.fwi_shots} shots "
                               f"per solver kind, checkpoints every {args.checkpoint} steps, L2 trace misfit "
                               + (f"with a {args.near_offset_mute}-cell near-offset mute" if args.near_offset_mute >= 0
                                  else "over every trace (no near-offset mute)"), "kinds": list(out)},
        "roofline": None, "cpu_baseline": None, "fwi": out,
    }


# ------------------------------------------------------------------ snapshots (drop-in time loop)
def bench_snapshots(args):
    """The drop-in's time loop (seidart_hip.cpmlfdtd, what the seidartfdtd CLI
    runs) on the 256^3 + 12-cell CPML model, FP64, writing FP32 snapshots of
    vx, vy, vz after every step (write_image3 files, 3 x 87.8 MB per step) to
    a tmpfs directory: device-packed FP32 snapshots, pinned double-buffered
    copies, writer threads overlapping the file writes with the next steps.
    The model comes from memory rather than from input files (reading the
    inputs is setup, outside the loop either way)."""
    import shutil
    import tempfile
    import types
    import numpy as np
    from seidart_hip import models
    from seidart_hip.cpmlfdtd import _time_loop
    from seidart_hip.solver import make_solver
    K = args.snapshot_steps
    n = args.n
    cfg = models.elastic3(nx=n, ny=n, nz=n, steps=K, src=(n // 2, n // 2, 20), nrec=min(256, n))
    base = "/dev/shm" if os.path.isdir("/dev/shm") else tempfile.gettempdir()
    work = tempfile.mkdtemp(prefix="sdt_snap_", dir=base)
    log(f"[snapshots] {cfg['shape']} {K} steps -> {work}")
    src = types.SimpleNamespace(time_steps=K, xind=n // 2, yind=n // 2, zind=20)
    s = make_solver(cfg, precision="fp64")
    try:
        s.run(0)
        s.synchronize()
        st = _time_loop(s, work, src, (("vx", "Vx"), ("vy", "Vy"), ("vz", "Vz")), True, True, 1)
        files = len(os.listdir(work))
    finally:
        s.close()
        shutil.rmtree(work, ignore_errors=True)
    ncell = int(np.prod(cfg["shape"]))
    return {"metric": "GCell-updates/s (3D elastic drop-in time loop with FP32 snapshots every step)",
            "value": round(ncell * K / st["loop_s"] / 1e9, 4), "unit": "GCell-updates/s", "steps": K,
            "ms_per_step": round(st["loop_s"] / K * 1e3, 3), "dtype": "f64",
            "config": {"workload": f"seismic3 {n}^3 + 12-cell CPML, vx vy vz FP32 snapshots every step to tmpfs "
                                   f"({base})", "files": files},
            "snapshot_GB_per_s": round(st["bytes"] / st["loop_s"] / 1e9, 2), "bytes_written": st["bytes"]}


# ------------------------------------------------------------------ other solvers
def bench_seismic3_iso(args):
    """The isotropic fast path: a smoothly varying isotropic 256^3 + 12-cell
    CPML model (every cell its own material, so no material table), timed
    with the three isotropic arrays (c11, c12, c44) and again with all 21
    stiffness arrays on the same model."""
    import numpy as np
    from seidart_hip import models

    K, W = args.steps, args.warmup
    n = args.n
    cfg = models.elastic3_iso(nx=n, ny=n, nz=n, steps=W + K, src=(n // 2, n // 2, 20), nrec=min(256, n))
    shape = cfg["shape"]
    ncell = int(np.prod(shape))
    w = 8 if args.precision == "fp64" else 4
    fast = timed_run(cfg, args, 1, isotropic=1)
    general = timed_run(cfg, args, 1, isotropic=0)
    if fast[5] != 1 or general[5] != 0 or fast[4] != 0:
        raise RuntimeError(f"isotropic path not selected as expected: {fast[4:6]}, {general[4:6]}")

    gam = [cfg["arrays"].get(g) for g in ("gamma_x", "gamma_y", "gamma_z", "gamma_xy", "gamma_xz", "gamma_yz")]
    distinct = []
    for g in gam:
        g = np.zeros(1) if g is None else g
        if not any(g.shape == h.shape and np.array_equal(g, h) for h in distinct):
            distinct.append(g)
    n_ig = len(distinct)

    def summary(run, layout):
        elapsed, ms_stress, ms_vel, launches = run[:4]
        path = el3_path_bytes(shape, cfg["cpml"], w, layout, n_ig=n_ig)
        avg_stress_s = ms_stress / max(launches, 1) / 1e3
        return {"value": round(ncell * K / elapsed / 1e9, 4), "ms_per_step": round(elapsed / K * 1e3, 4),
                "hbm_fraction_step": round(path["step"] * K / elapsed / 1e9 / HBM_PEAK_GBS, 4),
                "stress_pass_ms": round(avg_stress_s * 1e3, 4),
                "velocity_pass_ms": round(ms_vel / max(launches, 1), 4),
                "stress_achieved": round(path["stress"] / avg_stress_s / 1e9, 1), "stress_bytes": path["stress"]}

    f, g = summary(fast, "iso"), summary(general, "dense")
    return {
        "metric": "GCell-updates/s (3D elastic CPML-FDTD, isotropic per-cell model, "
                  + ("FP64)" if args.precision == "fp64" else "FP32)"),
        "value": f["value"], "unit": "GCell-updates/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": f["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (smooth isotropic model with 0.5 % per-cell jitter, Ricker source)",
        "config": {"workload": f"seismic3 isotropic {n}^3 + {cfg['cpml']}-cell CPML ({shape[0]}^3 padded), "
                               "per-cell arrays, attenuation", "cells": ncell, "precision": args.precision},
        "hbm_fraction_step": f["hbm_fraction_step"],
        "roofline": {"bound": "hbm", "kernel": "el3 stress pass, isotropic arrays",
                     "achieved": f["stress_achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(f["stress_achieved"] / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes": f["stress_bytes"], "avg_launch_ms": f["stress_pass_ms"]},
        "cpu_baseline": None, "distinct_attenuation_arrays": n_ig,
        "isotropic_path": f, "general_path_same_model": g,
        "speedup_vs_general_path": round(g["ms_per_step"] / f["ms_per_step"], 3),
    }


EL25 = dict(F=9, M=32, PSI=21, C=6)   # complex fields and memory variables (2 values each), real materials


def bench_seismic25(args):
    """The 2.5D elastic solver on a 2000 x 1000 (+10-cell CPML) x-z grid: a
    seeded layered model of rotated ice tensors (21 stiffness entries per
    cell) timed with the material table and with the per-cell arrays, and a
    smooth isotropic model timed with the isotropic arrays and with all 21."""
    import numpy as np
    from seidart_hip import models

    K, W = args.steps, args.warmup
    w = 8 if args.precision == "fp64" else 4
    nx, nz = 2000, 1000
    runs = {}
    layered = models.elastic25(nx=nx, nz=nz, steps=W + K, src=(nx // 2, 40), nrec=64)
    smooth = models.elastic25_iso(nx=nx, nz=nz, steps=W + K, src=(nx // 2, 40), nrec=64)
    for name, cfg, table, iso in (
            ("anisotropic", layered, args.material_table, 1),
            ("anisotropic_per_cell_arrays", layered, 0, 1),
            ("isotropic", smooth, 1, 1),
            ("isotropic_general_path", smooth, 1, 0)):
        elapsed, ms_s, ms_v, launches, nmat, used, _ = timed_run(cfg, args, table, isotropic=iso)
        shape = cfg["shape"]
        n = int(np.prod(shape))
        n_int = (shape[0] - 2 * cfg["cpml"]) * (shape[1] - 2 * cfg["cpml"])
        M = (3 + 6 + 5) if used else EL25["M"]
        step_bytes = w * (n * (2 * 2 * EL25["F"] + M) + (n - n_int) * (2 * 2 * EL25["PSI"] + EL25["C"]))
        runs[name] = {"value": round(n * K / elapsed / 1e9, 4), "ms_per_step": round(elapsed / K * 1e3, 4),
                      "hbm_fraction_step_contract": round(step_bytes * K / elapsed / 1e9 / HBM_PEAK_GBS, 4),
                      "stress_pass_ms": round(ms_s / max(launches, 1), 4),
                      "velocity_pass_ms": round(ms_v / max(launches, 1), 4), "isotropic_arrays": bool(used),
                      "materials_in_table": nmat, "step_bytes_contract": step_bytes}
        log(f"[seismic25] {name}: {runs[name]}")
    a = runs["anisotropic"]
    return {
        "metric": "GCell-updates/s (2.5D elastic CPML-FDTD, " + ("FP64)" if args.precision == "fp64" else "FP32)"),
        "value": a["value"], "unit": "GCell-updates/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": a["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded layered rotated-ice model; smooth isotropic model)",
        "config": {"workload": f"seismic25 {nx}x{nz} + 10-cell CPML, complex fields, ky=0.05",
                   "precision": args.precision,
                   "material_layout": f"table ({a['materials_in_table']} materials, 1-byte id per cell)"
                                      if a["materials_in_table"] else "per-cell arrays"},
        "roofline": {"bound": "hbm", "kernel": "whole step (contract byte model)", "achieved": None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                     "contract_frac": a["hbm_fraction_step_contract"], "traffic": None},
        "cpu_baseline": None, "runs": runs,
    }


EM25 = dict(F=6, M=12, PSI=8, C=6)   # complex E, H and memory variables; real coefficients


def bench_em25(args):
    """The 2.5D EM solver (electromag25) on a 2000 x 1000 (+10-cell CPML) x-z
    grid, ky = 5: seeded layers of SPD permittivity and conductivity, timed
    with the material table and with the 12 per-cell coefficient arrays."""
    import numpy as np
    from seidart_hip import models

    K, W = args.steps, args.warmup
    w = 8 if args.precision == "fp64" else 4
    nx, nz = 2000, 1000
    cfg = models.em25(nx=nx, nz=nz, steps=W + K, src=(nx // 2, 40), nrec=64, ky=5.0)
    shape = cfg["shape"]
    n = int(np.prod(shape))
    n_int = (shape[0] - 2 * cfg["cpml"]) * (shape[1] - 2 * cfg["cpml"])
    step_bytes = w * (n * (2 * 2 * EM25["F"] + EM25["M"]) + (n - n_int) * (2 * 2 * EM25["PSI"] + EM25["C"]))
    runs = {}
    for name, table in (("material_table", args.material_table), ("per_cell_arrays", 0)):
        elapsed, ms_h, ms_e, launches, nmat, _, _ = timed_run(cfg, args, table)
        runs[name] = {"value": round(n * K / elapsed / 1e9, 4), "ms_per_step": round(elapsed / K * 1e3, 4),
                      "hbm_fraction_step_contract": round(step_bytes * K / elapsed / 1e9 / HBM_PEAK_GBS, 4),
                      "h_pass_ms": round(ms_h / max(launches, 1), 4), "e_pass_ms": round(ms_e / max(launches, 1), 4),
                      "materials_in_table": nmat}
        log(f"[electromag25] {name}: {runs[name]}")
    a = runs["material_table"]
    return {
        "metric": "GCell-updates/s (2.5D EM CPML-FDTD, " + ("FP64)" if args.precision == "fp64" else "FP32)"),
        "value": a["value"], "unit": "GCell-updates/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": a["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded 4-layer SPD permittivity/conductivity, Ricker dipole)",
        "config": {"workload": f"electromag25 {nx}x{nz} + {cfg['cpml']}-cell CPML, complex fields, ky=5",
                   "precision": args.precision,
                   "material_layout": f"table ({a['materials_in_table']} materials, 1-byte id per cell)"
                                      if a["materials_in_table"] else "per-cell arrays"},
        "roofline": {"bound": "hbm", "kernel": "whole step (contract byte model)", "achieved": None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                     "contract_frac": a["hbm_fraction_step_contract"], "traffic": None,
                     "step_bytes_contract": step_bytes},
        "cpu_baseline": None, "runs": runs,
    }


# ------------------------------------------------------------------ sub-records
def _sub(args, **over):
    a = copy.copy(args)
    for k, v in over.items():
        setattr(a, k, v)
    return a


def _trim(line):
    """A sub-record: the full line without the keys shared with the main line."""
    return {k: v for k, v in line.items() if k not in ("n_gpus", "higher_is_better", "scaling", "vs_baseline")}


def subrecords(args, rec64):
    """The other benchmark configurations, each timed on its own, in one dict.
    A configuration that fails is recorded with its error (and the run goes
    on to the next)."""
    K2, W2 = max(args.steps, 100), max(args.warmup, 5)
    jobs = [
        ("seismic3_fp32", lambda: bench_el3(_sub(args, precision="fp32", cpu_steps=0, compare_layouts=0),
                                            rec64=rec64)[0]),
        ("seismic2_200x200_fp64", lambda: bench_seismic2(_sub(args, steps=1000, warmup=20, precision="fp64"))),
        ("electromag3_fp64", lambda: bench_em3(_sub(args, steps=K2, warmup=W2, precision="fp64"))),
        ("electromag3_fp32", lambda: bench_em3(_sub(args, steps=K2, warmup=W2, precision="fp32"))),
        ("batch2d_64shots_fp32", lambda: bench_batch2d(args)),
        ("fwi2d_600x300_16shots", lambda: bench_fwi2d(_sub(args, fd_cells=args.sub_fd_cells))),
        ("seismic3_snapshots_every_step", lambda: bench_snapshots(args)),
    ]
    want = set(args.subrecords.split(",")) if args.subrecords not in ("all", "none") else None
    out = {}
    for name, fn in jobs:
        if args.subrecords == "none" or (want is not None and not any(name.startswith(w) for w in want)):
            continue
        t0 = time.perf_counter()
        try:
            out[name] = _trim(fn())
        except Exception as e:   # recorded, not hidden: the line shows which configuration failed
            out[name] = {"error": f"{type(e).__name__}: {e}"}
            log(f"[{name}] FAILED: {e}")
        out[name]["wall_s"] = round(time.perf_counter() - t0, 1)
        log(f"[{name}] done in {out[name]['wall_s']} s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--n", type=int, default=256, help="unpadded cells per axis")
    ap.add_argument("--cpu-steps", type=int, default=16,
                    help="oracle steps for the 3D elastic CPU baseline (0 = skip every CPU baseline)")
    ap.add_argument("--em3-cpu-steps", type=int, default=8, help="oracle steps for the 3D EM CPU baseline")
    ap.add_argument("--zchunk", type=int, default=0, help="planes per z-marching workgroup (0 = auto)")
    ap.add_argument("--stagger", type=int, default=1, help="staggered array starts (1) or plain (0)")
    ap.add_argument("--tile-rows", type=int, default=0, help="rows per interior workgroup tile (0 = auto)")
    ap.add_argument("--material-table", type=int, default=1,
                    help="1: material ids + table when the model allows (default), 0: per-cell arrays")
    ap.add_argument("--fused", type=int, default=-1, choices=[-1, 0, 1],
                    help="3D EM: one kernel per step (1), one per half step (0), or by precision (-1)")
    ap.add_argument("--concurrent", type=int, default=0,
                    help="1: CPML-shell kernels on a second stream beside the interior kernels")
    ap.add_argument("--xcd-remap", type=int, default=0,
                    help="3D elastic block order: bit 0 interior, bit 1 shell (XCD-contiguous tile ranges)")
    ap.add_argument("--compare-layouts", type=int, default=1,
                    help="also time the per-cell-array layout and report it beside the main line")
    ap.add_argument("--subrecords", default="all",
                    help="seismic3 workload: the other configurations timed into the line's 'configs' "
                         "('all', 'none', or a comma list of name prefixes: seismic3_fp32, seismic2, "
                         "electromag3, batch2d, fwi2d, seismic3_snapshots)")
    ap.add_argument("--sub-fd-cells", type=int, default=8,
                    help="finite-difference cells of the FWI sub-record (--workload fwi2d: --fd-cells)")
    ap.add_argument("--snapshot-steps", type=int, default=20,
                    help="snapshots workload: steps of the drop-in time loop (3 snapshot files each)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="time a plain 2 GiB copy kernel (HBM counter calibration) and exit")
    ap.add_argument("--workload", default="seismic3",
                    choices=["seismic3", "seismic3iso", "seismic25", "seismic2", "electromag3", "electromag25",
                             "batch2d", "fwi2d", "snapshots"],
                    help="seismic3: the 3D elastic 256^3 case (default); electromag3: the 3D EM "
                         "384x384x192 case; seismic25 / electromag25: the 2.5D solvers on 2000x1000; "
                         "seismic2: the 2D elastic 200x200 case beside the CPU oracle; seismic3iso: an "
                         "isotropic 256^3 model; batch2d: uneven multi-shot 2D batches against serial "
                         "runs; fwi2d: batched FWI gradients")
    ap.add_argument("--shots", type=int, default=64, help="batch2d: shots per solver kind")
    ap.add_argument("--fwi-shots", type=int, default=16, help="fwi2d: shots per solver kind")
    ap.add_argument("--fwi-physics", default="both", choices=["both", "em2", "elastic2"])
    ap.add_argument("--fwi-steps", type=int, default=3000)
    ap.add_argument("--checkpoint", type=int, default=100, help="fwi2d: steps between saved forward states")
    ap.add_argument("--fd-cells", type=int, default=50, help="fwi2d: cells of the finite-difference check")
    ap.add_argument("--near-offset-mute", type=int, default=-1,
                    help="fwi2d: receivers within this many cells of a shot's source are left out of its "
                         "misfit (-1: keep every trace)")
    ap.add_argument("--fwi-serial", type=int, default=1,
                    help="fwi2d: also run the serial FP64 forward+adjoint and the checks (0 = batched only)")
    ap.add_argument("--batch-physics", default="both", choices=["both", "elastic2", "em2"])
    ap.add_argument("--batch-serial", type=int, default=1,
                    help="batch2d: also time the serial single-handle runs (0 = batched run only)")
    ap.add_argument("--batch-stencil", default="reference", choices=["reference", "corrected"],
                    help="batch2d elastic velocity differences")
    ap.add_argument("--batch-tilt", type=float, default=0.0,
                    help="batch2d elastic: largest symmetry-axis tilt in degrees; the reference's "
                         "uncentred c15/c35 coupling grows without bound over 1500-3000 steps for "
                         "tilted axes (with 0-30 degrees, shot 28 of the 64 passes the 1e25 blow-up "
                         "threshold at step 2685), so the default keeps the axes vertical")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (one time-stepped grid per GPU); "
              "--gpus must be 1", file=sys.stderr)
        sys.exit(2)

    single = {"electromag3": bench_em3, "batch2d": bench_batch2d, "fwi2d": bench_fwi2d,
              "seismic3iso": bench_seismic3_iso, "seismic25": bench_seismic25, "electromag25": bench_em25,
              "seismic2": bench_seismic2, "snapshots": bench_snapshots}
    if args.workload in single:
        print(json.dumps(single[args.workload](args)), flush=True)
        return
    t0 = time.perf_counter()
    line, rec = bench_el3(args)
    log(f"[seismic3] main line done in {time.perf_counter() - t0:.1f} s")
    if args.subrecords != "none":
        line["configs"] = subrecords(args, rec if args.precision == "fp64" else None)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
