"""Throughput of the 3D anisotropic elastic CPML-FDTD solver (seismic3) on one
MI355X: 256^3 cells + 12-cell CPML (280^3 padded), dx = 0.5 m, eight seeded
rotated-ice layers with attenuation, Ricker 100 Hz vertical force, 256
receivers, FP64 (the reference's arithmetic).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|fp32]

One step = one time step over the whole padded grid (stress kernel, velocity
kernel, source/receiver tail).  Prints one JSON line.  `value` is
GCell-updates/s over the K timed steps with every input resident in HBM.
`roofline` prices the dominant kernel (the fused stress kernel) by its
algorithmic bytes per launch against the 8 TB/s HBM peak, timed with HIP
events on the solver's stream during the timed region.  `cpu_baseline` runs
the CPU oracle (a restatement of cpmlfdtd.f08's seismic3 loops, OpenMP) on
the same model for a few steps on the host cores.
"""
from __future__ import annotations

import argparse
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

# per-cell counts of the algorithmic byte model (fields read+written once,
# material arrays read once; CPML cells add their live memory recursions read
# + written and the six coefficient arrays)
EL3 = dict(F=9, M=28, PSI=29, C=6)
EL3_STRESS = dict(reads_v=3, rw_sigma=6, mat=27, psi=20, C=6)
# isotropic-form arrays: c11, c12, c44, rho and six attenuation divisors per
# cell (the stress pass reads all but rho)
EL3_ISO_M, EL3_ISO_STRESS_M = 10, 9


def algorithmic_bytes(shape, cpml, w, table=False, iso=False, n_ig=6):
    """Bytes per step and per stress pass.  table=False is the model with one
    array per material parameter (M = 28 per cell); table=True replaces those
    by the one-byte material id read once per pass; iso=True counts the
    isotropic-form arrays (M = 10); n_ig < 6 counts that many distinct
    attenuation-divisor arrays (equal ones are held and read once)."""
    nx, ny, nz = shape
    n = nx * ny * nz
    n_int = max(nx - 2 * cpml, 0) * max(ny - 2 * cpml, 0) * max(nz - 2 * cpml, 0)
    n_pml = n - n_int
    M, Ms = (0, 0) if table else ((EL3_ISO_M, EL3_ISO_STRESS_M) if iso else (EL3["M"], EL3_STRESS["mat"]))
    if not table:
        M, Ms = M - (6 - n_ig), Ms - (6 - n_ig)
    step = w * (n * (2 * EL3["F"] + M) + n_pml * (2 * EL3["PSI"] + EL3["C"])) + (2 * n if table else 0)
    s = EL3_STRESS
    stress = w * (n * (s["reads_v"] + 2 * s["rw_sigma"] + Ms) + n_pml * (2 * s["psi"] + s["C"])) + (n if table else 0)
    return step, stress


def timed_run(cfg, args, material_table, isotropic=1, precision=None):
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
    s.set_arrays(cfg["arrays"])
    for comp, series in cfg["sources"].items():
        series = np.concatenate([np.asarray(series, dtype=np.float64), np.zeros(KT)])
        s.set_source(comp, series)
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
    rec = s.receivers()
    s.close()
    if not np.all(np.isfinite(rec)):
        raise RuntimeError("non-finite receiver traces")
    return elapsed, ms_stress, ms_vel, launches, nmat, iso, rec


def fp32_trace_error(cfg, args, rec32):
    """Relative L2 difference of the FP32 receiver traces from an FP64 run of
    the same workload and step count."""
    import numpy as np
    rec64 = timed_run(cfg, args, args.material_table, precision="fp64")[6]
    return float(f"{np.linalg.norm((rec32 - rec64).ravel()) / np.linalg.norm(rec64.ravel()):.3e}")


def measured_traffic():
    """HBM bytes per stress-pass launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration),
    or None when no summary is present."""
    path = os.path.join(ROOT, "profiles", "el3_seismic3_280cube_fp64_hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


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


EM3 = dict(F=6, M=12, PSI=12, C=6)


def bench_em3(args):
    """3D anisotropic EM (electromag3) on 384 x 384 x 192 + 10-cell CPML
    (404 x 404 x 212 padded), dx = 0.05 m, five seeded SPD permittivity /
    conductivity layers, 100 MHz Ricker Ex dipole, 128 receivers."""
    import numpy as np
    from seidart_hip import models
    K, W = args.steps, args.warmup
    cfg = models.em3(steps=W + K)
    shape = cfg["shape"]
    ncell = int(np.prod(shape))
    w = 8 if args.precision == "fp64" else 4
    elapsed, ms_h, ms_e, launches, nmat, _, rec = timed_run(cfg, args, args.material_table)
    nx, ny, nz = shape
    cp = cfg["cpml"]
    n_pml = ncell - (nx - 2 * cp) * (ny - 2 * cp) * (nz - 2 * cp)
    step_bytes = w * (ncell * (2 * EM3["F"] + EM3["M"]) + n_pml * (2 * EM3["PSI"] + EM3["C"]))
    # E pass: E read+written, H read, inverse permittivity and conductivity
    # (12), CPML cells 6 memory variables + 6 coefficients
    e_bytes = w * (ncell * (6 + 3 + 12) + n_pml * (12 + 6))
    avg_e = ms_e / max(launches, 1) / 1e3
    line = {
        "metric": f"GCell-updates/s (3D EM CPML-FDTD, {args.precision.upper()})",
        "value": round(ncell * K / elapsed / 1e9, 4), "unit": "GCell-updates/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": round(elapsed / K * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded 5-layer SPD permittivity/conductivity, Ricker Ex dipole)",
        "config": {"workload": f"electromag3 384x384x192 + {cp}-cell CPML ({nx}x{ny}x{nz} padded), dx=0.05 m",
                   "cells": ncell, "precision": args.precision,
                   "material_layout": f"table ({nmat} materials, 1-byte id per cell)" if nmat else "per-cell arrays"},
        "hbm_fraction_step": round(step_bytes * K / elapsed / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": {"bound": "hbm", "kernel": "em3 E update", "achieved": round(e_bytes / avg_e / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(e_bytes / avg_e / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "algorithmic_bytes": e_bytes, "avg_launch_ms": round(avg_e * 1e3, 4),
                     "h_kernel_avg_ms": round(ms_h / max(launches, 1), 4)},
        "cpu_baseline": None,
    }
    if args.precision == "fp32":
        line["fp32_vs_fp64_trace_rel_l2"] = fp32_trace_error(cfg, args, rec)
    print(json.dumps(line), flush=True)


# 2D per-cell counts of the same byte model (el2: 5 fields, 10 material
# words, 12 CPML memory variables, 6 coefficient arrays; em2: 3 fields, 7
# precomputed update coefficients, 4 memory variables)
EL2 = dict(F=5, M=10, PSI=12, C=6)
EM2 = dict(F=3, M=7, PSI=4, C=6)


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

    def rel_l2(a, b):
        return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))

    out = {}
    for physics in ("elastic2", "em2") if args.batch_physics == "both" else (args.batch_physics,):
        t_gen = time.perf_counter()
        cfgs = models.batch2d(physics, nshots=args.shots, stencil=args.batch_stencil,
                              tilt=(0.0, args.batch_tilt))
        t_gen = time.perf_counter() - t_gen
        cell_steps = sum(_cells_2d(c)[0] * c["time_steps"] for c in cfgs)
        P = EL2 if physics == "elastic2" else EM2
        bytes32 = 4 * sum((_cells_2d(c)[0] * (2 * P["F"] + P["M"]) + _cells_2d(c)[1] * (2 * P["PSI"] + P["C"]))
                          * c["time_steps"] for c in cfgs)
        print(f"[batch2d] {physics}: {len(cfgs)} shots, {cell_steps / 1e9:.2f} G cell-steps "
              f"(models {t_gen:.1f} s)", flush=True)

        shots = [make_solver(c, precision="fp32") for c in cfgs]
        b = Batch(shots)
        t0 = time.perf_counter()
        b.run()
        t_batch = time.perf_counter() - t0
        rows, tiles, xcd = b.info("tile_rows"), b.info("tiles"), b.info("xcd_order")
        b.close()
        rec32 = [s.receivers() for s in shots]
        for s in shots:
            s.close()
        print(f"[batch2d] {physics}: batched fp32 {t_batch:.3f} s (tile rows {rows}, xcd order {xcd})",
              flush=True)

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
            print(f"[batch2d] {physics}: serial {prec} {total:.3f} s", flush=True)
        if not all(np.isfinite(r).all() for r in rec32):
            raise RuntimeError("non-finite receiver traces in the batch")
        if not args.batch_serial:
            out[physics] = {"shots": len(cfgs), "cell_updates": cell_steps,
                            "batched_fp32": {"s": round(t_batch, 4),
                                             "GCell_per_s": round(cell_steps / t_batch / 1e9, 3)}}
            print(json.dumps(out[physics]), flush=True)
            continue
        errs = [rel_l2(a, b_) for a, b_ in zip(rec32, rec64)]
        out[physics] = {
            "shots": len(cfgs), "cell_updates": cell_steps,
            "batched_fp32": {"s": round(t_batch, 4), "GCell_per_s": round(cell_steps / t_batch / 1e9, 3),
                             "shots_per_s": round(len(cfgs) / t_batch, 3),
                             "hbm_fraction": round(bytes32 / t_batch / 1e9 / HBM_PEAK_GBS, 4),
                             "tile_rows": rows, "xcd_order": xcd, "tiles_per_launch": tiles},
            "serial_fp32": {"s": round(serial["fp32"], 4),
                            "GCell_per_s": round(cell_steps / serial["fp32"] / 1e9, 3)},
            "serial_fp64": {"s": round(serial["fp64"], 4),
                            "GCell_per_s": round(cell_steps / serial["fp64"] / 1e9, 3)},
            "speedup_vs_serial_fp32": round(serial["fp32"] / t_batch, 3),
            "speedup_vs_serial_fp64": round(serial["fp64"] / t_batch, 3),
            "fp32_vs_fp64_trace_rel_l2": {"max": float(f"{max(errs):.3e}"),
                                          "median": float(f"{float(np.median(errs)):.3e}")},
        }
    if not args.batch_serial:
        return
    main_p = next(iter(out))
    m = out[main_p]
    line = {
        "metric": "GCell-updates/s (batched uneven 2D shots, FP32)",
        "value": m["batched_fp32"]["GCell_per_s"], "unit": "GCell-updates/s", "n_gpus": 1,
        "steps": None, "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded 4-layer anisotropic models per shot, Ricker sources)",
        "config": {"workload": f"batch2d: {args.shots} shots per solver kind, nx 300-800, nz 150-400 "
                               "(+10-cell CPML), 1500-3000 steps each", "kinds": list(out),
                   "elastic_stencil": args.batch_stencil, "elastic_tilt_max_deg": args.batch_tilt},
        "roofline": {"bound": "hbm", "kernel": "whole batched step (byte model over every shot)",
                     "achieved": round(m["batched_fp32"]["hbm_fraction"] * HBM_PEAK_GBS, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": m["batched_fp32"]["hbm_fraction"],
                     "traffic": None},
        "cpu_baseline": None,
        "batch": out,
    }
    print(json.dumps(line), flush=True)


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

    def rel(a, b):
        return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))

    out = {}
    for physics in ("em2", "elastic2") if args.fwi_physics == "both" else (args.fwi_physics,):
        t0 = time.perf_counter()
        prob = models.fwi2d(physics, steps=args.fwi_steps, nshots=args.shots if args.shots != 64 else 16)
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
        true = [torch.tensor(prob["true"][n]) for n in names]
        start = [torch.tensor(prob["start"][n]) for n in names]
        with torch.no_grad():
            obs = fwi.batch_traces(physics, coefs(true), src_t, shots, precision="fp64")
        t_setup = time.perf_counter() - t0
        print(f"[fwi2d] {physics}: {ns} shots, {prob['cfg']['shape']}, {nt} steps, data {t_setup:.1f} s", flush=True)

        # batched FP32 gradient
        p32 = [x.clone().requires_grad_(True) for x in start]
        torch.cuda.synchronize() if torch.cuda.is_available() else None
        t0 = time.perf_counter()
        J32 = fwi.l2_misfit(fwi.batch_traces(physics, coefs(p32), src_t, shots, precision="fp32",
                                             checkpoint=args.checkpoint), obs)
        t_fwd32 = time.perf_counter() - t0
        J32.backward()
        t32 = time.perf_counter() - t0
        g32 = np.stack([x.grad.numpy() for x in p32])
        print(f"[fwi2d] {physics}: batched fp32 forward {t_fwd32:.3f} s, forward+adjoint {t32:.3f} s", flush=True)

        # serial FP64 forward + adjoint
        p64 = [x.clone().requires_grad_(True) for x in start]
        t0 = time.perf_counter()
        J64 = 0.0
        for q in range(ns if args.fwi_serial else 0):
            Jq = fwi.l2_misfit(fwi.traces(physics, coefs(p64), src_t[q], shots[q], precision="fp64",
                                          checkpoint=args.checkpoint), obs[q])
            Jq.backward()
            J64 += float(Jq.detach())
        t64 = time.perf_counter() - t0
        if not args.fwi_serial:
            print(f"[fwi2d] {physics}: batched fp32 only", flush=True)
            continue
        g64 = np.stack([x.grad.numpy() for x in p64])
        print(f"[fwi2d] {physics}: serial fp64 forward+adjoint {t64:.3f} s", flush=True)
        agree = {n: rel(g32[k], g64[k]) for k, n in enumerate(names) if np.any(g64[k])}
        # the same with the source near field muted (7 x 7 cells around each
        # source cell), where the FP32 wavefields' rounding concentrates
        mute = np.ones(g64.shape[1:], bool)
        for s in shots:
            si, sj = s["src"][0] + cp - 1, s["src"][1] + cp - 1
            mute[max(si - 3, 0):si + 4, max(sj - 3, 0):sj + 4] = False
        agree_muted = {n: rel(g32[k][mute], g64[k][mute]) for k, n in enumerate(names) if np.any(g64[k])}

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
                                                                         precision="fp64"), obs)))
                fdv.append((vals[0] - vals[1]) / (2 * h))
                adv.append(g64[fd_param, i, j])
                ad32.append(g32[fd_param, i, j])
            fdv, adv, ad32 = np.array(fdv), np.array(adv), np.array(ad32)
            fd = {"parameter": names[fd_param], "cells": len(cells), "step": "1e-4 relative, central",
                  "fp64_adjoint_vs_fd_rel_l2": float(f"{rel(adv, fdv):.3e}"),
                  "fp32_adjoint_vs_fd_rel_l2": float(f"{rel(ad32, fdv):.3e}"),
                  "s": round(time.perf_counter() - t0, 1)}
            print(f"[fwi2d] {physics}: finite differences {fd}", flush=True)
        out[physics] = {
            "shots": ns, "steps": nt, "grid": list(prob["cfg"]["shape"]), "checkpoint_every": args.checkpoint,
            "misfit": float(f"{float(J32.detach()):.6e}"), "misfit_fp64": float(f"{J64:.6e}"),
            "batched_fp32": {"s": round(t32, 4), "forward_s": round(t_fwd32, 4), "shots_per_s": round(ns / t32, 3)},
            "serial_fp64": {"s": round(t64, 4), "shots_per_s": round(ns / t64, 3)},
            "speedup": round(t64 / t32, 3),
            "fp32_vs_fp64_gradient_rel_l2": {k: float(f"{v:.3e}") for k, v in agree.items()},
            "fp32_vs_fp64_gradient_rel_l2_source_muted": {k: float(f"{v:.3e}") for k, v in agree_muted.items()},
            "finite_differences": fd,
        }
    if not args.fwi_serial:
        return
    total_shots = sum(v["shots"] for v in out.values())
    total_s = sum(v["batched_fp32"]["s"] for v in out.values())
    line = {
        "metric": "FWI gradient shots/s (2D, FP32 batched forward+adjoint; serial FP64 beside it)",
        "value": round(total_shots / total_s, 3), "unit": "shots/s", "n_gpus": 1, "steps": args.fwi_steps,
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded 6-layer anisotropic model + 10 Gaussian anomalies)",
        "config": {"workload": f"fwi2d: 600x300 + 10-cell CPML, {args.fwi_steps} steps, 16 shots per solver "
                               f"kind, checkpoints every {args.checkpoint} steps", "kinds": list(out)},
        "roofline": None, "cpu_baseline": None, "fwi": out,
    }
    print(json.dumps(line), flush=True)


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

    def summary(run, iso):
        elapsed, ms_stress, ms_vel, launches = run[:4]
        step_bytes, stress_bytes = algorithmic_bytes(shape, cfg["cpml"], w, iso=iso, n_ig=n_ig)
        avg_stress_s = ms_stress / max(launches, 1) / 1e3
        return {"value": round(ncell * K / elapsed / 1e9, 4), "ms_per_step": round(elapsed / K * 1e3, 4),
                "hbm_fraction_step": round(step_bytes * K / elapsed / 1e9 / HBM_PEAK_GBS, 4),
                "stress_pass_ms": round(avg_stress_s * 1e3, 4),
                "velocity_pass_ms": round(ms_vel / max(launches, 1), 4),
                "stress_achieved": round(stress_bytes / avg_stress_s / 1e9, 1), "stress_bytes": stress_bytes}

    f, g = summary(fast, True), summary(general, False)
    line = {
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
    print(json.dumps(line), flush=True)


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
                      "hbm_fraction_step": round(step_bytes * K / elapsed / 1e9 / HBM_PEAK_GBS, 4),
                      "stress_pass_ms": round(ms_s / max(launches, 1), 4),
                      "velocity_pass_ms": round(ms_v / max(launches, 1), 4), "isotropic_arrays": bool(used),
                      "materials_in_table": nmat, "step_bytes": step_bytes}
        print(f"[seismic25] {name}: {runs[name]}", flush=True)
    a = runs["anisotropic"]
    line = {
        "metric": "GCell-updates/s (2.5D elastic CPML-FDTD, " + ("FP64)" if args.precision == "fp64" else "FP32)"),
        "value": a["value"], "unit": "GCell-updates/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": a["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded layered rotated-ice model; smooth isotropic model)",
        "config": {"workload": f"seismic25 {nx}x{nz} + 10-cell CPML, complex fields, ky=0.05",
                   "precision": args.precision,
                   "material_layout": f"table ({a['materials_in_table']} materials, 1-byte id per cell)"
                                      if a["materials_in_table"] else "per-cell arrays"},
        "roofline": {"bound": "hbm", "kernel": "whole step (byte model)", "achieved": None, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": a["hbm_fraction_step"], "traffic": None},
        "cpu_baseline": None, "runs": runs,
    }
    print(json.dumps(line), flush=True)


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
                      "hbm_fraction_step": round(step_bytes * K / elapsed / 1e9 / HBM_PEAK_GBS, 4),
                      "h_pass_ms": round(ms_h / max(launches, 1), 4), "e_pass_ms": round(ms_e / max(launches, 1), 4),
                      "materials_in_table": nmat}
        print(f"[electromag25] {name}: {runs[name]}", flush=True)
    a = runs["material_table"]
    line = {
        "metric": "GCell-updates/s (2.5D EM CPML-FDTD, " + ("FP64)" if args.precision == "fp64" else "FP32)"),
        "value": a["value"], "unit": "GCell-updates/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": a["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded 4-layer SPD permittivity/conductivity, Ricker dipole)",
        "config": {"workload": f"electromag25 {nx}x{nz} + {cfg['cpml']}-cell CPML, complex fields, ky=5",
                   "precision": args.precision,
                   "material_layout": f"table ({a['materials_in_table']} materials, 1-byte id per cell)"
                                      if a["materials_in_table"] else "per-cell arrays"},
        "roofline": {"bound": "hbm", "kernel": "whole step (byte model)", "achieved": None, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": a["hbm_fraction_step"], "traffic": None, "step_bytes": step_bytes},
        "cpu_baseline": None, "runs": runs,
    }
    print(json.dumps(line), flush=True)


def bench_seismic2(args):
    """The reference's CPU-runnable case: 2D elastic, 200 x 200 (+10-cell
    CPML), two seeded tilted ice layers, 1000 steps.  About 10 MB per step,
    so the GPU step is bound by launch latency (captured step graphs), not
    by HBM; the C oracle runs the whole case beside it."""
    import numpy as np
    from seidart_hip import models

    K, W = args.steps, args.warmup
    cfg = models.elastic2(steps=W + K)
    shape = cfg["shape"]
    n = int(np.prod(shape))
    elapsed, ms_s, ms_v, launches, nmat, _, _ = timed_run(cfg, args, args.material_table)
    cpu = None
    if args.cpu_steps > 0:
        from oracle.oracle_ctypes import run_oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
        ref = models.elastic2(steps=K)
        t1 = time.perf_counter()
        run_oracle(ref, nthreads=threads)
        cpu_s = time.perf_counter() - t1
        cpu = {"value": round(n * K / cpu_s / 1e9, 5), "unit": "GCell-updates/s", "cores": threads, "kind": "port",
               "sample": f"the whole {K}-step {shape[0]}x{shape[1]} run through the C oracle "
                         f"(FP64, OpenMP, {threads} threads), {cpu_s:.2f} s"}
    line = {
        "metric": "GCell-updates/s (2D elastic CPML-FDTD, " + ("FP64)" if args.precision == "fp64" else "FP32)"),
        "value": round(n * K / elapsed / 1e9, 4), "unit": "GCell-updates/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(elapsed / K * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded 2-layer tilted ice model, Ricker source)",
        "config": {"workload": f"seismic2 200x200 + 10-cell CPML ({shape[0]}x{shape[1]} padded), {K} steps",
                   "precision": args.precision,
                   "material_layout": f"table ({nmat} materials, 1-byte id per cell)" if nmat else "per-cell arrays"},
        "roofline": {"bound": "launch latency", "achieved": None, "peak": None, "unit": None, "frac": None,
                     "traffic": None,
                     "note": "about 10 MB per step: 1.2 us at 8 TB/s, below one kernel launch"},
        "stress_pass_ms": round(ms_s / max(launches, 1), 4), "velocity_pass_ms": round(ms_v / max(launches, 1), 4),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--n", type=int, default=256, help="unpadded cells per axis")
    ap.add_argument("--cpu-steps", type=int, default=8, help="oracle steps for the CPU baseline, about 10 s on 16 host threads (0 = skip)")
    ap.add_argument("--zchunk", type=int, default=0, help="planes per z-marching workgroup (0 = auto)")
    ap.add_argument("--stagger", type=int, default=1, help="staggered array starts (1) or plain (0)")
    ap.add_argument("--tile-rows", type=int, default=0, help="rows per interior workgroup tile (0 = auto)")
    ap.add_argument("--material-table", type=int, default=1,
                    help="1: material ids + table when the model allows (default), 0: per-cell arrays")
    ap.add_argument("--concurrent", type=int, default=0,
                    help="1: CPML-shell kernels on a second stream beside the interior kernels")
    ap.add_argument("--xcd-remap", type=int, default=0,
                    help="3D elastic block order: bit 0 interior, bit 1 shell (XCD-contiguous tile ranges)")
    ap.add_argument("--compare-layouts", type=int, default=1,
                    help="also time the per-cell-array layout and report it beside the main line")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="time a plain 2 GiB copy kernel (HBM counter calibration) and exit")
    ap.add_argument("--workload", default="seismic3",
                    choices=["seismic3", "seismic3iso", "seismic25", "seismic2", "electromag3", "electromag25",
                             "batch2d", "fwi2d"],
                    help="seismic3: the 3D elastic 256^3 case (default); electromag3: the 3D EM "
                         "384x384x192 case; seismic25 / electromag25: the 2.5D solvers on 2000x1000; "
                         "seismic2: the 2D elastic 200x200 case beside the CPU oracle; seismic3iso: an "
                         "isotropic 256^3 model; batch2d: uneven multi-shot 2D batches against serial "
                         "runs; fwi2d: batched FWI gradients")
    ap.add_argument("--shots", type=int, default=64, help="batch2d: shots per solver kind (fwi2d: 16)")
    ap.add_argument("--fwi-physics", default="both", choices=["both", "em2", "elastic2"])
    ap.add_argument("--fwi-steps", type=int, default=3000)
    ap.add_argument("--checkpoint", type=int, default=100, help="fwi2d: steps between saved forward states")
    ap.add_argument("--fd-cells", type=int, default=50, help="fwi2d: cells of the finite-difference check")
    ap.add_argument("--fwi-serial", type=int, default=1,
                    help="fwi2d: also run the serial FP64 forward+adjoint and the checks (0 = batched only, "
                         "for profiling; prints no JSON line)")
    ap.add_argument("--batch-physics", default="both", choices=["both", "elastic2", "em2"])
    ap.add_argument("--batch-serial", type=int, default=1,
                    help="batch2d: also time the serial single-handle runs (0 = batched run only, "
                         "for profiling; prints the batched figures only)")
    ap.add_argument("--batch-stencil", default="reference", choices=["reference", "corrected"],
                    help="batch2d elastic velocity differences")
    ap.add_argument("--batch-tilt", type=float, default=0.0,
                    help="batch2d elastic: largest symmetry-axis tilt in degrees; the reference's "
                         "uncentred c15/c35 coupling grows without bound over 1500-3000 steps for "
                         "tilted axes, so the default keeps the axes vertical")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (one time-stepped grid per GPU); "
              "--gpus must be 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "electromag3":
        bench_em3(args)
        return
    if args.workload == "batch2d":
        bench_batch2d(args)
        return
    if args.workload == "fwi2d":
        bench_fwi2d(args)
        return
    if args.workload == "seismic3iso":
        bench_seismic3_iso(args)
        return
    if args.workload == "seismic25":
        bench_seismic25(args)
        return
    if args.workload == "electromag25":
        bench_em25(args)
        return
    if args.workload == "seismic2":
        bench_seismic2(args)
        return

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

    main_run = timed_run(cfg, args, args.material_table)
    dense_run = timed_run(cfg, args, 0) if (args.material_table and args.compare_layouts) else None
    elapsed, ms_stress, ms_vel, launches, nmat, _, rec = main_run

    value = ncell * K / elapsed / 1e9
    step_bytes, stress_bytes = algorithmic_bytes(shape, cfg["cpml"], w)
    table = bool(args.material_table) and nmat > 0
    step_bytes_path, stress_bytes_path = algorithmic_bytes(shape, cfg["cpml"], w, table=table)
    avg_stress_s = ms_stress / max(launches, 1) / 1e3
    avg_vel_s = ms_vel / max(launches, 1) / 1e3
    achieved = stress_bytes / avg_stress_s / 1e9
    achieved_path = stress_bytes_path / avg_stress_s / 1e9
    traffic = measured_traffic() if args.precision == "fp64" and n == 256 else None

    cpu = None
    if args.cpu_steps > 0:
        from oracle.oracle_ctypes import run_oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
        t1 = time.perf_counter()
        run_oracle(cfg, steps=args.cpu_steps, nthreads=threads)
        cpu_s = time.perf_counter() - t1
        cpu = {"value": round(ncell * args.cpu_steps / cpu_s / 1e9, 5), "unit": "GCell-updates/s",
               "cores": threads, "kind": "port",
               "sample": f"{args.cpu_steps} time steps of the same {shape[0]}x{shape[1]}x{shape[2]} model "
                         f"through the C oracle (FP64, OpenMP, {threads} threads), {cpu_s:.1f} s"}

    line = {
        "metric": "GCell-updates/s (3D elastic CPML-FDTD, FP64)" if args.precision == "fp64"
                  else "GCell-updates/s (3D elastic CPML-FDTD, FP32)",
        "value": round(value, 4),
        "unit": "GCell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(elapsed / K * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded 8-layer rotated-ice model, Ricker source)",
        "config": {"workload": f"seismic3 {n}^3 + {cfg['cpml']}-cell CPML ({shape[0]}^3 padded), "
                               f"dx=0.5 m, 8 layers, attenuation, 256 receivers",
                   "cells": ncell, "precision": args.precision,
                   "material_layout": f"table ({nmat} materials, 1-byte id per cell)" if table
                                      else "per-cell arrays",
                   "zchunk": args.zchunk, "stagger": args.stagger, "tile_rows": args.tile_rows,
                   "concurrent_shell": bool(args.concurrent)},
        "hbm_fraction_step": round(step_bytes * K / elapsed / 1e9 / HBM_PEAK_GBS, 4),
        "hbm_fraction_step_path_bytes": round(step_bytes_path * K / elapsed / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": {"bound": "hbm", "kernel": "el3 stress pass (interior + CPML-shell kernels)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes": stress_bytes,
                     "path_algorithmic_bytes": stress_bytes_path,
                     "path_achieved": round(achieved_path, 1),
                     "path_frac": round(achieved_path / HBM_PEAK_GBS, 4),
                     "avg_launch_ms": round(avg_stress_s * 1e3, 4),
                     "velocity_kernel_avg_ms": round(avg_vel_s * 1e3, 4)},
        "cpu_baseline": cpu,
        "setup_s": round(t_gen, 2),
    }
    if dense_run is not None:
        e2, s2, v2, l2, _, _, _ = dense_run
        line["per_cell_arrays_layout"] = {
            "value": round(ncell * K / e2 / 1e9, 4), "ms_per_step": round(e2 / K * 1e3, 4),
            "hbm_fraction_step": round(step_bytes * K / e2 / 1e9 / HBM_PEAK_GBS, 4),
            "stress_pass_ms": round(s2 / max(l2, 1), 4), "velocity_pass_ms": round(v2 / max(l2, 1), 4)}
    if args.precision == "fp32":
        line["fp32_vs_fp64_trace_rel_l2"] = fp32_trace_error(cfg, args, rec)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
