"""File-based drop-in of the reference solvers.

    seismic2(domain, source, density_method, SINGLE_OUTPUT)     cpmlfdtd.f08:57
    seismic25(domain, source)                                   cpmlfdtd.f08:600
    seismic3(domain, source, density_method, SINGLE_OUTPUT)     cpmlfdtd.f08:1341
    electromag2(domain, source, SINGLE_OUTPUT)                  cpmlfdtd.f08:2317
    electromag25(domain, source)                                cpmlfdtd.f08:2764
    electromag3(domain, source, SINGLE_OUTPUT)                  cpmlfdtd.f08:3353

Same argument order and meaning as the Fortran subroutines: `domain` holds
padded sizes (src/fortran/main.f08:69-71 pads before the call), every input
is read from the working directory under the reference's file names, and a
snapshot of every velocity / E component is written each step under the
reference's names and formats.  3D elastic also writes velocity_norm.dat and
stress_norm.dat (cpmlfdtd.f08:2282-2283).  The 2.5D solvers read KY,
SINGLE_OUTPUT, FDTD_LEGACY_OUTPUT, FDTD_STEPS_PER_BLOCK (and seismic25
DENSITY_METHOD) from the environment as the reference does and write zstd
blocks by default.

Differences from the reference, by design:
* blow-up raises seidart_hip.UnstableError (the reference prints and
  `stop`s with exit status 0, cpmlfdtd.f08:567);
* electromag3 writes its 3D snapshots with the write_image3 format; the
  reference's call to write_image stops at the end of step 1 for 3D grids
  (seidartio.f08:315-324);
* electromag25 with source type 'pw' raises ValueError: the reference sets up
  the plane-wave geometry but has no injection loop for it and adds its
  never-loaded point-source series instead (cpmlfdtd.f08:2937-2996,
  3260-3266).

Plane-wave sources: the elastic solvers take 'pw', 'pws', 'pw_s', 'pw-s'
(S) and 'pwp', 'pw_p', 'pw-p' (P) after trimming (plane_wave_source.f08:
20-35); electromag2 / electromag3 take exactly 'pw' (trailing blanks
ignored, Fortran string equality, cpmlfdtd.f08:2436, 3523) and run any other
type as a point source.  No source series is read for a plane wave.
"""
from __future__ import annotations

import dataclasses
import os

import numpy as np

from . import fortran_io as fio
from .solver import Solver

PLANE_WAVE_TYPES = {"pw", "pws", "pw_s", "pw-s", "pwp", "pw_p", "pw-p"}   # plane_wave_source.f08:20-27
P_WAVE_TYPES = {"pwp", "pw_p", "pw-p"}                                  # plane_wave_source.f08:29-35
WAVELETS = {"gaus0": 0, "gaus1": 1, "gaus2": 2}                         # plane_wave_source.f08:150-160


@dataclasses.dataclass
class Domain:
    """Domain_Type, src/fortran/seidart_types.f08:13-21."""
    dim: float
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    cpml: int
    nmats: int = 0
    image_file: str = ""


@dataclasses.dataclass
class Source:
    """Source_Type, src/fortran/seidart_types.f08:23-33."""
    dt: float
    time_steps: int
    x: float = 0.0
    y: float = 0.0
    z: float = 0.0
    xind: int = 1
    yind: int = 1
    zind: int = 1
    source_frequency: float = 0.0
    x_z_rotation: float = 0.0
    x_y_rotation: float = 0.0
    y_z_rotation: float = 0.0
    amplitude: float = 1.0
    source_type: str = ""
    source_wavelet: str = ""


def _read(workdir, name, shape):
    return fio.read_array(os.path.join(workdir, name + ".dat"), shape)


def _series(workdir, name, n):
    return fio.read_record(os.path.join(workdir, name + ".dat"), np.float64, n)


def _source_type(source):
    return (source.source_type or "").strip()


def _wavelet(source):
    """`select case` on the wavelet name; anything else is gaus0
    (plane_wave_source.f08:150-160)."""
    return WAVELETS.get((source.source_wavelet or "").rstrip(), 0)


def _plane_wave(s, source, kind):
    s.set_plane_wave(kind, (source.x_z_rotation, source.x_y_rotation, source.y_z_rotation),
                     source.source_frequency, source.amplitude, _wavelet(source))


def _elastic_plane_wave(s, source):
    """Elastic solvers: a plane wave when the trimmed source type is one of
    PLANE_WAVE_TYPES (is_plane_wave_source); True when it is."""
    st = _source_type(source)
    if st not in PLANE_WAVE_TYPES:
        return False
    _plane_wave(s, source, "p" if st in P_WAVE_TYPES else "s")
    return True


def _em_plane_wave(s, source):
    """EM solvers: a plane wave when source_type == 'pw' (Fortran equality:
    trailing blanks ignored); True when it is."""
    if (source.source_type or "").rstrip() != "pw":
        return False
    _plane_wave(s, source, "s")
    return True


def _cpml_inputs(s, workdir, shape):
    for stem in ("kappa", "alpha", "acoef", "bcoef"):
        for sfx in ("_cpml", "_half_cpml"):
            s.set_array(stem + sfx, _read(workdir, stem + sfx, shape))


class _SnapshotPipeline:
    """Per-step snapshots written while the next steps run: each snapshot
    step packs every component on the device (FP32 when single) into one of
    `depth` sets of pinned host buffers; the files of a set are written in
    16 MiB chunks by a pool of writer threads (positioned writes into the
    pre-sized file, then a rename) while the solver advances, and a set is
    reused once its files are written.  The files are those of
    write_image2/3 (seidartio.f08:366-426), byte for byte."""

    def __init__(self, s, workdir, source, names, single, is3d, depth=3, writers=None):
        from concurrent.futures import ThreadPoolExecutor
        from .solver import PinnedArray
        self.s, self.workdir, self.source, self.names = s, workdir, source, names
        self.single, self.is3d = single, is3d
        dt = np.float32 if single else np.float64
        self.sets = [[PinnedArray(s.ncell, dt) for _ in names] for _ in range(depth)]
        self.busy = [[] for _ in range(depth)]
        self.chunks = ThreadPoolExecutor(max_workers=writers or max(2, min(16, os.cpu_count() or 2)))
        self.pool = ThreadPoolExecutor(max_workers=len(names))   # per-file completion and rename
        self.count = 0
        self.bytes = 0

    def emit(self, it):
        q = self.count % len(self.sets)
        for f in self.busy[q]:
            f.result()
        bufs = self.sets[q]
        for (fname, _), b in zip(self.names, bufs):
            self.s.snapshot(fname, b.array, self.single)
        self.s.synchronize()
        src = self.source
        self.busy[q] = [self.pool.submit(fio.start_record(
                            os.path.join(self.workdir, fio.snapshot_name(
                                channel, it, src.xind, src.zind, src.yind if self.is3d else None)),
                            b.array, self.chunks))
                        for (_, channel), b in zip(self.names, bufs)]
        self.count += 1
        self.bytes += sum(b.array.nbytes + 8 for b in bufs)

    def close(self):
        try:
            for fs in self.busy:
                for f in fs:
                    f.result()
        finally:
            self.pool.shutdown()
            self.chunks.shutdown()


def _time_loop(s, workdir, source, names, single, is3d, snapshot_every):
    """Run every step, writing the snapshots of every `snapshot_every`-th
    step (0: none).  Returns {"loop_s", "snapshots", "bytes"}."""
    import time
    t0 = time.perf_counter()
    pipe = _SnapshotPipeline(s, workdir, source, names, single, is3d) if snapshot_every else None
    try:
        it = 0
        while it < source.time_steps:
            n = min(snapshot_every, source.time_steps - it) if snapshot_every else source.time_steps - it
            s.run(n)
            it += n
            if pipe and it % snapshot_every == 0:
                pipe.emit(it)
    finally:
        if pipe:
            pipe.close()
    return {"loop_s": time.perf_counter() - t0, "snapshots": pipe.count if pipe else 0,
            "bytes": pipe.bytes if pipe else 0}


def seismic2(domain, source, density_method="arithmetic", SINGLE_OUTPUT=True, *, workdir=".",
             precision="fp64", device=0, snapshot_every=1):
    """2D elastic; inputs c11 c13 c15 c33 c35 c55 rho gamma_{x,z,xz} CPML
    initialconditionV{x,z} and seismicsource{x,z} ('ac') or
    seismicsource{xx,xz,zz} (cpmlfdtd.f08:159-266); writes Vx/Vz snapshots."""
    st = _source_type(source)
    shape = (domain.nx, domain.nz)
    with Solver("elastic2", shape, (domain.dx, domain.dz), domain.cpml, source.dt, source.time_steps,
                (source.xind, source.zind), precision=precision, density_method=density_method.strip(),
                device=device) as s:
        for name in ("c11", "c13", "c15", "c33", "c35", "c55", "rho", "gamma_x", "gamma_z", "gamma_xz"):
            s.set_array(name, _read(workdir, name, shape))
        _cpml_inputs(s, workdir, shape)
        s.set_array("ic_x", _read(workdir, "initialconditionVx", shape))
        s.set_array("ic_z", _read(workdir, "initialconditionVz", shape))
        if not _elastic_plane_wave(s, source):
            comps = ("x", "z") if st == "ac" else ("xx", "xz", "zz")
            for c in comps:
                s.set_source(c, _series(workdir, "seismicsource" + c, source.time_steps))
        return _time_loop(s, workdir, source, (("vx", "Vx"), ("vz", "Vz")), SINGLE_OUTPUT, False, snapshot_every)


def seismic3(domain, source, density_method="arithmetic", SINGLE_OUTPUT=True, *, workdir=".",
             precision="fp64", device=0, snapshot_every=1):
    """3D elastic; inputs c11..c66 rho gamma_* CPML initialconditionV{x,y,z}
    and seismicsource{x,y,z} ('ac') or seismicsource{xx,xy,xz,yy,yz,zz}
    (cpmlfdtd.f08:1501-1645); writes Vx/Vy/Vz snapshots and the norm files."""
    st = _source_type(source)
    shape = (domain.nx, domain.ny, domain.nz)
    with Solver("elastic3", shape, (domain.dx, domain.dy, domain.dz), domain.cpml, source.dt,
                source.time_steps, (source.xind, source.yind, source.zind), precision=precision,
                density_method=density_method.strip(), device=device) as s:
        for name in ("c11", "c12", "c13", "c14", "c15", "c16", "c22", "c23", "c24", "c25", "c26",
                     "c33", "c34", "c35", "c36", "c44", "c45", "c46", "c55", "c56", "c66", "rho",
                     "gamma_x", "gamma_z", "gamma_y", "gamma_xz", "gamma_yz", "gamma_xy"):
            s.set_array(name, _read(workdir, name, shape))
        _cpml_inputs(s, workdir, shape)
        for c in "xyz":
            s.set_array("ic_" + c, _read(workdir, "initialconditionV" + c, shape))
        if not _elastic_plane_wave(s, source):
            comps = ("x", "y", "z") if st == "ac" else ("xx", "xy", "xz", "yy", "yz", "zz")
            for c in comps:
                s.set_source(c, _series(workdir, "seismicsource" + c, source.time_steps))
        stats = _time_loop(s, workdir, source, (("vx", "Vx"), ("vy", "Vy"), ("vz", "Vz")), SINGLE_OUTPUT, True,
                           snapshot_every)
        vn, sn = s.norms()
        fio.write_record(os.path.join(workdir, "velocity_norm.dat"), vn)
        fio.write_record(os.path.join(workdir, "stress_norm.dat"), sn)
        return stats


def electromag2(domain, source, SINGLE_OUTPUT=True, *, workdir=".", precision="fp64", device=0,
                snapshot_every=1):
    """2D EM; inputs e11 e13 e33 s11 s13 s33 CPML initialconditionE{x,z} and
    electromagneticsource{x,z} (cpmlfdtd.f08:2413-2506); writes Ex/Ez."""
    shape = (domain.nx, domain.nz)
    with Solver("em2", shape, (domain.dx, domain.dz), domain.cpml, source.dt, source.time_steps,
                (source.xind, source.zind), precision=precision, device=device) as s:
        for name in ("e11", "e13", "e33", "s11", "s13", "s33"):
            s.set_array(name, _read(workdir, name, shape))
        _cpml_inputs(s, workdir, shape)
        s.set_array("ic_x", _read(workdir, "initialconditionEx", shape))
        s.set_array("ic_z", _read(workdir, "initialconditionEz", shape))
        if not _em_plane_wave(s, source):
            for c in ("x", "z"):
                s.set_source(c, _series(workdir, "electromagneticsource" + c, source.time_steps))
        return _time_loop(s, workdir, source, (("ex", "Ex"), ("ez", "Ez")), SINGLE_OUTPUT, False, snapshot_every)


def electromag3(domain, source, SINGLE_OUTPUT=True, *, workdir=".", precision="fp64", device=0,
                snapshot_every=1):
    """3D EM; inputs e11..e33 s11..s33 CPML initialconditionE{x,y,z} and
    electromagneticsource{x,y,z} (cpmlfdtd.f08:3481-3618); writes Ex/Ey/Ez."""
    shape = (domain.nx, domain.ny, domain.nz)
    with Solver("em3", shape, (domain.dx, domain.dy, domain.dz), domain.cpml, source.dt,
                source.time_steps, (source.xind, source.yind, source.zind), precision=precision,
                device=device) as s:
        for name in ("e11", "e12", "e13", "e22", "e23", "e33", "s11", "s12", "s13", "s22", "s23", "s33"):
            s.set_array(name, _read(workdir, name, shape))
        _cpml_inputs(s, workdir, shape)
        for c in "xyz":
            s.set_array("ic_" + c, _read(workdir, "initialconditionE" + c, shape))
        if not _em_plane_wave(s, source):
            for c in "xyz":
                s.set_source(c, _series(workdir, "electromagneticsource" + c, source.time_steps))
        return _time_loop(s, workdir, source, (("ex", "Ex"), ("ey", "Ey"), ("ez", "Ez")), SINGLE_OUTPUT, True,
                          snapshot_every)


def _env_logical(name, default=True):
    """list-directed read of a LOGICAL (T/F, .TRUE./.FALSE.); unreadable or
    absent -> default (cpmlfdtd.f08:688-695)."""
    v = os.environ.get(name)
    if v is None:
        return default
    t = v.strip().lstrip(".").upper()
    if t.startswith("T"):
        return True
    if t.startswith("F"):
        return False
    return default


def _env_ky():
    """KY, 0 when absent or unreadable (cpmlfdtd.f08:697-703)."""
    try:
        return float(os.environ.get("KY", "0").strip().replace("d", "e").replace("D", "e"))
    except ValueError:
        return 0.0


def _env_density():
    """DENSITY_METHOD, harmonic by default and for unknown values
    (cpmlfdtd.f08:705-716)."""
    v = os.environ.get("DENSITY_METHOD", "").strip()
    return v if v in ("harmonic", "geometric", "arithmetic", "none") else "harmonic"


def seismic25(domain, source, *, workdir=".", precision="fp64", device=0, ky=None, density_method=None,
              SINGLE_OUTPUT=None, legacy_output=None, steps_per_block=None):
    """2.5D elastic; inputs c11..c66 rho gamma_* CPML, complex128
    initialconditionV{x,y,z} and seismicsource{x,y,z} ('ac') or
    seismicsource{xx,xy,xz,yy,yz,zz} (cpmlfdtd.f08:763-872).  Keyword
    arguments left as None take the reference's environment variables.
    Output: S25.<first>-<last>.<xind>.<yind>.<zind>.blk.zst blocks of the
    real parts of vx, vy, vz (default), or with FDTD_LEGACY_OUTPUT=TRUE
    per-step Vx/Vy/Vz images of the real parts (cpmlfdtd.f08:1295-1305)."""
    st = _source_type(source)
    ky = _env_ky() if ky is None else float(ky)
    density_method = _env_density() if density_method is None else density_method
    single = _env_logical("SINGLE_OUTPUT") if SINGLE_OUTPUT is None else bool(SINGLE_OUTPUT)
    if legacy_output is None:
        legacy_output = os.environ.get("FDTD_LEGACY_OUTPUT", "").strip() == "TRUE"
    shape = (domain.nx, domain.nz)
    writer = None
    if not legacy_output:
        spb = steps_per_block or fio.steps_per_block_2d(domain.nx, domain.nz, domain.cpml)
        writer = fio.BlockWriter2D(domain.nx, domain.nz, domain.cpml, spb, source.time_steps, "S25",
                                   source.xind, source.yind, source.zind, workdir)
    with Solver("elastic25", shape, (domain.dx, domain.dz), domain.cpml, source.dt, source.time_steps,
                (source.xind, source.zind), precision=precision, density_method=density_method,
                device=device, ky=ky) as s:
        for name in ("c11", "c12", "c13", "c14", "c15", "c16", "c22", "c23", "c24", "c25", "c26",
                     "c33", "c34", "c35", "c36", "c44", "c45", "c46", "c55", "c56", "c66", "rho",
                     "gamma_x", "gamma_z", "gamma_y", "gamma_xz", "gamma_yz", "gamma_xy"):
            s.set_array(name, _read(workdir, name, shape))
        _cpml_inputs(s, workdir, shape)
        for c in "xyz":
            s.set_array("ic_" + c, fio.read_array(os.path.join(workdir, "initialconditionV" + c + ".dat"),
                                                  shape, np.complex128))
        if not _elastic_plane_wave(s, source):
            comps = ("x", "y", "z") if st == "ac" else ("xx", "xy", "xz", "yy", "yz", "zz")
            for c in comps:
                s.set_source(c, _series(workdir, "seismicsource" + c, source.time_steps))
        for it in range(1, source.time_steps + 1):
            s.run(1)
            vx, vy, vz = s.field("vx"), s.field("vy"), s.field("vz")
            if writer is not None:
                writer.add(vx, vy, vz)
            if legacy_output:
                for a, ch in ((vx, "Vx"), (vy, "Vy"), (vz, "Vz")):
                    path = os.path.join(workdir, fio.snapshot_name(ch, it, source.xind, source.zind))
                    fio.write_record(path, a.real, np.float32 if single else np.float64)
        if writer is not None:
            writer.flush()
    return writer.written if writer is not None else []


def electromag25(domain, source, *, workdir=".", precision="fp64", device=0, ky=None, SINGLE_OUTPUT=None,
                 legacy_output=None, steps_per_block=None):
    """2.5D EM; inputs e11..e33 s11..s33 CPML, complex128
    initialconditionE{x,y,z} and electromagneticsource{x,y,z}
    (cpmlfdtd.f08:2896-3036).  Output: EM25.<first>-<last>.<xind>.<yind>.
    <zind>.blk.zst blocks of the real parts of Ex, Ey, Ez, or with
    FDTD_LEGACY_OUTPUT=TRUE per-step Ex/Ey/Ez images (cpmlfdtd.f08:3313-3327)."""
    if (source.source_type or "").rstrip() == "pw":
        raise ValueError("electromag25 has no plane-wave source: the reference sets up its geometry but injects "
                         "nothing and adds its never-loaded point-source series (cpmlfdtd.f08:2937-2996)")
    ky = _env_ky() if ky is None else float(ky)
    single = _env_logical("SINGLE_OUTPUT") if SINGLE_OUTPUT is None else bool(SINGLE_OUTPUT)
    if legacy_output is None:
        legacy_output = os.environ.get("FDTD_LEGACY_OUTPUT", "").strip() == "TRUE"
    shape = (domain.nx, domain.nz)
    writer = None
    if not legacy_output:
        spb = steps_per_block or fio.steps_per_block_2d(domain.nx, domain.nz, domain.cpml)
        writer = fio.BlockWriter2D(domain.nx, domain.nz, domain.cpml, spb, source.time_steps, "EM25",
                                   source.xind, source.yind, source.zind, workdir)
    with Solver("em25", shape, (domain.dx, domain.dz), domain.cpml, source.dt, source.time_steps,
                (source.xind, source.zind), precision=precision, device=device, ky=ky) as s:
        for name in ("e11", "e12", "e13", "e22", "e23", "e33", "s11", "s12", "s13", "s22", "s23", "s33"):
            s.set_array(name, _read(workdir, name, shape))
        _cpml_inputs(s, workdir, shape)
        for c in "xyz":
            s.set_array("ic_" + c, fio.read_array(os.path.join(workdir, "initialconditionE" + c + ".dat"),
                                                  shape, np.complex128))
        for c in "xyz":
            s.set_source(c, _series(workdir, "electromagneticsource" + c, source.time_steps))
        for it in range(1, source.time_steps + 1):
            s.run(1)
            ex, ey, ez = s.field("ex"), s.field("ey"), s.field("ez")
            if writer is not None:
                writer.add(ex, ey, ez)
            if legacy_output:
                for a, ch in ((ex, "Ex"), (ey, "Ey"), (ez, "Ez")):
                    path = os.path.join(workdir, fio.snapshot_name(ch, it, source.xind, source.zind))
                    fio.write_record(path, a.real, np.float32 if single else np.float64)
        if writer is not None:
            writer.flush()
    return writer.written if writer is not None else []
