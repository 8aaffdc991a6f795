"""In-memory solver handle over the C-ABI.

A `Solver` owns one device handle: inputs are numpy arrays of the padded grid
(Fortran order, shape (nx, nz) in 2D or (nx, ny, nz) in 3D), outputs are
numpy arrays of the same layout.  The file-based drop-in of the reference
(`seidart_hip.cpmlfdtd`) and the CLI are built on top of it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

PHYSICS = {"elastic2": L.ELASTIC2, "elastic25": L.ELASTIC25, "elastic3": L.ELASTIC3, "em2": L.EM2,
           "em25": L.EM25, "em3": L.EM3}


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class PinnedArray:
    """A 1-D numpy array over pinned (page-locked) host memory from
    sdt_host_alloc, freed when this object is collected.  Device-to-host
    snapshot copies into it run asynchronously."""

    def __init__(self, n, dtype):
        dtype = np.dtype(dtype)
        p = C.c_void_p()
        L.check(L.lib.sdt_host_alloc(int(n) * dtype.itemsize, C.byref(p)))
        self._p = p
        buf = (C.c_char * (int(n) * dtype.itemsize)).from_address(p.value)
        self.array = np.frombuffer(buf, dtype=dtype, count=int(n))

    def __del__(self):
        if getattr(self, "_p", None):
            self.array = None
            L.lib.sdt_host_free(self._p)
            self._p = None


def flatten_coefficients(coef):
    """(ncoef,) + grid per-cell coefficients -> the C-ABI's layout: one
    Fortran-order grid after the other, contiguous FP64."""
    c = np.asarray(coef, dtype=np.float64)
    return np.ascontiguousarray(np.stack([np.asfortranarray(x).ravel(order="F") for x in c]).ravel())


class Solver:
    """One simulation on one GPU.

    physics: "elastic2" | "elastic25" | "elastic3" | "em2" | "em25" | "em3"
    shape: padded (nx, nz) or (nx, ny, nz)
    spacing: (dx, dz) or (dx, dy, dz)
    src: 1-based unpadded source indices (xind, zind) or (xind, yind, zind),
         as Source_Type%xind.. (src/fortran/seidart_types.f08:23-33)
    ky: y wavenumber of the 2.5D solvers (the reference's KY environment
        variable, src/fortran/cpmlfdtd.f08:697-703, 2861-2867); their fields,
        initial conditions and receiver samples are complex128
    """

    def __init__(self, physics, shape, spacing, cpml, dt, time_steps, src, *,
                 precision="fp64", density_method="arithmetic", stencil="reference", device=0,
                 ky=0.0):
        if physics not in PHYSICS:
            raise ValueError(f"unknown physics {physics!r}")
        self.physics = physics
        self.is3d = physics.endswith("3")
        self.shape = tuple(int(s) for s in shape)
        if len(self.shape) != (3 if self.is3d else 2):
            raise ValueError("shape rank does not match physics")
        p = L.Params()
        p.abi_version = L.ABI_VERSION
        p.physics = PHYSICS[physics]
        p.precision = {"fp64": L.FP64, "fp32": L.FP32}[precision]
        if self.is3d:
            p.nx, p.ny, p.nz = self.shape
            p.dx, p.dy, p.dz = (float(s) for s in spacing)
            p.xind, p.yind, p.zind = (int(s) for s in src)
        else:
            p.nx, p.nz = self.shape
            p.ny = 1
            p.dx, p.dz = (float(s) for s in spacing)
            p.dy = 1.0
            p.xind, p.zind = (int(s) for s in src)
            p.yind = 1
        p.cpml = int(cpml)
        p.dt = float(dt)
        p.time_steps = int(time_steps)
        p.density_code = L.DENSITY_CODES[density_method]
        p.stencil = {"reference": L.STENCIL_REFERENCE, "corrected": L.STENCIL_CORRECTED}[stencil]
        p.device = int(device)
        p.ky = float(ky)
        self.params = p
        self.complex = physics in ("elastic25", "em25")
        self.ncell = int(np.prod(self.shape))
        self.time_steps = int(time_steps)
        self.fields = L.FIELD_IDS_ELASTIC if physics.startswith("elastic") else L.FIELD_IDS_EM
        self._h = C.c_void_p()
        L.check(L.lib.sdt_create(C.byref(p), C.byref(self._h)))
        self.nrec = 0

    # ---- lifetime ---------------------------------------------------------
    def close(self):
        """Free the handle's device memory.  Raises SdtError (and keeps the
        handle) while a live Batch holds it."""
        if self._h:
            L.check(L.lib.sdt_destroy(self._h), self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- inputs -----------------------------------------------------------
    def _grid_array(self, a):
        a = np.asarray(a, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"array shape {a.shape} != grid {self.shape}")
        return np.asfortranarray(a).ravel(order="F")

    def _complex_grid(self, a):
        a = np.asarray(a, dtype=np.complex128)
        if a.shape != self.shape:
            raise ValueError(f"array shape {a.shape} != grid {self.shape}")
        return np.ascontiguousarray(a.ravel(order="F")).view(np.float64)

    def set_array(self, name, a):
        """name: reference input file stem, e.g. 'c11', 'rho', 'kappa_cpml', 'ic_x'."""
        if self.complex and name.startswith("ic_"):
            flat = self._complex_grid(a)
        else:
            flat = np.ascontiguousarray(self._grid_array(a))
        L.check(L.lib.sdt_set_array(self._h, L.ARRAY_IDS[name], _dptr(flat), flat.size), self._h)

    def set_arrays(self, arrays):
        for k, v in arrays.items():
            self.set_array(k, v)

    def set_source(self, component, series):
        s = np.ascontiguousarray(np.asarray(series, dtype=np.float64).ravel())
        L.check(L.lib.sdt_set_source(self._h, L.SOURCE_IDS[component], _dptr(s), s.size), self._h)

    def set_coefficients(self, coef):
        """2D EM / 2D elastic: the per-cell update coefficients, shape
        (ncoef,) + grid (EM: aEx, bEx, bEz, dEx, eEx, dEz, eEz; elastic: c11,
        c13, c15, c33, c35, c55, rho, 1/(1+gamma_x dt), 1/(1+gamma_z dt),
        1/(1+gamma_xz dt)), in place of the model arrays; needs
        set_option('raw_source', 1).  A 1-D array is taken as the output of
        flatten_coefficients (shots sharing one model flatten it once)."""
        c = np.asarray(coef, dtype=np.float64)
        if c.ndim == 1:
            ncell = int(np.prod(self.shape))
            if c.size == 0 or c.size % ncell:
                raise ValueError(f"flat coefficient length {c.size} is not a multiple of {ncell} cells")
            flat, ncoef = np.ascontiguousarray(c), c.size // ncell
        else:
            if c.shape[1:] != self.shape:
                raise ValueError(f"coefficient shape {c.shape} does not match grid {self.shape}")
            flat, ncoef = flatten_coefficients(c), c.shape[0]
        L.check(L.lib.sdt_set_coefficients(self._h, _dptr(flat), flat.size), self._h)
        self.ncoef = ncoef

    def adjoint(self, trace_grad):
        """After run() over all time_steps with the 'checkpoint' option:
        trace_grad (time_steps, n_receivers, 3) = dJ/d(receivers()).  Returns
        (dJ/d coefficients (ncoef,) + grid, dJ/d source series (9, time_steps))."""
        g = np.ascontiguousarray(np.asarray(trace_grad, dtype=np.float64))
        if g.shape != (self.time_steps, self.nrec, 3):
            raise ValueError(f"trace_grad shape {g.shape} != {(self.time_steps, self.nrec, 3)}")
        nc = self.info("coefficients")
        cg = np.empty((nc, self.ncell), dtype=np.float64)
        sg = np.empty((L.NUM_SRC, self.time_steps), dtype=np.float64)
        L.check(L.lib.sdt_adjoint(self._h, _dptr(g), g.size, _dptr(cg), cg.size, _dptr(sg), sg.size), self._h)
        return np.stack([row.reshape(self.shape, order="F") for row in cg]), sg

    def set_plane_wave(self, kind, angles, f0, amplitude, wavelet=0):
        """A plane-wave source in place of the point source.  Elastic
        solvers: kind "s" (polarised across the direction: source types pw,
        pws, pw_s, pw-s) or "p" (along it: pwp, pw_p, pw-p).  2D and 3D EM:
        kind "s" only (source type pw; E across the direction, H = E / eta
        along p x e); the 2.5D EM solver has none.  angles = (x_z, x_y, y_z)
        rotations in degrees; wavelet 0 gaus0, 1 gaus1, 2 gaus2."""
        k = {"s": L.PW_S, "p": L.PW_P}[kind]
        axz, axy, ayz = (float(a) for a in angles)
        L.check(L.lib.sdt_set_plane_wave(self._h, k, axz, axy, ayz, float(f0), float(amplitude), int(wavelet)),
                self._h)

    def set_receivers(self, idx):
        """idx: (n, 3) 1-based unpadded (x, y, z) indices; y ignored in 2D."""
        r = np.ascontiguousarray(np.asarray(idx, dtype=np.int32).reshape(-1, 3))
        L.check(L.lib.sdt_set_receivers(self._h, r.shape[0], r.ctypes.data_as(C.POINTER(C.c_int32))),
                self._h)
        self.nrec = r.shape[0]

    # ---- time loop ----------------------------------------------------------
    def run(self, nsteps=None):
        if nsteps is None:
            nsteps = self.time_steps - self.steps_done
        failed = C.c_int32(0)
        rc = L.lib.sdt_run(self._h, int(nsteps), C.byref(failed))
        L.check(rc, self._h, failed.value)

    @property
    def steps_done(self):
        s = C.c_int32(0)
        L.check(L.lib.sdt_steps_done(self._h, C.byref(s)), self._h)
        return s.value

    def synchronize(self):
        L.check(L.lib.sdt_synchronize(self._h), self._h)

    # ---- outputs ------------------------------------------------------------
    def field(self, name):
        if self.complex:
            out = np.empty(self.ncell, dtype=np.complex128)
            flat = out.view(np.float64)
        else:
            out = flat = np.empty(self.ncell, dtype=np.float64)
        L.check(L.lib.sdt_get_field(self._h, self.fields[name], _dptr(flat), flat.size), self._h)
        return out.reshape(self.shape, order="F")

    def snapshot(self, name, out, single=True):
        """Queue a snapshot of field `name` (as the reference's write_image2/3
        write it: Fortran order, FP32 when single) into `out`, a contiguous
        float32 / float64 array of ncell values (pinned: see PinnedArray).
        The copy is asynchronous: `out` holds the snapshot after the next
        synchronize() or blocking call."""
        dt = np.float32 if single else np.float64
        if out.dtype != dt or out.size != self.ncell or not out.flags.c_contiguous:
            raise ValueError(f"snapshot buffer must be a contiguous {np.dtype(dt).name} array of {self.ncell} values")
        L.check(L.lib.sdt_snapshot(self._h, self.fields[name], 1 if single else 0,
                                   out.ctypes.data_as(C.c_void_p), out.size), self._h)
        return out

    def set_field(self, name, a):
        flat = self._complex_grid(a) if self.complex else np.ascontiguousarray(self._grid_array(a))
        L.check(L.lib.sdt_set_field(self._h, self.fields[name], _dptr(flat), flat.size), self._h)

    def receivers(self):
        """(steps_done, n_receivers, 3) array of the velocity / E components
        (complex128 for the 2.5D solver)."""
        n = self.steps_done
        out = np.empty((n, self.nrec, 3), dtype=np.complex128 if self.complex else np.float64)
        flat = out.view(np.float64)
        L.check(L.lib.sdt_get_receivers(self._h, _dptr(flat), flat.size), self._h)
        return out

    def norms(self):
        n = self.steps_done
        v = np.empty(n, dtype=np.float64)
        s = np.empty(n, dtype=np.float64)
        L.check(L.lib.sdt_get_norms(self._h, _dptr(v), _dptr(s), n), self._h)
        return v, s

    # ---- tuning -----------------------------------------------------------------
    def set_option(self, key, value):
        """key: 'zchunk' (planes per z-marching workgroup, 0 auto), 'graph'
        (0/1), 'stagger' (0/1), 'tile_rows' (0 auto, 4, 8, 16) or
        'material_table' (1 = one-byte material ids + table when the model
        has <= 256 distinct materials; 0 = per-cell arrays) or
        'concurrent' (1 = 3D elastic shell kernels on a second stream beside
        the interior kernels, 0 = sequential) or 'xcd_remap' (3D elastic block
        order, bit 0 interior, bit 1 shell: XCD-contiguous tile ranges) or
        'checkpoint' (2D: steps between forward states saved for adjoint())
        or 'raw_source' (1 = source series injected as given) or 'isotropic'
        (3D elastic without a material table: 1 = hold only c11, c12, c44 when
        every cell's stiffness has the isotropic form, 0 = all 21 arrays) or
        'fused' (3D EM: 1 = one kernel per step, 0 or -1 = one per half step,
        the default) or 'source_window' (2D EM in FP32: 1 = the cells within
        7 of the point source also updated in FP64, receivers there record
        the FP64 values, the default; 0 = FP32 throughout)."""
        k = {"zchunk": L.OPT_ZCHUNK, "graph": L.OPT_GRAPH, "stagger": L.OPT_STAGGER,
             "tile_rows": L.OPT_TILE_ROWS, "material_table": L.OPT_MATERIAL_TABLE,
             "concurrent": L.OPT_CONCURRENT, "xcd_remap": L.OPT_XCD_REMAP,
             "checkpoint": L.OPT_CHECKPOINT, "raw_source": L.OPT_RAW_SOURCE,
             "isotropic": L.OPT_ISOTROPIC, "fused": L.OPT_FUSED, "source_window": L.OPT_SOURCE_WINDOW}[key]
        L.check(L.lib.sdt_set_option(self._h, k, int(value)), self._h)

    def info(self, key):
        """'materials' (distinct materials in the table, 0 = per-cell arrays),
        'shell_tiles', 'device_bytes', 'coefficients', 'isotropic' (1 = the
        3D elastic isotropic arrays are in use), 'fused' (1 = the fused 3D
        EM step is in use) or 'cpml_tuples' (3D EM: distinct CPML coefficient
        tuples held as a dictionary, 0 = per-cell arrays); valid after the
        first run()."""
        k = {"materials": L.INFO_MATERIALS, "shell_tiles": L.INFO_SHELL_TILES,
             "device_bytes": L.INFO_DEVICE_BYTES, "coefficients": L.INFO_COEFFICIENTS,
             "isotropic": L.INFO_ISOTROPIC, "fused": L.INFO_FUSED, "cpml_tuples": L.INFO_CPML_TUPLES}[key]
        v = C.c_int64(0)
        L.check(L.lib.sdt_get_info(self._h, k, C.byref(v)), self._h)
        return v.value

    def materials(self):
        return self.info("materials")

    # ---- kernel timing --------------------------------------------------------
    def set_timing(self, on=True):
        L.check(L.lib.sdt_set_timing(self._h, 1 if on else 0), self._h)

    def kernel_times(self):
        a, b, n = C.c_double(0), C.c_double(0), C.c_int64(0)
        L.check(L.lib.sdt_get_kernel_times(self._h, C.byref(a), C.byref(b), C.byref(n)), self._h)
        return a.value, b.value, n.value


class Batch:
    """Many 2D shots advanced together on one GPU.

    `solvers`: prepared-or-not `Solver`s of one kind ("elastic2" or "em2")
    and one precision, each with its own grid, model, source, receivers and
    time_steps.  Every shot's inputs are fixed when the batch is made.  One
    launch per half step covers the tiles of all shots; the tile height is
    picked by timing the candidates on the first steps.  Results equal those
    of running each `Solver` alone; read them from the solvers afterwards.
    """

    def __init__(self, solvers):
        self.solvers = list(solvers)
        if not self.solvers:
            raise ValueError("empty batch")
        hs = (C.c_void_p * len(self.solvers))(*[s._h.value for s in self.solvers])
        self._b = C.c_void_p()
        L.check(L.lib.sdt_batch_create(hs, len(self.solvers), C.byref(self._b)))

    def run(self, nsteps=None):
        """Advance every shot by min(nsteps, its remaining steps); default:
        run all shots to their ends.  Raises UnstableError (with .shot and
        .step) when a shot blows up."""
        if nsteps is None:
            nsteps = max(s.time_steps - s.steps_done for s in self.solvers)
        shot, step = C.c_int32(-1), C.c_int32(0)
        rc = L.lib.sdt_batch_run(self._b, int(nsteps), C.byref(shot), C.byref(step))
        try:
            L.check(rc, None, step.value)
        except L.UnstableError as e:
            e.shot = shot.value
            raise

    def info(self, key):
        """'tile_rows' (rows per tile of the chosen shape, 0 before tuning),
        'tiles' (workgroups per launch) or 'xcd_order' (1 when the chosen
        tile list is ordered for XCD locality, -1 before tuning)."""
        k = {"tile_rows": L.BATCH_INFO_TILE_ROWS, "tiles": L.BATCH_INFO_TILES,
             "xcd_order": L.BATCH_INFO_XCD_ORDER}[key]
        v = C.c_int64(0)
        L.check(L.lib.sdt_batch_info(self._b, k, C.byref(v)))
        return v.value

    def close(self):
        if self._b:
            L.lib.sdt_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def make_solver(cfg, *, precision="fp64", device=0, options=None):
    """A `Solver` loaded with a configuration dict's model, sources and
    receivers (see seidart_hip.models), not yet run."""
    s = Solver(cfg["physics"], cfg["shape"], cfg["spacing"], cfg["cpml"], cfg["dt"],
               cfg["time_steps"], cfg["src"], precision=precision,
               density_method=cfg.get("density_method", "arithmetic"),
               stencil=cfg.get("stencil", "reference"), device=device, ky=cfg.get("ky", 0.0))
    try:
        for k, v in (options or {}).items():
            s.set_option(k, v)
        s.set_arrays(cfg["arrays"])
        for comp, series in cfg["sources"].items():
            s.set_source(comp, series)
        if len(cfg.get("receivers", ())):
            s.set_receivers(cfg["receivers"])
        pw = cfg.get("plane_wave")
        if pw:
            s.set_plane_wave(pw["kind"], pw["angles"], pw["f0"], pw["amplitude"], pw.get("wavelet", 0))
    except Exception:
        s.close()
        raise
    return s


def run_config(cfg, *, precision="fp64", steps=None, device=0, fields=(), options=None):
    """Run a synthetic configuration dict (see seidart_hip.models) and return
    (receivers, norms, {field: array})."""
    s = make_solver(cfg, precision=precision, device=device, options=options)
    try:
        s.run(steps if steps is not None else cfg["time_steps"])
        out = {f: s.field(f) for f in fields}
        return s.receivers(), s.norms(), out
    finally:
        s.close()
