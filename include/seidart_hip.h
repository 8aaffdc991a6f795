/*
 * seidart_hip.h -- C-ABI of libseidart_hip.so, the gfx950 HIP implementation of
 * SeidarT-Backend's anisotropic CPML finite-difference time-domain solvers.
 *
 * The reference has no C-ABI: each solver is a Fortran subroutine taking
 * derived types and reading its inputs from files in the working directory
 * (src/fortran/cpmlfdtd.f08:57 seismic2, :600 seismic25, :1341 seismic3,
 * :2317 electromag2, :2764 electromag25, :3353 electromag3).  The entry points below replace that contract one for
 * one:
 *
 *   Domain_Type / Source_Type (src/fortran/seidart_types.f08:13-33)
 *       -> sdt_params, passed to sdt_create
 *   material_rw2 / material_rw2c / material_rw3 reads of c11.dat ...
 *   bcoef_half_cpml.dat, initialcondition*.dat (src/fortran/seidartio.f08:
 *   225-292, used at cpmlfdtd.f08:159-174,253-266,763-872,1501-1645,
 *   2413-2506,3481-3618)
 *       -> sdt_set_array
 *   loadsource of seismicsource*.dat / electromagneticsource*.dat
 *   (src/fortran/seidartio.f08:161-183, used at cpmlfdtd.f08:230-237,
 *   1606-1617, 2476-2479, 3581-3585)
 *       -> sdt_set_source
 *   the environment variable KY read by seismic25 / electromag25
 *   (cpmlfdtd.f08:697-703, 2861-2867)
 *       -> sdt_params.ky
 *   the `do it = 1, time_steps` loops (cpmlfdtd.f08:311, 931, 1747, 2566,
 *   3689) -> sdt_run
 *   write_image2 / write_image3 snapshots (src/fortran/seidartio.f08:366-426)
 *       -> sdt_get_field (one call per step and component)
 *   write_array of velocity_norm.dat / stress_norm.dat (cpmlfdtd.f08:2282-2283)
 *       -> sdt_get_norms
 *   `stop 'code became unstable and blew up'` (cpmlfdtd.f08:567, 1292, 2263,
 *   2740, 3972) -> SDT_ERR_UNSTABLE returned with the step index
 *
 * Receiver recording (sdt_set_receivers / sdt_get_receivers) has no reference
 * counterpart: the SeidarT frontend samples the snapshot files instead; the
 * values recorded equal the snapshot values at the receiver cells.
 *
 * Conventions: arrays are Fortran order (x fastest, then y, then z), padded
 * sizes (nx = n_json + 2*cpml, src/fortran/main.f08:69-71), FP64 on the host
 * side.  2D solvers use the (x, z) plane with ny = 1.  Source and receiver
 * indices are 1-based and unpadded, as in the JSON (the solver adds cpml,
 * cpmlfdtd.f08:178-179, 1533-1535).  Every pointer is borrowed for the
 * duration of the call.  No entry point exits the process.  One HIP stream per
 * handle; handles are independent, a single handle is not thread-safe.
 *
 * Complex values (2.5D solvers, SDT_ELASTIC25 and SDT_EM25): fields, the initial-condition
 * arrays and receiver samples are complex128 stored as interleaved (re, im)
 * doubles, exactly numpy's complex128 memory layout; their counts are twice
 * the real ones.  Material, CPML and source arrays stay real.
 */
#ifndef SEIDART_HIP_H
#define SEIDART_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDT_ABI_VERSION 2

/* ---- status codes -------------------------------------------------------- */
enum {
    SDT_OK = 0,
    SDT_ERR_UNSTABLE = 1, /* max velocity/E norm exceeded 1e25
                             (constants.f08:12, cpmlfdtd.f08:567)       */
    SDT_ERR_ARG = 2,      /* bad argument                                */
    SDT_ERR_HIP = 3,      /* HIP runtime error                           */
    SDT_ERR_STATE = 4     /* call out of order (e.g. run before inputs)  */
};

/* ---- solver kinds (src/fortran/main.f08:90-119 dispatch) ---------------- */
enum {
    SDT_ELASTIC2 = 0,  /* seismic2     cpmlfdtd.f08:57   */
    SDT_ELASTIC25 = 1, /* seismic25    cpmlfdtd.f08:600  */
    SDT_ELASTIC3 = 2,  /* seismic3     cpmlfdtd.f08:1341 */
    SDT_EM2 = 3,       /* electromag2  cpmlfdtd.f08:2317 */
    SDT_EM25 = 4,      /* electromag25 cpmlfdtd.f08:2764 */
    SDT_EM3 = 5        /* electromag3  cpmlfdtd.f08:3353 */
};

enum { SDT_FP64 = 0, SDT_FP32 = 1 };

/* density averaging codes, identical to scalar_mean (averaging.f08:5-25) */
enum {
    SDT_DENSITY_HARMONIC = 1,
    SDT_DENSITY_GEOMETRIC = 2,
    SDT_DENSITY_ARITHMETIC = 3,
    SDT_DENSITY_NONE = 4
};

/* velocity-from-stress stencil: the reference's 3-term form
 * (cpmlfdtd.f08:405-406, 434-435, 2022-2024, 2063-2066, 2107-2111), or the
 * consistent 4-term staggered form (opt-in, not the reference's output). */
enum { SDT_STENCIL_REFERENCE = 0, SDT_STENCIL_CORRECTED = 1 };

/* ---- input array identifiers: one per reference input file ------------- */
enum {
    SDT_C11 = 0, SDT_C12, SDT_C13, SDT_C14, SDT_C15, SDT_C16,
    SDT_C22, SDT_C23, SDT_C24, SDT_C25, SDT_C26,
    SDT_C33, SDT_C34, SDT_C35, SDT_C36,
    SDT_C44, SDT_C45, SDT_C46,
    SDT_C55, SDT_C56,
    SDT_C66,                                   /* 0..20  c??.dat        */
    SDT_RHO = 21,                              /* rho.dat               */
    SDT_GAMMA_X, SDT_GAMMA_Y, SDT_GAMMA_Z,
    SDT_GAMMA_XY, SDT_GAMMA_XZ, SDT_GAMMA_YZ,  /* 22..27 gamma_*.dat    */
    SDT_KAPPA = 28, SDT_ALPHA, SDT_ACOEF, SDT_BCOEF,             /* *_cpml.dat      */
    SDT_KAPPA_HALF, SDT_ALPHA_HALF, SDT_ACOEF_HALF, SDT_BCOEF_HALF, /* *_half_cpml.dat */
    SDT_E11 = 36, SDT_E12, SDT_E13, SDT_E22, SDT_E23, SDT_E33,   /* e??.dat (relative) */
    SDT_S11 = 42, SDT_S12, SDT_S13, SDT_S22, SDT_S23, SDT_S33,   /* s??.dat (S/m)      */
    SDT_IC_X = 48, SDT_IC_Y, SDT_IC_Z,  /* initialconditionV?.dat / initialconditionE?.dat */
    SDT_NUM_ARRAYS = 51
};

/* ---- source series identifiers ------------------------------------------ */
enum {
    SDT_SRC_X = 0, SDT_SRC_Y, SDT_SRC_Z,             /* seismicsource?.dat, electromagneticsource?.dat */
    SDT_SRC_XX, SDT_SRC_XY, SDT_SRC_XZ,
    SDT_SRC_YY, SDT_SRC_YZ, SDT_SRC_ZZ,              /* seismicsource??.dat                            */
    SDT_NUM_SRC = 9
};

/* ---- field identifiers (state arrays) ------------------------------------ */
enum {
    SDT_F_VX = 0, SDT_F_VY, SDT_F_VZ,
    SDT_F_SXX, SDT_F_SYY, SDT_F_SZZ, SDT_F_SXY, SDT_F_SXZ, SDT_F_SYZ,
    SDT_F_EX = 0, SDT_F_EY = 1, SDT_F_EZ = 2,
    SDT_F_HX = 3, SDT_F_HY = 4, SDT_F_HZ = 5,
    SDT_NUM_FIELDS = 9
};

typedef struct sdt_params {
    int32_t abi_version;   /* must be SDT_ABI_VERSION                       */
    int32_t physics;       /* SDT_ELASTIC2 / 25 / 3, SDT_EM2 / SDT_EM3      */
    int32_t precision;     /* SDT_FP64 or SDT_FP32                          */
    int32_t nx, ny, nz;    /* padded sizes; ny = 1 for 2D                   */
    double dx, dy, dz;     /* Domain_Type%dx..dz                            */
    int32_t cpml;          /* Domain_Type%cpml                              */
    double dt;             /* Source_Type%dt                                */
    int32_t time_steps;    /* Source_Type%time_steps (length of the series) */
    int32_t xind, yind, zind; /* Source_Type%xind.. (1-based, unpadded)     */
    int32_t density_code;  /* SDT_DENSITY_*, elastic only                   */
    int32_t stencil;       /* SDT_STENCIL_*                                 */
    int32_t device;        /* HIP device ordinal                            */
    int32_t reserved0;     /* zero                                          */
    double ky;             /* 2.5D solvers: y wavenumber (env KY,
                              cpmlfdtd.f08:697-703); 0 otherwise             */
    int32_t reserved[6];   /* zero                                          */
} sdt_params;

typedef struct sdt_handle sdt_handle;

/* library-level queries */
int32_t sdt_version(void);
int32_t sdt_device_count(int32_t *count);
/* last error message of a handle, or of the calling thread when h is NULL */
const char *sdt_last_error(const sdt_handle *h);

/* lifetime; a handle that belongs to a live batch is not destroyed
 * (SDT_ERR_STATE) until sdt_batch_destroy releases it */
int32_t sdt_create(const sdt_params *p, sdt_handle **out);
int32_t sdt_destroy(sdt_handle *h);

/* inputs: count must equal nx*ny*nz (arrays) or time_steps (series); for
 * the 2.5D solvers the SDT_IC_* arrays take either 2*nx*nz doubles (complex128,
 * initialcondition[VE]?.dat read by material_rw2c) or nx*nz (real part only).
 * Arrays not set take the reference's defaults (kappa = 1, everything else 0,
 * cpmlfdtd.f08:240-247). */
int32_t sdt_set_array(sdt_handle *h, int32_t array_id, const double *host, int64_t count);
int32_t sdt_set_source(sdt_handle *h, int32_t component, const double *series, int64_t count);
/* n receivers, ijk = n triples of 1-based unpadded (x, y, z) indices (y ignored in 2D);
 * each receiver records the three velocity / E components every step */
int32_t sdt_set_receivers(sdt_handle *h, int32_t n, const int32_t *ijk);

/* advance nsteps time steps (continues from the current step).  On
 * SDT_ERR_UNSTABLE, *failed_step (if non-NULL) receives the 1-based step.
 * A handle that belongs to a live batch is advanced only by sdt_batch_run
 * (SDT_ERR_STATE here). */
int32_t sdt_run(sdt_handle *h, int32_t nsteps, int32_t *failed_step);
/* steps completed so far */
int32_t sdt_steps_done(const sdt_handle *h, int32_t *steps);

/* outputs (FP64 host copies; blocking); count = nx*ny*nz, or 2*nx*nz for the
 * complex fields of the 2.5D solvers */
int32_t sdt_get_field(sdt_handle *h, int32_t field_id, double *out, int64_t count);
int32_t sdt_set_field(sdt_handle *h, int32_t field_id, const double *host, int64_t count);
/* out = steps_done * n_receivers * 3 values, step-major (complex for the
 * 2.5D solvers, i.e. 6 doubles per receiver and step) */
int32_t sdt_get_receivers(sdt_handle *h, double *out, int64_t count);
/* per-step max norms: velocity (or E) and, for 3D elastic, stress; NULL skips */
int32_t sdt_get_norms(sdt_handle *h, double *vel, double *stress, int64_t count);

/* snapshot of one field, as write_image2/3 (seidartio.f08:366-426) write it
 * after every step: Fortran order, unpadded, FP32 when single = 1 (the
 * reference's SINGLE_OUTPUT default, main.f08:94-116) or FP64 when single =
 * 0, count = nx*ny*nz values into out (float or double).  The field is
 * packed and converted on the device into a staging buffer the handle
 * allocates on first use, then copied to out asynchronously on the handle's
 * stream: out must stay valid, and is complete, after the next
 * sdt_synchronize (or any blocking call on the handle).  Pinned host memory
 * (sdt_host_alloc) lets the copy overlap host work.  Real-valued solvers only
 * (the 2.5D solvers write zstd blocks instead). */
int32_t sdt_snapshot(sdt_handle *h, int32_t field_id, int32_t single, void *out, int64_t count);
/* sdt_snapshot for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc); no
 * counterpart in the reference.  Same arguments and output as sdt_snapshot,
 * but it allocates nothing: the field is packed through a fixed 16 MiB device
 * scratch (one per GPU, held until the process ends) in chunks of whole x-y
 * planes, and the call returns when the copy into out is complete.  The
 * handle must already be set up (its first sdt_run, which may take 0 steps;
 * SDT_ERR_STATE otherwise) and one x-y plane of the output (nx*ny*4 or
 * nx*ny*8 bytes) must fit the scratch (SDT_ERR_ARG otherwise).  Every call
 * on one GPU uses the same scratch: calls that overlap on one GPU (from
 * several host threads) race. */
int32_t sdt_snapshot_noalloc(sdt_handle *h, int32_t field_id, int32_t single, void *out, int64_t count);
/* pinned (page-locked) host memory for snapshot buffers; no reference
 * counterpart */
int32_t sdt_host_alloc(int64_t bytes, void **out);
int32_t sdt_host_free(void *p);

/* kernel timing with HIP events on the handle's stream.  When enabled, each
 * step records events around the stress (H) and velocity (E) kernels;
 * sdt_get_kernel_times returns the summed milliseconds and the number of
 * launches timed since enabling. */
int32_t sdt_set_timing(sdt_handle *h, int32_t enable);
int32_t sdt_get_kernel_times(sdt_handle *h, double *ms_stress, double *ms_velocity,
                             int64_t *launches);
/* block until all work queued on the handle's stream has finished */
int32_t sdt_synchronize(sdt_handle *h);

/* tuning options (no reference counterpart); set before the first step */
enum {
    SDT_OPT_ZCHUNK = 1,         /* planes per z-marching workgroup (0 = auto)            */
    SDT_OPT_GRAPH = 2,          /* 1 = replay captured hipGraphs of several steps when the
                                   grid is small (default), 0 = direct launches          */
    SDT_OPT_STAGGER = 3,        /* 1 = offset each grid array's start by a different
                                   multiple of 4352 B (default), 0 = plain hipMalloc     */
    SDT_OPT_TILE_ROWS = 4,      /* rows of the 3D interior workgroup tile: 0 auto, 4, 8, 16 */
    SDT_OPT_CONCURRENT = 6,     /* 3D elastic: 1 = run the CPML-shell kernels on a second stream
                                   concurrently with the interior kernels of the same half step,
                                   0 = one after the other on one stream (default; the two are
                                   equally fast at 280^3, where HBM is the shared limit)      */
    SDT_OPT_XCD_REMAP = 7,      /* 3D elastic block order: bit 0 interior kernels, bit 1 shell
                                   kernels; set = each XCD takes a contiguous range of tiles,
                                   clear = the dispatcher's round-robin order (default 0: the
                                   round-robin order measured faster at 280^3).  3D EM half-step
                                   kernels: bit 0 set = column order, each XCD runs tile columns
                                   through all planes in turn (default 0: measured slower)    */
    SDT_OPT_ISOTROPIC = 10,     /* 3D and 2.5D elastic: 1 = when the per-cell arrays are in use (no
                                   material table) and every cell's stiffness is isotropic in form
                                   (c11 = c22 = c33, c12 = c13 = c23, c44 = c55 = c66, all
                                   other entries 0), keep only c11, c12, c44 and read 3 arrays
                                   per cell instead of 21 (default; same arithmetic, results
                                   equal to the general path), 0 = always all 21 arrays      */
    SDT_OPT_FUSED = 11,         /* 3D EM: 1 = the H and E updates of a step run as one kernel
                                   (E, H and the H-update memory variables are kept in two
                                   buffers each, the current one swapped every step; a third
                                   fewer bytes per step), 0 = one kernel per half step, fields
                                   updated in place, -1 = the half-step kernels (default: the
                                   faster form in both precisions at 404 x 404 x 212)        */
    SDT_OPT_MATERIAL_TABLE = 5  /* every solver: 1 = when the model holds at most 256
                                   distinct material tuples (3D elastic: (c11..c66, gamma_*,
                                   rho); 2.5D and 2D elastic: (c11..c66, gamma_*), the density
                                   stays per cell; 2D EM: the seven update coefficients; 3D
                                   and 2.5D EM: (inverse permittivity, conductivity)), keep one byte of
                                   material id per cell and a table of the tuples instead of
                                   the per-cell arrays (default; same values, results equal to
                                   rounding; the 2D solvers keep per-cell arrays when an
                                   adjoint is planned), 0 = always the per-cell arrays     */
};
int32_t sdt_set_option(sdt_handle *h, int32_t key, int64_t value);

/* plane-wave source, replacing the point source: an analytic plane wave
 * (wavelet `wavelet` = 0 gaus0, 1 gaus1, 2 gaus2 of centre frequency f0,
 * times amplitude) travelling along the direction the three rotation angles
 * (degrees) give, added every step on the interior faces it enters through.
 * Elastic solvers (2D, 2.5D, 3D): polarised across the direction (SDT_PW_S:
 * source types 'pw', 'pws', 'pw_s', 'pw-s') or along it (SDT_PW_P: 'pwp',
 * 'pw_p', 'pw-p'), added to the velocities and (through the stiffness) the
 * stresses, at the model's largest Kelvin-form stiffness eigenvalue over its
 * smallest density as speed.  EM solvers (2D, 3D): source type 'pw' only
 * (SDT_PW_S), E along the polarisation and H = E / eta along p x e, at speed
 * c / sqrt(eps_min) and impedance eta = sqrt(mu0 / (eps0 eps_min)) of the
 * smallest relative-permittivity eigenvalue; the 2.5D EM solver has none
 * (SDT_ERR_ARG).  Replaces plane_wave_source.f08:20-230 with the injection
 * loops cpmlfdtd.f08:188-232 / 454-512 (seismic2), 802-832 / 1208-1262
 * (seismic25), 1548-1605 / 2133-2220 (seismic3), 2436-2475 / 2626-2673
 * (electromag2), 3523-3580 / 3833-3903 (electromag3).  Set before the first
 * step. */
enum { SDT_PW_S = 1, SDT_PW_P = 2 };
int32_t sdt_set_plane_wave(sdt_handle *h, int32_t kind, double angle_xz, double angle_xy, double angle_yz,
                           double f0, double amplitude, int32_t wavelet);

/* read-only facts about the device layout chosen at setup (no reference
 * counterpart); valid after the first sdt_run */
enum {
    SDT_INFO_MATERIALS = 1,     /* distinct materials in the material table, 0 = per-cell arrays */
    SDT_INFO_SHELL_TILES = 2,   /* 3D elastic CPML-shell tiles                                    */
    SDT_INFO_DEVICE_BYTES = 3,  /* device memory held by the handle's grid arrays                 */
    SDT_INFO_COEFFICIENTS = 4,  /* update coefficients per cell (2D EM 7, 2D elastic 10, else 0)  */
    SDT_INFO_ISOTROPIC = 5,     /* 1 = the isotropic arrays (c11, c12, c44) are in use (3D, 2.5D elastic) */
    SDT_INFO_FUSED = 6,         /* 1 = the fused 3D EM step is in use                      */
    SDT_INFO_CPML_TUPLES = 7    /* 3D EM and the 3D elastic CPML shell: distinct CPML
                                   coefficient tuples held as a one-byte id per cell and a
                                   table, 0 = six per-cell arrays (streams)               */
};
int32_t sdt_get_info(sdt_handle *h, int32_t key, int64_t *value);

/* ---- batched shots (2D elastic, 2D EM) ---------------------------------------
 * No reference counterpart: the reference runs one shot per process.  A batch
 * advances many prepared handles of one solver kind and precision (each with
 * its own grid size, model, source and number of time steps) together: one
 * kernel launch per half step covers the tiles of every shot, one more
 * handles every shot's sources and receivers, and the step loop is replayed
 * from captured hipGraphs.  The tile height is chosen by timing each
 * candidate on the first steps of the run.  Handles stay owned by the caller
 * and must outlive the batch; their fields, receivers and norms are read with
 * the usual calls afterwards.  A handle may appear once in one batch
 * (SDT_ERR_ARG for a repeat, SDT_ERR_STATE for a handle of another live
 * batch); while the batch lives, sdt_run, sdt_adjoint and sdt_destroy on its
 * handles return SDT_ERR_STATE. */
typedef struct sdt_batch sdt_batch;
int32_t sdt_batch_create(sdt_handle *const *handles, int32_t n, sdt_batch **out);
/* advance every shot by min(nsteps, its remaining steps); on a blow-up
 * returns SDT_ERR_UNSTABLE with the shot index and its 1-based step */
int32_t sdt_batch_run(sdt_batch *b, int32_t nsteps, int32_t *failed_shot, int32_t *failed_step);
/* TILE_ROWS: rows per tile of the chosen shape (0 before tuning); TILES:
 * workgroups per launch; XCD_ORDER: 1 when the chosen tile list is ordered
 * for XCD locality (-1 before tuning) */
enum { SDT_BATCH_INFO_TILE_ROWS = 1, SDT_BATCH_INFO_TILES = 2, SDT_BATCH_INFO_XCD_ORDER = 3 };
int32_t sdt_batch_info(const sdt_batch *b, int32_t key, int64_t *value);
int32_t sdt_batch_destroy(sdt_batch *b);

/* ---- adjoint-state gradients (2D EM, 2D elastic) ------------------------------
 * No reference counterpart: the reference has no inversion path.  These
 * entry points give the exact gradient of any function of the receiver
 * traces (through its gradient with respect to the traces) with respect to
 * the solver's per-cell update coefficients and to the source series, by the
 * discrete adjoint of the time stepping.
 *
 * Update coefficients (per cell, nx*nz each, in this order):
 *   2D EM:      aEx, bEx, bEz, dEx, eEx, dEz, eEz -- the 2x2 inverse
 *               permittivity and conductivity products of cpmlfdtd.f08:2519-2540
 *               (aEx = e33/det, bEx = -e13/det, bEz = e11/det, dEx = (e33 s11 -
 *               e13 s13)/det, eEx = (e33 s13 - e13 s33)/det, dEz = (e11 s13 -
 *               e13 s11)/det, eEz = (e11 s33 - e13 s13)/det, permittivity
 *               scaled by eps0)
 *   2D elastic: c11, c13, c15, c33, c35, c55, rho, 1/(1 + gamma_x dt),
 *               1/(1 + gamma_z dt), 1/(1 + gamma_xz dt)
 * Chain them to permittivity, conductivity, stiffness or attenuation on the
 * host (seidart_hip.fwi does so with torch autograd). */
enum {
    SDT_OPT_CHECKPOINT = 8,     /* steps between saved forward states for sdt_adjoint
                                   (0 = none, default); set before the first step      */
    SDT_OPT_RAW_SOURCE = 9,     /* 1 = source series are added to the fields as given
                                   (no dt / permittivity or density scaling); set before
                                   the first step                                         */
    SDT_OPT_SOURCE_WINDOW = 12  /* 2D EM in FP32: 1 = the cells within 7 of the point source
                                   (inside the CPML-free box, off the faces) are updated
                                   again in FP64 every step with FP64 coefficients, the
                                   source is added there in FP64 and receivers there
                                   record the FP64 values (default: FP32 rounding of the
                                   large near-source field otherwise dominates near-offset
                                   traces and FWI residuals), 0 = FP32 throughout; no
                                   effect in FP64; set before the first step              */
};
/* replace the model inputs by the update coefficients above (count = ncoef *
 * nx * nz, coefficient-major); needs SDT_OPT_RAW_SOURCE = 1 */
int32_t sdt_set_coefficients(sdt_handle *h, const double *coef, int64_t count);
/* After a forward run over all time_steps with checkpoints: trace_grad
 * (time_steps * n_receivers * 3, the layout of sdt_get_receivers) is the
 * gradient of the objective with respect to the traces.  Writes the gradient
 * with respect to the update coefficients (ncoef * nx * nz, coefficient-major)
 * and, when src_grad is non-null, with respect to the source series
 * (SDT_NUM_SRC * time_steps, the layout of sdt_set_source).  Fields,
 * receivers and norms read afterwards are those of the forward run. */
int32_t sdt_adjoint(sdt_handle *h, const double *trace_grad, int64_t tcount, double *coef_grad,
                    int64_t gcount, double *src_grad, int64_t scount);
/* the same for every shot of a batch (all shots run to their ends with
 * checkpoints, all of one time_steps); arrays per shot */
int32_t sdt_batch_adjoint(sdt_batch *b, const double *const *trace_grad, double *const *coef_grad,
                          double *const *src_grad);

#ifdef __cplusplus
}
#endif

#endif /* SEIDART_HIP_H */
