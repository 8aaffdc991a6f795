/*
 * oracle.h -- CPU restatement of SeidarT-Backend's CPML-FDTD time loops.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP solver
 * library: tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * load it; the product path never does.
 *
 * Array/field/source identifiers are shared with include/seidart_hip.h.
 * All arrays are Fortran order (x fastest), padded sizes, 1-based indices in
 * the comments as in the reference.
 */
#ifndef SEIDART_ORACLE_H
#define SEIDART_ORACLE_H

#include <stdint.h>
#include "../include/seidart_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_run {
    int32_t nx, ny, nz;          /* padded sizes, ny = 1 in 2D                    */
    double dx, dy, dz, dt;
    int32_t nsteps;              /* steps to run                                  */
    int32_t isrc, jsrc, ksrc;    /* padded 1-based source cell (xind+cpml, ...)   */
    int32_t density_code;        /* 1..4, averaging.f08:5-25                      */
    int32_t stencil;             /* SDT_STENCIL_REFERENCE / SDT_STENCIL_CORRECTED */
    const double *arr[SDT_NUM_ARRAYS];  /* NULL -> kappa*=1, others 0             */
    const double *src[SDT_NUM_SRC];     /* length nsteps, NULL -> zero            */
    int32_t nrec;
    const int32_t *rec;          /* nrec triples, padded 1-based (i, j, k)        */
    double *rec_out;             /* [nsteps][nrec][3] or NULL                     */
    const double *field_in[SDT_NUM_FIELDS]; /* optional initial state (overrides IC) */
    double *field_out[SDT_NUM_FIELDS];      /* final state or NULL                */
    double *velnorm;             /* [nsteps] or NULL                              */
    double *stressnorm;          /* [nsteps] or NULL (3D elastic only)            */
    int32_t unstable_step;       /* out: 0, or the 1-based step that blew up      */
    int32_t nthreads;            /* OpenMP threads, 0 = runtime default           */
    double ky;                   /* 2.5D elastic: y wavenumber (env KY)           */
    /* plane-wave source (replaces the point source); the EM solvers take
     * SDT_PW_S only (source type 'pw') */
    int32_t pw_kind;             /* 0 point source, SDT_PW_S, SDT_PW_P            */
    int32_t pw_wavelet;          /* 0 gaus0, 1 gaus1, 2 gaus2                     */
    int32_t cpml;                /* absorbing-layer width (face positions)        */
    double pw_angles[3];         /* x_z, x_y, y_z rotation (degrees)              */
    double pw_f0, pw_amp;
    double pw_speed;             /* background speed (cbackground), from the caller */
    double pw_eta;               /* EM: background impedance (eta), from the caller */
} oracle_run;

/* 2.5D elastic (seismic25): fields, field_in/field_out and the ic_x..ic_z
 * arrays are complex, stored as interleaved (re, im) pairs of 2*nx*nz
 * doubles; rec_out is [nsteps][nrec][6] = (re, im) of vx, vy, vz. */

int32_t oracle_seismic2(oracle_run *r);
int32_t oracle_seismic25(oracle_run *r);
/* 2.5D EM (electromag25): E and H complex, same conventions as seismic25;
 * rec_out = (re, im) of Ex, Ey, Ez */
int32_t oracle_electromag25(oracle_run *r);
int32_t oracle_seismic3(oracle_run *r);
int32_t oracle_electromag2(oracle_run *r);
int32_t oracle_electromag3(oracle_run *r);

/* scalar_mean(a, b, code), averaging.f08:5-25 */
double oracle_scalar_mean(double a, double b, int32_t code);
/* array_averaging2d(array, code, direction) in place, averaging.f08:27-80;
 * direction 0 '+i', 1 '-i', 2 '+j', 3 '-j' */
void oracle_array_averaging2d(double *a, int32_t ni, int32_t nj, int32_t code, int32_t direction);

#ifdef __cplusplus
}
#endif
#endif
