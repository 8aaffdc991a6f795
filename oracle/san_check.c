/*
 * san_check.c -- TEST INFRASTRUCTURE: runs every entry point of the CPU oracle
 * (oracle.c) on small seeded grids, built with AddressSanitizer and
 * UndefinedBehaviorSanitizer (oracle/Makefile target `san`), so that an
 * out-of-bounds access or undefined operation anywhere in the restated loops
 * aborts the run.
 *
 * Cases: each solver with every model array present, every source series,
 * receivers on the grid's corners and faces, both velocity stencils, the four
 * density codes, point sources near the low and high faces, given initial
 * fields and requested final fields, plane-wave sources of every kind the
 * solver takes, and the two averaging helpers in every direction.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static unsigned long long rng = 0x2545f4914f6cdd1dull;
static double unif(void)
{
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return (double)(rng >> 11) * (1.0 / 9007199254740992.0);
}

typedef struct {
    double *arr[SDT_NUM_ARRAYS];
    double *src[SDT_NUM_SRC];
    double *fin[SDT_NUM_FIELDS], *fout[SDT_NUM_FIELDS];
    double *rec_out, *vn, *sn;
    int32_t rec[6 * 3];
} bufs;

static double *fill(size_t n, double lo, double hi)
{
    double *a = (double *)malloc(n * sizeof(double));
    if (!a) exit(2);
    for (size_t e = 0; e < n; ++e) a[e] = lo + (hi - lo) * unif();
    return a;
}

/* model arrays sized exactly: complex (interleaved) initial conditions for
 * the 2.5D solvers, real otherwise */
static void setup(bufs *b, oracle_run *r, int nx, int ny, int nz, int nsteps, int cplx)
{
    const size_t n = (size_t)nx * ny * nz;
    memset(b, 0, sizeof(*b));
    for (int q = 0; q < SDT_NUM_ARRAYS; ++q) {
        double lo = 0.5, hi = 1.5;
        if (q <= SDT_C66) { lo = 1e9; hi = 2e9; }
        else if (q == SDT_RHO) { lo = 900.0; hi = 2600.0; }
        else if (q >= SDT_GAMMA_X && q <= SDT_GAMMA_YZ) { lo = 0.0; hi = 0.01; }
        else if (q == SDT_KAPPA || q == SDT_KAPPA_HALF) { lo = 1.0; hi = 2.0; }
        else if (q == SDT_ACOEF || q == SDT_ACOEF_HALF) { lo = -0.1; hi = 0.0; }
        else if (q == SDT_BCOEF || q == SDT_BCOEF_HALF) { lo = 0.5; hi = 1.0; }
        else if (q == SDT_E11 || q == SDT_E22 || q == SDT_E33) { lo = 2.0; hi = 9.0; }
        else if (q == SDT_E12 || q == SDT_E13 || q == SDT_E23) { lo = 0.0; hi = 0.3; }
        else if (q == SDT_S11 || q == SDT_S22 || q == SDT_S33) { lo = 0.0; hi = 1e-3; }
        else if (q == SDT_S12 || q == SDT_S13 || q == SDT_S23) { lo = 0.0; hi = 1e-4; }
        else if (q >= SDT_IC_X) { lo = -1e-3; hi = 1e-3; }
        const size_t len = (q >= SDT_IC_X && cplx) ? 2 * n : n;
        b->arr[q] = fill(len, lo, hi);
        r->arr[q] = b->arr[q];
    }
    for (int q = 0; q < SDT_NUM_SRC; ++q) {
        b->src[q] = fill((size_t)nsteps, -1.0, 1.0);
        r->src[q] = b->src[q];
    }
    for (int f = 0; f < SDT_NUM_FIELDS; ++f) {
        const size_t len = cplx ? 2 * n : n;
        b->fin[f] = fill(len, -1e-6, 1e-6);
        b->fout[f] = (double *)calloc(len, sizeof(double));
        r->field_out[f] = b->fout[f];
    }
    const int cells[6][3] = {{1, 1, 1}, {nx, ny, nz}, {1, ny, nz}, {nx, 1, 1}, {(nx + 1) / 2, (ny + 1) / 2, 1},
                             {(nx + 1) / 2, ny, (nz + 1) / 2}};
    for (int q = 0; q < 6; ++q)
        for (int c = 0; c < 3; ++c) b->rec[3 * q + c] = cells[q][c];
    r->nrec = 6;
    r->rec = b->rec;
    b->rec_out = (double *)calloc((size_t)nsteps * 6 * 6, sizeof(double));
    b->vn = (double *)calloc((size_t)nsteps, sizeof(double));
    b->sn = (double *)calloc((size_t)nsteps, sizeof(double));
    r->rec_out = b->rec_out;
    r->velnorm = b->vn;
    r->stressnorm = b->sn;
}

static void release(bufs *b)
{
    for (int q = 0; q < SDT_NUM_ARRAYS; ++q) free(b->arr[q]);
    for (int q = 0; q < SDT_NUM_SRC; ++q) free(b->src[q]);
    for (int f = 0; f < SDT_NUM_FIELDS; ++f) { free(b->fin[f]); free(b->fout[f]); }
    free(b->rec_out); free(b->vn); free(b->sn);
}

typedef int32_t (*entry)(oracle_run *);

static int run_case(const char *name, entry fn, int nx, int ny, int nz, int cplx, int em, int stencil, int density,
                    int src_hi, int with_ic, int pw_kind)
{
    const int nsteps = 12;
    oracle_run r;
    bufs b;
    memset(&r, 0, sizeof(r));
    setup(&b, &r, nx, ny, nz, nsteps, cplx);
    r.nx = nx; r.ny = ny; r.nz = nz;
    r.dx = 0.5; r.dy = ny > 1 ? 0.5 : 1.0; r.dz = 0.5;
    r.dt = em ? 1e-12 : 1e-5;
    r.nsteps = nsteps;
    r.isrc = src_hi ? nx - 2 : 3;
    r.jsrc = ny > 1 ? (src_hi ? ny - 2 : 3) : 1;
    r.ksrc = src_hi ? nz - 2 : 3;
    r.density_code = density;
    r.stencil = stencil;
    r.nthreads = 2;
    r.ky = cplx ? 3.0 : 0.0;
    r.cpml = 3;
    if (with_ic)
        for (int f = 0; f < SDT_NUM_FIELDS; ++f) r.field_in[f] = b.fin[f];
    if (pw_kind) {
        r.pw_kind = pw_kind;
        r.pw_wavelet = (pw_kind + stencil) % 3;
        r.pw_angles[0] = 25.0; r.pw_angles[1] = 10.0; r.pw_angles[2] = -15.0;
        r.pw_f0 = em ? 1e8 : 100.0;
        r.pw_amp = 1.0;
        r.pw_speed = em ? 1e8 : 2000.0;
        r.pw_eta = 200.0;
    }
    const int32_t rc = fn(&r);
    double acc = 0.0;
    for (int s = 0; s < nsteps; ++s) acc += b.vn[s];
    printf("%-20s %3dx%3dx%3d stencil %d density %d src %s ic %d pw %d: rc %d unstable %d sum|v| %.3e\n", name, nx,
           ny, nz, stencil, density, src_hi ? "high" : "low", with_ic, pw_kind, rc, r.unstable_step, acc);
    release(&b);
    return rc;
}

int main(void)
{
    int bad = 0;
    for (int v = 0; v < 4; ++v) {
        const int stencil = v & 1, density = 1 + v, hi = v >> 1, ic = (v + 1) & 1;
        bad |= run_case("seismic2", oracle_seismic2, 17, 1, 15, 0, 0, stencil, density, hi, ic, 0) != 0;
        bad |= run_case("seismic25", oracle_seismic25, 15, 1, 13, 1, 0, stencil, density, hi, ic, 0) != 0;
        bad |= run_case("seismic3", oracle_seismic3, 13, 12, 11, 0, 0, stencil, density, hi, ic, 0) != 0;
        bad |= run_case("electromag2", oracle_electromag2, 16, 1, 14, 0, 1, stencil, density, hi, ic, 0) != 0;
        bad |= run_case("electromag25", oracle_electromag25, 14, 1, 15, 1, 1, stencil, density, hi, ic, 0) != 0;
        bad |= run_case("electromag3", oracle_electromag3, 12, 13, 11, 0, 1, stencil, density, hi, ic, 0) != 0;
    }
    for (int kind = SDT_PW_S; kind <= SDT_PW_P; ++kind) {
        bad |= run_case("seismic2 pw", oracle_seismic2, 17, 1, 15, 0, 0, 0, 3, 0, 0, kind) != 0;
        bad |= run_case("seismic25 pw", oracle_seismic25, 15, 1, 13, 1, 0, 1, 3, 0, 0, kind) != 0;
        bad |= run_case("seismic3 pw", oracle_seismic3, 13, 12, 11, 0, 0, 0, 3, 0, 0, kind) != 0;
    }
    bad |= run_case("electromag2 pw", oracle_electromag2, 16, 1, 14, 0, 1, 0, 3, 0, 0, SDT_PW_S) != 0;
    bad |= run_case("electromag25 pw", oracle_electromag25, 14, 1, 15, 1, 1, 0, 3, 0, 0, SDT_PW_S) != 0;
    bad |= run_case("electromag3 pw", oracle_electromag3, 12, 13, 11, 0, 1, 0, 3, 0, 0, SDT_PW_S) != 0;

    /* scalar means and the in-place averaging sweeps */
    double m = 0.0;
    for (int code = 1; code <= 4; ++code) m += oracle_scalar_mean(2.0, 3.0, code);
    for (int dir = 0; dir < 4; ++dir)
        for (int code = 1; code <= 4; ++code) {
            const int ni = 1 + (dir * 4 + code) % 9, nj = 1 + (dir + 2 * code) % 7;
            double *a = fill((size_t)ni * nj, 1.0, 2.0);
            oracle_array_averaging2d(a, ni, nj, code, dir);
            m += a[0];
            free(a);
        }
    printf("averaging checksum %.6f\n", m);
    if (bad) {
        fprintf(stderr, "an oracle entry point returned an error\n");
        return 1;
    }
    printf("san_check: every oracle entry point ran clean\n");
    return 0;
}
