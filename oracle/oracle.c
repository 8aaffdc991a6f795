/*
 * oracle.c -- CPU restatement of SeidarT-Backend's CPML-FDTD time loops in
 * plain C (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libseidart_hip.so.  Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.
 *
 * Each loop nest below follows the reference nest cited beside it, operation
 * for operation and in the same evaluation order, including the reference's
 * quirks (truncated 3-term velocity stencils, per-nest coefficient sets, the
 * uncorrected dvx_dz of the sigma_xy nest, asymmetric source offsets, the
 * dt-less 2D velocity source).  Memory variables are kept as in the reference
 * (one array per nest and derivative, 45 in 3D), so that the GPU's
 * deduplicated recursions are checked against the undeduplicated form.
 *
 * Pinning (what fixes each part independently of this text) is listed in the
 * test files: tests/test_oracle_pins.py (linear-field patch tests, CPML memory
 * recursion closed form with distinct full and half sets, Christoffel plane
 * waves in the corrected stencil mode, Yee plane waves, EM one-step closed
 * forms, conductive decay, linearity and superposition, scalar_mean against
 * textbook means, blow-up threshold), tests/test_oracle_pins_25.py (2.5D:
 * ky = 0 against 2D, the y-mode oscillators, array_averaging2d impulse
 * responses), tests/test_oracle_pins_pw.py (plane waves), and
 * tests/test_oracle_stencil_pins.py (the reference-mode 3-term velocity
 * stencils of 2D, 2.5D and 3D differ from the pinned corrected mode by
 * exactly the omitted term).
 *
 * Parity unpinned (restated from the cited reference lines only; nothing
 * outside the reference fixes them and the reference cannot be built here):
 * which coefficient set (full or half CPML profile) each derivative of each
 * 3D and 2.5D nest takes, the uncorrected dvx_dz of 3D sigma_xy
 * (cpmlfdtd.f08:1858-1865), the asymmetric source offsets and the dt-less 2D
 * velocity source (519-523, 1268-1276, 2208-2219), and the assignment of the
 * smeared 2.5D densities to velocity terms (889-898, 1141-1198).
 *
 * Index helpers take the reference's 1-based (i, j, k).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* constants.f08:7-12 */
#define O_PI 3.141592653589793238462643
#define O_MU0 (4.0 * O_PI * 1.0e-7)
#define O_EPS0 8.85418782e-12
#define O_STABILITY 1.0e25

/* averaging.f08:5-25 */
double oracle_scalar_mean(double a, double b, int32_t code)
{
    switch (code) {
    case 1: return 2.0 * a * b / (a + b);
    case 2: return sqrt(a * b);
    case 3: return 0.5 * (a + b);
    case 4: return a;
    default: return NAN;
    }
}

typedef struct {
    int nx, ny, nz;
    size_t n;
} grid_t;

#define I2(i, j) ((size_t)((i) - 1) + (size_t)g.nx * (size_t)((j) - 1))
#define I3(i, j, k) ((size_t)((i) - 1) + (size_t)g.nx * ((size_t)((j) - 1) + (size_t)g.ny * (size_t)((k) - 1)))

static double *alloc_zero(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

static double *alloc_fill(size_t n, double v)
{
    double *p = (double *)malloc((n ? n : 1) * sizeof(double));
    for (size_t i = 0; i < n; ++i) p[i] = v;
    return p;
}

/* input array with the reference's initial fill (cpmlfdtd.f08:240-247: kappa
 * = 1, alpha/acoef/bcoef = 0) when not supplied */
static double *input_copy(const oracle_run *r, int id, size_t n)
{
    double fill = (id == SDT_KAPPA || id == SDT_KAPPA_HALF) ? 1.0 : 0.0;
    double *p = alloc_fill(n, fill);
    if (r->arr[id]) memcpy(p, r->arr[id], n * sizeof(double));
    return p;
}

static double src_at(const oracle_run *r, int c, int it)
{
    return r->src[c] ? r->src[c][it - 1] : 0.0;
}

static void set_threads(const oracle_run *r)
{
#ifdef _OPENMP
    if (r->nthreads > 0) omp_set_num_threads(r->nthreads);
#else
    (void)r;
#endif
}

static void record(const oracle_run *r, const grid_t g, int it, const double *f0, const double *f1,
                   const double *f2)
{
    if (!r->rec_out) return;
    for (int q = 0; q < r->nrec; ++q) {
        int i = r->rec[3 * q], j = r->rec[3 * q + 1], k = r->rec[3 * q + 2];
        size_t c = (g.ny == 1) ? I2(i, k) : I3(i, j, k);
        double *o = r->rec_out + ((size_t)(it - 1) * r->nrec + q) * 3;
        o[0] = f0 ? f0[c] : 0.0;
        o[1] = f1 ? f1[c] : 0.0;
        o[2] = f2 ? f2[c] : 0.0;
    }
}

static void init_field(const oracle_run *r, int fid, int ic_id, double *dst, size_t n)
{
    if (r->field_in[fid]) memcpy(dst, r->field_in[fid], n * sizeof(double));
    else if (ic_id >= 0 && r->arr[ic_id]) memcpy(dst, r->arr[ic_id], n * sizeof(double));
}

static void emit(oracle_run *r, int fid, const double *src, size_t n)
{
    if (r->field_out[fid]) memcpy(r->field_out[fid], src, n * sizeof(double));
}

/* 3-term (reference) or 4-term (corrected) velocity-from-stress differences.
 * Reference: cpmlfdtd.f08:405-406 (backward), 434-435 (forward). */
static inline double tminus(double s_m2, double s_m1, double s_0, double s_p1, int corrected)
{
    if (corrected) return 27.0 * (s_0 - s_m1) - (s_p1 - s_m2);
    return 27.0 * s_0 - 27.0 * s_m1 + s_m2;
}
static inline double tplus(double s_m1, double s_0, double s_p1, double s_p2, int corrected)
{
    if (corrected) return 27.0 * (s_p1 - s_0) - (s_p2 - s_m1);
    return -27.0 * s_0 + 27.0 * s_p1 - s_p2;
}

/* ------------------------------------------------------------------------
 * plane-wave source: plane_wave_source.f08:67-230, faces and reference point
 * as set up in cpmlfdtd.f08:188-232 (2D), 802-832 (2.5D), 1548-1605 (3D)
 * ------------------------------------------------------------------------ */
typedef struct pw_state {
    double p[3], e[3], r0[3], loc[3];
    int active[6];
    int face[3];       /* cell index of the face normal to x, y, z (0 = none) */
} pw_state;

static void pw_setup(const oracle_run *r, int three_d, pw_state *w)
{
    const double deg = 3.1415926535897932384626433832795 / 180.0;
    const double cx = cos(deg * r->pw_angles[2]), cy = cos(deg * r->pw_angles[0]), cz = cos(deg * r->pw_angles[1]);
    const double sx = sin(deg * r->pw_angles[2]), sy = sin(deg * r->pw_angles[0]), sz = sin(deg * r->pw_angles[1]);
    /* reshape fills column by column: Rx(2,3) = sx, Rx(3,2) = -sx, ... */
    double Rx[3][3] = {{1, 0, 0}, {0, cx, sx}, {0, -sx, cx}};
    double Ry[3][3] = {{cy, 0, -sy}, {0, 1, 0}, {sy, 0, cy}};
    double Rz[3][3] = {{cz, sz, 0}, {-sz, cz, 0}, {0, 0, 1}};
    double T[3][3], R[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T[i][j] = 0.0;
            for (int k = 0; k < 3; ++k) T[i][j] += Rz[i][k] * Rx[k][j];
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            R[i][j] = 0.0;
            for (int k = 0; k < 3; ++k) R[i][j] += Ry[i][k] * T[k][j];
        }
    double p[3] = {R[0][2], R[1][2], R[2][2]};
    double n = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    if (n > 0.0) for (int i = 0; i < 3; ++i) p[i] /= n;
    double t[3] = {R[0][0], R[1][0], R[2][0]};
    double dp = t[0] * p[0] + t[1] * p[1] + t[2] * p[2];
    for (int i = 0; i < 3; ++i) t[i] -= dp * p[i];
    n = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    if (n > 0.0) {
        for (int i = 0; i < 3; ++i) w->e[i] = t[i] / n;
    } else {
        if (fabs(p[0]) < 0.9) { t[0] = 1; t[1] = 0; t[2] = 0; } else { t[0] = 0; t[1] = 1; t[2] = 0; }
        dp = t[0] * p[0] + t[1] * p[1] + t[2] * p[2];
        for (int i = 0; i < 3; ++i) t[i] -= dp * p[i];
        n = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        for (int i = 0; i < 3; ++i) w->e[i] = n > 0.0 ? t[i] / n : 0.0;
    }
    for (int i = 0; i < 3; ++i) w->p[i] = p[i];
    if (r->pw_kind == SDT_PW_P) for (int i = 0; i < 3; ++i) w->e[i] = p[i];
    const double eps = 1.0e-6;   /* select_injection_faces */
    for (int a = 0; a < 3; ++a) {
        w->active[2 * a] = p[a] > eps;
        w->active[2 * a + 1] = p[a] < -eps;
    }
    const int n_[3] = {r->nx, r->ny, r->nz};
    const double h[3] = {r->dx, r->dy, r->dz};
    for (int a = 0; a < 3; ++a) {
        const double lo = h[a] * (1 + r->cpml), hi = h[a] * (n_[a] - r->cpml);
        w->loc[a] = 0.5 * (lo + hi);
        if (w->active[2 * a]) w->loc[a] = lo;
        if (w->active[2 * a + 1]) w->loc[a] = hi;
        w->r0[a] = w->loc[a];
        w->face[a] = (w->active[2 * a] || w->active[2 * a + 1]) ? (int)lround(w->loc[a] / h[a]) : 0;
    }
    if (!three_d) { w->r0[1] = 0.0; w->loc[1] = 0.0; w->face[1] = 0; }
}

/* amplitude of the wave at position x, time t: amp * wavelet(tau) with
 * tau = t - p.(x - r0) / v (plane_wave_source.f08:141-161, 180-183, 214-218) */
static double pw_signal(const oracle_run *r, const pw_state *w, const double x[3], double t)
{
    const double tau = t - (w->p[0] * (x[0] - w->r0[0]) + w->p[1] * (x[1] - w->r0[1]) + w->p[2] * (x[2] - w->r0[2])) /
                               r->pw_speed;
    const double f0 = r->pw_f0, t0 = 1.5 / f0;
    const double a = 3.1415926535897932384626433832795 * f0 * (tau - t0), g = exp(-a * a);
    double wv = g;
    if (r->pw_wavelet == 1) wv = -2.0 * a * g;
    else if (r->pw_wavelet == 2) wv = (1.0 - 2.0 * a * a) * g;
    return r->pw_amp * wv;
}

/* boundary_em_field, plane_wave_source.f08:163-194: E = s e, H = (s / eta)
 * (p x e) / |p x e| */
static void pw_em_field(const oracle_run *r, const pw_state *w, const double x[3], double t, double E[3], double H[3])
{
    const double s = pw_signal(r, w, x, t);
    const double *e = w->e, *p = w->p;
    double hh[3] = {p[1] * e[2] - p[2] * e[1], p[2] * e[0] - p[0] * e[2], p[0] * e[1] - p[1] * e[0]};
    const double hn = sqrt(hh[0] * hh[0] + hh[1] * hh[1] + hh[2] * hh[2]);
    if (hn > 0.0) for (int i = 0; i < 3; ++i) hh[i] /= hn;
    for (int i = 0; i < 3; ++i) {
        E[i] = s * e[i];
        H[i] = (s / r->pw_eta) * hh[i];
    }
}

/* boundary_seismic_field, plane_wave_source.f08:196-230 */
static void pw_field(const oracle_run *r, const pw_state *w, const double x[3], double t, double V[3], double st[6])
{
    const double s = pw_signal(r, w, x, t), sc = -s / r->pw_speed;
    const double *e = w->e, *p = w->p;
    for (int i = 0; i < 3; ++i) V[i] = s * e[i];
    st[0] = sc * e[0] * p[0];
    st[1] = sc * e[1] * p[1];
    st[2] = sc * e[2] * p[2];
    st[3] = sc * (e[1] * p[2] + e[2] * p[1]);
    st[4] = sc * (e[0] * p[2] + e[2] * p[0]);
    st[5] = sc * (e[0] * p[1] + e[1] * p[0]);
}

/* add_plane_wave_stress (cpmlfdtd.f08:16-50) into s[6] = xx yy zz yz xz xy */
static void pw_stress(const double *C21, const double st[6], double s[6])
{
    static const int ij[6][6] = {{0, 1, 2, 3, 4, 5},   {1, 6, 7, 8, 9, 10},   {2, 7, 11, 12, 13, 14},
                                 {3, 8, 12, 15, 16, 17}, {4, 9, 13, 16, 18, 19}, {5, 10, 14, 17, 19, 20}};
    for (int rw = 0; rw < 6; ++rw)
        for (int q = 0; q < 6; ++q) s[rw] = s[rw] + C21[ij[rw][q]] * st[q];
}

/* ========================================================================
 * seismic2: cpmlfdtd.f08:57-597 (time loop 311-576)
 * ======================================================================== */
int32_t oracle_seismic2(oracle_run *r)
{
    set_threads(r);
    grid_t g = {r->nx, 1, r->nz, (size_t)r->nx * (size_t)r->nz};
    const int nx = g.nx, nz = g.nz;
    const double dx = r->dx, dz = r->dz, dt = r->dt;
    const int dc = r->density_code, corr = r->stencil == SDT_STENCIL_CORRECTED;
    const size_t n = g.n;

    double *c11 = input_copy(r, SDT_C11, n), *c13 = input_copy(r, SDT_C13, n),
           *c15 = input_copy(r, SDT_C15, n), *c33 = input_copy(r, SDT_C33, n),
           *c35 = input_copy(r, SDT_C35, n), *c55 = input_copy(r, SDT_C55, n),
           *rho = input_copy(r, SDT_RHO, n);
    double *gamma_x = input_copy(r, SDT_GAMMA_X, n), *gamma_z = input_copy(r, SDT_GAMMA_Z, n),
           *gamma_xz = input_copy(r, SDT_GAMMA_XZ, n);
    double *kappa = input_copy(r, SDT_KAPPA, n), *acoef = input_copy(r, SDT_ACOEF, n),
           *bcoef = input_copy(r, SDT_BCOEF, n);
    double *kappa_half = input_copy(r, SDT_KAPPA_HALF, n),
           *acoef_half = input_copy(r, SDT_ACOEF_HALF, n),
           *bcoef_half = input_copy(r, SDT_BCOEF_HALF, n);

    double *vx = alloc_zero(n), *vz = alloc_zero(n);
    double *sigmaxx = alloc_zero(n), *sigmazz = alloc_zero(n), *sigmaxz = alloc_zero(n);
    double *m_dvx_dx1 = alloc_zero(n), *m_dvx_dz1 = alloc_zero(n), *m_dvz_dx1 = alloc_zero(n),
           *m_dvz_dz1 = alloc_zero(n), *m_dvx_dx2 = alloc_zero(n), *m_dvx_dz2 = alloc_zero(n),
           *m_dvz_dx2 = alloc_zero(n), *m_dvz_dz2 = alloc_zero(n);
    double *m_dsxx_dx = alloc_zero(n), *m_dszz_dz = alloc_zero(n), *m_dsxz_dx = alloc_zero(n),
           *m_dsxz_dz = alloc_zero(n);

    /* initialconditionVx/Vz (cpmlfdtd.f08:265-266); stresses start at 0 */
    init_field(r, SDT_F_VX, SDT_IC_X, vx, n);
    init_field(r, SDT_F_VZ, SDT_IC_Z, vz, n);
    init_field(r, SDT_F_SXX, -1, sigmaxx, n);
    init_field(r, SDT_F_SZZ, -1, sigmazz, n);
    init_field(r, SDT_F_SXZ, -1, sigmaxz, n);

    const int is = r->isrc, js = r->ksrc;
    r->unstable_step = 0;
    pw_state pw;
    if (r->pw_kind) pw_setup(r, 0, &pw);

    for (int it = 1; it <= r->nsteps; ++it) {
        /* sigma_xx, sigma_zz: cpmlfdtd.f08:324-353 */
#pragma omp parallel for schedule(static)
        for (int j = 3; j <= nz - 1; ++j)
            for (int i = 2; i <= nx - 2; ++i) {
                size_t c = I2(i, j);
                double value_dvx_dx = (27.0 * (vx[I2(i + 1, j)] - vx[c]) - (vx[I2(i + 2, j)] - vx[I2(i - 1, j)])) / (24.0 * dx);
                double value_dvz_dz = (27.0 * (vz[c] - vz[I2(i, j - 1)]) - (vz[I2(i, j + 1)] - vz[I2(i, j - 2)])) / (24.0 * dz);
                double value_dvz_dx = (27.0 * (vz[I2(i + 1, j)] - vz[c]) - (vz[I2(i + 2, j)] - vz[I2(i - 1, j)])) / (24.0 * dx);
                double value_dvx_dz = (27.0 * (vx[c] - vx[I2(i, j - 1)]) - (vx[I2(i, j + 1)] - vx[I2(i, j - 2)])) / (24.0 * dz);

                m_dvx_dx1[c] = bcoef_half[c] * m_dvx_dx1[c] + acoef_half[c] * value_dvx_dx;
                m_dvx_dz1[c] = bcoef_half[c] * m_dvx_dz1[c] + acoef_half[c] * value_dvx_dz;
                m_dvz_dz1[c] = bcoef[c] * m_dvz_dz1[c] + acoef[c] * value_dvz_dz;
                m_dvz_dx1[c] = bcoef[c] * m_dvz_dx1[c] + acoef[c] * value_dvz_dx;

                value_dvx_dx = value_dvx_dx / kappa_half[c] + m_dvx_dx1[c];
                value_dvx_dz = value_dvx_dz / kappa_half[c] + m_dvx_dz1[c];
                value_dvz_dz = value_dvz_dz / kappa[c] + m_dvz_dz1[c];
                value_dvz_dx = value_dvz_dx / kappa[c] + m_dvz_dx1[c];

                sigmaxx[c] = (sigmaxx[c] + (c11[c] * value_dvx_dx + c13[c] * value_dvz_dz +
                                            c15[c] * (value_dvz_dx + value_dvx_dz)) * dt) /
                             (1 + gamma_x[c] * dt);
                sigmazz[c] = (sigmazz[c] + (c13[c] * value_dvx_dx + c33[c] * value_dvz_dz +
                                            c35[c] * (value_dvz_dx + value_dvx_dz)) * dt) /
                             (1 + gamma_z[c] * dt);
            }
        /* sigma_xz: cpmlfdtd.f08:362-386 */
#pragma omp parallel for schedule(static)
        for (int j = 2; j <= nz - 2; ++j)
            for (int i = 3; i <= nx - 1; ++i) {
                size_t c = I2(i, j);
                double value_dvx_dx = (27.0 * (vx[c] - vx[I2(i - 1, j)]) - (vx[I2(i + 1, j)] - vx[I2(i - 2, j)])) / (24.0 * dx);
                double value_dvz_dz = (27.0 * (vz[I2(i, j + 1)] - vz[c]) - (vz[I2(i, j + 2)] - vz[I2(i, j - 1)])) / (24.0 * dz);
                double value_dvz_dx = (27.0 * (vz[c] - vz[I2(i - 1, j)]) - (vz[I2(i + 1, j)] - vz[I2(i - 2, j)])) / (24.0 * dx);
                double value_dvx_dz = (27.0 * (vx[I2(i, j + 1)] - vx[c]) - (vx[I2(i, j + 2)] - vx[I2(i, j - 1)])) / (24.0 * dz);

                m_dvx_dx2[c] = bcoef_half[c] * m_dvx_dx2[c] + acoef_half[c] * value_dvx_dx;
                m_dvx_dz2[c] = bcoef_half[c] * m_dvx_dz2[c] + acoef_half[c] * value_dvx_dz;
                m_dvz_dz2[c] = bcoef_half[c] * m_dvz_dz2[c] + acoef_half[c] * value_dvz_dz;
                m_dvz_dx2[c] = bcoef_half[c] * m_dvz_dx2[c] + acoef_half[c] * value_dvz_dx;

                value_dvx_dx = value_dvx_dx / kappa_half[c] + m_dvx_dx2[c];
                value_dvx_dz = value_dvx_dz / kappa_half[c] + m_dvx_dz2[c];
                value_dvz_dz = value_dvz_dz / kappa_half[c] + m_dvz_dz2[c];
                value_dvz_dx = value_dvz_dx / kappa_half[c] + m_dvz_dx2[c];

                sigmaxz[c] = (sigmaxz[c] + (c15[c] * value_dvx_dx + c35[c] * value_dvz_dz +
                                            c55[c] * (value_dvz_dx + value_dvx_dz)) * dt) /
                             (1 + gamma_xz[c] * dt);
            }
        /* vx: cpmlfdtd.f08:399-416 */
#pragma omp parallel for schedule(static)
        for (int j = 3; j <= nz - 1; ++j)
            for (int i = 3; i <= nx - 1; ++i) {
                size_t c = I2(i, j);
                double rhoxx = oracle_scalar_mean(rho[c], rho[I2(i - 1, j)], dc);
                double rhozx = oracle_scalar_mean(rho[c], rho[I2(i, j - 1)], dc);
                double d_sxx_dx = tminus(sigmaxx[I2(i - 2, j)], sigmaxx[I2(i - 1, j)], sigmaxx[c],
                                         sigmaxx[I2(i + 1, j)], corr) / (24.0 * dx);
                double d_sxz_dz = tminus(sigmaxz[I2(i, j - 2)], sigmaxz[I2(i, j - 1)], sigmaxz[c],
                                         sigmaxz[I2(i, j + 1)], corr) / (24.0 * dz);
                m_dsxx_dx[c] = bcoef[c] * m_dsxx_dx[c] + acoef[c] * d_sxx_dx;
                m_dsxz_dz[c] = bcoef[c] * m_dsxz_dz[c] + acoef[c] * d_sxz_dz;
                d_sxx_dx = d_sxx_dx / kappa[c] + m_dsxx_dx[c];
                d_sxz_dz = d_sxz_dz / kappa[c] + m_dsxz_dz[c];
                vx[c] = vx[c] + dt * (d_sxx_dx / rhoxx + d_sxz_dz / rhozx);
            }
        /* vz: cpmlfdtd.f08:428-446 */
#pragma omp parallel for schedule(static)
        for (int j = 2; j <= nz - 2; ++j)
            for (int i = 2; i <= nx - 2; ++i) {
                size_t c = I2(i, j);
                double rhoxz = oracle_scalar_mean(rho[c], rho[I2(i + 1, j)], dc);
                double rhozz = oracle_scalar_mean(rho[c], rho[I2(i, j + 1)], dc);
                double d_sxz_dx = tplus(sigmaxz[I2(i - 1, j)], sigmaxz[c], sigmaxz[I2(i + 1, j)],
                                        sigmaxz[I2(i + 2, j)], corr) / (24.0 * dx);
                double d_szz_dz = tplus(sigmazz[I2(i, j - 1)], sigmazz[c], sigmazz[I2(i, j + 1)],
                                        sigmazz[I2(i, j + 2)], corr) / (24.0 * dz);
                m_dsxz_dx[c] = bcoef_half[c] * m_dsxz_dx[c] + acoef_half[c] * d_sxz_dx;
                m_dszz_dz[c] = bcoef_half[c] * m_dszz_dz[c] + acoef_half[c] * d_szz_dz;
                d_sxz_dx = d_sxz_dx / kappa_half[c] + m_dsxz_dx[c];
                d_szz_dz = d_szz_dz / kappa_half[c] + m_dszz_dz[c];
                vz[c] = vz[c] + dt * (d_sxz_dx / rhoxz + d_szz_dz / rhozz);
            }
        if (r->pw_kind) {
            /* plane wave, cpmlfdtd.f08:454-512 */
            const double t = it * dt;
            for (int f = 0; f < 3; f += 2) {
                if (!pw.face[f]) continue;
                const int lo = r->cpml + 2, hi = (f == 0 ? nz : nx) - r->cpml - 1;
                for (int ip = lo; ip <= hi; ++ip) {
                    const int i = f == 0 ? pw.face[0] : ip, j = f == 0 ? ip : pw.face[2];
                    const double x[3] = {f == 0 ? pw.loc[0] : ip * dx, 0.0, f == 0 ? ip * dz : pw.loc[2]};
                    double V[3], st[6];
                    pw_field(r, &pw, x, t, V, st);
                    const size_t c = I2(i, j);
                    vx[c] = vx[c] + V[0];
                    vz[c] = vz[c] + V[2];
                    sigmaxx[c] = sigmaxx[c] + c11[c] * st[0] + c13[c] * st[2] + c15[c] * st[4];
                    sigmazz[c] = sigmazz[c] + c13[c] * st[0] + c33[c] * st[2] + c35[c] * st[4];
                    sigmaxz[c] = sigmaxz[c] + c15[c] * st[0] + c35[c] * st[2] + c55[c] * st[4];
                }
            }
        } else {
        /* point source, cpmlfdtd.f08:519-523 (velocity sources carry no dt) */
        sigmaxx[I2(is, js)] = sigmaxx[I2(is, js)] + src_at(r, SDT_SRC_XX, it) / rho[I2(is, js)];
        sigmaxz[I2(is + 1, js + 1)] = sigmaxz[I2(is + 1, js + 1)] + src_at(r, SDT_SRC_XZ, it) / rho[I2(is + 1, js + 1)];
        sigmazz[I2(is, js)] = sigmazz[I2(is, js)] + src_at(r, SDT_SRC_ZZ, it) / rho[I2(is, js)];
        vx[I2(is, js)] = vx[I2(is, js)] + src_at(r, SDT_SRC_X, it) / rho[I2(is, js)];
        vz[I2(is, js)] = vz[I2(is, js)] + src_at(r, SDT_SRC_Z, it) / rho[I2(is, js)];
        }
        /* Dirichlet edges, cpmlfdtd.f08:534-549 */
        for (int j = 1; j <= nz; ++j) {
            vx[I2(1, j)] = 0.0; vx[I2(nx, j)] = 0.0;
            vz[I2(1, j)] = 0.0; vz[I2(nx, j)] = 0.0;
        }
        for (int i = 1; i <= nx; ++i) {
            vx[I2(i, 1)] = 0.0; vx[I2(i, nz)] = 0.0;
            vz[I2(i, 1)] = 0.0; vz[I2(i, nz)] = 0.0;
        }
        /* velocity norm, cpmlfdtd.f08:552-567 */
        double velocnorm = 0.0;
#pragma omp parallel for reduction(max : velocnorm) schedule(static)
        for (int j = 1; j <= nz; ++j)
            for (int i = 1; i <= nx; ++i) {
                size_t c = I2(i, j);
                double v = sqrt(vx[c] * vx[c] + vz[c] * vz[c]);
                if (v > velocnorm) velocnorm = v;
            }
        if (r->velnorm) r->velnorm[it - 1] = velocnorm;
        record(r, g, it, vx, NULL, vz);
        if (velocnorm > O_STABILITY) { r->unstable_step = it; break; }
    }

    emit(r, SDT_F_VX, vx, n);
    emit(r, SDT_F_VZ, vz, n);
    emit(r, SDT_F_SXX, sigmaxx, n);
    emit(r, SDT_F_SZZ, sigmazz, n);
    emit(r, SDT_F_SXZ, sigmaxz, n);

    double *all[] = {c11, c13, c15, c33, c35, c55, rho, gamma_x, gamma_z, gamma_xz, kappa, acoef,
                     bcoef, kappa_half, acoef_half, bcoef_half, vx, vz, sigmaxx, sigmazz, sigmaxz,
                     m_dvx_dx1, m_dvx_dz1, m_dvz_dx1, m_dvz_dz1, m_dvx_dx2, m_dvx_dz2, m_dvz_dx2,
                     m_dvz_dz2, m_dsxx_dx, m_dszz_dz, m_dsxz_dx, m_dsxz_dz};
    for (size_t q = 0; q < sizeof(all) / sizeof(all[0]); ++q) free(all[q]);
    return r->unstable_step ? SDT_ERR_UNSTABLE : SDT_OK;
}

/* ========================================================================
 * seismic3: cpmlfdtd.f08:1341-2313 (time loop 1747-2277)
 * ======================================================================== */

/* 4th-order staggered differences of the velocity-to-stress nests,
 * cpmlfdtd.f08:1765-1775: forward D+ (f(i+1)-f(i) ...) and backward D-. */
#define DPX(f, i, j, k) ((27.0 * (f[I3(i + 1, j, k)] - f[I3(i, j, k)]) - (f[I3(i + 2, j, k)] - f[I3(i - 1, j, k)])) / (24.0 * dx))
#define DMX(f, i, j, k) ((27.0 * (f[I3(i, j, k)] - f[I3(i - 1, j, k)]) - (f[I3(i + 1, j, k)] - f[I3(i - 2, j, k)])) / (24.0 * dx))
#define DPY(f, i, j, k) ((27.0 * (f[I3(i, j + 1, k)] - f[I3(i, j, k)]) - (f[I3(i, j + 2, k)] - f[I3(i, j - 1, k)])) / (24.0 * dy))
#define DMY(f, i, j, k) ((27.0 * (f[I3(i, j, k)] - f[I3(i, j - 1, k)]) - (f[I3(i, j + 1, k)] - f[I3(i, j - 2, k)])) / (24.0 * dy))
#define DPZ(f, i, j, k) ((27.0 * (f[I3(i, j, k + 1)] - f[I3(i, j, k)]) - (f[I3(i, j, k + 2)] - f[I3(i, j, k - 1)])) / (24.0 * dz))
#define DMZ(f, i, j, k) ((27.0 * (f[I3(i, j, k)] - f[I3(i, j, k - 1)]) - (f[I3(i, j, k + 1)] - f[I3(i, j, k - 2)])) / (24.0 * dz))

/* memory recursion + correction, e.g. cpmlfdtd.f08:1777,1787:
 * psi = b*psi + a*d ; d = d/kappa + psi */
#define CPML(d, m, a, b, kap) \
    do { (m) = (b) * (m) + (a) * (d); (d) = (d) / (kap) + (m); } while (0)

int32_t oracle_seismic3(oracle_run *r)
{
    set_threads(r);
    grid_t g = {r->nx, r->ny, r->nz, (size_t)r->nx * (size_t)r->ny * (size_t)r->nz};
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const double dx = r->dx, dy = r->dy, dz = r->dz, dt = r->dt;
    const int dc = r->density_code, corr = r->stencil == SDT_STENCIL_CORRECTED;
    const size_t n = g.n;

    double *C[21];
    for (int q = 0; q < 21; ++q) C[q] = input_copy(r, SDT_C11 + q, n);
    double *c11 = C[0], *c12 = C[1], *c13 = C[2], *c14 = C[3], *c15 = C[4], *c16 = C[5],
           *c22 = C[6], *c23 = C[7], *c24 = C[8], *c25 = C[9], *c26 = C[10], *c33 = C[11],
           *c34 = C[12], *c35 = C[13], *c36 = C[14], *c44 = C[15], *c45 = C[16], *c46 = C[17],
           *c55 = C[18], *c56 = C[19], *c66 = C[20];
    double *rho = input_copy(r, SDT_RHO, n);
    double *gamma_x = input_copy(r, SDT_GAMMA_X, n), *gamma_y = input_copy(r, SDT_GAMMA_Y, n),
           *gamma_z = input_copy(r, SDT_GAMMA_Z, n), *gamma_xy = input_copy(r, SDT_GAMMA_XY, n),
           *gamma_xz = input_copy(r, SDT_GAMMA_XZ, n), *gamma_yz = input_copy(r, SDT_GAMMA_YZ, n);
    double *kappa = input_copy(r, SDT_KAPPA, n), *acoef = input_copy(r, SDT_ACOEF, n),
           *bcoef = input_copy(r, SDT_BCOEF, n);
    double *kappa_half = input_copy(r, SDT_KAPPA_HALF, n),
           *acoef_half = input_copy(r, SDT_ACOEF_HALF, n),
           *bcoef_half = input_copy(r, SDT_BCOEF_HALF, n);

    double *vx = alloc_zero(n), *vy = alloc_zero(n), *vz = alloc_zero(n);
    double *sxx = alloc_zero(n), *syy = alloc_zero(n), *szz = alloc_zero(n), *sxy = alloc_zero(n),
           *sxz = alloc_zero(n), *syz = alloc_zero(n);
    /* memory_dv?_d?{1..4}: index [nest][component][direction] */
    double *mv[4][3][3];
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 3; ++c) mv[a][b][c] = alloc_zero(n);
    double *m_sxx_dx = alloc_zero(n), *m_syy_dy = alloc_zero(n), *m_szz_dz = alloc_zero(n),
           *m_sxy_dx = alloc_zero(n), *m_sxy_dy = alloc_zero(n), *m_sxz_dx = alloc_zero(n),
           *m_sxz_dz = alloc_zero(n), *m_syz_dy = alloc_zero(n), *m_syz_dz = alloc_zero(n);

    /* initial conditions, cpmlfdtd.f08:1643-1645 */
    init_field(r, SDT_F_VX, SDT_IC_X, vx, n);
    init_field(r, SDT_F_VY, SDT_IC_Y, vy, n);
    init_field(r, SDT_F_VZ, SDT_IC_Z, vz, n);
    init_field(r, SDT_F_SXX, -1, sxx, n);
    init_field(r, SDT_F_SYY, -1, syy, n);
    init_field(r, SDT_F_SZZ, -1, szz, n);
    init_field(r, SDT_F_SXY, -1, sxy, n);
    init_field(r, SDT_F_SXZ, -1, sxz, n);
    init_field(r, SDT_F_SYZ, -1, syz, n);

    const int is = r->isrc, js = r->jsrc, ks = r->ksrc;
    r->unstable_step = 0;
    enum { X = 0, Y = 1, Z = 2 };
    pw_state pw;
    if (r->pw_kind) pw_setup(r, 1, &pw);

    for (int it = 1; it <= r->nsteps; ++it) {
        /* Loop 1, sigma_xx/yy/zz: cpmlfdtd.f08:1761-1818 */
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 3; k <= nz - 1; ++k)
            for (int j = 3; j <= ny - 1; ++j)
                for (int i = 2; i <= nx - 2; ++i) {
                    size_t c = I3(i, j, k);
                    double dvx_dx = DPX(vx, i, j, k), dvy_dx = DPX(vy, i, j, k), dvz_dx = DPX(vz, i, j, k);
                    double dvx_dy = DMY(vx, i, j, k), dvy_dy = DMY(vy, i, j, k), dvz_dy = DMY(vz, i, j, k);
                    double dvx_dz = DMZ(vx, i, j, k), dvy_dz = DMZ(vy, i, j, k), dvz_dz = DMZ(vz, i, j, k);
                    CPML(dvx_dx, mv[0][X][X][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(dvy_dx, mv[0][Y][X][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(dvz_dx, mv[0][Z][X][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(dvy_dy, mv[0][Y][Y][c], acoef[c], bcoef[c], kappa[c]);
                    CPML(dvx_dy, mv[0][X][Y][c], acoef[c], bcoef[c], kappa[c]);
                    CPML(dvz_dy, mv[0][Z][Y][c], acoef[c], bcoef[c], kappa[c]);
                    CPML(dvz_dz, mv[0][Z][Z][c], acoef[c], bcoef[c], kappa[c]);
                    CPML(dvx_dz, mv[0][X][Z][c], acoef[c], bcoef[c], kappa[c]);
                    CPML(dvy_dz, mv[0][Y][Z][c], acoef[c], bcoef[c], kappa[c]);
                    sxx[c] = (sxx[c] + (c11[c] * dvx_dx + c12[c] * dvy_dy + c13[c] * dvz_dz +
                                        c14[c] * (dvy_dz + dvz_dy) + c15[c] * (dvx_dz + dvz_dx) +
                                        c16[c] * (dvy_dx + dvx_dy)) * dt) / (1 + gamma_x[c] * dt);
                    syy[c] = (syy[c] + (c12[c] * dvx_dx + c22[c] * dvy_dy + c23[c] * dvz_dz +
                                        c24[c] * (dvy_dz + dvz_dy) + c25[c] * (dvx_dz + dvz_dx) +
                                        c26[c] * (dvy_dx + dvx_dy)) * dt) / (1 + gamma_y[c] * dt);
                    szz[c] = (szz[c] + (c13[c] * dvx_dx + c23[c] * dvy_dy + c33[c] * dvz_dz +
                                        c34[c] * (dvy_dz + dvz_dy) + c35[c] * (dvx_dz + dvz_dx) +
                                        c36[c] * (dvy_dx + dvx_dy)) * dt) / (1 + gamma_z[c] * dt);
                }
        /* Loop 2, sigma_xy: cpmlfdtd.f08:1832-1875.  dvx_dz's memory variable
         * is updated but dvx_dz itself is used uncorrected (1855, 1858-1865). */
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 3; k <= nz - 1; ++k)
            for (int j = 2; j <= ny - 2; ++j)
                for (int i = 3; i <= nx - 1; ++i) {
                    size_t c = I3(i, j, k);
                    double dvx_dx = DMX(vx, i, j, k), dvy_dx = DMX(vy, i, j, k), dvz_dx = DMX(vz, i, j, k);
                    double dvx_dy = DPY(vx, i, j, k), dvy_dy = DPY(vy, i, j, k), dvz_dy = DPY(vz, i, j, k);
                    double dvx_dz = DMZ(vx, i, j, k), dvy_dz = DMZ(vy, i, j, k), dvz_dz = DMZ(vz, i, j, k);
                    CPML(dvx_dx, mv[1][X][X][c], acoef[c], bcoef[c], kappa[c]);
                    CPML(dvy_dx, mv[1][Y][X][c], acoef[c], bcoef[c], kappa[c]);
                    CPML(dvz_dx, mv[1][Z][X][c], acoef[c], bcoef[c], kappa[c]);
                    CPML(dvy_dy, mv[1][Y][Y][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(dvx_dy, mv[1][X][Y][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(dvz_dy, mv[1][Z][Y][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(dvz_dz, mv[1][Z][Z][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    mv[1][X][Z][c] = bcoef_half[c] * mv[1][X][Z][c] + acoef_half[c] * dvx_dz;
                    CPML(dvy_dz, mv[1][Y][Z][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    sxy[c] = (sxy[c] + (c16[c] * dvx_dx + c26[c] * dvy_dy + c36[c] * dvz_dz +
                                        c46[c] * (dvy_dz + dvz_dy) + c56[c] * (dvx_dz + dvz_dx) +
                                        c66[c] * (dvy_dx + dvx_dy)) * dt) / (1 + gamma_xy[c] * dt);
                }
        /* Loop 3, sigma_xz: cpmlfdtd.f08:1889-1933 */
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 2; k <= nz - 2; ++k)
            for (int j = 3; j <= ny - 1; ++j)
                for (int i = 3; i <= nx - 1; ++i) {
                    size_t c = I3(i, j, k);
                    double dvx_dx = DMX(vx, i, j, k), dvy_dx = DMX(vy, i, j, k), dvz_dx = DMX(vz, i, j, k);
                    double dvx_dy = DMY(vx, i, j, k), dvy_dy = DMY(vy, i, j, k), dvz_dy = DMY(vz, i, j, k);
                    double dvx_dz = DPZ(vx, i, j, k), dvy_dz = DPZ(vy, i, j, k), dvz_dz = DPZ(vz, i, j, k);
                    CPML(dvx_dx, mv[2][X][X][c], acoef[c], bcoef[c], kappa[c]);
                    CPML(dvy_dx, mv[2][Y][X][c], acoef[c], bcoef[c], kappa[c]);
                    CPML(dvz_dx, mv[2][Z][X][c], acoef[c], bcoef[c], kappa[c]);
                    CPML(dvy_dy, mv[2][Y][Y][c], acoef[c], bcoef[c], kappa[c]);
                    CPML(dvx_dy, mv[2][X][Y][c], acoef[c], bcoef[c], kappa[c]);
                    CPML(dvz_dy, mv[2][Z][Y][c], acoef[c], bcoef[c], kappa[c]);
                    CPML(dvz_dz, mv[2][Z][Z][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(dvx_dz, mv[2][X][Z][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(dvy_dz, mv[2][Y][Z][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    sxz[c] = (sxz[c] + (c15[c] * dvx_dx + c25[c] * dvy_dy + c35[c] * dvz_dz +
                                        c45[c] * (dvy_dz + dvz_dy) + c55[c] * (dvx_dz + dvz_dx) +
                                        c56[c] * (dvy_dx + dvx_dy)) * dt) / (1 + gamma_xz[c] * dt);
                }
        /* Loop 4, sigma_yz: cpmlfdtd.f08:1949-1992 */
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 2; k <= nz - 2; ++k)
            for (int j = 2; j <= ny - 2; ++j)
                for (int i = 2; i <= nx - 2; ++i) {
                    size_t c = I3(i, j, k);
                    double dvx_dx = DPX(vx, i, j, k), dvy_dx = DPX(vy, i, j, k), dvz_dx = DPX(vz, i, j, k);
                    double dvx_dy = DPY(vx, i, j, k), dvy_dy = DPY(vy, i, j, k), dvz_dy = DPY(vz, i, j, k);
                    double dvx_dz = DPZ(vx, i, j, k), dvy_dz = DPZ(vy, i, j, k), dvz_dz = DPZ(vz, i, j, k);
                    CPML(dvx_dx, mv[3][X][X][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(dvy_dx, mv[3][Y][X][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(dvz_dx, mv[3][Z][X][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(dvy_dy, mv[3][Y][Y][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(dvx_dy, mv[3][X][Y][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(dvz_dy, mv[3][Z][Y][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(dvz_dz, mv[3][Z][Z][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(dvx_dz, mv[3][X][Z][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(dvy_dz, mv[3][Y][Z][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    syz[c] = (syz[c] + (c14[c] * dvx_dx + c24[c] * dvy_dy + c34[c] * dvz_dz +
                                        c44[c] * (dvy_dz + dvz_dy) + c45[c] * (dvx_dz + dvz_dx) +
                                        c46[c] * (dvy_dx + dvx_dy)) * dt) / (1 + gamma_yz[c] * dt);
                }
        /* vx: cpmlfdtd.f08:2014-2038 */
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 3; k <= nz - 1; ++k)
            for (int j = 3; j <= ny - 1; ++j)
                for (int i = 3; i <= nx - 1; ++i) {
                    size_t c = I3(i, j, k);
                    double rhoxx = oracle_scalar_mean(rho[c], rho[I3(i - 1, j, k)], dc);
                    double rhoyx = oracle_scalar_mean(rho[c], rho[c], dc);
                    double rhozx = oracle_scalar_mean(rho[c], rho[I3(i, j, k - 1)], dc);
                    double d1 = tminus(sxx[I3(i - 2, j, k)], sxx[I3(i - 1, j, k)], sxx[c], sxx[I3(i + 1, j, k)], corr) / (24.0 * dx);
                    double d2 = tminus(sxy[I3(i, j - 2, k)], sxy[I3(i, j - 1, k)], sxy[c], sxy[I3(i, j + 1, k)], corr) / (24.0 * dy);
                    double d3 = tminus(sxz[I3(i, j, k - 2)], sxz[I3(i, j, k - 1)], sxz[c], sxz[I3(i, j, k + 1)], corr) / (24.0 * dz);
                    CPML(d1, m_sxx_dx[c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(d2, m_sxy_dy[c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(d3, m_sxz_dz[c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    vx[c] = vx[c] + (d1 / rhoxx + d2 / rhoyx + d3 / rhozx) * dt;
                }
        /* vy: cpmlfdtd.f08:2055-2080 */
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 3; k <= nz - 1; ++k)
            for (int j = 2; j <= ny - 2; ++j)
                for (int i = 2; i <= nx - 2; ++i) {
                    size_t c = I3(i, j, k);
                    double rhoxy = oracle_scalar_mean(rho[c], rho[I3(i + 1, j, k)], dc);
                    double rhoyy = oracle_scalar_mean(rho[c], rho[c], dc);
                    double rhozy = oracle_scalar_mean(rho[c], rho[I3(i, j, k - 1)], dc);
                    double d1 = tplus(sxy[I3(i - 1, j, k)], sxy[c], sxy[I3(i + 1, j, k)], sxy[I3(i + 2, j, k)], corr) / (24.0 * dx);
                    double d2 = tplus(syy[I3(i, j - 1, k)], syy[c], syy[I3(i, j + 1, k)], syy[I3(i, j + 2, k)], corr) / (24.0 * dy);
                    double d3 = tminus(syz[I3(i, j, k - 2)], syz[I3(i, j, k - 1)], syz[c], syz[I3(i, j, k + 1)], corr) / (24.0 * dz);
                    CPML(d1, m_sxy_dx[c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(d2, m_syy_dy[c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(d3, m_syz_dz[c], acoef[c], bcoef[c], kappa[c]);
                    vy[c] = vy[c] + (d1 / rhoxy + d2 / rhoyy + d3 / rhozy) * dt;
                }
        /* vz: cpmlfdtd.f08:2098-2127 */
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 2; k <= nz - 2; ++k)
            for (int j = 3; j <= ny - 1; ++j)
                for (int i = 2; i <= nx - 2; ++i) {
                    size_t c = I3(i, j, k);
                    double rhoxz = oracle_scalar_mean(rho[c], rho[I3(i + 1, j, k)], dc);
                    double rhoyz = oracle_scalar_mean(rho[c], rho[c], dc);
                    double rhozz = oracle_scalar_mean(rho[c], rho[I3(i, j, k + 1)], dc);
                    double d1 = tplus(sxz[I3(i - 1, j, k)], sxz[c], sxz[I3(i + 1, j, k)], sxz[I3(i + 2, j, k)], corr) / (24.0 * dx);
                    double d2 = tminus(syz[I3(i, j - 2, k)], syz[I3(i, j - 1, k)], syz[c], syz[I3(i, j + 1, k)], corr) / (24.0 * dy);
                    double d3 = tplus(szz[I3(i, j, k - 1)], szz[c], szz[I3(i, j, k + 1)], szz[I3(i, j, k + 2)], corr) / (24.0 * dz);
                    CPML(d1, m_sxz_dx[c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    CPML(d2, m_syz_dy[c], acoef[c], bcoef[c], kappa[c]);
                    CPML(d3, m_szz_dz[c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                    vz[c] = vz[c] + (d1 / rhoxz + d2 / rhoyz + d3 / rhozz) * dt;
                }
        if (r->pw_kind) {
            /* plane wave, cpmlfdtd.f08:2133-2200: faces x, y, z in turn */
            const double t = it * dt;
            const int lo[3] = {r->cpml + 2, r->cpml + 2, r->cpml + 2};
            const int hi[3] = {nx - r->cpml - 1, ny - r->cpml - 1, nz - r->cpml - 1};
            for (int f = 0; f < 3; ++f) {
                if (!pw.face[f]) continue;
                const int a = f == 0 ? 1 : 0, b = f == 2 ? 1 : 2;
                for (int u = lo[a]; u <= hi[a]; ++u)
                    for (int w = lo[b]; w <= hi[b]; ++w) {
                        int ijk[3];
                        ijk[f] = pw.face[f];
                        ijk[a] = u;
                        ijk[b] = w;
                        double x[3] = {ijk[0] * r->dx, ijk[1] * r->dy, ijk[2] * r->dz};
                        x[f] = pw.loc[f];
                        double V[3], st[6];
                        pw_field(r, &pw, x, t, V, st);
                        const size_t c = I3(ijk[0], ijk[1], ijk[2]);
                        vx[c] = vx[c] + V[0];
                        vy[c] = vy[c] + V[1];
                        vz[c] = vz[c] + V[2];
                        double C21[21], s6[6] = {sxx[c], syy[c], szz[c], syz[c], sxz[c], sxy[c]};
                        for (int q = 0; q < 21; ++q) C21[q] = C[q][c];
                        pw_stress(C21, st, s6);
                        sxx[c] = s6[0]; syy[c] = s6[1]; szz[c] = s6[2];
                        syz[c] = s6[3]; sxz[c] = s6[4]; sxy[c] = s6[5];
                    }
            }
        } else {
        /* point source, cpmlfdtd.f08:2208-2219 */
        sxx[I3(is, js, ks)] = sxx[I3(is, js, ks)] + src_at(r, SDT_SRC_XX, it) / rho[I3(is, js, ks)];
        sxy[I3(is + 1, js + 1, ks)] = sxy[I3(is + 1, js + 1, ks)] + src_at(r, SDT_SRC_XY, it) / rho[I3(is + 1, js + 1, ks)];
        sxz[I3(is + 1, js, ks + 1)] = sxz[I3(is + 1, js, ks + 1)] + src_at(r, SDT_SRC_XZ, it) / rho[I3(is + 1, js, ks)];
        syy[I3(is, js, ks)] = syy[I3(is, js, ks)] + src_at(r, SDT_SRC_YY, it) / rho[I3(is, js, ks)];
        syz[I3(is, js + 1, ks + 1)] = syz[I3(is, js + 1, ks + 1)] + src_at(r, SDT_SRC_YZ, it) / rho[I3(is, js, ks + 1)];
        szz[I3(is, js, ks)] = szz[I3(is, js, ks)] + src_at(r, SDT_SRC_ZZ, it) / rho[I3(is, js, ks)];
        vx[I3(is + 1, js, ks)] = vx[I3(is + 1, js, ks)] + src_at(r, SDT_SRC_X, it) * dt / rho[I3(is + 1, js, ks)];
        vy[I3(is, js + 1, ks)] = vy[I3(is, js + 1, ks)] + src_at(r, SDT_SRC_Y, it) * dt / rho[I3(is, js, ks)];
        vz[I3(is, js, ks + 1)] = vz[I3(is, js, ks + 1)] + src_at(r, SDT_SRC_Z, it) * dt / rho[I3(is, js, ks + 1)];
        }
        /* Dirichlet faces, cpmlfdtd.f08:2226-2248 */
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    if (i == 1 || i == nx || j == 1 || j == ny || k == 1 || k == nz) {
                        size_t c = I3(i, j, k);
                        vx[c] = 0.0; vy[c] = 0.0; vz[c] = 0.0;
                    }
                }
        /* norms, cpmlfdtd.f08:2255-2258 */
        double velocnorm = 0.0, stressnorm = 0.0;
#pragma omp parallel for collapse(2) reduction(max : velocnorm, stressnorm) schedule(static)
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    size_t c = I3(i, j, k);
                    double v = sqrt(vx[c] * vx[c] + vy[c] * vy[c] + vz[c] * vz[c]);
                    double s = sqrt(sxx[c] * sxx[c] + syy[c] * syy[c] + szz[c] * szz[c] +
                                    2.0 * (sxy[c] * sxy[c] + syz[c] * syz[c] + sxz[c] * sxz[c]));
                    if (v > velocnorm) velocnorm = v;
                    if (s > stressnorm) stressnorm = s;
                }
        if (r->velnorm) r->velnorm[it - 1] = velocnorm;
        if (r->stressnorm) r->stressnorm[it - 1] = stressnorm;
        record(r, g, it, vx, vy, vz);
        if (velocnorm > O_STABILITY) { r->unstable_step = it; break; }
    }

    emit(r, SDT_F_VX, vx, n); emit(r, SDT_F_VY, vy, n); emit(r, SDT_F_VZ, vz, n);
    emit(r, SDT_F_SXX, sxx, n); emit(r, SDT_F_SYY, syy, n); emit(r, SDT_F_SZZ, szz, n);
    emit(r, SDT_F_SXY, sxy, n); emit(r, SDT_F_SXZ, sxz, n); emit(r, SDT_F_SYZ, syz, n);

    for (int q = 0; q < 21; ++q) free(C[q]);
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 3; ++c) free(mv[a][b][c]);
    double *all[] = {rho, gamma_x, gamma_y, gamma_z, gamma_xy, gamma_xz, gamma_yz, kappa, acoef,
                     bcoef, kappa_half, acoef_half, bcoef_half, vx, vy, vz, sxx, syy, szz, sxy,
                     sxz, syz, m_sxx_dx, m_syy_dy, m_szz_dz, m_sxy_dx, m_sxy_dy, m_sxz_dx,
                     m_sxz_dz, m_syz_dy, m_syz_dz};
    for (size_t q = 0; q < sizeof(all) / sizeof(all[0]); ++q) free(all[q]);
    return r->unstable_step ? SDT_ERR_UNSTABLE : SDT_OK;
}

/* ========================================================================
 * electromag2: cpmlfdtd.f08:2317-2760 (time loop 2566-2748)
 * ======================================================================== */
int32_t oracle_electromag2(oracle_run *r)
{
    set_threads(r);
    grid_t g = {r->nx, 1, r->nz, (size_t)r->nx * (size_t)r->nz};
    const int nx = g.nx, nz = g.nz;
    const double dx = r->dx, dz = r->dz, dt = r->dt;
    const size_t n = g.n;

    double *eps11 = input_copy(r, SDT_E11, n), *eps13 = input_copy(r, SDT_E13, n),
           *eps33 = input_copy(r, SDT_E33, n);
    double *sig11 = input_copy(r, SDT_S11, n), *sig13 = input_copy(r, SDT_S13, n),
           *sig33 = input_copy(r, SDT_S33, n);
    double *kappa = input_copy(r, SDT_KAPPA, n), *acoef = input_copy(r, SDT_ACOEF, n),
           *bcoef = input_copy(r, SDT_BCOEF, n);
    double *kappa_half = input_copy(r, SDT_KAPPA_HALF, n),
           *acoef_half = input_copy(r, SDT_ACOEF_HALF, n),
           *bcoef_half = input_copy(r, SDT_BCOEF_HALF, n);
    double *aEx = alloc_zero(n), *bEx = alloc_zero(n), *dEx = alloc_zero(n), *eEx = alloc_zero(n),
           *aEz = alloc_zero(n), *bEz = alloc_zero(n), *dEz = alloc_zero(n), *eEz = alloc_zero(n);
    double *epsilon11 = alloc_zero(n), *epsilon13 = alloc_zero(n), *epsilon33 = alloc_zero(n);
    double *Ex = alloc_zero(n), *Ez = alloc_zero(n), *Hy = alloc_zero(n);
    double *Ex_old = alloc_zero(n), *Ez_old = alloc_zero(n);
    double *m_dEx_dz = alloc_zero(n), *m_dEz_dx = alloc_zero(n), *m_dHy_dx = alloc_zero(n),
           *m_dHy_dz = alloc_zero(n);

    init_field(r, SDT_F_EX, SDT_IC_X, Ex, n);   /* cpmlfdtd.f08:2505-2506 */
    init_field(r, SDT_F_EZ, SDT_IC_Z, Ez, n);
    init_field(r, SDT_F_HY, -1, Hy, n);

    /* coefficients, cpmlfdtd.f08:2516-2540 */
    const double dbHy = dt / O_MU0, daHy = 1.0;
    for (size_t c = 0; c < n; ++c) {
        epsilon11[c] = eps11[c] * O_EPS0;
        epsilon13[c] = eps13[c] * O_EPS0;
        epsilon33[c] = eps33[c] * O_EPS0;
        double det = epsilon11[c] * epsilon33[c] - epsilon13[c] * epsilon13[c];
        if (det <= 1.0e-18 * O_EPS0 * O_EPS0) det = 1.0e-18 * O_EPS0 * O_EPS0;
        aEx[c] = epsilon33[c] / det;
        bEx[c] = -epsilon13[c] / det;
        dEx[c] = (epsilon33[c] * sig11[c] - epsilon13[c] * sig13[c]) / det;
        eEx[c] = (epsilon33[c] * sig13[c] - epsilon13[c] * sig33[c]) / det;
        aEz[c] = -epsilon13[c] / det;
        bEz[c] = epsilon11[c] / det;
        dEz[c] = (epsilon11[c] * sig13[c] - epsilon13[c] * sig11[c]) / det;
        eEz[c] = (epsilon11[c] * sig33[c] - epsilon13[c] * sig13[c]) / det;
    }
    memcpy(Ex_old, Ex, n * sizeof(double));
    memcpy(Ez_old, Ez, n * sizeof(double));

    const int is = r->isrc, js = r->ksrc;
    pw_state pw;
    if (r->pw_kind) pw_setup(r, 0, &pw);
    r->unstable_step = 0;
    for (int it = 1; it <= r->nsteps; ++it) {
        /* Hy: cpmlfdtd.f08:2576-2593 */
#pragma omp parallel for schedule(static)
        for (int j = 1; j <= nz - 1; ++j)
            for (int i = 1; i <= nx - 1; ++i) {
                size_t c = I2(i, j);
                double dEx_dz = (Ex[I2(i, j + 1)] - Ex[c]) / dz;
                double dEz_dx = (Ez[I2(i + 1, j)] - Ez[c]) / dx;
                m_dEx_dz[c] = bcoef[c] * m_dEx_dz[c] + acoef[c] * dEx_dz;
                m_dEz_dx[c] = bcoef_half[c] * m_dEz_dx[c] + acoef_half[c] * dEz_dx;
                dEz_dx = dEz_dx / kappa_half[c] + m_dEz_dx[c];
                dEx_dz = dEx_dz / kappa[c] + m_dEx_dz[c];
                Hy[c] = daHy * Hy[c] + dbHy * (dEz_dx - dEx_dz);
            }
        /* Ex, Ez: cpmlfdtd.f08:2605-2620 */
#pragma omp parallel for schedule(static)
        for (int j = 2; j <= nz; ++j)
            for (int i = 2; i <= nx; ++i) {
                size_t c = I2(i, j);
                double dHy_dz = (Hy[c] - Hy[I2(i, j - 1)]) / dz;
                double dHy_dx = (Hy[c] - Hy[I2(i - 1, j)]) / dx;
                m_dHy_dx[c] = bcoef_half[c] * m_dHy_dx[c] + acoef_half[c] * dHy_dx;
                m_dHy_dz[c] = bcoef[c] * m_dHy_dz[c] + acoef[c] * dHy_dz;
                dHy_dz = dHy_dz / kappa[c] + m_dHy_dz[c];
                dHy_dx = dHy_dx / kappa_half[c] + m_dHy_dx[c];
                Ez[c] = Ez_old[c] + (-aEz[c] * dHy_dz + bEz[c] * dHy_dx - dEz[c] * Ex_old[c] - eEz[c] * Ez_old[c]) * dt;
                Ex[c] = Ex_old[c] + (-aEx[c] * dHy_dz + bEx[c] * dHy_dx - dEx[c] * Ex_old[c] - eEx[c] * Ez_old[c]) * dt;
            }
        if (r->pw_kind) {
            /* plane wave, cpmlfdtd.f08:2626-2673 */
            const double t = it * dt;
            for (int f = 0; f < 3; f += 2) {
                if (!pw.face[f]) continue;
                const int lo = r->cpml + 2, hi = (f == 0 ? nz : nx) - r->cpml - 1;
                for (int ip = lo; ip <= hi; ++ip) {
                    const int i = f == 0 ? pw.face[0] : ip, j = f == 0 ? ip : pw.face[2];
                    const double x[3] = {f == 0 ? pw.loc[0] : ip * dx, 0.0, f == 0 ? ip * dz : pw.loc[2]};
                    double E[3], H[3];
                    pw_em_field(r, &pw, x, t, E, H);
                    const size_t c = I2(i, j);
                    Ex[c] = Ex[c] + E[0];
                    Ez[c] = Ez[c] + E[2];
                    Hy[c] = Hy[c] + H[1];
                }
            }
        } else {
        /* point source, cpmlfdtd.f08:2678-2681 */
        Ex[I2(is, js)] = Ex[I2(is, js)] + src_at(r, SDT_SRC_X, it) * dt / epsilon11[I2(is, js)];
        Ez[I2(is, js)] = Ez[I2(is, js)] + src_at(r, SDT_SRC_Z, it) * dt / epsilon33[I2(is, js)];
        }
        /* Dirichlet edges, cpmlfdtd.f08:2692-2711 */
        for (int j = 1; j <= nz; ++j) {
            Ex[I2(1, j)] = 0.0; Ex[I2(nx, j)] = 0.0;
            Ez[I2(1, j)] = 0.0; Ez[I2(nx, j)] = 0.0;
            Hy[I2(1, j)] = 0.0; Hy[I2(nx, j)] = 0.0;
        }
        for (int i = 1; i <= nx; ++i) {
            Ex[I2(i, 1)] = 0.0; Ex[I2(i, nz)] = 0.0;
            Ez[I2(i, 1)] = 0.0; Ez[I2(i, nz)] = 0.0;
            Hy[I2(i, 1)] = 0.0; Hy[I2(i, nz)] = 0.0;
        }
        /* E_old = E, cpmlfdtd.f08:2716-2721 */
        memcpy(Ex_old, Ex, n * sizeof(double));
        memcpy(Ez_old, Ez, n * sizeof(double));
        /* norm, cpmlfdtd.f08:2726-2740 */
        double velocnorm = 0.0;
#pragma omp parallel for reduction(max : velocnorm) schedule(static)
        for (int j = 1; j <= nz; ++j)
            for (int i = 1; i <= nx; ++i) {
                size_t c = I2(i, j);
                double v = sqrt(Ex[c] * Ex[c] + Ez[c] * Ez[c]);
                if (v > velocnorm) velocnorm = v;
            }
        if (r->velnorm) r->velnorm[it - 1] = velocnorm;
        record(r, g, it, Ex, NULL, Ez);
        if (velocnorm > O_STABILITY) { r->unstable_step = it; break; }
    }
    emit(r, SDT_F_EX, Ex, n); emit(r, SDT_F_EZ, Ez, n); emit(r, SDT_F_HY, Hy, n);

    double *all[] = {eps11, eps13, eps33, sig11, sig13, sig33, kappa, acoef, bcoef, kappa_half,
                     acoef_half, bcoef_half, aEx, bEx, dEx, eEx, aEz, bEz, dEz, eEz, epsilon11,
                     epsilon13, epsilon33, Ex, Ez, Hy, Ex_old, Ez_old, m_dEx_dz, m_dEz_dx,
                     m_dHy_dx, m_dHy_dz};
    for (size_t q = 0; q < sizeof(all) / sizeof(all[0]); ++q) free(all[q]);
    return r->unstable_step ? SDT_ERR_UNSTABLE : SDT_OK;
}

/* ========================================================================
 * electromag3: cpmlfdtd.f08:3353-4001 (time loop 3689-3978)
 * ======================================================================== */
int32_t oracle_electromag3(oracle_run *r)
{
    set_threads(r);
    grid_t g = {r->nx, r->ny, r->nz, (size_t)r->nx * (size_t)r->ny * (size_t)r->nz};
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const double dx = r->dx, dy = r->dy, dz = r->dz, dt = r->dt;
    const size_t n = g.n;

    double *eps11 = input_copy(r, SDT_E11, n), *eps12 = input_copy(r, SDT_E12, n),
           *eps13 = input_copy(r, SDT_E13, n), *eps22 = input_copy(r, SDT_E22, n),
           *eps23 = input_copy(r, SDT_E23, n), *eps33 = input_copy(r, SDT_E33, n);
    double *sig11 = input_copy(r, SDT_S11, n), *sig12 = input_copy(r, SDT_S12, n),
           *sig13 = input_copy(r, SDT_S13, n), *sig22 = input_copy(r, SDT_S22, n),
           *sig23 = input_copy(r, SDT_S23, n), *sig33 = input_copy(r, SDT_S33, n);
    double *kappa = input_copy(r, SDT_KAPPA, n), *acoef = input_copy(r, SDT_ACOEF, n),
           *bcoef = input_copy(r, SDT_BCOEF, n);
    double *kappa_half = input_copy(r, SDT_KAPPA_HALF, n),
           *acoef_half = input_copy(r, SDT_ACOEF_HALF, n),
           *bcoef_half = input_copy(r, SDT_BCOEF_HALF, n);
    double *aEx = alloc_zero(n), *bEx = alloc_zero(n), *cEx = alloc_zero(n), *bEy = alloc_zero(n),
           *cEy = alloc_zero(n), *cEz = alloc_zero(n);
    double *Ex = alloc_zero(n), *Ey = alloc_zero(n), *Ez = alloc_zero(n), *Hx = alloc_zero(n),
           *Hy = alloc_zero(n), *Hz = alloc_zero(n);
    double *Ex_old = alloc_zero(n), *Ey_old = alloc_zero(n), *Ez_old = alloc_zero(n);
    double *m_dEx_dy = alloc_zero(n), *m_dEy_dx = alloc_zero(n), *m_dEx_dz = alloc_zero(n),
           *m_dEz_dx = alloc_zero(n), *m_dEz_dy = alloc_zero(n), *m_dEy_dz = alloc_zero(n);
    double *m_dHz_dx = alloc_zero(n), *m_dHx_dz = alloc_zero(n), *m_dHz_dy = alloc_zero(n),
           *m_dHy_dz = alloc_zero(n), *m_dHx_dy = alloc_zero(n), *m_dHy_dx = alloc_zero(n);

    init_field(r, SDT_F_EX, SDT_IC_X, Ex, n);   /* cpmlfdtd.f08:3616-3618 */
    init_field(r, SDT_F_EY, SDT_IC_Y, Ey, n);
    init_field(r, SDT_F_EZ, SDT_IC_Z, Ez, n);
    init_field(r, SDT_F_HX, -1, Hx, n);
    init_field(r, SDT_F_HY, -1, Hy, n);
    init_field(r, SDT_F_HZ, -1, Hz, n);

    /* cpmlfdtd.f08:3510-3518 and 3640-3661 */
    const double dbH = dt / O_MU0, daH = 1.0;
    for (size_t c = 0; c < n; ++c) {
        eps11[c] = eps11[c] * O_EPS0; eps12[c] = eps12[c] * O_EPS0; eps13[c] = eps13[c] * O_EPS0;
        eps22[c] = eps22[c] * O_EPS0; eps23[c] = eps23[c] * O_EPS0; eps33[c] = eps33[c] * O_EPS0;
        double det = eps11[c] * (eps22[c] * eps33[c] - eps23[c] * eps23[c]) -
                     eps12[c] * (eps12[c] * eps33[c] - eps23[c] * eps13[c]) +
                     eps13[c] * (eps12[c] * eps23[c] - eps22[c] * eps13[c]);
        aEx[c] = (eps22[c] * eps33[c] - eps23[c] * eps23[c]) / det;
        bEx[c] = -(eps12[c] * eps33[c] - eps23[c] * eps13[c]) / det;
        cEx[c] = (eps12[c] * eps23[c] - eps22[c] * eps13[c]) / det;
        bEy[c] = (eps11[c] * eps33[c] - eps13[c] * eps13[c]) / det;
        cEy[c] = -(eps11[c] * eps23[c] - eps12[c] * eps13[c]) / det;
        cEz[c] = (eps11[c] * eps22[c] - eps12[c] * eps12[c]) / det;
    }
    /* aEy = bEx, aEz = cEx, bEz = cEy (cpmlfdtd.f08:3655-3660) */
    double *aEy = bEx, *aEz = cEx, *bEz = cEy;
    memcpy(Ex_old, Ex, n * sizeof(double));
    memcpy(Ey_old, Ey, n * sizeof(double));
    memcpy(Ez_old, Ez, n * sizeof(double));

    const int is = r->isrc, js = r->jsrc, ks = r->ksrc;
    pw_state pw;
    if (r->pw_kind) pw_setup(r, 1, &pw);
    r->unstable_step = 0;
    for (int it = 1; it <= r->nsteps; ++it) {
        /* Hx: cpmlfdtd.f08:3700-3718 */
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 1; k <= nz - 1; ++k)
            for (int i = 1; i <= nx - 1; ++i)
                for (int j = 1; j <= ny - 1; ++j) {
                    size_t c = I3(i, j, k);
                    double dEz_dy = (Ez[I3(i, j + 1, k)] - Ez[c]) / dy;
                    double dEy_dz = (Ey[I3(i, j, k + 1)] - Ey[c]) / dz;
                    m_dEy_dz[c] = bcoef_half[c] * m_dEy_dz[c] + acoef_half[c] * dEy_dz;
                    m_dEz_dy[c] = bcoef_half[c] * m_dEz_dy[c] + acoef_half[c] * dEz_dy;
                    dEz_dy = dEz_dy / kappa_half[c] + m_dEz_dy[c];
                    dEy_dz = dEy_dz / kappa_half[c] + m_dEy_dz[c];
                    Hx[c] = daH * Hx[c] + dbH * (dEy_dz - dEz_dy);
                }
        /* Hy: cpmlfdtd.f08:3730-3749 */
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 1; k <= nz - 1; ++k)
            for (int i = 1; i <= nx - 1; ++i)
                for (int j = 1; j <= ny - 1; ++j) {
                    size_t c = I3(i, j, k);
                    double dEx_dz = (Ex[I3(i, j, k + 1)] - Ex[c]) / dz;
                    double dEz_dx = (Ez[I3(i + 1, j, k)] - Ez[c]) / dx;
                    m_dEx_dz[c] = bcoef[c] * m_dEx_dz[c] + acoef[c] * dEx_dz;
                    m_dEz_dx[c] = bcoef[c] * m_dEz_dx[c] + acoef[c] * dEz_dx;
                    dEx_dz = dEx_dz / kappa[c] + m_dEx_dz[c];
                    dEz_dx = dEz_dx / kappa[c] + m_dEz_dx[c];
                    Hy[c] = daH * Hy[c] + dbH * (dEz_dx - dEx_dz);
                }
        /* Hz: cpmlfdtd.f08:3761-3778 (k starts at 2) */
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 2; k <= nz - 1; ++k)
            for (int i = 1; i <= nx - 1; ++i)
                for (int j = 1; j <= ny - 1; ++j) {
                    size_t c = I3(i, j, k);
                    double dEx_dy = (Ex[I3(i, j + 1, k)] - Ex[c]) / dy;
                    double dEy_dx = (Ey[I3(i + 1, j, k)] - Ey[c]) / dx;
                    m_dEx_dy[c] = bcoef[c] * m_dEx_dy[c] + acoef[c] * dEx_dy;
                    m_dEy_dx[c] = bcoef[c] * m_dEy_dx[c] + acoef[c] * dEy_dx;
                    dEx_dy = dEx_dy / kappa[c] + m_dEx_dy[c];
                    dEy_dx = dEy_dx / kappa[c] + m_dEy_dx[c];
                    Hz[c] = daH * Hz[c] + dbH * (dEx_dy - dEy_dx);
                }
        /* E: cpmlfdtd.f08:3795-3828 */
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 2; k <= nz - 1; ++k)
            for (int i = 2; i <= nx - 1; ++i)
                for (int j = 2; j <= ny - 1; ++j) {
                    size_t c = I3(i, j, k);
                    double dHz_dy = (Hz[c] - Hz[I3(i, j - 1, k)]) / dy;
                    double dHy_dz = (Hy[c] - Hy[I3(i, j, k - 1)]) / dz;
                    double dHz_dx = (Hz[c] - Hz[I3(i - 1, j, k)]) / dx;
                    double dHx_dz = (Hx[c] - Hx[I3(i, j, k - 1)]) / dz;
                    double dHy_dx = (Hy[c] - Hy[I3(i - 1, j, k)]) / dx;
                    double dHx_dy = (Hx[c] - Hx[I3(i, j - 1, k)]) / dy;
                    m_dHz_dy[c] = bcoef_half[c] * m_dHz_dy[c] + acoef_half[c] * dHz_dy;
                    m_dHy_dz[c] = bcoef[c] * m_dHy_dz[c] + acoef[c] * dHy_dz;
                    m_dHz_dx[c] = bcoef_half[c] * m_dHz_dx[c] + acoef_half[c] * dHz_dx;
                    m_dHx_dz[c] = bcoef_half[c] * m_dHx_dz[c] + acoef_half[c] * dHx_dz;
                    m_dHx_dy[c] = bcoef_half[c] * m_dHx_dy[c] + acoef_half[c] * dHx_dy;
                    m_dHy_dx[c] = bcoef_half[c] * m_dHy_dx[c] + acoef_half[c] * dHy_dx;
                    dHz_dy = dHz_dy / kappa_half[c] + m_dHz_dy[c];
                    dHy_dz = dHy_dz / kappa[c] + m_dHy_dz[c];
                    dHz_dx = dHz_dx / kappa_half[c] + m_dHz_dx[c];
                    dHx_dz = dHx_dz / kappa_half[c] + m_dHx_dz[c];
                    dHx_dy = dHx_dy / kappa_half[c] + m_dHx_dy[c];
                    dHy_dx = dHy_dx / kappa_half[c] + m_dHy_dx[c];
                    double rhs_x = dHz_dy - dHy_dz - (sig11[c] * Ex_old[c] + sig12[c] * Ey_old[c] + sig13[c] * Ez_old[c]);
                    double rhs_y = dHx_dz - dHz_dx - (sig12[c] * Ex_old[c] + sig22[c] * Ey_old[c] + sig23[c] * Ez_old[c]);
                    double rhs_z = dHy_dx - dHx_dy - (sig13[c] * Ex_old[c] + sig23[c] * Ey_old[c] + sig33[c] * Ez_old[c]);
                    Ex[c] = Ex_old[c] + (aEx[c] * rhs_x + bEx[c] * rhs_y + cEx[c] * rhs_z) * dt;
                    Ey[c] = Ey_old[c] + (aEy[c] * rhs_x + bEy[c] * rhs_y + cEy[c] * rhs_z) * dt;
                    Ez[c] = Ez_old[c] + (aEz[c] * rhs_x + bEz[c] * rhs_y + cEz[c] * rhs_z) * dt;
                }
        if (r->pw_kind) {
            /* plane wave, cpmlfdtd.f08:3833-3903: faces x, y, z, cells
             * cpml+3 .. n-cpml-2 along the face */
            const double t = it * dt;
            const int nn[3] = {nx, ny, nz};
            const double hh[3] = {dx, dy, dz};
            for (int f = 0; f < 3; ++f) {
                if (!pw.face[f]) continue;
                const int a = f == 0 ? 1 : 0, b = f == 2 ? 1 : 2;
                for (int ia = r->cpml + 3; ia <= nn[a] - r->cpml - 2; ++ia)
                    for (int ib = r->cpml + 3; ib <= nn[b] - r->cpml - 2; ++ib) {
                        int ijk[3];
                        double x[3];
                        ijk[f] = pw.face[f];
                        x[f] = pw.loc[f];
                        ijk[a] = ia;
                        x[a] = ia * hh[a];
                        ijk[b] = ib;
                        x[b] = ib * hh[b];
                        double E[3], H[3];
                        pw_em_field(r, &pw, x, t, E, H);
                        const size_t c = I3(ijk[0], ijk[1], ijk[2]);
                        Ex[c] = Ex[c] + E[0]; Ey[c] = Ey[c] + E[1]; Ez[c] = Ez[c] + E[2];
                        Hx[c] = Hx[c] + H[0]; Hy[c] = Hy[c] + H[1]; Hz[c] = Hz[c] + H[2];
                    }
            }
        } else {
        /* point source, cpmlfdtd.f08:3909-3914 (eps already scaled by eps0) */
        Ex[I3(is, js, ks)] = Ex[I3(is, js, ks)] + src_at(r, SDT_SRC_X, it) * dt / eps11[I3(is, js, ks)];
        Ey[I3(is, js, ks)] = Ey[I3(is, js, ks)] + src_at(r, SDT_SRC_Y, it) * dt / eps22[I3(is, js, ks)];
        Ez[I3(is, js, ks)] = Ez[I3(is, js, ks)] + src_at(r, SDT_SRC_Z, it) * dt / eps33[I3(is, js, ks)];
        }
        /* Dirichlet faces for E and H, cpmlfdtd.f08:3921-3961 */
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i)
                    if (i == 1 || i == nx || j == 1 || j == ny || k == 1 || k == nz) {
                        size_t c = I3(i, j, k);
                        Ex[c] = 0.0; Ey[c] = 0.0; Ez[c] = 0.0;
                        Hx[c] = 0.0; Hy[c] = 0.0; Hz[c] = 0.0;
                    }
        /* E_old = E, cpmlfdtd.f08:3963-3965 */
        memcpy(Ex_old, Ex, n * sizeof(double));
        memcpy(Ey_old, Ey, n * sizeof(double));
        memcpy(Ez_old, Ez, n * sizeof(double));
        /* norm, cpmlfdtd.f08:3971-3972 */
        double velocnorm = 0.0;
#pragma omp parallel for collapse(2) reduction(max : velocnorm) schedule(static)
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    size_t c = I3(i, j, k);
                    double v = sqrt(Ex[c] * Ex[c] + Ey[c] * Ey[c] + Ez[c] * Ez[c]);
                    if (v > velocnorm) velocnorm = v;
                }
        if (r->velnorm) r->velnorm[it - 1] = velocnorm;
        record(r, g, it, Ex, Ey, Ez);
        if (velocnorm > O_STABILITY) { r->unstable_step = it; break; }
    }
    emit(r, SDT_F_EX, Ex, n); emit(r, SDT_F_EY, Ey, n); emit(r, SDT_F_EZ, Ez, n);
    emit(r, SDT_F_HX, Hx, n); emit(r, SDT_F_HY, Hy, n); emit(r, SDT_F_HZ, Hz, n);

    double *all[] = {eps11, eps12, eps13, eps22, eps23, eps33, sig11, sig12, sig13, sig22, sig23,
                     sig33, kappa, acoef, bcoef, kappa_half, acoef_half, bcoef_half, aEx, bEx, cEx,
                     bEy, cEy, cEz, Ex, Ey, Ez, Hx, Hy, Hz, Ex_old, Ey_old, Ez_old, m_dEx_dy,
                     m_dEy_dx, m_dEx_dz, m_dEz_dx, m_dEz_dy, m_dEy_dz, m_dHz_dx, m_dHx_dz,
                     m_dHz_dy, m_dHy_dz, m_dHx_dy, m_dHy_dx};
    for (size_t q = 0; q < sizeof(all) / sizeof(all[0]); ++q) free(all[q]);
    return r->unstable_step ? SDT_ERR_UNSTABLE : SDT_OK;
}

/* ========================================================================
 * array_averaging2d: averaging.f08:27-80.  The sweeps run in place in the
 * reference's loop order, so '+i' and '+j' average with the neighbour that was
 * already overwritten (a recursive smear) while '-i' and '-j' average with the
 * neighbour that is still original.
 * ======================================================================== */
void oracle_array_averaging2d(double *a, int32_t ni, int32_t nj, int32_t code, int32_t direction)
{
#define A2(i, j) a[(size_t)((i) - 1) + (size_t)ni * (size_t)((j) - 1)]
    switch (direction) {
    case 0:
        for (int j = 1; j <= nj; ++j)
            for (int i = 2; i <= ni; ++i) A2(i, j) = oracle_scalar_mean(A2(i, j), A2(i - 1, j), code);
        break;
    case 1:
        for (int j = 1; j <= nj; ++j)
            for (int i = 1; i <= ni - 1; ++i) A2(i, j) = oracle_scalar_mean(A2(i, j), A2(i + 1, j), code);
        break;
    case 2:
        for (int j = 2; j <= nj; ++j)
            for (int i = 1; i <= ni; ++i) A2(i, j) = oracle_scalar_mean(A2(i, j), A2(i, j - 1), code);
        break;
    case 3:
        for (int j = 1; j <= nj - 1; ++j)
            for (int i = 1; i <= ni; ++i) A2(i, j) = oracle_scalar_mean(A2(i, j), A2(i, j + 1), code);
        break;
    }
#undef A2
}

/* ========================================================================
 * seismic25: cpmlfdtd.f08:600-1337 (time loop 931-1305).  Complex fields as
 * (re, im) pairs; the y derivative is i*ky*f, i.e. (-ky*im, ky*re), which is
 * what (0,1)*ky*f evaluates to (cpmlfdtd.f08:946-948).  Real coefficients
 * multiply and divide both parts.  Memory variables are kept per nest as in
 * the reference (24 stress-side + 6 velocity-side complex arrays).
 * ======================================================================== */
typedef struct { double re, im; } cplx_t;

static inline cplx_t cx(double re, double im) { cplx_t z = {re, im}; return z; }
static inline cplx_t cadd(cplx_t a, cplx_t b) { return cx(a.re + b.re, a.im + b.im); }
static inline cplx_t crmul(double s, cplx_t a) { return cx(s * a.re, s * a.im); }
static inline cplx_t cdivr(cplx_t a, double s) { return cx(a.re / s, a.im / s); }
static inline cplx_t cmulr(cplx_t a, double s) { return cx(a.re * s, a.im * s); }
static inline cplx_t ciky(double ky, cplx_t a) { return cx(-ky * a.im, ky * a.re); }
/* 4th-order staggered difference (27(f1-f0) - (f2-fm1)) on both parts */
static inline cplx_t cd4(cplx_t fm1, cplx_t f0, cplx_t f1, cplx_t f2, double h24)
{
    return cx((27.0 * (f1.re - f0.re) - (f2.re - fm1.re)) / h24, (27.0 * (f1.im - f0.im) - (f2.im - fm1.im)) / h24);
}
/* psi = b*psi + a*d ; d = d/kappa + psi */
static inline cplx_t ccpml(cplx_t d, cplx_t *m, double a, double b, double kap)
{
    m->re = b * m->re + a * d.re;
    m->im = b * m->im + a * d.im;
    return cx(d.re / kap + m->re, d.im / kap + m->im);
}
static inline cplx_t ctminus(cplx_t sm2, cplx_t sm1, cplx_t s0, cplx_t sp1, int corr, double h24)
{
    return cx(tminus(sm2.re, sm1.re, s0.re, sp1.re, corr) / h24, tminus(sm2.im, sm1.im, s0.im, sp1.im, corr) / h24);
}
static inline cplx_t ctplus(cplx_t sm1, cplx_t s0, cplx_t sp1, cplx_t sp2, int corr, double h24)
{
    return cx(tplus(sm1.re, s0.re, sp1.re, sp2.re, corr) / h24, tplus(sm1.im, s0.im, sp1.im, sp2.im, corr) / h24);
}
/* (s + (C(1)*d1 + C(2)*d2 + C(3)*d3 + C(4)*(yz) + C(5)*(xz) + C(6)*(xy))*dt)/(1+gamma*dt),
 * cpmlfdtd.f08:974-987 */
static inline cplx_t cstress(cplx_t s, double k1, double k2, double k3, double k4, double k5, double k6,
                             cplx_t dxx, cplx_t dyy, cplx_t dzz, cplx_t yz, cplx_t xz, cplx_t xy, double dt,
                             double gam)
{
    cplx_t t = crmul(k1, dxx);
    t = cadd(t, crmul(k2, dyy));
    t = cadd(t, crmul(k3, dzz));
    t = cadd(t, crmul(k4, yz));
    t = cadd(t, crmul(k5, xz));
    t = cadd(t, crmul(k6, xy));
    return cdivr(cadd(s, cmulr(t, dt)), 1 + gam * dt);
}

int32_t oracle_seismic25(oracle_run *r)
{
    set_threads(r);
    grid_t g = {r->nx, 1, r->nz, (size_t)r->nx * (size_t)r->nz};
    const int nx = g.nx, nz = g.nz;
    const double dx = r->dx, dz = r->dz, dt = r->dt, ky = r->ky;
    const double hx = 24.0 * dx, hz = 24.0 * dz;
    const int dc = r->density_code, corr = r->stencil == SDT_STENCIL_CORRECTED;
    const size_t n = g.n;

    double *C[21];
    for (int q = 0; q < 21; ++q) C[q] = input_copy(r, SDT_C11 + q, n);
    double *c11 = C[0], *c12 = C[1], *c13 = C[2], *c14 = C[3], *c15 = C[4], *c16 = C[5], *c22 = C[6],
           *c23 = C[7], *c24 = C[8], *c25 = C[9], *c26 = C[10], *c33 = C[11], *c34 = C[12], *c35 = C[13],
           *c36 = C[14], *c44 = C[15], *c45 = C[16], *c46 = C[17], *c55 = C[18], *c56 = C[19], *c66 = C[20];
    double *rho = input_copy(r, SDT_RHO, n);
    double *gamma_x = input_copy(r, SDT_GAMMA_X, n), *gamma_y = input_copy(r, SDT_GAMMA_Y, n),
           *gamma_z = input_copy(r, SDT_GAMMA_Z, n), *gamma_xy = input_copy(r, SDT_GAMMA_XY, n),
           *gamma_xz = input_copy(r, SDT_GAMMA_XZ, n), *gamma_yz = input_copy(r, SDT_GAMMA_YZ, n);
    double *kappa = input_copy(r, SDT_KAPPA, n), *acoef = input_copy(r, SDT_ACOEF, n),
           *bcoef = input_copy(r, SDT_BCOEF, n);
    double *kappa_half = input_copy(r, SDT_KAPPA_HALF, n), *acoef_half = input_copy(r, SDT_ACOEF_HALF, n),
           *bcoef_half = input_copy(r, SDT_BCOEF_HALF, n);

    /* density averaging, cpmlfdtd.f08:889-898 */
    double *rhoxx = input_copy(r, SDT_RHO, n), *rhozx = input_copy(r, SDT_RHO, n),
           *rhoxy = input_copy(r, SDT_RHO, n), *rhozy = input_copy(r, SDT_RHO, n),
           *rhoxz = input_copy(r, SDT_RHO, n), *rhozz = input_copy(r, SDT_RHO, n);
    const double *rhoyx = rho, *rhoyy = rho, *rhoyz = rho;
    oracle_array_averaging2d(rhoxx, nx, nz, dc, 0);
    oracle_array_averaging2d(rhozx, nx, nz, dc, 2);
    oracle_array_averaging2d(rhoxy, nx, nz, dc, 1);
    oracle_array_averaging2d(rhozy, nx, nz, dc, 2);
    oracle_array_averaging2d(rhoxz, nx, nz, dc, 1);
    oracle_array_averaging2d(rhozz, nx, nz, dc, 3);

    cplx_t *F[9];
    for (int f = 0; f < 9; ++f) F[f] = (cplx_t *)calloc(n ? n : 1, sizeof(cplx_t));
    cplx_t *vx = F[SDT_F_VX], *vy = F[SDT_F_VY], *vz = F[SDT_F_VZ], *sxx = F[SDT_F_SXX],
           *syy = F[SDT_F_SYY], *szz = F[SDT_F_SZZ], *sxy = F[SDT_F_SXY], *sxz = F[SDT_F_SXZ],
           *syz = F[SDT_F_SYZ];
    /* initialconditionV{x,y,z} (complex128, cpmlfdtd.f08:870-872) */
    for (int f = 0; f < 9; ++f) {
        const double *in = r->field_in[f] ? r->field_in[f] : (f < 3 ? r->arr[SDT_IC_X + f] : NULL);
        if (in) memcpy(F[f], in, n * sizeof(cplx_t));
    }
    /* memory arrays: [nest 0..3][x: vx,vy,vz | z: vx,vy,vz] and 6 velocity-side */
    cplx_t *M[4][6], *MV[6];
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 6; ++b) M[a][b] = (cplx_t *)calloc(n ? n : 1, sizeof(cplx_t));
    for (int b = 0; b < 6; ++b) MV[b] = (cplx_t *)calloc(n ? n : 1, sizeof(cplx_t));

    const int is = r->isrc, ks = r->ksrc;
    r->unstable_step = 0;
    cplx_t *V[3] = {vx, vy, vz};
    pw_state pw;
    if (r->pw_kind) pw_setup(r, 0, &pw);

    for (int it = 1; it <= r->nsteps; ++it) {
        /* L1 sigma_xx, sigma_yy, sigma_zz: x forward half, z backward full (cpmlfdtd.f08:941-984) */
#pragma omp parallel for schedule(static)
        for (int k = 3; k <= nz - 1; ++k)
            for (int i = 2; i <= nx - 2; ++i) {
                const size_t c = I2(i, k);
                cplx_t ddx[3], ddy[3], ddz[3];
                for (int m = 0; m < 3; ++m) {
                    const cplx_t *v = V[m];
                    ddx[m] = cd4(v[I2(i - 1, k)], v[c], v[I2(i + 1, k)], v[I2(i + 2, k)], hx);
                    ddy[m] = ciky(ky, v[c]);
                    ddz[m] = cd4(v[I2(i, k - 2)], v[I2(i, k - 1)], v[c], v[I2(i, k + 1)], hz);
                }
                for (int m = 0; m < 3; ++m) ddx[m] = ccpml(ddx[m], &M[0][m][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                for (int m = 0; m < 3; ++m) ddz[m] = ccpml(ddz[m], &M[0][3 + m][c], acoef[c], bcoef[c], kappa[c]);
                const cplx_t yz = cadd(ddz[1], ddy[2]), xz = cadd(ddz[0], ddx[2]), xy = cadd(ddx[1], ddy[0]);
                sxx[c] = cstress(sxx[c], c11[c], c12[c], c13[c], c14[c], c15[c], c16[c], ddx[0], ddy[1], ddz[2], yz, xz, xy, dt, gamma_x[c]);
                syy[c] = cstress(syy[c], c12[c], c22[c], c23[c], c24[c], c25[c], c26[c], ddx[0], ddy[1], ddz[2], yz, xz, xy, dt, gamma_y[c]);
                szz[c] = cstress(szz[c], c13[c], c23[c], c33[c], c34[c], c35[c], c36[c], ddx[0], ddy[1], ddz[2], yz, xz, xy, dt, gamma_z[c]);
            }
        /* L2 sigma_xy: x backward full, z backward half (cpmlfdtd.f08:998-1027) */
#pragma omp parallel for schedule(static)
        for (int k = 3; k <= nz - 1; ++k)
            for (int i = 3; i <= nx - 1; ++i) {
                const size_t c = I2(i, k);
                cplx_t ddx[3], ddy[3], ddz[3];
                for (int m = 0; m < 3; ++m) {
                    const cplx_t *v = V[m];
                    ddx[m] = cd4(v[I2(i - 2, k)], v[I2(i - 1, k)], v[c], v[I2(i + 1, k)], hx);
                    ddy[m] = ciky(ky, v[c]);
                    ddz[m] = cd4(v[I2(i, k - 2)], v[I2(i, k - 1)], v[c], v[I2(i, k + 1)], hz);
                }
                for (int m = 0; m < 3; ++m) ddx[m] = ccpml(ddx[m], &M[1][m][c], acoef[c], bcoef[c], kappa[c]);
                for (int m = 0; m < 3; ++m) ddz[m] = ccpml(ddz[m], &M[1][3 + m][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                const cplx_t yz = cadd(ddz[1], ddy[2]), xz = cadd(ddz[0], ddx[2]), xy = cadd(ddx[1], ddy[0]);
                sxy[c] = cstress(sxy[c], c16[c], c26[c], c36[c], c46[c], c56[c], c66[c], ddx[0], ddy[1], ddz[2], yz, xz, xy, dt, gamma_xy[c]);
            }
        /* L3 sigma_xz: x backward full, z forward half (cpmlfdtd.f08:1041-1070) */
#pragma omp parallel for schedule(static)
        for (int k = 2; k <= nz - 2; ++k)
            for (int i = 3; i <= nx - 1; ++i) {
                const size_t c = I2(i, k);
                cplx_t ddx[3], ddy[3], ddz[3];
                for (int m = 0; m < 3; ++m) {
                    const cplx_t *v = V[m];
                    ddx[m] = cd4(v[I2(i - 2, k)], v[I2(i - 1, k)], v[c], v[I2(i + 1, k)], hx);
                    ddy[m] = ciky(ky, v[c]);
                    ddz[m] = cd4(v[I2(i, k - 1)], v[c], v[I2(i, k + 1)], v[I2(i, k + 2)], hz);
                }
                for (int m = 0; m < 3; ++m) ddx[m] = ccpml(ddx[m], &M[2][m][c], acoef[c], bcoef[c], kappa[c]);
                for (int m = 0; m < 3; ++m) ddz[m] = ccpml(ddz[m], &M[2][3 + m][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                const cplx_t yz = cadd(ddz[1], ddy[2]), xz = cadd(ddz[0], ddx[2]), xy = cadd(ddx[1], ddy[0]);
                sxz[c] = cstress(sxz[c], c15[c], c25[c], c35[c], c45[c], c55[c], c56[c], ddx[0], ddy[1], ddz[2], yz, xz, xy, dt, gamma_xz[c]);
            }
        /* L4 sigma_yz: x forward half, z forward half (cpmlfdtd.f08:1084-1113) */
#pragma omp parallel for schedule(static)
        for (int k = 2; k <= nz - 2; ++k)
            for (int i = 2; i <= nx - 2; ++i) {
                const size_t c = I2(i, k);
                cplx_t ddx[3], ddy[3], ddz[3];
                for (int m = 0; m < 3; ++m) {
                    const cplx_t *v = V[m];
                    ddx[m] = cd4(v[I2(i - 1, k)], v[c], v[I2(i + 1, k)], v[I2(i + 2, k)], hx);
                    ddy[m] = ciky(ky, v[c]);
                    ddz[m] = cd4(v[I2(i, k - 1)], v[c], v[I2(i, k + 1)], v[I2(i, k + 2)], hz);
                }
                for (int m = 0; m < 3; ++m) ddx[m] = ccpml(ddx[m], &M[3][m][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                for (int m = 0; m < 3; ++m) ddz[m] = ccpml(ddz[m], &M[3][3 + m][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                const cplx_t yz = cadd(ddz[1], ddy[2]), xz = cadd(ddz[0], ddx[2]), xy = cadd(ddx[1], ddy[0]);
                syz[c] = cstress(syz[c], c14[c], c24[c], c34[c], c44[c], c45[c], c46[c], ddx[0], ddy[1], ddz[2], yz, xz, xy, dt, gamma_yz[c]);
            }
        /* vx: T- x sigma_xx (half), i ky sigma_xy, T- z sigma_xz (half) (cpmlfdtd.f08:1130-1143) */
#pragma omp parallel for schedule(static)
        for (int k = 3; k <= nz - 1; ++k)
            for (int i = 3; i <= nx - 1; ++i) {
                const size_t c = I2(i, k);
                cplx_t d1 = ctminus(sxx[I2(i - 2, k)], sxx[I2(i - 1, k)], sxx[c], sxx[I2(i + 1, k)], corr, hx);
                const cplx_t d2 = ciky(ky, sxy[c]);
                cplx_t d3 = ctminus(sxz[I2(i, k - 2)], sxz[I2(i, k - 1)], sxz[c], sxz[I2(i, k + 1)], corr, hz);
                d1 = ccpml(d1, &MV[0][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                d3 = ccpml(d3, &MV[1][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                vx[c] = cadd(vx[c], cmulr(cadd(cadd(cdivr(d1, rhoxx[c]), cdivr(d2, rhoyx[c])), cdivr(d3, rhozx[c])), dt));
            }
        /* vy: T+ x sigma_xy (half), i ky sigma_yy, T- z sigma_yz (full) (cpmlfdtd.f08:1159-1172) */
#pragma omp parallel for schedule(static)
        for (int k = 3; k <= nz - 1; ++k)
            for (int i = 2; i <= nx - 2; ++i) {
                const size_t c = I2(i, k);
                cplx_t d1 = ctplus(sxy[I2(i - 1, k)], sxy[c], sxy[I2(i + 1, k)], sxy[I2(i + 2, k)], corr, hx);
                const cplx_t d2 = ciky(ky, syy[c]);
                cplx_t d3 = ctminus(syz[I2(i, k - 2)], syz[I2(i, k - 1)], syz[c], syz[I2(i, k + 1)], corr, hz);
                d1 = ccpml(d1, &MV[2][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                d3 = ccpml(d3, &MV[3][c], acoef[c], bcoef[c], kappa[c]);
                vy[c] = cadd(vy[c], cmulr(cadd(cadd(cdivr(d1, rhoxy[c]), cdivr(d2, rhoyy[c])), cdivr(d3, rhozy[c])), dt));
            }
        /* vz: T+ x sigma_xz (half), i ky sigma_yz, T+ z sigma_zz (half) (cpmlfdtd.f08:1187-1200) */
#pragma omp parallel for schedule(static)
        for (int k = 2; k <= nz - 2; ++k)
            for (int i = 2; i <= nx - 2; ++i) {
                const size_t c = I2(i, k);
                cplx_t d1 = ctplus(sxz[I2(i - 1, k)], sxz[c], sxz[I2(i + 1, k)], sxz[I2(i + 2, k)], corr, hx);
                const cplx_t d2 = ciky(ky, syz[c]);
                cplx_t d3 = ctplus(szz[I2(i, k - 1)], szz[c], szz[I2(i, k + 1)], szz[I2(i, k + 2)], corr, hz);
                d1 = ccpml(d1, &MV[4][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                d3 = ccpml(d3, &MV[5][c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                vz[c] = cadd(vz[c], cmulr(cadd(cadd(cdivr(d1, rhoxz[c]), cdivr(d2, rhoyz[c])), cdivr(d3, rhozz[c])), dt));
            }
        if (r->pw_kind) {
            /* plane wave, cpmlfdtd.f08:1208-1262: x face then z face, real parts */
            for (int f = 0; f < 3; f += 2) {
                if (!pw.face[f]) continue;
                const int lo = r->cpml + 2, hi = (f == 0 ? nz : nx) - r->cpml - 1;
                for (int ip = lo; ip <= hi; ++ip) {
                    const int i = f == 0 ? pw.face[0] : ip, k = f == 0 ? ip : pw.face[2];
                    const double x[3] = {f == 0 ? pw.loc[0] : ip * dx, 0.0, f == 0 ? ip * dz : pw.loc[2]};
                    double Vv[3], st[6];
                    pw_field(r, &pw, x, it * dt, Vv, st);
                    const size_t c = I2(i, k);
                    vx[c].re = vx[c].re + Vv[0];
                    vy[c].re = vy[c].re + Vv[1];
                    vz[c].re = vz[c].re + Vv[2];
                    double C21[21], s6[6] = {0, 0, 0, 0, 0, 0};
                    for (int q = 0; q < 21; ++q) C21[q] = C[q][c];
                    pw_stress(C21, st, s6);
                    sxx[c].re = sxx[c].re + s6[0];
                    syy[c].re = syy[c].re + s6[1];
                    szz[c].re = szz[c].re + s6[2];
                    syz[c].re = syz[c].re + s6[3];
                    sxz[c].re = sxz[c].re + s6[4];
                    sxy[c].re = sxy[c].re + s6[5];
                }
            }
        } else {
        /* point source, cpmlfdtd.f08:1268-1276 (real series add to the real part) */
        sxx[I2(is, ks)].re += src_at(r, SDT_SRC_XX, it) / rho[I2(is, ks)];
        sxy[I2(is + 1, ks)].re += src_at(r, SDT_SRC_XY, it) / rho[I2(is + 1, ks)];
        sxz[I2(is + 1, ks + 1)].re += src_at(r, SDT_SRC_XZ, it) / rho[I2(is + 1, ks)];
        syy[I2(is, ks)].re += src_at(r, SDT_SRC_YY, it) / rho[I2(is, ks)];
        syz[I2(is, ks + 1)].re += src_at(r, SDT_SRC_YZ, it) / rho[I2(is, ks + 1)];
        szz[I2(is, ks)].re += src_at(r, SDT_SRC_ZZ, it) / rho[I2(is, ks)];
        vx[I2(is + 1, ks)].re += src_at(r, SDT_SRC_X, it) * dt / rho[I2(is + 1, ks)];
        vy[I2(is, ks)].re += src_at(r, SDT_SRC_Y, it) * dt / rho[I2(is, ks)];
        vz[I2(is, ks + 1)].re += src_at(r, SDT_SRC_Z, it) * dt / rho[I2(is, ks + 1)];
        }
        /* Dirichlet edges, cpmlfdtd.f08:1280-1285 */
        for (int m = 0; m < 3; ++m) {
            for (int k = 1; k <= nz; ++k) { V[m][I2(1, k)] = cx(0, 0); V[m][I2(nx, k)] = cx(0, 0); }
            for (int i = 1; i <= nx; ++i) { V[m][I2(i, 1)] = cx(0, 0); V[m][I2(i, nz)] = cx(0, 0); }
        }
        /* velocity norm, cpmlfdtd.f08:1291-1292 */
        double velocnorm = 0.0;
#pragma omp parallel for reduction(max : velocnorm) schedule(static)
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) {
                const size_t c = I2(i, k);
                const double s = (vx[c].re * vx[c].re + vx[c].im * vx[c].im) + (vy[c].re * vy[c].re + vy[c].im * vy[c].im) +
                                 (vz[c].re * vz[c].re + vz[c].im * vz[c].im);
                const double v = sqrt(s);
                if (v > velocnorm) velocnorm = v;
            }
        if (r->velnorm) r->velnorm[it - 1] = velocnorm;
        if (r->rec_out)
            for (int q = 0; q < r->nrec; ++q) {
                const size_t c = I2(r->rec[3 * q], r->rec[3 * q + 2]);
                double *o = r->rec_out + ((size_t)(it - 1) * r->nrec + q) * 6;
                for (int m = 0; m < 3; ++m) { o[2 * m] = V[m][c].re; o[2 * m + 1] = V[m][c].im; }
            }
        if (velocnorm > O_STABILITY) { r->unstable_step = it; break; }
    }

    for (int f = 0; f < 9; ++f)
        if (r->field_out[f]) memcpy(r->field_out[f], F[f], n * sizeof(cplx_t));

    for (int q = 0; q < 21; ++q) free(C[q]);
    double *all[] = {rho, gamma_x, gamma_y, gamma_z, gamma_xy, gamma_xz, gamma_yz, kappa, acoef, bcoef,
                     kappa_half, acoef_half, bcoef_half, rhoxx, rhozx, rhoxy, rhozy, rhoxz, rhozz};
    for (size_t q = 0; q < sizeof(all) / sizeof(all[0]); ++q) free(all[q]);
    for (int f = 0; f < 9; ++f) free(F[f]);
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 6; ++b) free(M[a][b]);
    for (int b = 0; b < 6; ++b) free(MV[b]);
    return r->unstable_step ? SDT_ERR_UNSTABLE : SDT_OK;
}

/* ========================================================================
 * electromag25: cpmlfdtd.f08:2764-3350 (time loop 3095-3330).  Complex E and
 * H on the (x, z) plane, y derivatives i*ky*f; 4th-order staggered
 * differences; H-side memory variables with the half coefficient set, E-side
 * with the full set (3108-3228); sigma applied to E_old (3209-3225).
 * ======================================================================== */
int32_t oracle_electromag25(oracle_run *r)
{
    set_threads(r);
    grid_t g = {r->nx, 1, r->nz, (size_t)r->nx * (size_t)r->nz};
    const int nx = g.nx, nz = g.nz;
    const double dx = r->dx, dz = r->dz, dt = r->dt, ky = r->ky;
    const double hx = 24.0 * dx, hz = 24.0 * dz;
    const size_t n = g.n;
    const double dbH = dt / O_MU0;

    double *eps[6], *sig[6];
    for (int q = 0; q < 6; ++q) {
        eps[q] = input_copy(r, SDT_E11 + q, n);
        sig[q] = input_copy(r, SDT_S11 + q, n);
    }
    double *eps11 = eps[0], *eps12 = eps[1], *eps13 = eps[2], *eps22 = eps[3], *eps23 = eps[4], *eps33 = eps[5];
    double *sig11 = sig[0], *sig12 = sig[1], *sig13 = sig[2], *sig22 = sig[3], *sig23 = sig[4], *sig33 = sig[5];
    double *kappa = input_copy(r, SDT_KAPPA, n), *acoef = input_copy(r, SDT_ACOEF, n),
           *bcoef = input_copy(r, SDT_BCOEF, n);
    double *kappa_half = input_copy(r, SDT_KAPPA_HALF, n), *acoef_half = input_copy(r, SDT_ACOEF_HALF, n),
           *bcoef_half = input_copy(r, SDT_BCOEF_HALF, n);
    /* eps scaled to absolute values, adjugate/determinant (cpmlfdtd.f08:3040-3066) */
    double *aEx = alloc_zero(n), *bEx = alloc_zero(n), *cEx = alloc_zero(n), *bEy = alloc_zero(n),
           *cEy = alloc_zero(n), *cEz = alloc_zero(n);
    for (size_t c = 0; c < n; ++c) {
        for (int q = 0; q < 6; ++q) eps[q][c] *= O_EPS0;
        const double det = eps11[c] * (eps22[c] * eps33[c] - eps23[c] * eps23[c]) -
                           eps12[c] * (eps12[c] * eps33[c] - eps23[c] * eps13[c]) +
                           eps13[c] * (eps12[c] * eps23[c] - eps22[c] * eps13[c]);
        aEx[c] = (eps22[c] * eps33[c] - eps23[c] * eps23[c]) / det;
        bEx[c] = -(eps12[c] * eps33[c] - eps23[c] * eps13[c]) / det;
        cEx[c] = (eps12[c] * eps23[c] - eps22[c] * eps13[c]) / det;
        bEy[c] = (eps11[c] * eps33[c] - eps13[c] * eps13[c]) / det;
        cEy[c] = -(eps11[c] * eps23[c] - eps12[c] * eps13[c]) / det;
        cEz[c] = (eps11[c] * eps22[c] - eps12[c] * eps12[c]) / det;
    }
    const double *aEy = bEx, *aEz = cEx, *bEz = cEy;

    cplx_t *F[6];
    for (int f = 0; f < 6; ++f) F[f] = (cplx_t *)calloc(n ? n : 1, sizeof(cplx_t));
    cplx_t *Ex = F[SDT_F_EX], *Ey = F[SDT_F_EY], *Ez = F[SDT_F_EZ], *Hx = F[SDT_F_HX], *Hy = F[SDT_F_HY],
           *Hz = F[SDT_F_HZ];
    for (int f = 0; f < 6; ++f) {
        const double *in = r->field_in[f] ? r->field_in[f] : (f < 3 ? r->arr[SDT_IC_X + f] : NULL);
        if (in) memcpy(F[f], in, n * sizeof(cplx_t));
    }
    cplx_t *Exo = (cplx_t *)calloc(n ? n : 1, sizeof(cplx_t)), *Eyo = (cplx_t *)calloc(n ? n : 1, sizeof(cplx_t)),
           *Ezo = (cplx_t *)calloc(n ? n : 1, sizeof(cplx_t));
    memcpy(Exo, Ex, n * sizeof(cplx_t));
    memcpy(Eyo, Ey, n * sizeof(cplx_t));
    memcpy(Ezo, Ez, n * sizeof(cplx_t));
    cplx_t *m_dEy_dx = (cplx_t *)calloc(n ? n : 1, sizeof(cplx_t)), *m_dEy_dz = (cplx_t *)calloc(n ? n : 1, sizeof(cplx_t)),
           *m_dEx_dz = (cplx_t *)calloc(n ? n : 1, sizeof(cplx_t)), *m_dEz_dx = (cplx_t *)calloc(n ? n : 1, sizeof(cplx_t)),
           *m_dHz_dx = (cplx_t *)calloc(n ? n : 1, sizeof(cplx_t)), *m_dHx_dz = (cplx_t *)calloc(n ? n : 1, sizeof(cplx_t)),
           *m_dHy_dz = (cplx_t *)calloc(n ? n : 1, sizeof(cplx_t)), *m_dHy_dx = (cplx_t *)calloc(n ? n : 1, sizeof(cplx_t));

    const int is = r->isrc, ks = r->ksrc;
    r->unstable_step = 0;
    for (int it = 1; it <= r->nsteps; ++it) {
        /* Hx, cpmlfdtd.f08:3106-3116 */
#pragma omp parallel for schedule(static)
        for (int k = 2; k <= nz - 2; ++k)
            for (int i = 2; i <= nx - 1; ++i) {
                const size_t c = I2(i, k);
                cplx_t d = cd4(Ey[I2(i, k - 1)], Ey[c], Ey[I2(i, k + 1)], Ey[I2(i, k + 2)], hz);
                d = ccpml(d, &m_dEy_dz[c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                const cplx_t t = ciky(ky, Ez[c]);
                Hx[c] = cx(Hx[c].re - dbH * (t.re - d.re), Hx[c].im - dbH * (t.im - d.im));
            }
        /* Hy, cpmlfdtd.f08:3131-3146 */
#pragma omp parallel for schedule(static)
        for (int k = 2; k <= nz - 2; ++k)
            for (int i = 2; i <= nx - 2; ++i) {
                const size_t c = I2(i, k);
                cplx_t dxz = cd4(Ex[I2(i, k - 1)], Ex[c], Ex[I2(i, k + 1)], Ex[I2(i, k + 2)], hz);
                cplx_t dzx = cd4(Ez[I2(i - 1, k)], Ez[c], Ez[I2(i + 1, k)], Ez[I2(i + 2, k)], hx);
                dxz = ccpml(dxz, &m_dEx_dz[c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                dzx = ccpml(dzx, &m_dEz_dx[c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                Hy[c] = cx(Hy[c].re - dbH * (dxz.re - dzx.re), Hy[c].im - dbH * (dxz.im - dzx.im));
            }
        /* Hz, cpmlfdtd.f08:3160-3170 */
#pragma omp parallel for schedule(static)
        for (int k = 2; k <= nz - 1; ++k)
            for (int i = 2; i <= nx - 2; ++i) {
                const size_t c = I2(i, k);
                cplx_t d = cd4(Ey[I2(i - 1, k)], Ey[c], Ey[I2(i + 1, k)], Ey[I2(i + 2, k)], hx);
                d = ccpml(d, &m_dEy_dx[c], acoef_half[c], bcoef_half[c], kappa_half[c]);
                const cplx_t t = ciky(ky, Ex[c]);
                Hz[c] = cx(Hz[c].re - dbH * (d.re - t.re), Hz[c].im - dbH * (d.im - t.im));
            }
        /* E, cpmlfdtd.f08:3190-3228 */
#pragma omp parallel for schedule(static)
        for (int k = 3; k <= nz - 2; ++k)
            for (int i = 3; i <= nx - 2; ++i) {
                const size_t c = I2(i, k);
                const cplx_t dHz_dy = ciky(ky, Hz[c]);
                cplx_t dHy_dz = cd4(Hy[I2(i, k - 2)], Hy[I2(i, k - 1)], Hy[c], Hy[I2(i, k + 1)], hz);
                cplx_t dHx_dz = cd4(Hx[I2(i, k - 2)], Hx[I2(i, k - 1)], Hx[c], Hx[I2(i, k + 1)], hz);
                cplx_t dHz_dx = cd4(Hz[I2(i - 2, k)], Hz[I2(i - 1, k)], Hz[c], Hz[I2(i + 1, k)], hx);
                cplx_t dHy_dx = cd4(Hy[I2(i - 2, k)], Hy[I2(i - 1, k)], Hy[c], Hy[I2(i + 1, k)], hx);
                const cplx_t dHx_dy = ciky(ky, Hx[c]);
                dHy_dz = ccpml(dHy_dz, &m_dHy_dz[c], acoef[c], bcoef[c], kappa[c]);
                dHx_dz = ccpml(dHx_dz, &m_dHx_dz[c], acoef[c], bcoef[c], kappa[c]);
                dHz_dx = ccpml(dHz_dx, &m_dHz_dx[c], acoef[c], bcoef[c], kappa[c]);
                dHy_dx = ccpml(dHy_dx, &m_dHy_dx[c], acoef[c], bcoef[c], kappa[c]);
                const cplx_t eo[3] = {Exo[c], Eyo[c], Ezo[c]};
                cplx_t rx, ry, rz;
                rx.re = dHz_dy.re - dHy_dz.re - (sig11[c] * eo[0].re + sig12[c] * eo[1].re + sig13[c] * eo[2].re);
                rx.im = dHz_dy.im - dHy_dz.im - (sig11[c] * eo[0].im + sig12[c] * eo[1].im + sig13[c] * eo[2].im);
                ry.re = dHx_dz.re - dHz_dx.re - (sig12[c] * eo[0].re + sig22[c] * eo[1].re + sig23[c] * eo[2].re);
                ry.im = dHx_dz.im - dHz_dx.im - (sig12[c] * eo[0].im + sig22[c] * eo[1].im + sig23[c] * eo[2].im);
                rz.re = dHy_dx.re - dHx_dy.re - (sig13[c] * eo[0].re + sig23[c] * eo[1].re + sig33[c] * eo[2].re);
                rz.im = dHy_dx.im - dHx_dy.im - (sig13[c] * eo[0].im + sig23[c] * eo[1].im + sig33[c] * eo[2].im);
                Ex[c] = cx(eo[0].re + (aEx[c] * rx.re + bEx[c] * ry.re + cEx[c] * rz.re) * dt,
                           eo[0].im + (aEx[c] * rx.im + bEx[c] * ry.im + cEx[c] * rz.im) * dt);
                Ey[c] = cx(eo[1].re + (aEy[c] * rx.re + bEy[c] * ry.re + cEy[c] * rz.re) * dt,
                           eo[1].im + (aEy[c] * rx.im + bEy[c] * ry.im + cEy[c] * rz.im) * dt);
                Ez[c] = cx(eo[2].re + (aEz[c] * rx.re + bEz[c] * ry.re + cEz[c] * rz.re) * dt,
                           eo[2].im + (aEz[c] * rx.im + bEz[c] * ry.im + cEz[c] * rz.im) * dt);
            }
        /* point source, cpmlfdtd.f08:3236-3241 (real part) */
        Ex[I2(is, ks)].re += src_at(r, SDT_SRC_X, it) * dt / eps11[I2(is, ks)];
        Ey[I2(is, ks)].re += src_at(r, SDT_SRC_Y, it) * dt / eps22[I2(is, ks)];
        Ez[I2(is, ks)].re += src_at(r, SDT_SRC_Z, it) * dt / eps33[I2(is, ks)];
        /* Dirichlet edges on E and H, cpmlfdtd.f08:3246-3274 */
        for (int f = 0; f < 6; ++f) {
            for (int k = 1; k <= nz; ++k) { F[f][I2(1, k)] = cx(0, 0); F[f][I2(nx, k)] = cx(0, 0); }
            for (int i = 1; i <= nx; ++i) { F[f][I2(i, 1)] = cx(0, 0); F[f][I2(i, nz)] = cx(0, 0); }
        }
        memcpy(Exo, Ex, n * sizeof(cplx_t));
        memcpy(Eyo, Ey, n * sizeof(cplx_t));
        memcpy(Ezo, Ez, n * sizeof(cplx_t));
        /* norm, cpmlfdtd.f08:3309-3310 */
        double velocnorm = 0.0;
#pragma omp parallel for reduction(max : velocnorm) schedule(static)
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) {
                const size_t c = I2(i, k);
                const double s = (Ex[c].re * Ex[c].re + Ex[c].im * Ex[c].im) + (Ey[c].re * Ey[c].re + Ey[c].im * Ey[c].im) +
                                 (Ez[c].re * Ez[c].re + Ez[c].im * Ez[c].im);
                const double v = sqrt(s);
                if (v > velocnorm) velocnorm = v;
            }
        if (r->velnorm) r->velnorm[it - 1] = velocnorm;
        if (r->rec_out)
            for (int q = 0; q < r->nrec; ++q) {
                const size_t c = I2(r->rec[3 * q], r->rec[3 * q + 2]);
                double *o = r->rec_out + ((size_t)(it - 1) * r->nrec + q) * 6;
                for (int m = 0; m < 3; ++m) { o[2 * m] = F[m][c].re; o[2 * m + 1] = F[m][c].im; }
            }
        if (velocnorm > O_STABILITY) { r->unstable_step = it; break; }
    }
    for (int f = 0; f < 6; ++f)
        if (r->field_out[f]) memcpy(r->field_out[f], F[f], n * sizeof(cplx_t));

    for (int q = 0; q < 6; ++q) { free(eps[q]); free(sig[q]); }
    double *all[] = {kappa, acoef, bcoef, kappa_half, acoef_half, bcoef_half, aEx, bEx, cEx, bEy, cEy, cEz};
    for (size_t q = 0; q < sizeof(all) / sizeof(all[0]); ++q) free(all[q]);
    for (int f = 0; f < 6; ++f) free(F[f]);
    cplx_t *cm[] = {Exo, Eyo, Ezo, m_dEy_dx, m_dEy_dz, m_dEx_dz, m_dEz_dx, m_dHz_dx, m_dHx_dz, m_dHy_dz, m_dHy_dx};
    for (size_t q = 0; q < sizeof(cm) / sizeof(cm[0]); ++q) free(cm[q]);
    return r->unstable_step ? SDT_ERR_UNSTABLE : SDT_OK;
}
