// pw.hip -- plane-wave sources: an analytic plane wave (plane_wave_source.f08)
// injected every step on the interior faces its direction selects, in place
// of the point source.  Elastic (seismic2, seismic25, seismic3:
// cpmlfdtd.f08:188-232 / 454-512, 802-832 / 1208-1262, 1548-1605 /
// 2133-2220) and electromagnetic (electromag2, electromag3: 2436-2475 /
// 2626-2673, 3523-3580 / 3833-3903).
//
// Host: propagation and polarisation vectors from the three rotation angles,
// the active faces and reference point, and the background speed over the
// model: sqrt(max Kelvin-form stiffness eigenvalue / min density), or
// c / sqrt(min permittivity eigenvalue) with the matching impedance (a cyclic
// Jacobi eigensolver; cells repeating the previous cell's tensor reuse its
// eigenvalue).  Device: one thread per face cell evaluates the wave and adds
// the velocity and C : strain (or E and H) to the fields; a second pass,
// after every face is done, folds the injected cells' velocity / E (and 3D
// stress) magnitudes into the step's norms, which the fused kernels skip for
// these cells.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "sdt_device.h"
#include "sdt_internal.h"

namespace sdt {

// ---- host: geometry -----------------------------------------------------------
// rotation of plane_wave_source.f08:67-98, R = Ry (Rz Rx), matrices filled
// column by column as the reference's reshape does
static void rotmat(double axz, double axy, double ayz, double R[3][3])
{
    const double d = 3.1415926535897932384626433832795 / 180.0;
    const double cx = cos(d * ayz), cy = cos(d * axz), cz = cos(d * axy);
    const double sx = sin(d * ayz), sy = sin(d * axz), sz = sin(d * axy);
    const double Rx[3][3] = {{1, 0, 0}, {0, cx, sx}, {0, -sx, cx}};
    const double Ry[3][3] = {{cy, 0, -sy}, {0, 1, 0}, {sy, 0, cy}};
    const double Rz[3][3] = {{cz, sz, 0}, {-sz, cz, 0}, {0, 0, 1}};
    double T[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += Rz[i][k] * Rx[k][j];
            T[i][j] = s;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += Ry[i][k] * T[k][j];
            R[i][j] = s;
        }
}

static double norm3(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// direction_polarization_vector (plane_wave_source.f08:100-139): p = R z,
// e = R x made transverse to p (fallback x or y when R x is parallel to p)
static void direction_vectors(double axz, double axy, double ayz, double p[3], double e[3])
{
    double R[3][3];
    rotmat(axz, axy, ayz, R);
    for (int i = 0; i < 3; ++i) p[i] = R[i][2];
    double n = norm3(p);
    if (n > 0.0)
        for (int i = 0; i < 3; ++i) p[i] /= n;
    double t[3] = {R[0][0], R[1][0], R[2][0]};
    double dp = t[0] * p[0] + t[1] * p[1] + t[2] * p[2];
    for (int i = 0; i < 3; ++i) t[i] -= dp * p[i];
    n = norm3(t);
    if (n > 0.0) {
        for (int i = 0; i < 3; ++i) e[i] = t[i] / n;
        return;
    }
    if (fabs(p[0]) < 0.9) { t[0] = 1; t[1] = 0; t[2] = 0; }
    else { t[0] = 0; t[1] = 1; t[2] = 0; }
    dp = t[0] * p[0] + t[1] * p[1] + t[2] * p[2];
    for (int i = 0; i < 3; ++i) t[i] -= dp * p[i];
    n = norm3(t);
    for (int i = 0; i < 3; ++i) e[i] = n > 0.0 ? t[i] / n : 0.0;
}

// eigenvalues of a symmetric n x n matrix (n <= 6) on its diagonal after
// cyclic Jacobi rotations
static void jacobi_diagonalise(double A[6][6], int n)
{
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < n; ++i) {
            diag += A[i][i] * A[i][i];
            for (int j = i + 1; j < n; ++j) off += A[i][j] * A[i][j];
        }
        if (off <= 1e-34 * diag || off == 0.0) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[p][q];
                if (apq == 0.0) continue;
                const double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {   // columns p, q
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {   // rows p, q
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
            }
    }
}

static double jacobi_max_eig(double A[6][6], int n)
{
    jacobi_diagonalise(A, n);
    double m = A[0][0];
    for (int i = 1; i < n; ++i) m = A[i][i] > m ? A[i][i] : m;
    return m;
}

static double jacobi_min_eig(double A[6][6], int n)
{
    jacobi_diagonalise(A, n);
    double m = A[0][0];
    for (int i = 1; i < n; ++i) m = A[i][i] < m ? A[i][i] : m;
    return m;
}

// Kelvin (Mandel) form of the Voigt stiffness: rows/columns 4..6 scaled by
// sqrt(2) (tensor_operations.f08:57-132); 2D keeps the (1, 3, 5) block, whose
// eigenvalues with three zeros are those of the reference's 6 x 6
// (tensor_operations.f08:157-181)
static double max_kelvin_eig(const double *c21, bool two_d)
{
    static const int ij[6][6] = {{0, 1, 2, 3, 4, 5},   {1, 6, 7, 8, 9, 10},   {2, 7, 11, 12, 13, 14},
                                 {3, 8, 12, 15, 16, 17}, {4, 9, 13, 16, 18, 19}, {5, 10, 14, 17, 19, 20}};
    const double s[6] = {1, 1, 1, sqrt(2.0), sqrt(2.0), sqrt(2.0)};
    double A[6][6] = {};
    if (two_d) {
        const int idx[3] = {0, 2, 4};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) A[a][b] = s[idx[a]] * s[idx[b]] * c21[ij[idx[a]][idx[b]]];
        const double m = jacobi_max_eig(A, 3);
        return m > 0.0 ? m : 0.0;
    }
    for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b) A[a][b] = s[a] * s[b] * c21[ij[a][b]];
    return jacobi_max_eig(A, 6);
}

// Permittivity bounds (relative).  2D: the reference takes eigenvalues 2..3
// (ascending) of [[e11, e13, 0], [e13, e33, 0], [0, 0, 0]]
// (tensor_operations.f08:135-154), so a cell's smallest is the middle one of
// {0, lambda-, lambda+}.  3D: the smallest eigenvalue of the full tensor
// (tensor_operations.f08:240-262).
static double min_eps_eig2(const double *e)   // e = e11, e13, e33
{
    const double m = 0.5 * (e[0] + e[2]), h = 0.5 * (e[0] - e[2]);
    const double r = sqrt(h * h + e[1] * e[1]);
    const double lo = m - r, hi = m + r;
    return lo >= 0.0 ? lo : (hi <= 0.0 ? hi : 0.0);
}

static double min_eps_eig3(const double *e)   // e = e11 e12 e13 e22 e23 e33
{
    double A[6][6] = {};
    A[0][0] = e[0];
    A[0][1] = A[1][0] = e[1];
    A[0][2] = A[2][0] = e[2];
    A[1][1] = e[3];
    A[1][2] = A[2][1] = e[4];
    A[2][2] = e[5];
    return jacobi_min_eig(A, 3);
}

// Reduction (max or min) of eval(tensor) over all cells, the tensor's
// components read from `ncol` input arrays (null = 0); threads over chunks of
// cells, each reusing the previous cell's value when the tensor repeats.
template <typename F>
static double cell_reduce(const double *const *col, int ncol, int64_t n, bool want_max, F eval)
{
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<double> best((size_t)nt, 0.0);
    std::vector<char> any((size_t)nt, 0);
    std::vector<std::thread> th;
    for (int w = 0; w < nt; ++w)
        th.emplace_back([&, w]() {
            const int64_t b = n * w / nt, e = n * (w + 1) / nt;
            double prev[21], cur[21], val = 0.0, m = 0.0;
            bool have = false;
            for (int64_t c = b; c < e; ++c) {
                for (int q = 0; q < ncol; ++q) cur[q] = col[q] ? col[q][c] : 0.0;
                if (!have || memcmp(cur, prev, sizeof(double) * (size_t)ncol) != 0) {
                    val = eval(cur);
                    memcpy(prev, cur, sizeof(double) * (size_t)ncol);
                    if (!have) m = val;
                    have = true;
                }
                m = want_max ? (val > m ? val : m) : (val < m ? val : m);
            }
            best[(size_t)w] = m;
            any[(size_t)w] = have;
        });
    for (auto &t : th) t.join();
    bool have = false;
    double m = 0.0;
    for (int w = 0; w < nt; ++w) {
        if (!any[(size_t)w]) continue;
        const double v = best[(size_t)w];
        m = !have ? v : (want_max ? std::max(m, v) : std::min(m, v));
        have = true;
    }
    return m;
}

int plane_wave_setup(Device &d, const PlaneWaveSpec &spec, const double *const *in)
{
    PlaneWave &w = d.pw;
    w = PlaneWave();
    const bool em = d.physics == SDT_EM2 || d.physics == SDT_EM3;
    const bool d3 = d.physics == SDT_ELASTIC3 || d.physics == SDT_EM3;
    double p[3], e[3];
    direction_vectors(spec.angle_xz, spec.angle_xy, spec.angle_yz, p, e);
    if (spec.kind == SDT_PW_P)
        for (int i = 0; i < 3; ++i) e[i] = p[i];
    const double eps = 1.0e-6;   // select_injection_faces, plane_wave_source.f08:46-65
    const int n[3] = {d.nx, d.ny, d.nz};
    const double h[3] = {d.dx, d.dy, d.dz};
    const int cp = d.cpml_width;
    // cell ranges along a face: cpml+2 .. n-cpml-1, except the 3D EM solver's
    // cpml+3 .. n-cpml-2 (cpmlfdtd.f08:3569-3574)
    const int inset = d.physics == SDT_EM3 ? 1 : 0;
    for (int a = 0; a < 3; ++a) {
        const double lo = h[a] * (1 + cp), hi = h[a] * (n[a] - cp), mid = 0.5 * (lo + hi);
        w.loc[a] = mid;
        w.face[a] = 0;
        const bool axis = a != 1 || d3;   // the 2D solvers have no y faces
        if (axis && p[a] > eps) w.loc[a] = lo;
        if (axis && p[a] < -eps) w.loc[a] = hi;
        if (axis && (p[a] > eps || p[a] < -eps)) w.face[a] = (int)lround(w.loc[a] / h[a]);
        w.r0[a] = (a == 1 && !d3) ? 0.0 : w.loc[a];
        w.lo[a] = cp + 2 + inset;
        w.hi[a] = n[a] - cp - 1 - inset;
        w.h[a] = h[a];
        w.p[a] = p[a];
        w.e[a] = e[a];
    }
    if (!d3) {
        w.lo[1] = w.hi[1] = 1;
        w.loc[1] = 0.0;   // r(2) = 0 in the 2D and 2.5D injection loops
    }
    if (em) {
        // boundary_em_field (plane_wave_source.f08:163-194): speed and
        // impedance of the smallest permittivity eigenvalue over the model
        // (cpmlfdtd.f08:2472-2475, 3577-3580); H along p x e normalised
        double emin;
        if (d3) {
            const double *col[6] = {in[SDT_E11], in[SDT_E12], in[SDT_E13], in[SDT_E22], in[SDT_E23], in[SDT_E33]};
            emin = cell_reduce(col, 6, d.ncell, false, min_eps_eig3);
        } else {
            const double *col[3] = {in[SDT_E11], in[SDT_E13], in[SDT_E33]};
            emin = cell_reduce(col, 3, d.ncell, false, min_eps_eig2);
        }
        if (!(emin > 0.0)) return SDT_ERR_ARG;
        w.v = kClight / sqrt(emin);
        w.eta = sqrt(kMu0 / (kEps0 * emin));
        w.em = 1;
        const double c[3] = {p[1] * e[2] - p[2] * e[1], p[2] * e[0] - p[0] * e[2], p[0] * e[1] - p[1] * e[0]};
        const double cn = norm3(c);
        for (int i = 0; i < 3; ++i) w.hh[i] = cn > 0.0 ? c[i] / cn : c[i];
    } else {
        double rho_min = 0.0;
        bool first = true;
        for (int64_t c = 0; c < d.ncell; ++c) {
            const double r = in[SDT_RHO] ? in[SDT_RHO][c] : 0.0;
            if (first || r < rho_min) rho_min = r;
            first = false;
        }
        const double *col[21];
        for (int q = 0; q < 21; ++q) col[q] = in[SDT_C11 + q];
        const bool two_d = d.physics == SDT_ELASTIC2;
        const double lmax = cell_reduce(col, 21, d.ncell, true, [two_d](const double *c21) {
            return max_kelvin_eig(c21, two_d);
        });
        if (!(rho_min > 0.0) || !(lmax > 0.0)) return SDT_ERR_ARG;
        w.v = sqrt(lmax / rho_min);
    }
    w.f0 = spec.f0;
    w.amp = spec.amplitude;
    w.dt = d.dt;
    w.wavelet = spec.wavelet;
    w.on = w.face[0] || w.face[1] || w.face[2];
    // the plane wave replaces the point source
    TailArgs &t = d.tail;
    t.nterms = 0;
    t.nvfix = 0;
    t.nsfix = 0;
    for (int q = 0; q < kMaxNormFix; ++q) t.vfix[q] = t.sfix[q] = -1;
    return SDT_OK;
}

// ---- device ----------------------------------------------------------------
template <typename T>
struct PwArgs {
    PlaneWave w;
    T *v[3];              // vx vy vz (vy null in 2D)
    T *s[6];              // sxx syy szz sxy sxz syz (syy, sxy, syz null in 2D)
    const T *C[21];       // per-cell stiffness (Voigt upper triangle), or null
    const uint8_t *mid;   // material ids (3D table), or null
    const T *tab;         // material rows (c11..c66 first)
    int tab_row;
    int nx, ny, nz, pitch;
    int64_t xoff;
    int cplx;             // 1: complex fields, the wave adds to the real parts
    int dim;              // 2: seismic2 (strain 1, 3, 5 with six constants), 3: full
    int three_d;          // 1: grid edges on all six sides (y too)
    const int *d_it;
    unsigned long long *vnorm, *snorm;
};

__device__ __forceinline__ double pw_wavelet(double tau, double f0, int type)
{
    const double t0 = 1.5 / f0;
    const double a = 3.1415926535897932384626433832795 * f0 * (tau - t0);
    const double g = exp(-a * a);
    if (type == 1) return -2.0 * a * g;
    if (type == 2) return (1.0 - 2.0 * a * a) * g;
    return g;
}

// face f's q-th cell: (i, j, k) 1-based
__device__ __forceinline__ bool pw_cell(const PlaneWave &w, int f, int q, int &i, int &j, int &k)
{
    const int a = f == 0 ? 1 : 0, b = f == 2 ? 1 : 2;   // the two axes spanning face f
    const int na = w.hi[a] - w.lo[a] + 1, nb = w.hi[b] - w.lo[b] + 1;
    if (q >= na * nb) return false;
    int ijk[3];
    ijk[f] = w.face[f];
    ijk[a] = w.lo[a] + q % na;
    ijk[b] = w.lo[b] + q / na;
    i = ijk[0];
    j = ijk[1];
    k = ijk[2];
    return true;
}

template <typename T>
__device__ __forceinline__ int64_t pw_index(const PwArgs<T> &A, int i, int j, int k)
{
    return A.xoff + (int64_t)(i - 1) + (int64_t)A.pitch * ((int64_t)(j - 1) + (int64_t)A.ny * (k - 1));
}

template <typename T>
__device__ __forceinline__ void pw_add(const PwArgs<T> &A, T *f, int64_t c, double x)
{
    T *p = f + (c << A.cplx);
    *p = (T)((double)*p + x);
}

template <typename T>
__global__ void __launch_bounds__(256) pw_inject_kernel(PwArgs<T> A, int f)
{
    const PlaneWave &w = A.w;
    int i, j, k;
    if (!pw_cell(w, f, blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
    const int64_t c = pw_index(A, i, j, k);
    // A face on the grid edge (no absorbing layer): the Dirichlet step that
    // follows the injection in the reference zeroes the velocities (E and H)
    // there again, which the solvers' kernels have already done.
    const bool edge = i == 1 || i == A.nx || k == 1 || k == A.nz || (A.three_d && (j == 1 || j == A.ny));
    // boundary_seismic_field / boundary_em_field (plane_wave_source.f08:163-230)
    const double t = (double)(*A.d_it) * w.dt;
    // position of the cell: the face coordinate on the face's axis, index *
    // spacing on the others (y = 0 in 2D and 2.5D)
    const double r[3] = {f == 0 ? w.loc[0] : (double)i * w.h[0],
                         (f == 1 || !A.three_d) ? w.loc[1] : (double)j * w.h[1],
                         f == 2 ? w.loc[2] : (double)k * w.h[2]};
    const double tau = t - (w.p[0] * (r[0] - w.r0[0]) + w.p[1] * (r[1] - w.r0[1]) + w.p[2] * (r[2] - w.r0[2])) / w.v;
    const double s = w.amp * pw_wavelet(tau, w.f0, w.wavelet);
    if (w.em) {   // E += s e, H += (s / eta) hh (v = E fields, s[0..2] = H fields)
        if (edge) return;
        const double sh = s / w.eta;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            if (A.v[m]) pw_add(A, A.v[m], c, s * w.e[m]);
            if (A.s[m]) pw_add(A, A.s[m], c, sh * w.hh[m]);
        }
        return;
    }
    const double sc = -s / w.v;
    const double *e = w.e, *p = w.p;
    const double st[6] = {sc * e[0] * p[0], sc * e[1] * p[1], sc * e[2] * p[2], sc * (e[1] * p[2] + e[2] * p[1]),
                          sc * (e[0] * p[2] + e[2] * p[0]), sc * (e[0] * p[1] + e[1] * p[0])};
    double C[21];
    if (A.mid) {
        const T *row = A.tab + (int64_t)A.mid[c] * A.tab_row;
#pragma unroll
        for (int q = 0; q < 21; ++q) C[q] = (double)row[q];
    } else {
#pragma unroll
        for (int q = 0; q < 21; ++q) C[q] = A.C[q] ? (double)A.C[q][c] : 0.0;
    }
    if (!edge) {
        pw_add(A, A.v[0], c, s * e[0]);
        if (A.v[1]) pw_add(A, A.v[1], c, s * e[1]);
        pw_add(A, A.v[2], c, s * e[2]);
    }
    if (A.dim == 2) {   // cpmlfdtd.f08:473-481
        pw_add(A, A.s[0], c, C[0] * st[0] + C[2] * st[2] + C[4] * st[4]);
        pw_add(A, A.s[2], c, C[2] * st[0] + C[11] * st[2] + C[13] * st[4]);
        pw_add(A, A.s[4], c, C[4] * st[0] + C[13] * st[2] + C[18] * st[4]);
        return;
    }
    // add_plane_wave_stress (cpmlfdtd.f08:16-50): the rows of the Voigt
    // matrix (upper-triangle slots of C) times the strain, in the fields'
    // order xx yy zz yz xz xy
    pw_add(A, A.s[0], c, C[0] * st[0] + C[1] * st[1] + C[2] * st[2] + C[3] * st[3] + C[4] * st[4] + C[5] * st[5]);
    pw_add(A, A.s[1], c, C[1] * st[0] + C[6] * st[1] + C[7] * st[2] + C[8] * st[3] + C[9] * st[4] + C[10] * st[5]);
    pw_add(A, A.s[2], c,
           C[2] * st[0] + C[7] * st[1] + C[11] * st[2] + C[12] * st[3] + C[13] * st[4] + C[14] * st[5]);
    pw_add(A, A.s[5], c,
           C[3] * st[0] + C[8] * st[1] + C[12] * st[2] + C[15] * st[3] + C[16] * st[4] + C[17] * st[5]);
    pw_add(A, A.s[4], c,
           C[4] * st[0] + C[9] * st[1] + C[13] * st[2] + C[16] * st[3] + C[18] * st[4] + C[19] * st[5]);
    pw_add(A, A.s[3], c,
           C[5] * st[0] + C[10] * st[1] + C[14] * st[2] + C[17] * st[3] + C[19] * st[4] + C[20] * st[5]);
}

// norms of the injected cells after every face is done (the fused kernels
// skipped them)
template <typename T>
__global__ void __launch_bounds__(256) pw_norm_kernel(PwArgs<T> A, int f)
{
    const PlaneWave &w = A.w;
    int i, j, k;
    double vn = 0.0, sn = 0.0;
    if (pw_cell(w, f, blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) {
        const int64_t c = pw_index(A, i, j, k);
        for (int m = 0; m < 3; ++m) {
            if (!A.v[m]) continue;
            for (int part = 0; part <= A.cplx; ++part) {
                const double x = (double)A.v[m][(c << A.cplx) + part];
                vn += x * x;
            }
        }
        vn = sqrt(vn);
        if (A.snorm) {
            double a[6];
#pragma unroll
            for (int m = 0; m < 6; ++m) a[m] = (double)A.s[m][c];
            sn = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + 2.0 * (a[3] * a[3] + a[4] * a[4] + a[5] * a[5]));
        }
    }
    block_max_to(&A.vnorm[*A.d_it - 1], vn);
    if (A.snorm) block_max_to(&A.snorm[*A.d_it - 1], sn);
}

template <typename T>
static void launch_plane_wave_t(Device &d)
{
    PwArgs<T> A;
    A.w = d.pw;
    const bool d3 = d.physics == SDT_ELASTIC3;
    for (int m = 0; m < 3; ++m) A.v[m] = (T *)d.field[SDT_F_VX + m];
    const int sf[6] = {SDT_F_SXX, SDT_F_SYY, SDT_F_SZZ, SDT_F_SXY, SDT_F_SXZ, SDT_F_SYZ};
    for (int m = 0; m < 6; ++m) A.s[m] = (T *)d.field[sf[m]];
    for (int q = 0; q < 21; ++q) A.C[q] = nullptr;
    A.mid = nullptr;
    A.tab = nullptr;
    A.tab_row = 0;
    A.dim = 3;
    if (d.pw.em) {   // E fields in v, H fields in s[0..2]
        for (int m = 0; m < 3; ++m) {
            A.v[m] = (T *)d.field[SDT_F_EX + m];
            A.s[m] = (T *)d.field[SDT_F_HX + m];
            A.s[m + 3] = nullptr;
        }
        if (d.physics == SDT_EM2) A.v[1] = A.s[0] = A.s[2] = nullptr;   // Ex, Ez, Hy
    } else if (d.physics == SDT_ELASTIC2) {
        A.v[1] = nullptr;
        A.s[1] = A.s[3] = A.s[5] = nullptr;
        if (d.mat_table) {   // rows hold c11 c13 c15 c33 c35 c55 in their Voigt slots
            A.mid = d.mid;
            A.tab = (const T *)d.tab;
            A.tab_row = kEl2TabRow;
        } else {
            const int slot[6] = {0, 2, 4, 11, 13, 18};   // c11 c13 c15 c33 c35 c55
            for (int q = 0; q < 6; ++q) A.C[slot[q]] = (const T *)d.mat[q];
        }
        A.dim = 2;
    } else if (d.mat_table) {   // 3D or 2.5D elastic: stiffness from the material table
        A.mid = d.mid;
        A.tab = (const T *)d.tab;
        A.tab_row = d3 ? kEl3TabRow : kEl25TabRow;
        A.dim = d3 ? 3 : 25;
    } else if (d.mat_iso) {   // isotropic form: slots of c11, c12, c44 repeat, the rest are 0
        static const int slot[21] = {0, 1, 1, -1, -1, -1, 0, 1, -1, -1, -1, 0, -1, -1, -1, 15, -1, -1, 15, -1, 15};
        for (int q = 0; q < 21; ++q) A.C[q] = slot[q] < 0 ? nullptr : (const T *)d.mat[slot[q]];
        A.dim = d3 ? 3 : 25;
    } else {
        for (int q = 0; q < 21; ++q) A.C[q] = (const T *)d.mat[q];
        A.dim = d3 ? 3 : 25;
    }
    A.nx = d.nx;
    A.ny = d.ny;
    A.nz = d.nz;
    A.pitch = d.pitch;
    A.xoff = d.xoff;
    A.cplx = d.ncomp == 2 ? 1 : 0;
    A.three_d = d.physics == SDT_ELASTIC3 || d.physics == SDT_EM3;
    A.d_it = d.d_it;
    A.vnorm = d.norms;
    A.snorm = d3 ? d.norms + d.time_steps : nullptr;
    for (int pass = 0; pass < 2; ++pass)
        for (int f = 0; f < 3; ++f) {
            if (!d.pw.face[f]) continue;
            const int a = f == 0 ? 1 : 0, b = f == 2 ? 1 : 2;
            const int cells = (d.pw.hi[a] - d.pw.lo[a] + 1) * (d.pw.hi[b] - d.pw.lo[b] + 1);
            if (cells <= 0) continue;
            const dim3 grid((cells + 255) / 256), block(256);
            if (pass == 0) hipLaunchKernelGGL(pw_inject_kernel<T>, grid, block, 0, d.stream, A, f);
            else hipLaunchKernelGGL(pw_norm_kernel<T>, grid, block, 0, d.stream, A, f);
        }
}

void launch_plane_wave(Device &d)
{
    if (!d.pw.on) return;
    if (d.precision == SDT_FP64) launch_plane_wave_t<double>(d);
    else launch_plane_wave_t<float>(d);
}

}  // namespace sdt
