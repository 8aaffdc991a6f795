// em25.hip -- 2.5D anisotropic electromagnetic CPML-FDTD (electromag25,
// cpmlfdtd.f08:2764-3350) on gfx950.
//
// Complex E and H on the (x, z) plane with exp(i ky y) dependence: y
// derivatives are i*ky*f = (-ky*im, ky*re); x and z use 4th-order staggered
// differences.  H is updated first (E-side memory variables, half CPML set),
// then E from curl H with the inverse permittivity and the conductivity
// applied to E_old (H-side memory variables, full set).  E_old == E at each
// step start (3282-3284), so the E update runs in place.  H is never written
// on the faces by the reference's nests and is zeroed there every step, so it
// stays 0 there; E faces are written 0 by the E kernel.
//
// Layout as the 2D solvers: padded rows, one thread per cell, 64 x 4 tiles,
// interior tiles without memory variables.  Complex values are interleaved
// (re, im).
#include <math.h>

#include <vector>

#include "sdt_device.h"
#include "sdt_internal.h"

namespace sdt {

template <typename T>
struct alignas(2 * sizeof(T)) Cz {
    T re, im;
};

template <typename T>
__device__ __forceinline__ Cz<T> czk(T re, T im)
{
    Cz<T> z;
    z.re = re;
    z.im = im;
    return z;
}
// i*ky*f
template <typename T>
__device__ __forceinline__ Cz<T> ikyz(T ky, Cz<T> f) { return czk(-ky * f.im, ky * f.re); }
// 27(f1 - f0) - (f2 - fm1), scaled by r = 1/(24 h)
template <typename T>
__device__ __forceinline__ Cz<T> d4z(Cz<T> fm1, Cz<T> f0, Cz<T> f1, Cz<T> f2, T r)
{
    return czk((T(27) * (f1.re - f0.re) - (f2.re - fm1.re)) * r, (T(27) * (f1.im - f0.im) - (f2.im - fm1.im)) * r);
}
// psi = b*psi + a*d ; d = d/kappa + psi
template <typename T>
__device__ __forceinline__ Cz<T> fixz(Cz<T> d, Cz<T> *psi, int64_t c, T a, T b, T ik)
{
    Cz<T> m = psi[c];
    m.re = b * m.re + a * d.re;
    m.im = b * m.im + a * d.im;
    psi[c] = m;
    return czk(d.re * ik + m.re, d.im * ik + m.im);
}

// memory-variable slots
enum { Q_EY_DZ = 0, Q_EX_DZ, Q_EZ_DX, Q_EY_DX, Q_HY_DZ, Q_HX_DZ, Q_HZ_DX, Q_HY_DX, Q_TOTAL };

template <typename T>
struct Em25Args {
    Cz<T> *E[3], *H[3];
    const T *aEx, *bEx, *cEx, *bEy, *cEy, *cEz;   // aEy = bEx, aEz = cEx, bEz = cEy
    const T *s11, *s12, *s13, *s22, *s23, *s33;
    const T *a, *b, *ik, *ah, *bh, *ikh;
    Cz<T> *m[Q_TOTAL];
    const uint8_t *flags;
    int box[6];
    int nx, nz, tiles_x, pitch;
    int64_t xoff;
    T rdx24, rdz24, dt, dbH, ky;
    const int *d_it;
    unsigned long long *vnorm;
    int64_t src_cell;
    // material table (at most 256 distinct (inverse permittivity,
    // conductivity) tuples): id per padded cell, rows of kEm25TabRow holding
    // aEx bEx cEx bEy cEy cEz s11 s12 s13 s22 s23 s33; null: per-cell arrays
    const uint8_t *mid;
    const T *tab;
};

template <typename T>
using em25_const_row = __attribute__((address_space(4))) const T *;

// The twelve E-update coefficients of a cell: from its table row (TAB;
// scalar loads when the wavefront's cells share one material, per-lane cached
// loads otherwise) or from the per-cell arrays.
template <typename T, bool TAB>
__device__ __forceinline__ void em25_coefs(const Em25Args<T> &A, int64_t c, int m, T (&co)[12])
{
    if (TAB) {
        const int m0 = __builtin_amdgcn_readfirstlane(m);
        if (__all(m == m0)) {
            const em25_const_row<T> p = (em25_const_row<T>)A.tab + m0 * kEm25TabRow;
#pragma unroll
            for (int q = 0; q < 12; ++q) co[q] = p[q];
        } else {
            const T *p = A.tab + m * kEm25TabRow;
#pragma unroll
            for (int q = 0; q < 12; ++q) co[q] = p[q];
        }
        return;
    }
    const T *const arr[12] = {A.aEx, A.bEx, A.cEx, A.bEy, A.cEy, A.cEz, A.s11, A.s12, A.s13, A.s22, A.s23, A.s33};
#pragma unroll
    for (int q = 0; q < 12; ++q) co[q] = arr[q][c];
}

// Hx (cpmlfdtd.f08:3106-3116), Hy (3131-3146), Hz (3160-3170)
template <typename T, bool PML>
__device__ __forceinline__ void em25_h_cell(const Em25Args<T> &A, int i, int k)
{
    const int nx = A.nx, nz = A.nz;
    const int64_t c = A.xoff + (int64_t)(i - 1) + (int64_t)A.pitch * (k - 1);
    const int64_t sz = A.pitch;
    const Cz<T> *Ex = A.E[0], *Ey = A.E[1], *Ez = A.E[2];
    T ah = 0, bh = 0, ikh = 1;
    if (PML) {
        ah = A.ah[c]; bh = A.bh[c]; ikh = A.ikh[c];
    }
    const T dbH = A.dbH;
    if (k >= 2 && k <= nz - 2 && i >= 2 && i <= nx - 1) {
        Cz<T> d = d4z(Ey[c - sz], Ey[c], Ey[c + sz], Ey[c + 2 * sz], A.rdz24);
        if (PML) d = fixz(d, A.m[Q_EY_DZ], c, ah, bh, ikh);
        const Cz<T> t = ikyz(A.ky, Ez[c]);
        const Cz<T> h = A.H[0][c];
        A.H[0][c] = czk(h.re - dbH * (t.re - d.re), h.im - dbH * (t.im - d.im));
    }
    if (k >= 2 && k <= nz - 2 && i >= 2 && i <= nx - 2) {
        Cz<T> dxz = d4z(Ex[c - sz], Ex[c], Ex[c + sz], Ex[c + 2 * sz], A.rdz24);
        Cz<T> dzx = d4z(Ez[c - 1], Ez[c], Ez[c + 1], Ez[c + 2], A.rdx24);
        if (PML) {
            dxz = fixz(dxz, A.m[Q_EX_DZ], c, ah, bh, ikh);
            dzx = fixz(dzx, A.m[Q_EZ_DX], c, ah, bh, ikh);
        }
        const Cz<T> h = A.H[1][c];
        A.H[1][c] = czk(h.re - dbH * (dxz.re - dzx.re), h.im - dbH * (dxz.im - dzx.im));
    }
    if (k >= 2 && k <= nz - 1 && i >= 2 && i <= nx - 2) {
        Cz<T> d = d4z(Ey[c - 1], Ey[c], Ey[c + 1], Ey[c + 2], A.rdx24);
        if (PML) d = fixz(d, A.m[Q_EY_DX], c, ah, bh, ikh);
        const Cz<T> t = ikyz(A.ky, Ex[c]);
        const Cz<T> h = A.H[2][c];
        A.H[2][c] = czk(h.re - dbH * (d.re - t.re), h.im - dbH * (d.im - t.im));
    }
}

template <typename T>
__global__ void __launch_bounds__(256) em25_h_kernel(Em25Args<T> A)
{
    const int i = blockIdx.x * 64 + threadIdx.x + 1;
    const int k = blockIdx.y * 4 + threadIdx.y + 1;
    if (i > A.nx || k > A.nz) return;
    if (!in_box(A.box, i, 1, k) || A.flags[blockIdx.x + (size_t)A.tiles_x * blockIdx.y]) em25_h_cell<T, true>(A, i, k);
    else em25_h_cell<T, false>(A, i, k);
}

// E (cpmlfdtd.f08:3190-3228), Dirichlet faces (3246-3258) and the norm
// (3309) without the source cell
template <typename T, bool PML, bool TAB>
__device__ __forceinline__ double em25_e_cell(const Em25Args<T> &A, int i, int k)
{
    const int nx = A.nx, nz = A.nz;
    const int64_t c = A.xoff + (int64_t)(i - 1) + (int64_t)A.pitch * (k - 1);
    const int64_t sz = A.pitch;
    Cz<T> ex = A.E[0][c], ey = A.E[1][c], ez = A.E[2][c];
    if (i == 1 || i == nx || k == 1 || k == nz) {
        ex = ey = ez = czk(T(0), T(0));
        A.E[0][c] = ex; A.E[1][c] = ey; A.E[2][c] = ez;
    } else if (k >= 3 && k <= nz - 2 && i >= 3 && i <= nx - 2) {
        const int mat = TAB ? (int)A.mid[c] : 0;   // first: the row loads wait on it
        const Cz<T> *Hx = A.H[0], *Hy = A.H[1], *Hz = A.H[2];
        const T rx = A.rdx24, rz = A.rdz24;
        const Cz<T> dHz_dy = ikyz(A.ky, Hz[c]);
        Cz<T> dHy_dz = d4z(Hy[c - 2 * sz], Hy[c - sz], Hy[c], Hy[c + sz], rz);
        Cz<T> dHx_dz = d4z(Hx[c - 2 * sz], Hx[c - sz], Hx[c], Hx[c + sz], rz);
        Cz<T> dHz_dx = d4z(Hz[c - 2], Hz[c - 1], Hz[c], Hz[c + 1], rx);
        Cz<T> dHy_dx = d4z(Hy[c - 2], Hy[c - 1], Hy[c], Hy[c + 1], rx);
        const Cz<T> dHx_dy = ikyz(A.ky, Hx[c]);
        if (PML) {
            const T a = A.a[c], b = A.b[c], ik = A.ik[c];
            dHy_dz = fixz(dHy_dz, A.m[Q_HY_DZ], c, a, b, ik);
            dHx_dz = fixz(dHx_dz, A.m[Q_HX_DZ], c, a, b, ik);
            dHz_dx = fixz(dHz_dx, A.m[Q_HZ_DX], c, a, b, ik);
            dHy_dx = fixz(dHy_dx, A.m[Q_HY_DX], c, a, b, ik);
        }
        T co[12];
        em25_coefs<T, TAB>(A, c, mat, co);
        const T aEx = co[0], bEx = co[1], cEx = co[2], bEy = co[3], cEy = co[4], cEz = co[5];
        const T s11 = co[6], s12 = co[7], s13 = co[8], s22 = co[9], s23 = co[10], s33 = co[11];
        const T dt = A.dt;
        const T rxr = dHz_dy.re - dHy_dz.re - (s11 * ex.re + s12 * ey.re + s13 * ez.re);
        const T rxi = dHz_dy.im - dHy_dz.im - (s11 * ex.im + s12 * ey.im + s13 * ez.im);
        const T ryr = dHx_dz.re - dHz_dx.re - (s12 * ex.re + s22 * ey.re + s23 * ez.re);
        const T ryi = dHx_dz.im - dHz_dx.im - (s12 * ex.im + s22 * ey.im + s23 * ez.im);
        const T rzr = dHy_dx.re - dHx_dy.re - (s13 * ex.re + s23 * ey.re + s33 * ez.re);
        const T rzi = dHy_dx.im - dHx_dy.im - (s13 * ex.im + s23 * ey.im + s33 * ez.im);
        ex = czk(ex.re + (aEx * rxr + bEx * ryr + cEx * rzr) * dt, ex.im + (aEx * rxi + bEx * ryi + cEx * rzi) * dt);
        ey = czk(ey.re + (bEx * rxr + bEy * ryr + cEy * rzr) * dt, ey.im + (bEx * rxi + bEy * ryi + cEy * rzi) * dt);
        ez = czk(ez.re + (cEx * rxr + cEy * ryr + cEz * rzr) * dt, ez.im + (cEx * rxi + cEy * ryi + cEz * rzi) * dt);
        A.E[0][c] = ex; A.E[1][c] = ey; A.E[2][c] = ez;
    }
    if (c == A.src_cell) return 0.0;
    const double s = ((double)ex.re * ex.re + (double)ex.im * ex.im) + ((double)ey.re * ey.re + (double)ey.im * ey.im) +
                     ((double)ez.re * ez.re + (double)ez.im * ez.im);
    return sqrt(s);
}

template <typename T, bool TAB>
__global__ void __launch_bounds__(256) em25_e_kernel(Em25Args<T> A)
{
    const int i = blockIdx.x * 64 + threadIdx.x + 1;
    const int k = blockIdx.y * 4 + threadIdx.y + 1;
    double nrm = 0.0;
    if (i <= A.nx && k <= A.nz) {
        if (!in_box(A.box, i, 1, k) || A.flags[blockIdx.x + (size_t)A.tiles_x * blockIdx.y])
            nrm = em25_e_cell<T, true, TAB>(A, i, k);
        else
            nrm = em25_e_cell<T, false, TAB>(A, i, k);
    }
    block_max_to(&A.vnorm[*A.d_it - 1], nrm);
}

template <typename T>
static Em25Args<T> em25_args(Device &d)
{
    Em25Args<T> A;
    for (int q = 0; q < 3; ++q) {
        A.E[q] = (Cz<T> *)d.field[SDT_F_EX + q];
        A.H[q] = (Cz<T> *)d.field[SDT_F_HX + q];
    }
    const T **mp[] = {&A.aEx, &A.bEx, &A.cEx, &A.bEy, &A.cEy, &A.cEz,
                      &A.s11, &A.s12, &A.s13, &A.s22, &A.s23, &A.s33};
    for (int q = 0; q < 12; ++q) *mp[q] = (const T *)d.mat[q];
    A.a = (const T *)d.cpml[0];
    A.b = (const T *)d.cpml[1];
    A.ik = (const T *)d.cpml[2];
    A.ah = (const T *)d.cpml[3];
    A.bh = (const T *)d.cpml[4];
    A.ikh = (const T *)d.cpml[5];
    for (int q = 0; q < Q_TOTAL; ++q) A.m[q] = (Cz<T> *)d.mem[q];
    A.flags = d.tile_flags;
    for (int q = 0; q < 6; ++q) A.box[q] = d.box[q];
    A.nx = d.nx;
    A.nz = d.nz;
    A.pitch = d.pitch;
    A.xoff = d.xoff;
    A.tiles_x = d.tiles_x;
    A.rdx24 = (T)(1.0 / (24.0 * d.dx));
    A.rdz24 = (T)(1.0 / (24.0 * d.dz));
    A.dt = (T)d.dt;
    A.dbH = (T)(d.dt / kMu0);
    A.ky = (T)d.ky;
    A.d_it = d.d_it;
    A.vnorm = d.norms;
    A.src_cell = d.tail.vfix[0];
    A.mid = d.mat_table ? d.mid : nullptr;
    A.tab = d.mat_table ? (const T *)d.tab : nullptr;
    return A;
}

template <typename T>
static void em25_step_t(Device &d, bool timed)
{
    const Em25Args<T> A = em25_args<T>(d);
    const dim3 grid(d.tiles_x, d.tiles_z), block(64, 4);
    if (timed) (void)hipEventRecord(d.ev[0], d.stream);
    hipLaunchKernelGGL(em25_h_kernel<T>, grid, block, 0, d.stream, A);
    if (timed) (void)hipEventRecord(d.ev[1], d.stream);
    if (d.mat_table) hipLaunchKernelGGL((em25_e_kernel<T, true>), grid, block, 0, d.stream, A);
    else hipLaunchKernelGGL((em25_e_kernel<T, false>), grid, block, 0, d.stream, A);
    if (timed) (void)hipEventRecord(d.ev[2], d.stream);
    launch_tail(d);
}

void em25_step(Device &d, bool timed)
{
    if (d.precision == SDT_FP64) em25_step_t<double>(d, timed);
    else em25_step_t<float>(d, timed);
}

// Inverse permittivity by adjugate/determinant of the absolute permittivity
// (cpmlfdtd.f08:3040-3066), FP64 on the host, six unique entries; the six
// conductivities as given.  Per-cell arrays, or table rows with one byte of
// id per cell when the model has at most 256 distinct tuples.
int em25_setup(Device &d, const double *const *in)
{
    d.tile_x = 64;
    d.tile_y = 1;
    d.tile_z = 4;
    d.ncomp = 2;
    int rc = compute_tile_flags(d, in);
    if (rc) return rc;
    // the kernels' 64-wide tiles start at column 1: put it on a line boundary
    set_layout(d, 1);
    const int64_t n = d.ncell;
    auto val = [&](int id, int64_t c) { return in[id] ? in[id][c] : 0.0; };
    std::vector<double> co[6];
    for (auto &v : co) v.resize((size_t)n);
    for (int64_t c = 0; c < n; ++c) {
        const double e11 = val(SDT_E11, c) * kEps0, e12 = val(SDT_E12, c) * kEps0, e13 = val(SDT_E13, c) * kEps0;
        const double e22 = val(SDT_E22, c) * kEps0, e23 = val(SDT_E23, c) * kEps0, e33 = val(SDT_E33, c) * kEps0;
        const double det = e11 * (e22 * e33 - e23 * e23) - e12 * (e12 * e33 - e23 * e13) +
                           e13 * (e12 * e23 - e22 * e13);
        co[0][(size_t)c] = (e22 * e33 - e23 * e23) / det;
        co[1][(size_t)c] = -(e12 * e33 - e23 * e13) / det;
        co[2][(size_t)c] = (e12 * e23 - e22 * e13) / det;
        co[3][(size_t)c] = (e11 * e33 - e13 * e13) / det;
        co[4][(size_t)c] = -(e11 * e23 - e12 * e13) / det;
        co[5][(size_t)c] = (e11 * e22 - e12 * e12) / det;
    }
    // few distinct materials (layered / blocky models): one-byte ids and a
    // table of (inverse permittivity, conductivity) rows
    if (d.mat_table_opt) {
        rc = build_material_table(d, kEm25TabRow, [&](int64_t c, double *r) {
            for (int q = 0; q < 6; ++q) r[q] = co[q][(size_t)c];
            for (int q = 0; q < 6; ++q) r[6 + q] = val(SDT_S11 + q, c);
        });
        if (rc) return rc;
    }
    d.nmat = 0;
    if (!d.mat_table) {
        for (int q = 0; q < 6; ++q) {
            rc = upload_grid(d, co[q].data(), &d.mat[d.nmat++]);
            if (rc) return rc;
        }
        std::vector<double> zero;
        for (int id = SDT_S11; id <= SDT_S33; ++id) {
            if (!in[id] && zero.empty()) zero.assign((size_t)n, 0.0);
            rc = upload_grid(d, in[id] ? in[id] : zero.data(), &d.mat[d.nmat++]);
            if (rc) return rc;
        }
    }
    d.nmem = Q_TOTAL;
    for (int q = 0; q < d.nmem; ++q) {
        rc = alloc_grid_zeroed(d, &d.mem[q], 2);
        if (rc) return rc;
    }
    for (int f = SDT_F_EX; f <= SDT_F_HZ; ++f) {
        if (f <= SDT_F_EZ && in[SDT_IC_X + f]) rc = upload_grid(d, in[SDT_IC_X + f], &d.field[f], 2);
        else rc = alloc_grid_zeroed(d, &d.field[f], 2);
        if (rc) return rc;
    }

    // point source, cpmlfdtd.f08:3236-3241: E += src * dt / (eps0 * eps_ii)
    const int is = d.isrc, ks = d.ksrc, nx = d.nx, nz = d.nz;
    const int64_t hc = (int64_t)(is - 1) + (int64_t)nx * (ks - 1);
    const int64_t c = cell_index(d, is, 1, ks);
    const int face = (is == 1 || is == nx || ks == 1 || ks == nz);
    TailArgs &t = d.tail;
    t.nterms = 0;
    t.terms[t.nterms++] = {c, SDT_F_EX, SDT_SRC_X, d.dt, val(SDT_E11, hc) * kEps0, face};
    t.terms[t.nterms++] = {c, SDT_F_EY, SDT_SRC_Y, d.dt, val(SDT_E22, hc) * kEps0, face};
    t.terms[t.nterms++] = {c, SDT_F_EZ, SDT_SRC_Z, d.dt, val(SDT_E33, hc) * kEps0, face};
    t.nvfix = 1;
    t.vfix[0] = c;
    t.nvcomp = 3;
    t.vcomp[0] = SDT_F_EX; t.vcomp[1] = SDT_F_EY; t.vcomp[2] = SDT_F_EZ;
    t.stress_norm = 0;
    t.nsfix = 0;
    t.rec_field[0] = SDT_F_EX; t.rec_field[1] = SDT_F_EY; t.rec_field[2] = SDT_F_EZ;
    t.cplx = 1;
    return upload_cpml_full(d, in);
}

}  // namespace sdt
