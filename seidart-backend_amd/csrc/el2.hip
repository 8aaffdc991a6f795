// el2.hip -- 2D anisotropic elastic CPML-FDTD (seismic2, cpmlfdtd.f08:57-597)
// on gfx950.
//
// Layout: every array is (nx, nz) Fortran order, x fastest, in HBM, rows
// padded to whole 128-byte lines.  One thread per cell; a workgroup of four
// wavefronts covers 64 line-aligned columns and a band of rows, each
// wavefront marching its own rows.  A wavefront row whose cells all have
// identity CPML coefficients runs the interior body (no memory variables, no
// a/b/kappa reads); other rows run the CPML body.
#include <math.h>

#include <vector>

#include "sdt_device.h"
#include "sdt_internal.h"

namespace sdt {

template <typename T>
struct El2Args {
    T *vx, *vz, *sxx, *szz, *sxz;
    const T *c11, *c13, *c15, *c33, *c35, *c55, *rho, *igx, *igz, *igxz;
    const T *a, *b, *ik, *ah, *bh, *ikh;
    T *m[12];
    const uint8_t *flags;
    int box[6];
    int nx, nz, tiles_x, pitch;
    int64_t xoff;
    T rdx24, rdz24, dt;
    int density_code, corrected;
    const int *d_it;
    unsigned long long *vnorm;
    int64_t src_cell;   // excluded from the fused norm (the tail adds it)
    int nsteps;         // time_steps of the run (batched shots stop there)
    // segment recomputation for the adjoint: per step (0-based t), slot
    // t - store_first of `store` ([slot][kEl2Fwd][stride]) holds the
    // CPML-corrected derivatives and pre-attenuation stresses of every cell
    // (zero outside each update's region): 0 dvx/dx, 1 dvz/dz, 2 shear
    // (sigma_xx/zz region), 3 sigma_xx and 4 sigma_zz before the 1/(1 +
    // gamma dt) factor, 5 dvx/dx, 6 dvz/dz, 7 dvz/dx + dvx/dz (sigma_xz
    // region), 8 sigma_xz before the factor, 9 dsxx/dx, 10 dsxz/dz (vx
    // region), 11 dsxz/dx, 12 dszz/dz (vz region)
    T *store;
    int64_t stride;
    int store_first;
    PwFaces pwf;        // plane-wave cells, skipped by the fused norm
    // material table (a model with at most 256 distinct (stiffness,
    // attenuation) tuples and no adjoint): id per padded cell, rows of
    // kEl2TabRow with c11 c13 c15 c33 c35 c55 in their Voigt slots (0 2 4 11
    // 13 18) and the divisors x, z, xz at 21..23; null: per-cell arrays.
    // The density stays a per-cell array.
    const uint8_t *mid;
    const T *tab;
};
constexpr int kEl2Fwd = 13;

template <typename T>
using el2_const_row = __attribute__((address_space(4))) const T *;

// Stiffness and divisors of a cell from its table row: scalar loads when the
// wavefront's cells share one material, per-lane cached loads otherwise.
template <typename T>
__device__ __forceinline__ void el2_row(const T *tab, int m, T &k11, T &k13, T &k15, T &k33, T &k35, T &k55, T &gx,
                                        T &gz, T &gxz)
{
    const int m0 = __builtin_amdgcn_readfirstlane(m);
    if (__all(m == m0)) {
        const el2_const_row<T> p = (el2_const_row<T>)tab + m0 * kEl2TabRow;
        k11 = p[0]; k13 = p[2]; k15 = p[4]; k33 = p[11]; k35 = p[13]; k55 = p[18];
        gx = p[21]; gz = p[22]; gxz = p[23];
    } else {
        const T *p = tab + m * kEl2TabRow;
        k11 = p[0]; k13 = p[2]; k15 = p[4]; k33 = p[11]; k35 = p[13]; k55 = p[18];
        gx = p[21]; gz = p[22]; gxz = p[23];
    }
}

// memory recursion psi = b*psi + a*d ; d = d/kappa + psi (cpmlfdtd.f08:332-340)
template <typename T>
__device__ __forceinline__ T cpml_fix(T d, T &psi, T a, T b, T ik)
{
    psi = b * psi + a * d;
    return d * ik + psi;
}

// sigma_xx, sigma_zz (cpmlfdtd.f08:324-353) and sigma_xz (362-386).  Every
// operand is read before the first store: the compiler cannot tell the
// arrays apart, so a store ahead of a load would serialise the memory
// round trips.
template <typename T, bool PML, bool STORE, bool TAB>
__device__ __forceinline__ void el2_stress_cell(const El2Args<T> &A, int i, int j)
{
    const int nx = A.nx, nz = A.nz;
    const bool r1 = i >= 2 && i <= nx - 2 && j >= 3 && j <= nz - 1;   // sigma_xx, sigma_zz
    const bool r2 = i >= 3 && i <= nx - 1 && j >= 2 && j <= nz - 2;   // sigma_xz
    const int64_t c = A.xoff + (int64_t)(i - 1) + (int64_t)A.pitch * (j - 1);
    T *st = nullptr;
    if (STORE) {
        st = A.store + (int64_t)(*A.d_it - 1 - A.store_first) * kEl2Fwd * A.stride + c;
        for (int q = 0; q < 9; ++q) st[q * A.stride] = T(0);
    }
    if (!r1 && !r2) return;
    // TAB: the launch may meet a table (compiled out otherwise); uniform over
    // the launch's shot; recording runs (the adjoint's) never hold a table
    const bool tab = TAB && !STORE && A.mid != nullptr;
    const int mat = tab ? (int)A.mid[c] : 0;
    const T *vx = A.vx, *vz = A.vz;
    const int64_t sx = 1, sz = A.pitch;
    // stencil operands: x0 = vx(i,j), xl/xr = vx(i-+1,j), xd/xu = vx(i,j-+1), ...
    const T x0 = vx[c], xl = vx[c - sx], xr = vx[c + sx], xd = vx[c - sz], xu = vx[c + sz];
    const T z0 = vz[c], zl = vz[c - sx], zr = vz[c + sx], zd = vz[c - sz], zu = vz[c + sz];
    T xr2 = 0, xd2 = 0, zr2 = 0, zd2 = 0, xl2 = 0, xu2 = 0, zl2 = 0, zu2 = 0;
    T s_xx = 0, s_zz = 0, s_xz = 0, k11 = 0, k13 = 0, k15 = 0, k33 = 0, k35 = 0, k55 = 0;
    T gx = 0, gz = 0, gxz = 0;
    T ah = 0, bh = 0, ikh = 0, a = 0, b = 0, ik = 0, m[8] = {};
    if (r1) {
        xr2 = vx[c + 2 * sx];
        xd2 = vx[c - 2 * sz];
        zr2 = vz[c + 2 * sx];
        zd2 = vz[c - 2 * sz];
        s_xx = A.sxx[c];
        s_zz = A.szz[c];
        if (!tab) {
            k11 = ldnt(A.c11 + c);
            k13 = ldnt(A.c13 + c);
            k33 = ldnt(A.c33 + c);
            gx = ldnt(A.igx + c);
            gz = ldnt(A.igz + c);
        }
    }
    if (r2) {
        xl2 = vx[c - 2 * sx];
        xu2 = vx[c + 2 * sz];
        zl2 = vz[c - 2 * sx];
        zu2 = vz[c + 2 * sz];
        s_xz = A.sxz[c];
        if (!tab) {
            k55 = ldnt(A.c55 + c);
            gxz = ldnt(A.igxz + c);
        }
    }
    if (!tab) {
        k15 = ldnt(A.c15 + c);
        k35 = ldnt(A.c35 + c);
    }
    if (PML) {
        ah = ldnt(A.ah + c);
        bh = ldnt(A.bh + c);
        ikh = ldnt(A.ikh + c);
        if (r1) {
            a = ldnt(A.a + c);
            b = ldnt(A.b + c);
            ik = ldnt(A.ik + c);
            for (int q = 0; q < 4; ++q) m[q] = ldnt(A.m[q] + c);
        }
        if (r2)
            for (int q = 4; q < 8; ++q) m[q] = ldnt(A.m[q] + c);
    }
    if (tab) el2_row(A.tab, mat, k11, k13, k15, k33, k35, k55, gx, gz, gxz);
    if (r1) {
        T dvx_dx = (T(27) * (xr - x0) - (xr2 - xl)) * A.rdx24;
        T dvz_dz = (T(27) * (z0 - zd) - (zu - zd2)) * A.rdz24;
        T dvz_dx = (T(27) * (zr - z0) - (zr2 - zl)) * A.rdx24;
        T dvx_dz = (T(27) * (x0 - xd) - (xu - xd2)) * A.rdz24;
        if (PML) {
            dvx_dx = cpml_fix(dvx_dx, m[0], ah, bh, ikh);
            dvx_dz = cpml_fix(dvx_dz, m[1], ah, bh, ikh);
            dvz_dz = cpml_fix(dvz_dz, m[2], a, b, ik);
            dvz_dx = cpml_fix(dvz_dx, m[3], a, b, ik);
        }
        const T shear = dvz_dx + dvx_dz;
        const T pxx = s_xx + (k11 * dvx_dx + k13 * dvz_dz + k15 * shear) * A.dt;
        const T pzz = s_zz + (k13 * dvx_dx + k33 * dvz_dz + k35 * shear) * A.dt;
        s_xx = pxx * gx;
        s_zz = pzz * gz;
        if (STORE) {
            st[0] = dvx_dx;
            st[A.stride] = dvz_dz;
            st[2 * A.stride] = shear;
            st[3 * A.stride] = pxx;
            st[4 * A.stride] = pzz;
        }
    }
    if (r2) {
        T dvx_dx = (T(27) * (x0 - xl) - (xr - xl2)) * A.rdx24;
        T dvz_dz = (T(27) * (zu - z0) - (zu2 - zd)) * A.rdz24;
        T dvz_dx = (T(27) * (z0 - zl) - (zr - zl2)) * A.rdx24;
        T dvx_dz = (T(27) * (xu - x0) - (xu2 - xd)) * A.rdz24;
        if (PML) {
            dvx_dx = cpml_fix(dvx_dx, m[4], ah, bh, ikh);
            dvx_dz = cpml_fix(dvx_dz, m[5], ah, bh, ikh);
            dvz_dz = cpml_fix(dvz_dz, m[6], ah, bh, ikh);
            dvz_dx = cpml_fix(dvz_dx, m[7], ah, bh, ikh);
        }
        const T pxz = s_xz + (k15 * dvx_dx + k35 * dvz_dz + k55 * (dvz_dx + dvx_dz)) * A.dt;
        s_xz = pxz * gxz;
        if (STORE) {
            st[5 * A.stride] = dvx_dx;
            st[6 * A.stride] = dvz_dz;
            st[7 * A.stride] = dvz_dx + dvx_dz;
            st[8 * A.stride] = pxz;
        }
    }
    if (r1) {
        A.sxx[c] = s_xx;
        A.szz[c] = s_zz;
        if (PML)
            for (int q = 0; q < 4; ++q) A.m[q][c] = m[q];
    }
    if (r2) {
        A.sxz[c] = s_xz;
        if (PML)
            for (int q = 4; q < 8; ++q) A.m[q][c] = m[q];
    }
}

// 3-term (reference) / 4-term (corrected) velocity stencils on operand values,
// cpmlfdtd.f08:405-406 and 434-435: tminus over f(c), f(c-s), f(c-2s), f(c+s);
// tplus over f(c), f(c+s), f(c+2s), f(c-s).
template <typename T>
__device__ __forceinline__ T tminus(T f0, T fm1, T fm2, T fp1, int corr)
{
    return corr ? T(27) * (f0 - fm1) - (fp1 - fm2) : T(27) * f0 - T(27) * fm1 + fm2;
}
template <typename T>
__device__ __forceinline__ T tplus(T f0, T fp1, T fp2, T fm1, int corr)
{
    return corr ? T(27) * (fp1 - f0) - (fp2 - fm1) : -T(27) * f0 + T(27) * fp1 - fp2;
}

// vx (cpmlfdtd.f08:399-416), vz (428-446), Dirichlet edges (534-549) and the
// squared velocity magnitude for the per-step norm (552-567), 0 at the source
// cell.  Loads first, as in the stress update.  The norm is the square root
// of the maximum squared magnitude (sqrt is monotone and correctly rounded,
// so this equals the maximum of the magnitudes).
template <typename T, bool PML, bool STORE>
__device__ __forceinline__ double el2_velocity_cell(const El2Args<T> &A, int i, int j)
{
    const int nx = A.nx, nz = A.nz;
    const int64_t c = A.xoff + (int64_t)(i - 1) + (int64_t)A.pitch * (j - 1);
    const int64_t sx = 1, sz = A.pitch;
    const int corr = A.corrected;
    const bool edge = i == 1 || i == nx || j == 1 || j == nz;
    const bool rx = i >= 3 && i <= nx - 1 && j >= 3 && j <= nz - 1;
    const bool rz = i >= 2 && i <= nx - 2 && j >= 2 && j <= nz - 2;
    T vx = A.vx[c], vz = A.vz[c];
    T r = 0, rxl = 0, rxd = 0, rzr = 0, rzu = 0;
    T f[8] = {}, g[8] = {};
    T a = 0, b = 0, ik = 0, ah = 0, bh = 0, ikh = 0, m[4] = {};
    // Every load is issued, whatever the cell's region: outside a region its
    // neighbour loads read the cell's own element instead (always in the
    // allocation) and the values go unused, so the compiler can batch the
    // loads instead of guarding them branch by branch.
    auto nx_ = [&](bool ok, int64_t off) { return ok ? c + off : c; };
    r = ldnt(A.rho + c);
    rxl = A.rho[nx_(rx, -sx)];
    rxd = A.rho[nx_(rx, -sz)];
    f[0] = A.sxx[c];
    f[1] = A.sxx[nx_(rx, -sx)];
    f[2] = A.sxx[nx_(rx, -2 * sx)];
    f[3] = corr ? A.sxx[nx_(rx, sx)] : T(0);
    f[4] = A.sxz[c];
    f[5] = A.sxz[nx_(rx, -sz)];
    f[6] = A.sxz[nx_(rx, -2 * sz)];
    f[7] = corr ? A.sxz[nx_(rx, sz)] : T(0);
    rzr = A.rho[nx_(rz, sx)];
    rzu = A.rho[nx_(rz, sz)];
    g[0] = f[4];
    g[1] = A.sxz[nx_(rz, sx)];
    g[2] = A.sxz[nx_(rz, 2 * sx)];
    g[3] = corr ? A.sxz[nx_(rz, -sx)] : T(0);
    g[4] = A.szz[c];
    g[5] = A.szz[nx_(rz, sz)];
    g[6] = A.szz[nx_(rz, 2 * sz)];
    g[7] = corr ? A.szz[nx_(rz, -sz)] : T(0);
    if (PML) {
        a = ldnt(A.a + c);
        b = ldnt(A.b + c);
        ik = ldnt(A.ik + c);
        ah = ldnt(A.ah + c);
        bh = ldnt(A.bh + c);
        ikh = ldnt(A.ikh + c);
        m[0] = ldnt(A.m[8] + c);
        m[1] = ldnt(A.m[9] + c);
        m[2] = ldnt(A.m[10] + c);
        m[3] = ldnt(A.m[11] + c);
    }
    if (edge) {
        vx = T(0);
        vz = T(0);
    } else {
        if (rx) {
            const T rhoxx = scalar_mean(r, rxl, A.density_code);
            const T rhozx = scalar_mean(r, rxd, A.density_code);
            T d1 = tminus(f[0], f[1], f[2], f[3], corr) * A.rdx24;
            T d2 = tminus(f[4], f[5], f[6], f[7], corr) * A.rdz24;
            if (PML) {
                d1 = cpml_fix(d1, m[0], a, b, ik);
                d2 = cpml_fix(d2, m[1], a, b, ik);
            }
            vx = vx + A.dt * (d1 / rhoxx + d2 / rhozx);
            if (STORE) {
                T *st = A.store + (int64_t)(*A.d_it - 1 - A.store_first) * kEl2Fwd * A.stride + c;
                st[9 * A.stride] = d1;
                st[10 * A.stride] = d2;
            }
        }
        if (rz) {
            const T rhoxz = scalar_mean(r, rzr, A.density_code);
            const T rhozz = scalar_mean(r, rzu, A.density_code);
            T d1 = tplus(g[0], g[1], g[2], g[3], corr) * A.rdx24;
            T d2 = tplus(g[4], g[5], g[6], g[7], corr) * A.rdz24;
            if (PML) {
                d1 = cpml_fix(d1, m[2], ah, bh, ikh);
                d2 = cpml_fix(d2, m[3], ah, bh, ikh);
            }
            vz = vz + A.dt * (d1 / rhoxz + d2 / rhozz);
            if (STORE) {
                T *st = A.store + (int64_t)(*A.d_it - 1 - A.store_first) * kEl2Fwd * A.stride + c;
                st[11 * A.stride] = d1;
                st[12 * A.stride] = d2;
            }
        }
    }
    if (STORE) {   // zero outside the update regions
        T *st = A.store + (int64_t)(*A.d_it - 1 - A.store_first) * kEl2Fwd * A.stride + c;
        if (edge || !rx) st[9 * A.stride] = st[10 * A.stride] = T(0);
        if (edge || !rz) st[11 * A.stride] = st[12 * A.stride] = T(0);
    }
    if (edge || rx) A.vx[c] = vx;
    if (edge || rz) A.vz[c] = vz;
    if (PML && rx) {
        A.m[8][c] = m[0];
        A.m[9][c] = m[1];
    }
    if (PML && rz) {
        A.m[10][c] = m[2];
        A.m[11][c] = m[3];
    }
    if (c == A.src_cell || pw_face(A.pwf, i, 1, j)) return 0.0;
    const double dvx = (double)vx, dvz = (double)vz;
    return dvx * dvx + dvz * dvz;
}

template <typename T>
static El2Args<T> el2_args(Device &d)

{
    El2Args<T> A;
    A.vx = (T *)d.field[SDT_F_VX];
    A.vz = (T *)d.field[SDT_F_VZ];
    A.sxx = (T *)d.field[SDT_F_SXX];
    A.szz = (T *)d.field[SDT_F_SZZ];
    A.sxz = (T *)d.field[SDT_F_SXZ];
    const T **mp[] = {&A.c11, &A.c13, &A.c15, &A.c33, &A.c35, &A.c55, &A.rho, &A.igx, &A.igz, &A.igxz};
    for (int q = 0; q < 10; ++q) *mp[q] = (const T *)d.mat[q];
    A.a = (const T *)d.cpml[0];
    A.b = (const T *)d.cpml[1];
    A.ik = (const T *)d.cpml[2];
    A.ah = (const T *)d.cpml[3];
    A.bh = (const T *)d.cpml[4];
    A.ikh = (const T *)d.cpml[5];
    for (int q = 0; q < 12; ++q) A.m[q] = (T *)d.mem[q];
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
    A.density_code = d.density_code;
    A.corrected = d.stencil == SDT_STENCIL_CORRECTED;
    A.d_it = d.d_it;
    A.vnorm = d.norms;
    A.src_cell = d.tail.vfix[0];
    A.nsteps = d.time_steps;
    A.store = (T *)d.adj.seg_fwd;
    A.stride = d.ncell_dev;
    A.store_first = d.adj.store_first;
    A.pwf = d.pw.faces();
    A.mid = d.mat_table ? d.mid : nullptr;
    A.tab = d.mat_table ? (const T *)d.tab : nullptr;
    return A;
}

// ---- tiles ------------------------------------------------------------------
// A tile is 64 consecutive elements of a padded row (line-aligned: element
// e = xoff + i - 1 of a row, tile x index e / 64) by kTile2dWaves * R rows;
// wavefront w marches rows R*w .. R*w + R - 1 of the band.  Single-shot
// launches index tiles by blockIdx; batched launches read (shot, tile x,
// band) from a tile list covering every shot.  The CPML body runs for a
// wavefront row when any of its cells needs it (outside the interior box or
// in a 64 x 4 block with non-identity coefficients); identity cells give the
// same values through either body.
template <typename T>
__device__ __forceinline__ bool el2_pml(const El2Args<T> &A, int i, int j)
{
    return !in_box(A.box, i, 1, j) || A.flags[((i - 1) >> 6) + (size_t)A.tiles_x * ((j - 1) >> 2)];
}

template <typename T, int R, bool STORE, bool TAB>
__device__ __forceinline__ void el2_stress_tile(const El2Args<T> &A, int tx, int band)
{
    const int i = tx * 64 + (int)threadIdx.x - (int)A.xoff + 1;
    const bool valid = i >= 1 && i <= A.nx;
    const int j0 = (band * kTile2dWaves + (int)threadIdx.y) * R + 1;
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        if (j > A.nz) break;
        if (__any(valid && el2_pml(A, i, j))) {
            if (valid) el2_stress_cell<T, true, STORE, TAB>(A, i, j);
        } else if (valid) {
            el2_stress_cell<T, false, STORE, TAB>(A, i, j);
        }
    }
}

template <typename T, int R, bool STORE>
__device__ __forceinline__ double el2_velocity_tile(const El2Args<T> &A, int tx, int band)
{
    const int i = tx * 64 + (int)threadIdx.x - (int)A.xoff + 1;
    const bool valid = i >= 1 && i <= A.nx;
    const int j0 = (band * kTile2dWaves + (int)threadIdx.y) * R + 1;
    double nrm = 0.0;
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        if (j > A.nz) break;
        double v = 0.0;
        if (__any(valid && el2_pml(A, i, j))) {
            if (valid) v = el2_velocity_cell<T, true, STORE>(A, i, j);
        } else if (valid) {
            v = el2_velocity_cell<T, false, STORE>(A, i, j);
        }
        nrm = (v > nrm || v != v) ? v : nrm;
    }
    return nrm;
}

template <typename T, int R, bool STORE, bool TAB>
__global__ void __launch_bounds__(64 * kTile2dWaves) el2_stress_kernel(El2Args<T> A)
{
    el2_stress_tile<T, R, STORE, TAB>(A, blockIdx.x, blockIdx.y);
}

template <typename T, int R, bool STORE>
__global__ void __launch_bounds__(64 * kTile2dWaves) el2_velocity_kernel(El2Args<T> A)
{
    const double nrm = sqrt(el2_velocity_tile<T, R, STORE>(A, blockIdx.x, blockIdx.y));
    block_max_to(&A.vnorm[*A.d_it - 1], nrm);
}

template <typename T, int R, bool STORE, bool TAB>
__global__ void __launch_bounds__(64 * kTile2dWaves) el2_stress_batch_kernel(const El2Args<T> *args,
                                                                             const int4 *tiles)
{
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;   // padding tile
    const El2Args<T> A = args[t.x];
    if (*A.d_it > A.nsteps) return;   // finished shot (uniform per workgroup)
    el2_stress_tile<T, R, STORE, TAB>(A, t.y, t.z);
}

template <typename T, int R, bool STORE>
__global__ void __launch_bounds__(64 * kTile2dWaves) el2_velocity_batch_kernel(const El2Args<T> *args,
                                                                               const int4 *tiles)
{
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;   // padding tile
    const El2Args<T> A = args[t.x];
    if (*A.d_it > A.nsteps) return;
    const double nrm = sqrt(el2_velocity_tile<T, R, STORE>(A, t.y, t.z));
    block_max_to(&A.vnorm[*A.d_it - 1], nrm);
}

template <typename T>
static int el2_batch_args_t(Batch &b, Device *const *shots, int n)
{
    std::vector<El2Args<T>> a((size_t)n);
    b.any_table = false;
    for (int s = 0; s < n; ++s) {
        a[(size_t)s] = el2_args<T>(*shots[s]);
        b.any_table = b.any_table || a[(size_t)s].mid != nullptr;
    }
    if (!b.args && hipMalloc(&b.args, a.size() * sizeof(El2Args<T>)) != hipSuccess) return SDT_ERR_HIP;
    if (hipMemcpy(b.args, a.data(), a.size() * sizeof(El2Args<T>), hipMemcpyHostToDevice) != hipSuccess)
        return SDT_ERR_HIP;
    return SDT_OK;
}

int el2_batch_args(Batch &b, Device *const *shots, int n)
{
    return b.precision == SDT_FP64 ? el2_batch_args_t<double>(b, shots, n) : el2_batch_args_t<float>(b, shots, n);
}

template <typename T, int R>
static void el2_batch_launch_t(Batch &b, int which)
{
    const El2Args<T> *a = (const El2Args<T> *)b.args;
    const dim3 grid(b.ntiles[which]), block(64, kTile2dWaves);
    if (b.store) {
        hipLaunchKernelGGL((el2_stress_batch_kernel<T, R, true, false>), grid, block, 0, b.stream, a, b.tiles[which]);
        hipLaunchKernelGGL((el2_velocity_batch_kernel<T, R, true>), grid, block, 0, b.stream, a, b.tiles[which]);
    } else {
        if (b.any_table)
            hipLaunchKernelGGL((el2_stress_batch_kernel<T, R, false, true>), grid, block, 0, b.stream, a,
                               b.tiles[which]);
        else
            hipLaunchKernelGGL((el2_stress_batch_kernel<T, R, false, false>), grid, block, 0, b.stream, a,
                               b.tiles[which]);
        hipLaunchKernelGGL((el2_velocity_batch_kernel<T, R, false>), grid, block, 0, b.stream, a, b.tiles[which]);
    }
}

template <typename T>
static void el2_batch_launch_p(Batch &b, int which)
{
    switch (kBatchShapes[which].rows) {
    case 1: el2_batch_launch_t<T, 1>(b, which); break;
    case 8: el2_batch_launch_t<T, 8>(b, which); break;
    default: el2_batch_launch_t<T, 4>(b, which); break;
    }
}

void el2_batch_launch(Batch &b, int which)
{
    if (b.precision == SDT_FP64) el2_batch_launch_p<double>(b, which);
    else el2_batch_launch_p<float>(b, which);
}

template <typename T, int R>
static void el2_step_r(Device &d, bool timed);

template <typename T>
static void el2_step_t(Device &d, bool timed)
{
    if (small_grid_2d(d)) el2_step_r<T, 1>(d, timed);
    else el2_step_r<T, kSingleRows2d>(d, timed);
}

template <typename T, int R>
static void el2_step_r(Device &d, bool timed)
{
    const El2Args<T> A = el2_args<T>(d);
    const dim3 grid(tiles_x_2d(d), tiles_z_2d(d, R)), block(64, kTile2dWaves);
    if (timed) (void)hipEventRecord(d.ev[0], d.stream);
    if (d.adj.store)
        hipLaunchKernelGGL((el2_stress_kernel<T, R, true, false>), grid, block, 0, d.stream, A);
    else if (d.mat_table)
        hipLaunchKernelGGL((el2_stress_kernel<T, R, false, true>), grid, block, 0, d.stream, A);
    else
        hipLaunchKernelGGL((el2_stress_kernel<T, R, false, false>), grid, block, 0, d.stream, A);
    if (timed) (void)hipEventRecord(d.ev[1], d.stream);
    if (d.adj.store)
        hipLaunchKernelGGL((el2_velocity_kernel<T, R, true>), grid, block, 0, d.stream, A);
    else
        hipLaunchKernelGGL((el2_velocity_kernel<T, R, false>), grid, block, 0, d.stream, A);
    if (timed) (void)hipEventRecord(d.ev[2], d.stream);
    launch_plane_wave(d);
    launch_tail(d);
}

// ---- adjoint state (discrete adjoint of the step above) ----------------------
// Step t: sigma' = S(sigma^t, v^t, psi) (stress pass), v' = V(v^t, sigma',
// psi) (velocity pass), then the point sources add to sigma' and v'
// (sigma^{t+1}, v^{t+1}) and the receivers read v^{t+1}.  Its transpose, per
// step, latest first:
//   tail     trace-gradient row t into lam(v^{t+1}) at the receivers;
//   V pass   lam(v^{t+1}) completed with the S-pass transfer terms of step
//            t+1 (recorded at the source cell and in the segment store),
//            pushed through V: own-cell lam(v^t) part, lam(psi_V), transfer
//            terms P, Q, U, V (the derivative adjoints, scaled by
//            1/(24 dx) or 1/(24 dz)) to lam(sigma');
//   S pass   lam(sigma^{t+1}) (recorded at the source cells) plus the
//            gathered V transfer terms = lam(sigma'), stored, pushed through
//            S: own-cell lam(sigma^t), lam(psi_S), transfer terms to
//            lam(v^t) gathered by the next V pass.
// Gradients (per cell, summed over steps in FP64 by the correlation pass):
//   c11: dt igx lxx a, c13: dt (igx lxx b + igz lzz a), c15: dt (igx lxx s +
//   igxz lxz p), c33: dt igz lzz b, c35: dt (igz lzz s + igxz lxz q), c55:
//   dt igxz lxz (r + u), igx: lxx pxx, igz: lzz pzz, igxz: lxz pxz, and for
//   the density the four sums lvx d1x, lvx d2x, lvz d1z, lvz d2z, turned
//   into d/drho through the face averages once all segments are done.
constexpr int kEl2Adj = 5;    // stored adjoint values: lam sxx', szz', sxz', vx', vz'
constexpr int kEl2Acc = 14;   // FP64 accumulators: the 10 coefficients + 4 density sums

template <typename T>
struct El2AdjArgs {
    El2Args<T> F;
    T *lvx, *lvz, *lxx, *lzz, *lxz;     // own-cell parts of the field adjoints
    T *lm[12];                          // lam of the memory variables (forward order)
    T *sx[8];                           // S transfers: Ax Bz Ez Fx (sigma_xx/zz region), Cx Dz Gz Hx (sigma_xz)
    T *vt[4];                           // V transfers: P Q (vx region), U V (vz region)
    T *adj;                             // segment store [slot][kEl2Adj][stride]
    double *src_lam;                    // [time_steps][kAdjSrcSlots]
    double *grad;                       // [kEl2Acc][stride]
    int64_t src_cell2;                  // cell (is + 1, js + 1) of the sigma_xz source
};

// neighbour value of a transfer array, 0 outside the grid.  The load is
// unconditional (of the cell's own element when the neighbour is outside),
// so the compiler can issue all of a cell's neighbour loads together instead
// of one guarded load per branch.
template <typename T>
__device__ __forceinline__ T el2_nb(const T *f, const El2Args<T> &A, int64_t c, int i, int j, int di, int dj)
{
    const int ii = i + di, jj = j + dj;
    const bool in = ii >= 1 && ii <= A.nx && jj >= 1 && jj <= A.nz;
    const T v = f[in ? c + di + (int64_t)dj * A.pitch : c];
    return in ? v : T(0);
}

// CPML transpose: forward d' = d ik + psi', psi' = b psi + a d
template <typename T>
__device__ __forceinline__ T cpml_fix_adj(T ld, T &lpsi, T a, T b, T ik)
{
    const T L = lpsi + ld;
    lpsi = b * L;
    return ld * ik + a * L;
}

template <typename T, bool PML>
__device__ __forceinline__ void el2_adj_v_cell(const El2AdjArgs<T> &B, int i, int j, int t, int first)
{
    const El2Args<T> &A = B.F;
    const int nx = A.nx, nz = A.nz;
    const int64_t c = A.xoff + (int64_t)(i - 1) + (int64_t)A.pitch * (j - 1);
    // complete lam(v^{t+1}) with the transposed stress-pass derivatives of step t+1
    T gx = B.lvx[c], gz = B.lvz[c];
    {
        T *const *S = B.sx;
        gx += el2_nb(S[0], A, c, i, j, 1, 0) - T(27) * S[0][c] + T(27) * el2_nb(S[0], A, c, i, j, -1, 0) -
              el2_nb(S[0], A, c, i, j, -2, 0);
        gx += T(27) * S[1][c] - T(27) * el2_nb(S[1], A, c, i, j, 0, 1) - el2_nb(S[1], A, c, i, j, 0, -1) +
              el2_nb(S[1], A, c, i, j, 0, 2);
        gx += T(27) * S[4][c] - T(27) * el2_nb(S[4], A, c, i, j, 1, 0) - el2_nb(S[4], A, c, i, j, -1, 0) +
              el2_nb(S[4], A, c, i, j, 2, 0);
        gx += T(27) * el2_nb(S[5], A, c, i, j, 0, -1) - T(27) * S[5][c] - el2_nb(S[5], A, c, i, j, 0, -2) +
              el2_nb(S[5], A, c, i, j, 0, 1);
        gz += el2_nb(S[3], A, c, i, j, 1, 0) - T(27) * S[3][c] + T(27) * el2_nb(S[3], A, c, i, j, -1, 0) -
              el2_nb(S[3], A, c, i, j, -2, 0);
        gz += T(27) * S[2][c] - T(27) * el2_nb(S[2], A, c, i, j, 0, 1) - el2_nb(S[2], A, c, i, j, 0, -1) +
              el2_nb(S[2], A, c, i, j, 0, 2);
        gz += T(27) * S[7][c] - T(27) * el2_nb(S[7], A, c, i, j, 1, 0) - el2_nb(S[7], A, c, i, j, -1, 0) +
              el2_nb(S[7], A, c, i, j, 2, 0);
        gz += T(27) * el2_nb(S[6], A, c, i, j, 0, -1) - T(27) * S[6][c] - el2_nb(S[6], A, c, i, j, 0, -2) +
              el2_nb(S[6], A, c, i, j, 0, 1);
    }
    if (c == A.src_cell) {
        B.src_lam[kAdjSrcSlots * (int64_t)t + 3] = (double)gx;
        B.src_lam[kAdjSrcSlots * (int64_t)t + 4] = (double)gz;
    }
    T *as = B.adj + (int64_t)(t - first) * kEl2Adj * A.stride + c;
    as[3 * A.stride] = gx;
    as[4 * A.stride] = gz;
    const bool edge = i == 1 || i == nx || j == 1 || j == nz;
    const bool rx = i >= 3 && i <= nx - 1 && j >= 3 && j <= nz - 1;
    const bool rz = i >= 2 && i <= nx - 2 && j >= 2 && j <= nz - 2;
    T P = 0, Q = 0, U = 0, V = 0;
    if (rx || rz) {
        const T r = A.rho[c];
        if (rx) {
            const T rhoxx = scalar_mean(r, A.rho[c - 1], A.density_code);
            const T rhozx = scalar_mean(r, A.rho[c - A.pitch], A.density_code);
            T l1 = A.dt * gx / rhoxx, l2 = A.dt * gx / rhozx;
            if (PML) {
                const T a = ldnt(A.a + c), b = ldnt(A.b + c), ik = ldnt(A.ik + c);
                l1 = cpml_fix_adj(l1, B.lm[8][c], a, b, ik);
                l2 = cpml_fix_adj(l2, B.lm[9][c], a, b, ik);
            }
            P = l1 * A.rdx24;
            Q = l2 * A.rdz24;
        }
        if (rz) {
            const T rhoxz = scalar_mean(r, A.rho[c + 1], A.density_code);
            const T rhozz = scalar_mean(r, A.rho[c + A.pitch], A.density_code);
            T l1 = A.dt * gz / rhoxz, l2 = A.dt * gz / rhozz;
            if (PML) {
                const T ah = ldnt(A.ah + c), bh = ldnt(A.bh + c), ikh = ldnt(A.ikh + c);
                l1 = cpml_fix_adj(l1, B.lm[10][c], ah, bh, ikh);
                l2 = cpml_fix_adj(l2, B.lm[11][c], ah, bh, ikh);
            }
            U = l1 * A.rdx24;
            V = l2 * A.rdz24;
        }
    }
    B.vt[0][c] = P;
    B.vt[1][c] = Q;
    B.vt[2][c] = U;
    B.vt[3][c] = V;
    B.lvx[c] = edge ? T(0) : gx;
    B.lvz[c] = edge ? T(0) : gz;
}

template <typename T, bool PML>
__device__ __forceinline__ void el2_adj_s_cell(const El2AdjArgs<T> &B, int i, int j, int t, int first)
{
    const El2Args<T> &A = B.F;
    const int nx = A.nx, nz = A.nz;
    const int64_t c = A.xoff + (int64_t)(i - 1) + (int64_t)A.pitch * (j - 1);
    const int corr = A.corrected;
    T lxx = B.lxx[c], lzz = B.lzz[c], lxz = B.lxz[c];
    if (c == A.src_cell) {
        B.src_lam[kAdjSrcSlots * (int64_t)t] = (double)lxx;
        B.src_lam[kAdjSrcSlots * (int64_t)t + 2] = (double)lzz;
    }
    if (c == B.src_cell2) B.src_lam[kAdjSrcSlots * (int64_t)t + 1] = (double)lxz;
    // lam(sigma') = lam(sigma^{t+1}) + transposed velocity-pass differences
    {
        T *const *V = B.vt;
        lxx += T(27) * V[0][c] - T(27) * el2_nb(V[0], A, c, i, j, 1, 0) + el2_nb(V[0], A, c, i, j, 2, 0);
        lxz += T(27) * V[1][c] - T(27) * el2_nb(V[1], A, c, i, j, 0, 1) + el2_nb(V[1], A, c, i, j, 0, 2);
        lxz += -T(27) * V[2][c] + T(27) * el2_nb(V[2], A, c, i, j, -1, 0) - el2_nb(V[2], A, c, i, j, -2, 0);
        lzz += -T(27) * V[3][c] + T(27) * el2_nb(V[3], A, c, i, j, 0, -1) - el2_nb(V[3], A, c, i, j, 0, -2);
        if (corr) {
            lxx -= el2_nb(V[0], A, c, i, j, -1, 0);
            lxz -= el2_nb(V[1], A, c, i, j, 0, -1);
            lxz += el2_nb(V[2], A, c, i, j, 1, 0);
            lzz += el2_nb(V[3], A, c, i, j, 0, 1);
        }
    }
    T *as = B.adj + (int64_t)(t - first) * kEl2Adj * A.stride + c;
    as[0] = lxx;
    as[A.stride] = lzz;
    as[2 * A.stride] = lxz;
    const bool r1 = i >= 2 && i <= nx - 2 && j >= 3 && j <= nz - 1;
    const bool r2 = i >= 3 && i <= nx - 1 && j >= 2 && j <= nz - 2;
    T s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0, s6 = 0, s7 = 0;
    if (r1) {
        const T gx = ldnt(A.igx + c), gz = ldnt(A.igz + c);
        const T ax = A.dt * gx * lxx, az = A.dt * gz * lzz;
        T lux = ax * ldnt(A.c11 + c) + az * ldnt(A.c13 + c);
        T lwz = ax * ldnt(A.c13 + c) + az * ldnt(A.c33 + c);
        const T lsh = ax * ldnt(A.c15 + c) + az * ldnt(A.c35 + c);
        T lwx = lsh, luz = lsh;
        if (PML) {
            const T ah = ldnt(A.ah + c), bh = ldnt(A.bh + c), ikh = ldnt(A.ikh + c);
            const T a = ldnt(A.a + c), b = ldnt(A.b + c), ik = ldnt(A.ik + c);
            lux = cpml_fix_adj(lux, B.lm[0][c], ah, bh, ikh);
            luz = cpml_fix_adj(luz, B.lm[1][c], ah, bh, ikh);
            lwz = cpml_fix_adj(lwz, B.lm[2][c], a, b, ik);
            lwx = cpml_fix_adj(lwx, B.lm[3][c], a, b, ik);
        }
        s0 = lux * A.rdx24;   // D+x vx
        s1 = luz * A.rdz24;   // D-z vx
        s2 = lwz * A.rdz24;   // D-z vz
        s3 = lwx * A.rdx24;   // D+x vz
        lxx *= gx;
        lzz *= gz;
    }
    if (r2) {
        const T g = ldnt(A.igxz + c);
        const T al = A.dt * g * lxz;
        T lux = al * ldnt(A.c15 + c), lwz = al * ldnt(A.c35 + c);
        T lwx = al * ldnt(A.c55 + c), luz = lwx;
        if (PML) {
            const T ah = ldnt(A.ah + c), bh = ldnt(A.bh + c), ikh = ldnt(A.ikh + c);
            lux = cpml_fix_adj(lux, B.lm[4][c], ah, bh, ikh);
            luz = cpml_fix_adj(luz, B.lm[5][c], ah, bh, ikh);
            lwz = cpml_fix_adj(lwz, B.lm[6][c], ah, bh, ikh);
            lwx = cpml_fix_adj(lwx, B.lm[7][c], ah, bh, ikh);
        }
        s4 = lux * A.rdx24;   // D-x vx
        s5 = luz * A.rdz24;   // D+z vx
        s6 = lwz * A.rdz24;   // D+z vz
        s7 = lwx * A.rdx24;   // D-x vz
        lxz *= g;
    }
    B.sx[0][c] = s0;
    B.sx[1][c] = s1;
    B.sx[2][c] = s2;
    B.sx[3][c] = s3;
    B.sx[4][c] = s4;
    B.sx[5][c] = s5;
    B.sx[6][c] = s6;
    B.sx[7][c] = s7;
    B.lxx[c] = lxx;
    B.lzz[c] = lzz;
    B.lxz[c] = lxz;
}

template <typename T, int R, int PASS>   // PASS 0: velocity adjoint, 1: stress adjoint
__device__ __forceinline__ void el2_adj_tile(const El2AdjArgs<T> &B, int tx, int band, int t, int first)
{
    const El2Args<T> &A = B.F;
    const int i = tx * 64 + (int)threadIdx.x - (int)A.xoff + 1;
    const bool valid = i >= 1 && i <= A.nx;
    const int j0 = (band * kTile2dWaves + (int)threadIdx.y) * R + 1;
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        if (j > A.nz) break;
        if (__any(valid && el2_pml(A, i, j))) {
            if (valid) {
                if (PASS == 0) el2_adj_v_cell<T, true>(B, i, j, t, first);
                else el2_adj_s_cell<T, true>(B, i, j, t, first);
            }
        } else if (valid) {
            if (PASS == 0) el2_adj_v_cell<T, false>(B, i, j, t, first);
            else el2_adj_s_cell<T, false>(B, i, j, t, first);
        }
    }
}

// per-cell FP64 sums over the stored segment, folded into the gradients
template <typename T, int R>
__device__ __forceinline__ void el2_correlate_tile(const El2AdjArgs<T> &B, int tx, int band, int n)
{
    const El2Args<T> &A = B.F;
    const int i = tx * 64 + (int)threadIdx.x - (int)A.xoff + 1;
    if (i < 1 || i > A.nx) return;
    const int j0 = (band * kTile2dWaves + (int)threadIdx.y) * R + 1;
    const int64_t st = A.stride;
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        if (j > A.nz) break;
        const int64_t c = A.xoff + (int64_t)(i - 1) + (int64_t)A.pitch * (j - 1);
        double S[16] = {};
        for (int s = 0; s < n; ++s) {
            const T *f = A.store + (int64_t)s * kEl2Fwd * st + c;
            const T *l = B.adj + (int64_t)s * kEl2Adj * st + c;
            const double lxx = ldnt(l), lzz = ldnt(l + st), lxz = ldnt(l + 2 * st);
            const double lvx = ldnt(l + 3 * st), lvz = ldnt(l + 4 * st);
            const double a = ldnt(f), b = ldnt(f + st), sh = ldnt(f + 2 * st), pxx = ldnt(f + 3 * st),
                         pzz = ldnt(f + 4 * st);
            const double p = ldnt(f + 5 * st), q = ldnt(f + 6 * st), ru = ldnt(f + 7 * st), pxz = ldnt(f + 8 * st);
            S[0] += lxx * a;
            S[1] += lxx * b;
            S[2] += lxx * sh;
            S[3] += lzz * a;
            S[4] += lzz * b;
            S[5] += lzz * sh;
            S[6] += lxx * pxx;
            S[7] += lzz * pzz;
            S[8] += lxz * p;
            S[9] += lxz * q;
            S[10] += lxz * ru;
            S[11] += lxz * pxz;
            S[12] += lvx * (double)ldnt(f + 9 * st);
            S[13] += lvx * (double)ldnt(f + 10 * st);
            S[14] += lvz * (double)ldnt(f + 11 * st);
            S[15] += lvz * (double)ldnt(f + 12 * st);
        }
        const double dt = (double)A.dt, gx = (double)A.igx[c], gz = (double)A.igz[c], g = (double)A.igxz[c];
        double *G = B.grad + c;
        G[0] += dt * gx * S[0];
        G[st] += dt * (gx * S[1] + gz * S[3]);
        G[2 * st] += dt * (gx * S[2] + g * S[8]);
        G[3 * st] += dt * gz * S[4];
        G[4 * st] += dt * (gz * S[5] + g * S[9]);
        G[5 * st] += dt * g * S[10];
        G[7 * st] += S[6];
        G[8 * st] += S[7];
        G[9 * st] += S[11];
        for (int q = 0; q < 4; ++q) G[(10 + q) * st] += S[12 + q];
    }
}

// d/drho from the density sums: rho enters the face averages of vx
// (rho(i-1), rho(i); rho(j-1), rho(j)) and vz (rho(i), rho(i+1); rho(j),
// rho(j+1)) as d1 / mean, so d v'/d mean = -dt d1 / mean^2
__device__ __forceinline__ double mean_d(double a, double b, int code, bool first)
{
    switch (code) {
    case 1: { const double s = a + b, o = first ? b : a; return 2.0 * o * o / (s * s); }
    case 2: return 0.5 * sqrt(first ? b / a : a / b);
    case 3: return 0.5;
    default: return first ? 1.0 : 0.0;
    }
}

template <typename T>
__device__ __forceinline__ double mean_v(double a, double b, int code)
{
    switch (code) {
    case 1: return 2.0 * a * b / (a + b);
    case 2: return sqrt(a * b);
    case 3: return 0.5 * (a + b);
    default: return a;
    }
}

template <typename T, int R>
__device__ __forceinline__ void el2_rho_grad_tile(const El2AdjArgs<T> &B, int tx, int band)
{
    const El2Args<T> &A = B.F;
    const int i = tx * 64 + (int)threadIdx.x - (int)A.xoff + 1;
    if (i < 1 || i > A.nx) return;
    const int j0 = (band * kTile2dWaves + (int)threadIdx.y) * R + 1;
    const int64_t st = A.stride, pz = A.pitch;
    const int code = A.density_code;
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        if (j > A.nz) break;
        const int64_t c = A.xoff + (int64_t)(i - 1) + pz * (j - 1);
        const double *X = B.grad + 10 * st;
        const double rc = (double)A.rho[c];
        double g = 0.0;
        auto term = [&](const double *x, int64_t cc, int ii, int jj, int64_t other, bool first) {
            if (ii < 1 || ii > A.nx || jj < 1 || jj > A.nz) return;
            const double sum = x[cc];
            if (sum == 0.0) return;
            const double a = (double)A.rho[cc], b = (double)A.rho[other];
            const double m = mean_v<T>(a, b, code);
            g -= sum / (m * m) * mean_d(a, b, code, first);
        };
        // vx at c uses (rho(c), rho(c-1)) and (rho(c), rho(c-pz)); vx at c+1 and c+pz use rho(c) second
        if (i >= 2) term(X, c, i, j, c - 1, true);
        if (i + 1 <= A.nx) term(X, c + 1, i + 1, j, c, false);
        if (j >= 2) term(X + st, c, i, j, c - pz, true);
        if (j + 1 <= A.nz) term(X + st, c + pz, i, j + 1, c, false);
        // vz at c uses (rho(c), rho(c+1)) and (rho(c), rho(c+pz)); vz at c-1 and c-pz use rho(c) second
        if (i + 1 <= A.nx) term(X + 2 * st, c, i, j, c + 1, true);
        if (i >= 2) term(X + 2 * st, c - 1, i - 1, j, c, false);
        if (j + 1 <= A.nz) term(X + 3 * st, c, i, j, c + pz, true);
        if (j >= 2) term(X + 3 * st, c - pz, i, j - 1, c, false);
        (void)rc;
        B.grad[6 * st + c] += (double)A.dt * g;
    }
}

template <typename T, int R, int PASS>
__global__ void __launch_bounds__(64 * kTile2dWaves) el2_adj_kernel(El2AdjArgs<T> B, int t, int first)
{
    el2_adj_tile<T, R, PASS>(B, blockIdx.x, blockIdx.y, t, first);
}

template <typename T, int R>
__global__ void __launch_bounds__(64 * kTile2dWaves) el2_correlate_kernel(El2AdjArgs<T> B, int n)
{
    el2_correlate_tile<T, R>(B, blockIdx.x, blockIdx.y, n);
}

template <typename T, int R>
__global__ void __launch_bounds__(64 * kTile2dWaves) el2_rho_grad_kernel(El2AdjArgs<T> B)
{
    el2_rho_grad_tile<T, R>(B, blockIdx.x, blockIdx.y);
}

template <typename T, int R, int PASS>
__global__ void __launch_bounds__(64 * kTile2dWaves) el2_adj_batch_kernel(const El2AdjArgs<T> *args,
                                                                          const int4 *tiles, int t, int first)
{
    const int4 q = tiles[blockIdx.x];
    if (q.x < 0) return;
    el2_adj_tile<T, R, PASS>(args[q.x], q.y, q.z, t, first);
}

template <typename T, int R>
__global__ void __launch_bounds__(64 * kTile2dWaves) el2_correlate_batch_kernel(const El2AdjArgs<T> *args,
                                                                                const int4 *tiles, int n)
{
    const int4 q = tiles[blockIdx.x];
    if (q.x < 0) return;
    el2_correlate_tile<T, R>(args[q.x], q.y, q.z, n);
}

template <typename T, int R>
__global__ void __launch_bounds__(64 * kTile2dWaves) el2_rho_grad_batch_kernel(const El2AdjArgs<T> *args,
                                                                               const int4 *tiles)
{
    const int4 q = tiles[blockIdx.x];
    if (q.x < 0) return;
    el2_rho_grad_tile<T, R>(args[q.x], q.y, q.z);
}

template <typename T>
static El2AdjArgs<T> el2_adj_args(Device &d)
{
    El2AdjArgs<T> B;
    B.F = el2_args<T>(d);
    T **l = (T **)d.adj.lam;
    B.lvx = l[0];
    B.lvz = l[1];
    B.lxx = l[2];
    B.lzz = l[3];
    B.lxz = l[4];
    for (int q = 0; q < 12; ++q) B.lm[q] = l[5 + q];
    for (int q = 0; q < 8; ++q) B.sx[q] = l[17 + q];
    for (int q = 0; q < 4; ++q) B.vt[q] = l[25 + q];
    B.adj = (T *)d.adj.seg_adj;
    B.src_lam = d.adj.src_lam;
    B.grad = d.adj.grad;
    B.src_cell2 = cell_index(d, d.isrc + 1, 1, d.ksrc + 1);
    return B;
}

int el2_adj_setup(Device &d)
{
    d.adj.nlam = 29;
    for (int q = 0; q < d.adj.nlam; ++q)
        if (!d.adj.lam[q] && alloc_grid_zeroed(d, &d.adj.lam[q]) != SDT_OK) return SDT_ERR_HIP;
    return SDT_OK;
}

template <typename T>
static void el2_adj_step_t(Device &d, int t)
{
    constexpr int R = kSingleRows2d;
    const El2AdjArgs<T> B = el2_adj_args<T>(d);
    void *lam_field[3] = {B.lvx, nullptr, B.lvz};
    launch_adj_tail(d, t, lam_field);
    const dim3 grid(tiles_x_2d(d), tiles_z_2d(d, R)), block(64, kTile2dWaves);
    hipLaunchKernelGGL((el2_adj_kernel<T, R, 0>), grid, block, 0, d.stream, B, t, d.adj.store_first);
    hipLaunchKernelGGL((el2_adj_kernel<T, R, 1>), grid, block, 0, d.stream, B, t, d.adj.store_first);
}

void el2_adj_step(Device &d, int t)
{
    if (d.precision == SDT_FP64) el2_adj_step_t<double>(d, t);
    else el2_adj_step_t<float>(d, t);
}

template <typename T>
static void el2_adj_correlate_t(Device &d, int n)
{
    constexpr int R = kSingleRows2d;
    const El2AdjArgs<T> B = el2_adj_args<T>(d);
    const dim3 grid(tiles_x_2d(d), tiles_z_2d(d, R)), block(64, kTile2dWaves);
    hipLaunchKernelGGL((el2_correlate_kernel<T, R>), grid, block, 0, d.stream, B, n);
}

void el2_adj_correlate(Device &d, int first, int n)
{
    (void)first;
    if (d.precision == SDT_FP64) el2_adj_correlate_t<double>(d, n);
    else el2_adj_correlate_t<float>(d, n);
}

template <typename T>
static void el2_adj_finish_t(Device &d)
{
    constexpr int R = kSingleRows2d;
    const El2AdjArgs<T> B = el2_adj_args<T>(d);
    const dim3 grid(tiles_x_2d(d), tiles_z_2d(d, R)), block(64, kTile2dWaves);
    hipLaunchKernelGGL((el2_rho_grad_kernel<T, R>), grid, block, 0, d.stream, B);
}

void el2_adj_finish(Device &d)
{
    if (d.precision == SDT_FP64) el2_adj_finish_t<double>(d);
    else el2_adj_finish_t<float>(d);
}

template <typename T>
static int el2_batch_adj_args_t(Batch &b, Device *const *shots, int n)
{
    std::vector<El2AdjArgs<T>> a((size_t)n);
    std::vector<AdjTailArgs> tails((size_t)n);
    for (int s = 0; s < n; ++s) {
        a[(size_t)s] = el2_adj_args<T>(*shots[s]);
        AdjTailArgs &t = tails[(size_t)s];
        t.lam[0] = a[(size_t)s].lvx;
        t.lam[1] = nullptr;
        t.lam[2] = a[(size_t)s].lvz;
        t.rec_cell = shots[s]->rec_cell;
        t.tg = shots[s]->adj.trace_grad;
        t.nrec = shots[s]->nrec;
    }
    if (!b.adj_args && hipMalloc(&b.adj_args, a.size() * sizeof(El2AdjArgs<T>)) != hipSuccess) return SDT_ERR_HIP;
    if (!b.adj_tails && hipMalloc((void **)&b.adj_tails, tails.size() * sizeof(AdjTailArgs)) != hipSuccess)
        return SDT_ERR_HIP;
    if (hipMemcpy(b.adj_args, a.data(), a.size() * sizeof(El2AdjArgs<T>), hipMemcpyHostToDevice) != hipSuccess)
        return SDT_ERR_HIP;
    if (hipMemcpy(b.adj_tails, tails.data(), tails.size() * sizeof(AdjTailArgs), hipMemcpyHostToDevice) != hipSuccess)
        return SDT_ERR_HIP;
    return SDT_OK;
}

int el2_batch_adj_args(Batch &b, Device *const *shots, int n)
{
    return b.precision == SDT_FP64 ? el2_batch_adj_args_t<double>(b, shots, n)
                                   : el2_batch_adj_args_t<float>(b, shots, n);
}

// what: 0 adjoint step, 1 correlation, 2 density gradient
template <typename T, int R>
static void el2_batch_adj_launch(Batch &b, int which, int what, int t, int first, int n)
{
    const El2AdjArgs<T> *a = (const El2AdjArgs<T> *)b.adj_args;
    const dim3 grid(b.ntiles[which]), block(64, kTile2dWaves);
    if (what == 0) {
        hipLaunchKernelGGL((el2_adj_batch_kernel<T, R, 0>), grid, block, 0, b.stream, a, b.tiles[which], t, first);
        hipLaunchKernelGGL((el2_adj_batch_kernel<T, R, 1>), grid, block, 0, b.stream, a, b.tiles[which], t, first);
    } else if (what == 1) {
        hipLaunchKernelGGL((el2_correlate_batch_kernel<T, R>), grid, block, 0, b.stream, a, b.tiles[which], n);
    } else {
        hipLaunchKernelGGL((el2_rho_grad_batch_kernel<T, R>), grid, block, 0, b.stream, a, b.tiles[which]);
    }
}

template <typename T>
static void el2_batch_adj_p(Batch &b, int which, int what, int t, int first, int n)
{
    switch (kBatchShapes[which].rows) {
    case 1: el2_batch_adj_launch<T, 1>(b, which, what, t, first, n); break;
    case 8: el2_batch_adj_launch<T, 8>(b, which, what, t, first, n); break;
    default: el2_batch_adj_launch<T, 4>(b, which, what, t, first, n); break;
    }
}

void el2_batch_adj_step(Batch &b, int which, int t, int first)
{
    if (b.precision == SDT_FP64) el2_batch_adj_p<double>(b, which, 0, t, first, 0);
    else el2_batch_adj_p<float>(b, which, 0, t, first, 0);
}

void el2_batch_adj_correlate(Batch &b, int which, int n)
{
    if (b.precision == SDT_FP64) el2_batch_adj_p<double>(b, which, 1, 0, 0, n);
    else el2_batch_adj_p<float>(b, which, 1, 0, 0, n);
}

void el2_batch_adj_finish(Batch &b, int which)
{
    if (b.precision == SDT_FP64) el2_batch_adj_p<double>(b, which, 2, 0, 0, 0);
    else el2_batch_adj_p<float>(b, which, 2, 0, 0, 0);
}

void el2_step(Device &d, bool timed)
{
    if (d.precision == SDT_FP64) el2_step_t<double>(d, timed);
    else el2_step_t<float>(d, timed);
}

// Material layout on the device: c11 c13 c15 c33 c35 c55 rho and the three
// attenuation divisors precomputed as 1/(1 + gamma*dt) in FP64.  Without an
// adjoint and with at most 256 distinct (stiffness, divisor) tuples, those
// nine come from a material table (one byte of id per cell) and only the
// density is a per-cell array (slot 6).
int el2_setup(Device &d, const double *const *in)
{
    d.tile_x = 64;
    d.tile_y = 1;
    d.tile_z = 4;
    int rc = compute_tile_flags(d, in);
    if (rc) return rc;
    set_layout(d);
    const int64_t n = d.ncell;
    const int ids[] = {SDT_C11, SDT_C13, SDT_C15, SDT_C33, SDT_C35, SDT_C55, SDT_RHO};
    d.nmat = 0;
    std::vector<double> zero((size_t)n, 0.0);
    if (d.coef_in) {   // the ten coefficients given directly
        for (int q = 0; q < 10; ++q) {
            rc = upload_grid(d, d.coef_in + (size_t)q * n, &d.mat[d.nmat++]);
            if (rc) return rc;
        }
    } else {
        const int gids[] = {SDT_GAMMA_X, SDT_GAMMA_Z, SDT_GAMMA_XZ};
        auto divisor = [&](int q, int64_t c) { return 1.0 / (1.0 + (in[gids[q]] ? in[gids[q]][c] : 0.0) * d.dt); };
        if (d.mat_table_opt && d.adj.every == 0) {
            static const int voigt[6] = {0, 2, 4, 11, 13, 18};
            rc = build_material_table(d, kEl2TabRow, [&](int64_t c, double *r) {
                for (int q = 0; q < 6; ++q) r[voigt[q]] = in[ids[q]] ? in[ids[q]][c] : 0.0;
                for (int q = 0; q < 3; ++q) r[21 + q] = divisor(q, c);
            });
            if (rc) return rc;
        }
        for (int q = 0; q < 7; ++q, ++d.nmat) {
            if (d.mat_table && q != 6) continue;   // only the density beside the table
            rc = upload_grid(d, in[ids[q]] ? in[ids[q]] : zero.data(), &d.mat[d.nmat]);
            if (rc) return rc;
        }
        std::vector<double> inv((size_t)n);
        for (int q = 0; q < 3; ++q, ++d.nmat) {
            if (d.mat_table) continue;
            for (int64_t c = 0; c < n; ++c) inv[(size_t)c] = divisor(q, c);
            rc = upload_grid(d, inv.data(), &d.mat[d.nmat]);
            if (rc) return rc;
        }
    }
    d.nmem = 12;
    for (int q = 0; q < d.nmem; ++q) {
        rc = alloc_grid_zeroed(d, &d.mem[q]);
        if (rc) return rc;
    }
    const int fids[] = {SDT_F_VX, SDT_F_VZ, SDT_F_SXX, SDT_F_SZZ, SDT_F_SXZ};
    for (int f : fids) {
        if (f == SDT_F_VX && in[SDT_IC_X]) rc = upload_grid(d, in[SDT_IC_X], &d.field[f]);
        else if (f == SDT_F_VZ && in[SDT_IC_Z]) rc = upload_grid(d, in[SDT_IC_Z], &d.field[f]);
        else rc = alloc_grid_zeroed(d, &d.field[f]);
        if (rc) return rc;
    }

    // point sources, cpmlfdtd.f08:519-523; indices (is, js) are padded 1-based
    const int is = d.isrc, js = d.ksrc, nx = d.nx, nz = d.nz;
    auto cell = [&](int i, int j) { return cell_index(d, i, 1, j); };
    const double *rho = in[SDT_RHO];
    auto rho_at = [&](int i, int j) { return rho ? rho[(int64_t)(i - 1) + (int64_t)nx * (j - 1)] : 0.0; };
    const int face = (is == 1 || is == nx || js == 1 || js == nz);
    TailArgs &t = d.tail;
    t.nterms = 0;
    t.terms[t.nterms++] = {cell(is, js), SDT_F_SXX, SDT_SRC_XX, 1.0, rho_at(is, js), 0};
    t.terms[t.nterms++] = {cell(is + 1, js + 1), SDT_F_SXZ, SDT_SRC_XZ, 1.0, rho_at(is + 1, js + 1), 0};
    t.terms[t.nterms++] = {cell(is, js), SDT_F_SZZ, SDT_SRC_ZZ, 1.0, rho_at(is, js), 0};
    t.terms[t.nterms++] = {cell(is, js), SDT_F_VX, SDT_SRC_X, 1.0, rho_at(is, js), face};
    t.terms[t.nterms++] = {cell(is, js), SDT_F_VZ, SDT_SRC_Z, 1.0, rho_at(is, js), face};
    t.nvfix = 1;
    t.vfix[0] = cell(is, js);
    t.nvcomp = 2;
    t.vcomp[0] = SDT_F_VX;
    t.vcomp[1] = SDT_F_VZ;
    t.stress_norm = 0;
    t.nsfix = 0;
    t.rec_field[0] = SDT_F_VX;
    t.rec_field[1] = -1;
    t.rec_field[2] = SDT_F_VZ;

    // adjoint layout: checkpoint = fields + memory variables; 10 coefficients
    // (+ 4 density sums); per step kEl2Fwd forward and kEl2Adj adjoint values
    Adjoint &a = d.adj;
    a.nstate = 0;
    for (int f : fids) a.state[a.nstate++] = d.field[f];
    for (int q = 0; q < d.nmem; ++q) a.state[a.nstate++] = d.mem[q];
    a.ncoef = 10;
    a.ngrad = kEl2Acc;
    a.nfwd = kEl2Fwd;
    a.nadj = kEl2Adj;
    return upload_cpml_full(d, in);
}

}  // namespace sdt
