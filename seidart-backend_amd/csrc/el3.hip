// el3.hip -- 3D anisotropic elastic CPML-FDTD (seismic3,
// cpmlfdtd.f08:1341-2313) on gfx950.
//
// The reference's four stress nests (L1 sigma_xx/yy/zz, L2 sigma_xy, L3
// sigma_xz, L4 sigma_yz; cpmlfdtd.f08:1761-1992) run as one fused stress
// kernel; its three velocity nests (2014-2127) as one fused velocity kernel
// that also applies the Dirichlet faces (2226-2248) and the velocity norm
// (2255).  Each cell evaluates only the nests whose index range contains it.
//
// Memory variables are deduplicated: a recursion psi <- b psi + a d depends
// only on its own cell's history of d, so the reference arrays that share
// (velocity component, direction, one-sided stencil, coefficient set) hold
// identical values wherever they are updated, and one array updated on the
// union of their ranges is exact.  Stress side, per component: x {D+ half,
// D- full}, y {D- full, D+ half}, z {D- full, D- half, D+ half} = 21 arrays;
// vx's (z, D-, half) recursion (memory_dvx_dz2, 1855) is written but never
// read by the reference (1858-1865) and is not kept.  Velocity side: 9.
#include <math.h>

#include <vector>

#include "el3_common.h"

namespace sdt {

// Velocity values of one column at planes k-2..k+2 (zq[comp][0..4]); the
// x and y neighbours are read from memory (L1/L2 hits within a workgroup).
template <typename T, bool PML>
__device__ __forceinline__ double el3_stress_cell(const El3Args<T> &A, int i, int j, int k, const T (&zq)[3][5])
{
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const int64_t sx = 1, sy = nx, sz = (int64_t)nx * ny;
    const int64_t c = (int64_t)(i - 1) + sy * (j - 1) + sz * (k - 1);
    const bool L1 = i >= 2 && i <= nx - 2 && j >= 3 && j <= ny - 1 && k >= 3 && k <= nz - 1;
    const bool L2 = i >= 3 && i <= nx - 1 && j >= 2 && j <= ny - 2 && k >= 3 && k <= nz - 1;
    const bool L3 = i >= 3 && i <= nx - 1 && j >= 3 && j <= ny - 1 && k >= 2 && k <= nz - 2;
    const bool L4 = i >= 2 && i <= nx - 2 && j >= 2 && j <= ny - 2 && k >= 2 && k <= nz - 2;
    T d[3][K_NKIND];
    T a = 0, b = 0, ik = 1, ah = 0, bh = 0, ikh = 1;
    if (PML && (L1 || L2 || L3 || L4)) {
        a = ldnt(&A.a[c]); b = ldnt(&A.b[c]); ik = ldnt(&A.ik[c]);
        ah = ldnt(&A.ah[c]); bh = ldnt(&A.bh[c]); ikh = ldnt(&A.ikh[c]);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const T *v = A.v[q];
        T *const *m = A.m + q * K_NKIND;
        const T v0 = zq[q][2];
        if (L1 || L4) {   // D+x, half set
            T x = (T(27) * (v[c + sx] - v0) - (v[c + 2 * sx] - v[c - sx])) * A.rd24[0];
            if (PML) x = cpml_fix3(x, &m[K_XPH][c], ah, bh, ikh);
            d[q][K_XPH] = x;
        }
        if (L2 || L3) {   // D-x, full set
            T x = (T(27) * (v0 - v[c - sx]) - (v[c + sx] - v[c - 2 * sx])) * A.rd24[0];
            if (PML) x = cpml_fix3(x, &m[K_XMF][c], a, b, ik);
            d[q][K_XMF] = x;
        }
        if (L1 || L3) {   // D-y, full set
            T x = (T(27) * (v0 - v[c - sy]) - (v[c + sy] - v[c - 2 * sy])) * A.rd24[1];
            if (PML) x = cpml_fix3(x, &m[K_YMF][c], a, b, ik);
            d[q][K_YMF] = x;
        }
        if (L2 || L4) {   // D+y, half set
            T x = (T(27) * (v[c + sy] - v0) - (v[c + 2 * sy] - v[c - sy])) * A.rd24[1];
            if (PML) x = cpml_fix3(x, &m[K_YPH][c], ah, bh, ikh);
            d[q][K_YPH] = x;
        }
        const T zm = (T(27) * (zq[q][2] - zq[q][1]) - (zq[q][3] - zq[q][0])) * A.rd24[2];
        if (L1) {         // D-z, full set
            T x = zm;
            if (PML) x = cpml_fix3(x, &m[K_ZMF][c], a, b, ik);
            d[q][K_ZMF] = x;
        }
        if (L2) {         // D-z, half set; vx's enters sigma_xy uncorrected (1858-1865)
            T x = zm;
            if (PML && q != 0) x = cpml_fix3(x, &m[K_ZMH][c], ah, bh, ikh);
            d[q][K_ZMH] = x;
        }
        if (L3 || L4) {   // D+z, half set
            T x = (T(27) * (zq[q][3] - zq[q][2]) - (zq[q][4] - zq[q][1])) * A.rd24[2];
            if (PML) x = cpml_fix3(x, &m[K_ZPH][c], ah, bh, ikh);
            d[q][K_ZPH] = x;
        }
    }
    const T dt = A.dt;
    const T *const *C = A.C;
    T sxx = ldnt(&A.s[0][c]), syy = ldnt(&A.s[1][c]), szz = ldnt(&A.s[2][c]), sxy = ldnt(&A.s[3][c]),
      sxz = ldnt(&A.s[4][c]), syz = ldnt(&A.s[5][c]);
    if (L1) {
        const T dxx = d[0][K_XPH], dyy = d[1][K_YMF], dzz = d[2][K_ZMF];
        const T yz = d[1][K_ZMF] + d[2][K_YMF], xz = d[0][K_ZMF] + d[2][K_XPH], xy = d[1][K_XPH] + d[0][K_YMF];
        sxx = stress_update(sxx, C, c, 0, 1, 2, 3, 4, 5, dxx, dyy, dzz, yz, xz, xy, dt, ldnt(&A.ig[0][c]));
        syy = stress_update(syy, C, c, 1, 6, 7, 8, 9, 10, dxx, dyy, dzz, yz, xz, xy, dt, ldnt(&A.ig[1][c]));
        szz = stress_update(szz, C, c, 2, 7, 11, 12, 13, 14, dxx, dyy, dzz, yz, xz, xy, dt, ldnt(&A.ig[2][c]));
        stnt(&A.s[0][c], sxx); stnt(&A.s[1][c], syy); stnt(&A.s[2][c], szz);
    }
    if (L2) {
        const T dxx = d[0][K_XMF], dyy = d[1][K_YPH], dzz = d[2][K_ZMH];
        const T yz = d[1][K_ZMH] + d[2][K_YPH], xz = d[0][K_ZMH] + d[2][K_XMF], xy = d[1][K_XMF] + d[0][K_YPH];
        sxy = stress_update(sxy, C, c, 5, 10, 14, 17, 19, 20, dxx, dyy, dzz, yz, xz, xy, dt, ldnt(&A.ig[3][c]));
        stnt(&A.s[3][c], sxy);
    }
    if (L3) {
        const T dxx = d[0][K_XMF], dyy = d[1][K_YMF], dzz = d[2][K_ZPH];
        const T yz = d[1][K_ZPH] + d[2][K_YMF], xz = d[0][K_ZPH] + d[2][K_XMF], xy = d[1][K_XMF] + d[0][K_YMF];
        sxz = stress_update(sxz, C, c, 4, 9, 13, 16, 18, 19, dxx, dyy, dzz, yz, xz, xy, dt, ldnt(&A.ig[4][c]));
        stnt(&A.s[4][c], sxz);
    }
    if (L4) {
        const T dxx = d[0][K_XPH], dyy = d[1][K_YPH], dzz = d[2][K_ZPH];
        const T yz = d[1][K_ZPH] + d[2][K_YPH], xz = d[0][K_ZPH] + d[2][K_XPH], xy = d[1][K_XPH] + d[0][K_YPH];
        syz = stress_update(syz, C, c, 3, 8, 12, 15, 16, 17, dxx, dyy, dzz, yz, xz, xy, dt, ldnt(&A.ig[5][c]));
        stnt(&A.s[5][c], syz);
    }
    if (c == A.sfix[0] || c == A.sfix[1] || c == A.sfix[2] || c == A.sfix[3]) return 0.0;
    const double a0 = sxx, a1 = syy, a2 = szz, a3 = sxy, a4 = syz, a5 = sxz;
    return sqrt(a0 * a0 + a1 * a1 + a2 * a2 + 2.0 * (a3 * a3 + a4 * a4 + a5 * a5));
}

template <typename T>
__device__ __forceinline__ T zload(const T *f, int64_t col, int64_t sz, int k, int nz)
{
    return (k >= 1 && k <= nz) ? f[col + sz * (k - 1)] : T(0);
}

// One thread per cell (64 x 4 x 1 workgroups): every z neighbour read from memory.
template <typename T>
__global__ void __launch_bounds__(256) el3_stress_kernel(El3Args<T> A)
{
    const int i = blockIdx.x * 64 + threadIdx.x + 1;
    const int j = blockIdx.y * 4 + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    double nrm = 0.0;
    if (i <= A.nx && j <= A.ny) {
        const int64_t sz = (int64_t)A.nx * A.ny, col = (int64_t)(i - 1) + (int64_t)A.nx * (j - 1);
        T zq[3][5];
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int s = 0; s < 5; ++s) zq[q][s] = zload(A.v[q], col, sz, k - 2 + s, A.nz);
        const size_t t = blockIdx.x + (size_t)A.tiles_x * (blockIdx.y + (size_t)A.tiles_y * blockIdx.z);
        if (!in_box(A.box, i, j, k) || A.flags[t]) nrm = el3_stress_cell<T, true>(A, i, j, k, zq);
        else nrm = el3_stress_cell<T, false>(A, i, j, k, zq);
    }
    block_max_to(&A.snorm[*A.d_it - 1], nrm);
}

// zs[0..2] = sigma_xz, sigma_yz, sigma_zz at planes k-2..k+2 of the column;
// zr = rho at k-1, k, k+1.
template <typename T, bool PML>
__device__ __forceinline__ double el3_velocity_cell(const El3Args<T> &A, int i, int j, int k, const T (&zs)[3][5],
                                                    const T (&zr)[3])
{
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const int64_t sx = 1, sy = nx, sz = (int64_t)nx * ny;
    const int64_t c = (int64_t)(i - 1) + sy * (j - 1) + sz * (k - 1);
    T vx = ldnt(&A.v[0][c]), vy = ldnt(&A.v[1][c]), vz = ldnt(&A.v[2][c]);
    if (i == 1 || i == nx || j == 1 || j == ny || k == 1 || k == nz) {
        vx = vy = vz = T(0);
        stnt(&A.v[0][c], vx); stnt(&A.v[1][c], vy); stnt(&A.v[2][c], vz);
    } else {
        const int dc = A.density_code, corr = A.corrected;
        const T r = zr[1];
        const T ryy = scalar_mean(r, r, dc);
        T a = 0, b = 0, ik = 1, ah = 0, bh = 0, ikh = 1;
        if (PML) {
            a = ldnt(&A.a[c]); b = ldnt(&A.b[c]); ik = ldnt(&A.ik[c]);
            ah = ldnt(&A.ah[c]); bh = ldnt(&A.bh[c]); ikh = ldnt(&A.ikh[c]);
        }
        const T *sxx = A.s[0], *syy = A.s[1], *sxy = A.s[3], *sxz = A.s[4], *syz = A.s[5];
        const T sxz0 = zs[0][2], syz0 = zs[1][2], szz0 = zs[2][2];
        if (i >= 3 && i <= nx - 1 && j >= 3 && j <= ny - 1 && k >= 3 && k <= nz - 1) {
            // cpmlfdtd.f08:2014-2038
            const T rx = scalar_mean(r, A.rho[c - sx], dc), rz = scalar_mean(r, zr[0], dc);
            const T sxy0 = sxy[c];
            T d1 = tminus5(sxx[c - 2 * sx], sxx[c - sx], sxx[c], corr ? sxx[c + sx] : T(0), corr) * A.rd24[0];
            T d2 = tminus5(sxy[c - 2 * sy], sxy[c - sy], sxy0, corr ? sxy[c + sy] : T(0), corr) * A.rd24[1];
            T d3 = tminus5(zs[0][0], zs[0][1], sxz0, zs[0][3], corr) * A.rd24[2];
            if (PML) {
                d1 = cpml_fix3(d1, &A.m[M_SXX_X][c], ah, bh, ikh);
                d2 = cpml_fix3(d2, &A.m[M_SXY_Y][c], ah, bh, ikh);
                d3 = cpml_fix3(d3, &A.m[M_SXZ_Z][c], ah, bh, ikh);
            }
            vx = vx + (d1 / rx + d2 / ryy + d3 / rz) * A.dt;
            stnt(&A.v[0][c], vx);
        }
        if (i >= 2 && i <= nx - 2 && j >= 2 && j <= ny - 2 && k >= 3 && k <= nz - 1) {
            // cpmlfdtd.f08:2055-2080
            const T rx = scalar_mean(r, A.rho[c + sx], dc), rz = scalar_mean(r, zr[0], dc);
            T d1 = tplus5(corr ? sxy[c - sx] : T(0), sxy[c], sxy[c + sx], sxy[c + 2 * sx], corr) * A.rd24[0];
            T d2 = tplus5(corr ? syy[c - sy] : T(0), syy[c], syy[c + sy], syy[c + 2 * sy], corr) * A.rd24[1];
            T d3 = tminus5(zs[1][0], zs[1][1], syz0, zs[1][3], corr) * A.rd24[2];
            if (PML) {
                d1 = cpml_fix3(d1, &A.m[M_SXY_X][c], ah, bh, ikh);
                d2 = cpml_fix3(d2, &A.m[M_SYY_Y][c], ah, bh, ikh);
                d3 = cpml_fix3(d3, &A.m[M_SYZ_Z][c], a, b, ik);
            }
            vy = vy + (d1 / rx + d2 / ryy + d3 / rz) * A.dt;
            stnt(&A.v[1][c], vy);
        }
        if (i >= 2 && i <= nx - 2 && j >= 3 && j <= ny - 1 && k >= 2 && k <= nz - 2) {
            // cpmlfdtd.f08:2098-2127
            const T rx = scalar_mean(r, A.rho[c + sx], dc), rz = scalar_mean(r, zr[2], dc);
            T d1 = tplus5(corr ? sxz[c - sx] : T(0), sxz0, sxz[c + sx], sxz[c + 2 * sx], corr) * A.rd24[0];
            T d2 = tminus5(syz[c - 2 * sy], syz[c - sy], syz0, corr ? syz[c + sy] : T(0), corr) * A.rd24[1];
            T d3 = tplus5(zs[2][1], szz0, zs[2][3], zs[2][4], corr) * A.rd24[2];
            if (PML) {
                d1 = cpml_fix3(d1, &A.m[M_SXZ_X][c], ah, bh, ikh);
                d2 = cpml_fix3(d2, &A.m[M_SYZ_Y][c], a, b, ik);
                d3 = cpml_fix3(d3, &A.m[M_SZZ_Z][c], ah, bh, ikh);
            }
            vz = vz + (d1 / rx + d2 / ryy + d3 / rz) * A.dt;
            stnt(&A.v[2][c], vz);
        }
    }
    if (c == A.vfix[0] || c == A.vfix[1] || c == A.vfix[2]) return 0.0;
    const double a0 = vx, a1 = vy, a2 = vz;
    return sqrt(a0 * a0 + a1 * a1 + a2 * a2);
}

template <typename T>
__global__ void __launch_bounds__(256) el3_velocity_kernel(El3Args<T> A)
{
    const int i = blockIdx.x * 64 + threadIdx.x + 1;
    const int j = blockIdx.y * 4 + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    double nrm = 0.0;
    if (i <= A.nx && j <= A.ny) {
        const int64_t sz = (int64_t)A.nx * A.ny, col = (int64_t)(i - 1) + (int64_t)A.nx * (j - 1);
        T zs[3][5], zr[3];
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            zs[0][s] = zload(A.s[4], col, sz, k - 2 + s, A.nz);
            zs[1][s] = zload(A.s[5], col, sz, k - 2 + s, A.nz);
            zs[2][s] = zload(A.s[2], col, sz, k - 2 + s, A.nz);
        }
#pragma unroll
        for (int s = 0; s < 3; ++s) zr[s] = zload(A.rho, col, sz, k - 1 + s, A.nz);
        const size_t t = blockIdx.x + (size_t)A.tiles_x * (blockIdx.y + (size_t)A.tiles_y * blockIdx.z);
        if (!in_box(A.box, i, j, k) || A.flags[t]) nrm = el3_velocity_cell<T, true>(A, i, j, k, zs, zr);
        else nrm = el3_velocity_cell<T, false>(A, i, j, k, zs, zr);
    }
    block_max_to(&A.vnorm[*A.d_it - 1], nrm);
}

template <typename T>
static El3Args<T> el3_args(Device &d)
{
    El3Args<T> A;
    for (int q = 0; q < 3; ++q) A.v[q] = (T *)d.field[SDT_F_VX + q];
    for (int q = 0; q < 6; ++q) A.s[q] = (T *)d.field[SDT_F_SXX + q];
    for (int q = 0; q < 21; ++q) A.C[q] = (const T *)d.mat[q];
    A.rho = (const T *)d.mat[21];
    for (int q = 0; q < 6; ++q) A.ig[q] = (const T *)d.mat[22 + q];
    A.a = (const T *)d.cpml[0];
    A.b = (const T *)d.cpml[1];
    A.ik = (const T *)d.cpml[2];
    A.ah = (const T *)d.cpml[3];
    A.bh = (const T *)d.cpml[4];
    A.ikh = (const T *)d.cpml[5];
    for (int q = 0; q < M_TOTAL; ++q) A.m[q] = (T *)d.mem[q];
    A.flags = d.tile_flags;
    for (int q = 0; q < 6; ++q) A.box[q] = d.box[q];
    A.nx = d.nx; A.ny = d.ny; A.nz = d.nz;
    A.tiles_x = d.tiles_x; A.tiles_y = d.tiles_y;
    A.rd24[0] = (T)(1.0 / (24.0 * d.dx));
    A.rd24[1] = (T)(1.0 / (24.0 * d.dy));
    A.rd24[2] = (T)(1.0 / (24.0 * d.dz));
    A.dt = (T)d.dt;
    A.density_code = d.density_code;
    A.corrected = d.stencil == SDT_STENCIL_CORRECTED;
    A.d_it = d.d_it;
    A.vnorm = d.norms;
    A.snorm = d.norms + d.time_steps;
    for (int q = 0; q < 3; ++q) A.vfix[q] = d.tail.vfix[q];
    for (int q = 0; q < 4; ++q) A.sfix[q] = d.tail.sfix[q];
    return A;
}

template <typename T>
static void el3_step_t(Device &d, bool timed)
{
    const El3Args<T> A = el3_args<T>(d);
    const bool fast = d.variant != 1;
    const dim3 grid(d.tiles_x, d.tiles_y, d.tiles_z), block(64, 4, 1);
    if (timed) (void)hipEventRecord(d.ev[0], d.stream);
    if (fast) el3_fast_stress<T>(d, A);
    else hipLaunchKernelGGL(el3_stress_kernel<T>, grid, block, 0, d.stream, A);
    if (timed) (void)hipEventRecord(d.ev[1], d.stream);
    if (fast) el3_fast_velocity<T>(d, A);
    else hipLaunchKernelGGL(el3_velocity_kernel<T>, grid, block, 0, d.stream, A);
    if (timed) (void)hipEventRecord(d.ev[2], d.stream);
    launch_tail(d);
}

void el3_step(Device &d, bool timed)
{
    if (d.precision == SDT_FP64) el3_step_t<double>(d, timed);
    else el3_step_t<float>(d, timed);
}

int el3_setup(Device &d, const double *const *in)
{
    const int64_t n = d.ncell;
    std::vector<double> zero;
    d.nmat = 0;
    for (int id = SDT_C11; id <= SDT_RHO; ++id) {
        if (!in[id] && zero.empty()) zero.assign((size_t)n, 0.0);
        int rc = upload_converted(d, in[id] ? in[id] : zero.data(), &d.mat[d.nmat++], n);
        if (rc) return rc;
    }
    std::vector<double> inv((size_t)n);
    const int gids[] = {SDT_GAMMA_X, SDT_GAMMA_Y, SDT_GAMMA_Z, SDT_GAMMA_XY, SDT_GAMMA_XZ, SDT_GAMMA_YZ};
    for (int id : gids) {
        for (int64_t c = 0; c < n; ++c) inv[(size_t)c] = 1.0 / (1.0 + (in[id] ? in[id][c] : 0.0) * d.dt);
        int rc = upload_converted(d, inv.data(), &d.mat[d.nmat++], n);
        if (rc) return rc;
    }
    d.nmem = M_TOTAL;
    for (int q = 0; q < d.nmem; ++q) {
        if (q == 0 * K_NKIND + K_ZMH) continue;   // dead recursion, never touched
        int rc = alloc_zeroed(d, &d.mem[q], n);
        if (rc) return rc;
    }
    for (int f = SDT_F_VX; f <= SDT_F_SYZ; ++f) {
        int rc;
        if (f <= SDT_F_VZ && in[SDT_IC_X + f]) rc = upload_converted(d, in[SDT_IC_X + f], &d.field[f], n);
        else rc = alloc_zeroed(d, &d.field[f], n);
        if (rc) return rc;
    }

    // point sources, cpmlfdtd.f08:2208-2219
    const int is = d.isrc, js = d.jsrc, ks = d.ksrc, nx = d.nx, ny = d.ny, nz = d.nz;
    auto cell = [&](int i, int j, int k) {
        return (int64_t)(i - 1) + (int64_t)nx * ((int64_t)(j - 1) + (int64_t)ny * (k - 1));
    };
    auto face = [&](int i, int j, int k) {
        return (int)(i == 1 || i == nx || j == 1 || j == ny || k == 1 || k == nz);
    };
    const double *rho = in[SDT_RHO];
    auto R = [&](int i, int j, int k) { return rho ? rho[cell(i, j, k)] : 0.0; };
    TailArgs &t = d.tail;
    t.nterms = 0;
    t.terms[t.nterms++] = {cell(is, js, ks), SDT_F_SXX, SDT_SRC_XX, 1.0, R(is, js, ks), 0};
    t.terms[t.nterms++] = {cell(is + 1, js + 1, ks), SDT_F_SXY, SDT_SRC_XY, 1.0, R(is + 1, js + 1, ks), 0};
    t.terms[t.nterms++] = {cell(is + 1, js, ks + 1), SDT_F_SXZ, SDT_SRC_XZ, 1.0, R(is + 1, js, ks), 0};
    t.terms[t.nterms++] = {cell(is, js, ks), SDT_F_SYY, SDT_SRC_YY, 1.0, R(is, js, ks), 0};
    t.terms[t.nterms++] = {cell(is, js + 1, ks + 1), SDT_F_SYZ, SDT_SRC_YZ, 1.0, R(is, js, ks + 1), 0};
    t.terms[t.nterms++] = {cell(is, js, ks), SDT_F_SZZ, SDT_SRC_ZZ, 1.0, R(is, js, ks), 0};
    t.terms[t.nterms++] = {cell(is + 1, js, ks), SDT_F_VX, SDT_SRC_X, d.dt, R(is + 1, js, ks), face(is + 1, js, ks)};
    t.terms[t.nterms++] = {cell(is, js + 1, ks), SDT_F_VY, SDT_SRC_Y, d.dt, R(is, js, ks), face(is, js + 1, ks)};
    t.terms[t.nterms++] = {cell(is, js, ks + 1), SDT_F_VZ, SDT_SRC_Z, d.dt, R(is, js, ks + 1), face(is, js, ks + 1)};
    t.nvfix = 3;
    t.vfix[0] = cell(is + 1, js, ks);
    t.vfix[1] = cell(is, js + 1, ks);
    t.vfix[2] = cell(is, js, ks + 1);
    t.nvcomp = 3;
    t.vcomp[0] = SDT_F_VX; t.vcomp[1] = SDT_F_VY; t.vcomp[2] = SDT_F_VZ;
    t.stress_norm = 1;
    t.nsfix = 4;
    t.sfix[0] = cell(is, js, ks);
    t.sfix[1] = cell(is + 1, js + 1, ks);
    t.sfix[2] = cell(is + 1, js, ks + 1);
    t.sfix[3] = cell(is, js + 1, ks + 1);
    t.rec_field[0] = SDT_F_VX; t.rec_field[1] = SDT_F_VY; t.rec_field[2] = SDT_F_VZ;

    d.tile_x = 64;
    d.tile_y = 4;
    d.tile_z = 1;
    int rc = compute_tile_flags(d, in);
    if (rc) return rc;
    // The branch-free interior kernels need every nest active on the box
    // (cpmlfdtd.f08:1761-2127 ranges intersect to [3, n-2] per axis) and no
    // CPML cell inside it; otherwise every cell takes the shell kernels.
    const int lim[3] = {d.nx, d.ny, d.nz};
    for (int a = 0; a < 3; ++a) {
        d.box[2 * a] = d.box[2 * a] < 3 ? 3 : d.box[2 * a];
        d.box[2 * a + 1] = d.box[2 * a + 1] > lim[a] - 2 ? lim[a] - 2 : d.box[2 * a + 1];
    }
    if (d.nflagged > 0) {
        for (int a = 0; a < 3; ++a) { d.box[2 * a] = 1; d.box[2 * a + 1] = 0; }
    }
    return SDT_OK;
}

}  // namespace sdt
