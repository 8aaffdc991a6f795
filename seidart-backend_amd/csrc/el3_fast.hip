// el3_fast.hip -- the 3D elastic step as two kernel pairs:
//
//  * interior kernels over the verified interior box (every cell an identity
//    CPML cell and inside all seven nest ranges of cpmlfdtd.f08:1761-2127):
//    no memory variables, no CPML coefficients, no range tests.  A 64 x 4
//    workgroup walks `kc` planes along z; each thread keeps its column's five
//    stencil planes in registers (z-neighbours fetched once per sweep) and
//    reads the x / y neighbours of the current plane through L1/L2.  Every
//    load of a plane is unconditional so the compiler issues them together.
//
//  * shell kernels over the cells outside the box (the CPML slabs and the
//    Dirichlet faces), enumerated compactly (z slabs, then y slabs, then x
//    slabs).  Neighbour indices are clamped into the grid so every load is
//    unconditional; nest ranges only mask the stores.
//
// Both write disjoint cells of the same arrays and fold into the same per-step
// norm slots, so their order within a step does not matter.
#include <math.h>

#include "el3_common.h"

namespace sdt {

template <typename T>
__device__ __forceinline__ T dp4(T fm1, T f0, T fp1, T fp2, T r)
{
    return (T(27) * (fp1 - f0) - (fp2 - fm1)) * r;
}
template <typename T>
__device__ __forceinline__ T dm4(T fm2, T fm1, T f0, T fp1, T r)
{
    return (T(27) * (f0 - fm1) - (fp1 - fm2)) * r;
}

template <typename T>
__device__ __forceinline__ T supd(T s, const T *const *C, int64_t c, int i0, int i1, int i2, int i3, int i4,
                                  int i5, T dxx, T dyy, T dzz, T yz, T xz, T xy, T dt, T ig)
{
    return (s + (ldnt(&C[i0][c]) * dxx + ldnt(&C[i1][c]) * dyy + ldnt(&C[i2][c]) * dzz +
                 ldnt(&C[i3][c]) * yz + ldnt(&C[i4][c]) * xz + ldnt(&C[i5][c]) * xy) * dt) * ig;
}

__device__ __forceinline__ double fmax_nan(double a, double b) { return (b > a || b != b) ? b : a; }

// ------------------------------------------------------------------ interior
template <typename T>
__global__ void __launch_bounds__(256) el3_int_stress_kernel(El3Args<T> A, int kc)
{
    const int i0 = A.box[0], j0 = A.box[2];
    const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
    const unsigned L = xcd_remap(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), nb);
    const unsigned bx = L % gridDim.x, by = (L / gridDim.x) % gridDim.y, bz = L / (gridDim.x * gridDim.y);
    const int i = i0 + bx * 64 + threadIdx.x;
    const int j = j0 + by * 4 + threadIdx.y;
    const int kb = A.box[4] + bz * kc;
    const int ke = min(kb + kc - 1, A.box[5]);
    const bool act = i <= A.box[1] && j <= A.box[3];
    const int64_t sy = A.nx, sz = (int64_t)A.nx * A.ny;
    int64_t c = (int64_t)((act ? i : i0) - 1) + sy * ((act ? j : j0) - 1) + sz * (kb - 1);
    const T rx = A.rd24[0], ry = A.rd24[1], rz = A.rd24[2], dt = A.dt;
    const T *const *C = A.C;
    T q[3][5];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int s = 0; s < 4; ++s) q[m][s] = A.v[m][c + (s - 2) * sz];
    double nrm = 0.0;
    for (int k = kb; k <= ke; ++k, c += sz) {
        T dxp[3], dxm[3], dyp[3], dym[3], dzp[3], dzm[3];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const T *v = A.v[m];
            q[m][4] = v[c + 2 * sz];
            const T xm2 = v[c - 2], xm1 = v[c - 1], xp1 = v[c + 1], xp2 = v[c + 2];
            const T ym2 = v[c - 2 * sy], ym1 = v[c - sy], yp1 = v[c + sy], yp2 = v[c + 2 * sy];
            const T v0 = q[m][2];
            dxp[m] = dp4(xm1, v0, xp1, xp2, rx);
            dxm[m] = dm4(xm2, xm1, v0, xp1, rx);
            dyp[m] = dp4(ym1, v0, yp1, yp2, ry);
            dym[m] = dm4(ym2, ym1, v0, yp1, ry);
            dzp[m] = dp4(q[m][1], v0, q[m][3], q[m][4], rz);
            dzm[m] = dm4(q[m][0], q[m][1], v0, q[m][3], rz);
        }
        // nests (cpmlfdtd.f08:1761-1992); without CPML the full and half
        // coefficient sets leave a derivative unchanged
        T sxx, syy, szz, sxy, sxz, syz;
        {   // L1: D+x, D-y, D-z
            const T yz = dym[2] + dzm[1], xz = dzm[0] + dxp[2], xy = dxp[1] + dym[0];
            sxx = supd(ldnt(&A.s[0][c]), C, c, 0, 1, 2, 3, 4, 5, dxp[0], dym[1], dzm[2], yz, xz, xy, dt, ldnt(&A.ig[0][c]));
            syy = supd(ldnt(&A.s[1][c]), C, c, 1, 6, 7, 8, 9, 10, dxp[0], dym[1], dzm[2], yz, xz, xy, dt, ldnt(&A.ig[1][c]));
            szz = supd(ldnt(&A.s[2][c]), C, c, 2, 7, 11, 12, 13, 14, dxp[0], dym[1], dzm[2], yz, xz, xy, dt, ldnt(&A.ig[2][c]));
        }
        {   // L2: D-x, D+y, D-z
            const T yz = dzm[1] + dyp[2], xz = dzm[0] + dxm[2], xy = dxm[1] + dyp[0];
            sxy = supd(ldnt(&A.s[3][c]), C, c, 5, 10, 14, 17, 19, 20, dxm[0], dyp[1], dzm[2], yz, xz, xy, dt, ldnt(&A.ig[3][c]));
        }
        {   // L3: D-x, D-y, D+z
            const T yz = dzp[1] + dym[2], xz = dzp[0] + dxm[2], xy = dxm[1] + dym[0];
            sxz = supd(ldnt(&A.s[4][c]), C, c, 4, 9, 13, 16, 18, 19, dxm[0], dym[1], dzp[2], yz, xz, xy, dt, ldnt(&A.ig[4][c]));
        }
        {   // L4: D+x, D+y, D+z
            const T yz = dzp[1] + dyp[2], xz = dzp[0] + dxp[2], xy = dxp[1] + dyp[0];
            syz = supd(ldnt(&A.s[5][c]), C, c, 3, 8, 12, 15, 16, 17, dxp[0], dyp[1], dzp[2], yz, xz, xy, dt, ldnt(&A.ig[5][c]));
        }
        if (act) {
            stnt(&A.s[0][c], sxx); stnt(&A.s[1][c], syy); stnt(&A.s[2][c], szz);
            stnt(&A.s[3][c], sxy); stnt(&A.s[4][c], sxz); stnt(&A.s[5][c], syz);
            const bool fix = c == A.sfix[0] || c == A.sfix[1] || c == A.sfix[2] || c == A.sfix[3];
            const double a0 = sxx, a1 = syy, a2 = szz, a3 = sxy, a4 = syz, a5 = sxz;
            const double n = sqrt(a0 * a0 + a1 * a1 + a2 * a2 + 2.0 * (a3 * a3 + a4 * a4 + a5 * a5));
            nrm = fix ? nrm : fmax_nan(nrm, n);
        }
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            q[m][0] = q[m][1]; q[m][1] = q[m][2]; q[m][2] = q[m][3]; q[m][3] = q[m][4];
        }
    }
    block_max_to(&A.snorm[*A.d_it - 1], nrm);
}

template <typename T, bool CORR>
__global__ void __launch_bounds__(256) el3_int_velocity_kernel(El3Args<T> A, int kc)
{
    const int i0 = A.box[0], j0 = A.box[2];
    const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
    const unsigned L = xcd_remap(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), nb);
    const unsigned bx = L % gridDim.x, by = (L / gridDim.x) % gridDim.y, bz = L / (gridDim.x * gridDim.y);
    const int i = i0 + bx * 64 + threadIdx.x;
    const int j = j0 + by * 4 + threadIdx.y;
    const int kb = A.box[4] + bz * kc;
    const int ke = min(kb + kc - 1, A.box[5]);
    const bool act = i <= A.box[1] && j <= A.box[3];
    const int64_t sy = A.nx, sz = (int64_t)A.nx * A.ny;
    int64_t c = (int64_t)((act ? i : i0) - 1) + sy * ((act ? j : j0) - 1) + sz * (kb - 1);
    const T rx = A.rd24[0], ry = A.rd24[1], rz = A.rd24[2], dt = A.dt;
    const int dc = A.density_code;
    const T *sxx = A.s[0], *syy = A.s[1], *szz = A.s[2], *sxy = A.s[3], *sxz = A.s[4], *syz = A.s[5];
    T qxz[5], qyz[5], qzz[5], qr[3];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        qxz[s] = sxz[c + (s - 2) * sz];
        qyz[s] = syz[c + (s - 2) * sz];
        qzz[s] = szz[c + (s - 2) * sz];
    }
    qr[0] = A.rho[c - sz];
    qr[1] = A.rho[c];
    double nrm = 0.0;
    for (int k = kb; k <= ke; ++k, c += sz) {
        qxz[4] = sxz[c + 2 * sz];
        qyz[4] = syz[c + 2 * sz];
        qzz[4] = szz[c + 2 * sz];
        qr[2] = A.rho[c + sz];
        const T r = qr[1], rxm = A.rho[c - 1], rxp = A.rho[c + 1];
        const T sxx_m2 = sxx[c - 2], sxx_m1 = sxx[c - 1], sxx_0 = sxx[c], sxx_p1 = CORR ? sxx[c + 1] : T(0);
        const T sxy_ym2 = sxy[c - 2 * sy], sxy_ym1 = sxy[c - sy], sxy_0 = sxy[c], sxy_yp1 = CORR ? sxy[c + sy] : T(0);
        const T sxy_xm1 = CORR ? sxy[c - 1] : T(0), sxy_xp1 = sxy[c + 1], sxy_xp2 = sxy[c + 2];
        const T syy_m1 = CORR ? syy[c - sy] : T(0), syy_0 = syy[c], syy_p1 = syy[c + sy], syy_p2 = syy[c + 2 * sy];
        const T sxz_xm1 = CORR ? sxz[c - 1] : T(0), sxz_xp1 = sxz[c + 1], sxz_xp2 = sxz[c + 2];
        const T syz_ym2 = syz[c - 2 * sy], syz_ym1 = syz[c - sy], syz_yp1 = CORR ? syz[c + sy] : T(0);
        T vx = ldnt(&A.v[0][c]), vy = ldnt(&A.v[1][c]), vz = ldnt(&A.v[2][c]);
        const T ryy = scalar_mean(r, r, dc);
        {   // vx: cpmlfdtd.f08:2014-2038
            const T d1 = tminus5(sxx_m2, sxx_m1, sxx_0, sxx_p1, CORR) * rx;
            const T d2 = tminus5(sxy_ym2, sxy_ym1, sxy_0, sxy_yp1, CORR) * ry;
            const T d3 = tminus5(qxz[0], qxz[1], qxz[2], qxz[3], CORR) * rz;
            vx = vx + (d1 / scalar_mean(r, rxm, dc) + d2 / ryy + d3 / scalar_mean(r, qr[0], dc)) * dt;
        }
        {   // vy: cpmlfdtd.f08:2055-2080
            const T d1 = tplus5(sxy_xm1, sxy_0, sxy_xp1, sxy_xp2, CORR) * rx;
            const T d2 = tplus5(syy_m1, syy_0, syy_p1, syy_p2, CORR) * ry;
            const T d3 = tminus5(qyz[0], qyz[1], qyz[2], qyz[3], CORR) * rz;
            vy = vy + (d1 / scalar_mean(r, rxp, dc) + d2 / ryy + d3 / scalar_mean(r, qr[0], dc)) * dt;
        }
        {   // vz: cpmlfdtd.f08:2098-2127
            const T d1 = tplus5(sxz_xm1, qxz[2], sxz_xp1, sxz_xp2, CORR) * rx;
            const T d2 = tminus5(syz_ym2, syz_ym1, qyz[2], syz_yp1, CORR) * ry;
            const T d3 = tplus5(qzz[1], qzz[2], qzz[3], qzz[4], CORR) * rz;
            vz = vz + (d1 / scalar_mean(r, rxp, dc) + d2 / ryy + d3 / scalar_mean(r, qr[2], dc)) * dt;
        }
        if (act) {
            stnt(&A.v[0][c], vx); stnt(&A.v[1][c], vy); stnt(&A.v[2][c], vz);
            const bool fix = c == A.vfix[0] || c == A.vfix[1] || c == A.vfix[2];
            const double a0 = vx, a1 = vy, a2 = vz;
            nrm = fix ? nrm : fmax_nan(nrm, sqrt(a0 * a0 + a1 * a1 + a2 * a2));
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            qxz[s] = qxz[s + 1]; qyz[s] = qyz[s + 1]; qzz[s] = qzz[s + 1];
        }
        qr[0] = qr[1]; qr[1] = qr[2];
    }
    block_max_to(&A.vnorm[*A.d_it - 1], nrm);
}

// --------------------------------------------------------------------- shell
struct ShellMap {
    int b[6];                 // box (possibly empty: b[1] < b[0] etc.)
    int64_t nzs, nys, nxs;    // cells in the z slabs, y slabs, x slabs
    int nxi, nyi, nzi;        // box extents (0 when empty)
};

// shell cell number t -> (i, j, k), 1-based
__device__ __forceinline__ void shell_cell(const ShellMap &S, int nx, int ny, int64_t t, int &i, int &j, int &k)
{
    if (t < S.nzs) {
        const int64_t plane = (int64_t)nx * ny;
        const int p = (int)(t / plane);
        const int64_t rem = t - (int64_t)p * plane;
        k = p < S.b[4] - 1 ? p + 1 : p + 1 + S.nzi;
        j = (int)(rem / nx) + 1;
        i = (int)(rem - (int64_t)(j - 1) * nx) + 1;
        return;
    }
    t -= S.nzs;
    if (t < S.nys) {
        const int row = (int)(t / nx);
        i = (int)(t - (int64_t)row * nx) + 1;
        const int nyo = ny - S.nyi;
        const int kk = row / nyo, jr = row - kk * nyo;
        k = S.b[4] + kk;
        j = jr < S.b[2] - 1 ? jr + 1 : jr + 1 + S.nyi;
        return;
    }
    t -= S.nys;
    const int nxo = nx - S.nxi;
    const int col = (int)(t % nxo);
    const int64_t rest = t / nxo;
    j = S.b[2] + (int)(rest % S.nyi);
    k = S.b[4] + (int)(rest / S.nyi);
    i = col < S.b[0] - 1 ? col + 1 : col + 1 + S.nxi;
}

template <typename T>
__device__ __forceinline__ T cpml_upd(T d, const T *psi_p, T a, T b, T ik, T &psi_out)
{
    psi_out = b * ldnt(psi_p) + a * d;
    return d * ik + psi_out;
}

template <typename T>
__global__ void __launch_bounds__(256) el3_shell_stress_kernel(El3Args<T> A, ShellMap S, int64_t nshell)
{
    const int64_t t = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    double nrm = 0.0;
    if (t < nshell) {
        const int nx = A.nx, ny = A.ny, nz = A.nz;
        int i, j, k;
        shell_cell(S, nx, ny, t, i, j, k);
        const int64_t sy = nx, sz = (int64_t)nx * ny;
        const int64_t c = (int64_t)(i - 1) + sy * (j - 1) + sz * (k - 1);
        const bool L1 = i >= 2 && i <= nx - 2 && j >= 3 && j <= ny - 1 && k >= 3 && k <= nz - 1;
        const bool L2 = i >= 3 && i <= nx - 1 && j >= 2 && j <= ny - 2 && k >= 3 && k <= nz - 1;
        const bool L3 = i >= 3 && i <= nx - 1 && j >= 3 && j <= ny - 1 && k >= 2 && k <= nz - 2;
        const bool L4 = i >= 2 && i <= nx - 2 && j >= 2 && j <= ny - 2 && k >= 2 && k <= nz - 2;
        // clamped neighbour offsets (loads stay inside the grid; results of
        // out-of-range stencils are never stored)
        const int64_t xm2 = (int64_t)(max(i - 2, 1) - i), xm1 = (int64_t)(max(i - 1, 1) - i);
        const int64_t xp1 = (int64_t)(min(i + 1, nx) - i), xp2 = (int64_t)(min(i + 2, nx) - i);
        const int64_t ym2 = (max(j - 2, 1) - j) * sy, ym1 = (max(j - 1, 1) - j) * sy;
        const int64_t yp1 = (min(j + 1, ny) - j) * sy, yp2 = (min(j + 2, ny) - j) * sy;
        const int64_t zm2 = (max(k - 2, 1) - k) * sz, zm1 = (max(k - 1, 1) - k) * sz;
        const int64_t zp1 = (min(k + 1, nz) - k) * sz, zp2 = (min(k + 2, nz) - k) * sz;
        const T rx = A.rd24[0], ry = A.rd24[1], rz = A.rd24[2], dt = A.dt;
        const T a = ldnt(&A.a[c]), b = ldnt(&A.b[c]), ik = ldnt(&A.ik[c]);
        const T ah = ldnt(&A.ah[c]), bh = ldnt(&A.bh[c]), ikh = ldnt(&A.ikh[c]);
        T d[3][K_NKIND];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const T *v = A.v[m];
            T *const *mm = A.m + m * K_NKIND;
            const T v0 = v[c];
            const T dxp = dp4(v[c + xm1], v0, v[c + xp1], v[c + xp2], rx);
            const T dxm = dm4(v[c + xm2], v[c + xm1], v0, v[c + xp1], rx);
            const T dyp = dp4(v[c + ym1], v0, v[c + yp1], v[c + yp2], ry);
            const T dym = dm4(v[c + ym2], v[c + ym1], v0, v[c + yp1], ry);
            const T dzp = dp4(v[c + zm1], v0, v[c + zp1], v[c + zp2], rz);
            const T dzm = dm4(v[c + zm2], v[c + zm1], v0, v[c + zp1], rz);
            T p0, p1, p2, p3, p4, p5 = T(0), p6;
            d[m][K_XPH] = cpml_upd(dxp, &mm[K_XPH][c], ah, bh, ikh, p0);
            d[m][K_XMF] = cpml_upd(dxm, &mm[K_XMF][c], a, b, ik, p1);
            d[m][K_YMF] = cpml_upd(dym, &mm[K_YMF][c], a, b, ik, p2);
            d[m][K_YPH] = cpml_upd(dyp, &mm[K_YPH][c], ah, bh, ikh, p3);
            d[m][K_ZMF] = cpml_upd(dzm, &mm[K_ZMF][c], a, b, ik, p4);
            // vx's z derivative enters sigma_xy uncorrected (cpmlfdtd.f08:1858-1865)
            d[m][K_ZMH] = m == 0 ? dzm : cpml_upd(dzm, &mm[K_ZMH][c], ah, bh, ikh, p5);
            d[m][K_ZPH] = cpml_upd(dzp, &mm[K_ZPH][c], ah, bh, ikh, p6);
            if (L1 || L4) stnt(&mm[K_XPH][c], p0);
            if (L2 || L3) stnt(&mm[K_XMF][c], p1);
            if (L1 || L3) stnt(&mm[K_YMF][c], p2);
            if (L2 || L4) stnt(&mm[K_YPH][c], p3);
            if (L1) stnt(&mm[K_ZMF][c], p4);
            if (L2 && m != 0) stnt(&mm[K_ZMH][c], p5);
            if (L3 || L4) stnt(&mm[K_ZPH][c], p6);
        }
        const T *const *C = A.C;
        T sxx = ldnt(&A.s[0][c]), syy = ldnt(&A.s[1][c]), szz = ldnt(&A.s[2][c]);
        T sxy = ldnt(&A.s[3][c]), sxz = ldnt(&A.s[4][c]), syz = ldnt(&A.s[5][c]);
        {
            const T dxx = d[0][K_XPH], dyy = d[1][K_YMF], dzz = d[2][K_ZMF];
            const T yz = d[1][K_ZMF] + d[2][K_YMF], xz = d[0][K_ZMF] + d[2][K_XPH], xy = d[1][K_XPH] + d[0][K_YMF];
            const T n0 = supd(sxx, C, c, 0, 1, 2, 3, 4, 5, dxx, dyy, dzz, yz, xz, xy, dt, ldnt(&A.ig[0][c]));
            const T n1 = supd(syy, C, c, 1, 6, 7, 8, 9, 10, dxx, dyy, dzz, yz, xz, xy, dt, ldnt(&A.ig[1][c]));
            const T n2 = supd(szz, C, c, 2, 7, 11, 12, 13, 14, dxx, dyy, dzz, yz, xz, xy, dt, ldnt(&A.ig[2][c]));
            if (L1) { sxx = n0; syy = n1; szz = n2; stnt(&A.s[0][c], sxx); stnt(&A.s[1][c], syy); stnt(&A.s[2][c], szz); }
        }
        {
            const T dxx = d[0][K_XMF], dyy = d[1][K_YPH], dzz = d[2][K_ZMH];
            const T yz = d[1][K_ZMH] + d[2][K_YPH], xz = d[0][K_ZMH] + d[2][K_XMF], xy = d[1][K_XMF] + d[0][K_YPH];
            const T n = supd(sxy, C, c, 5, 10, 14, 17, 19, 20, dxx, dyy, dzz, yz, xz, xy, dt, ldnt(&A.ig[3][c]));
            if (L2) { sxy = n; stnt(&A.s[3][c], sxy); }
        }
        {
            const T dxx = d[0][K_XMF], dyy = d[1][K_YMF], dzz = d[2][K_ZPH];
            const T yz = d[1][K_ZPH] + d[2][K_YMF], xz = d[0][K_ZPH] + d[2][K_XMF], xy = d[1][K_XMF] + d[0][K_YMF];
            const T n = supd(sxz, C, c, 4, 9, 13, 16, 18, 19, dxx, dyy, dzz, yz, xz, xy, dt, ldnt(&A.ig[4][c]));
            if (L3) { sxz = n; stnt(&A.s[4][c], sxz); }
        }
        {
            const T dxx = d[0][K_XPH], dyy = d[1][K_YPH], dzz = d[2][K_ZPH];
            const T yz = d[1][K_ZPH] + d[2][K_YPH], xz = d[0][K_ZPH] + d[2][K_XPH], xy = d[1][K_XPH] + d[0][K_YPH];
            const T n = supd(syz, C, c, 3, 8, 12, 15, 16, 17, dxx, dyy, dzz, yz, xz, xy, dt, ldnt(&A.ig[5][c]));
            if (L4) { syz = n; stnt(&A.s[5][c], syz); }
        }
        const bool fix = c == A.sfix[0] || c == A.sfix[1] || c == A.sfix[2] || c == A.sfix[3];
        const double a0 = sxx, a1 = syy, a2 = szz, a3 = sxy, a4 = syz, a5 = sxz;
        nrm = fix ? 0.0 : sqrt(a0 * a0 + a1 * a1 + a2 * a2 + 2.0 * (a3 * a3 + a4 * a4 + a5 * a5));
    }
    block_max_to(&A.snorm[*A.d_it - 1], nrm);
}

template <typename T, bool CORR>
__global__ void __launch_bounds__(256) el3_shell_velocity_kernel(El3Args<T> A, ShellMap S, int64_t nshell)
{
    const int64_t t = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    double nrm = 0.0;
    if (t < nshell) {
        const int nx = A.nx, ny = A.ny, nz = A.nz;
        int i, j, k;
        shell_cell(S, nx, ny, t, i, j, k);
        const int64_t sy = nx, sz = (int64_t)nx * ny;
        const int64_t c = (int64_t)(i - 1) + sy * (j - 1) + sz * (k - 1);
        T vx = ldnt(&A.v[0][c]), vy = ldnt(&A.v[1][c]), vz = ldnt(&A.v[2][c]);
        if (i == 1 || i == nx || j == 1 || j == ny || k == 1 || k == nz) {
            // Dirichlet faces, cpmlfdtd.f08:2226-2248
            vx = vy = vz = T(0);
            stnt(&A.v[0][c], vx); stnt(&A.v[1][c], vy); stnt(&A.v[2][c], vz);
        } else {
            const bool Lx = i >= 3 && i <= nx - 1 && j >= 3 && j <= ny - 1 && k >= 3 && k <= nz - 1;
            const bool Ly = i >= 2 && i <= nx - 2 && j >= 2 && j <= ny - 2 && k >= 3 && k <= nz - 1;
            const bool Lz = i >= 2 && i <= nx - 2 && j >= 3 && j <= ny - 1 && k >= 2 && k <= nz - 2;
            const int64_t xm2 = (int64_t)(max(i - 2, 1) - i), xm1 = -1, xp1 = 1;
            const int64_t xp2 = (int64_t)(min(i + 2, nx) - i);
            const int64_t ym2 = (max(j - 2, 1) - j) * sy, ym1 = -sy, yp1 = sy, yp2 = (min(j + 2, ny) - j) * sy;
            const int64_t zm2 = (max(k - 2, 1) - k) * sz, zm1 = -sz, zp1 = sz, zp2 = (min(k + 2, nz) - k) * sz;
            const T rx = A.rd24[0], ry = A.rd24[1], rz = A.rd24[2], dt = A.dt;
            const int dc = A.density_code;
            const T *sxx = A.s[0], *syy = A.s[1], *szz = A.s[2], *sxy = A.s[3], *sxz = A.s[4], *syz = A.s[5];
            const T a = ldnt(&A.a[c]), b = ldnt(&A.b[c]), ik = ldnt(&A.ik[c]);
            const T ah = ldnt(&A.ah[c]), bh = ldnt(&A.bh[c]), ikh = ldnt(&A.ikh[c]);
            const T r = A.rho[c];
            const T ryy = scalar_mean(r, r, dc);
            T p0, p1, p2;
            {   // vx: cpmlfdtd.f08:2014-2038
                T d1 = tminus5(sxx[c + xm2], sxx[c + xm1], sxx[c], CORR ? sxx[c + xp1] : T(0), CORR) * rx;
                T d2 = tminus5(sxy[c + ym2], sxy[c + ym1], sxy[c], CORR ? sxy[c + yp1] : T(0), CORR) * ry;
                T d3 = tminus5(sxz[c + zm2], sxz[c + zm1], sxz[c], CORR ? sxz[c + zp1] : T(0), CORR) * rz;
                d1 = cpml_upd(d1, &A.m[M_SXX_X][c], ah, bh, ikh, p0);
                d2 = cpml_upd(d2, &A.m[M_SXY_Y][c], ah, bh, ikh, p1);
                d3 = cpml_upd(d3, &A.m[M_SXZ_Z][c], ah, bh, ikh, p2);
                const T n = vx + (d1 / scalar_mean(r, A.rho[c + xm1], dc) + d2 / ryy +
                                  d3 / scalar_mean(r, A.rho[c + zm1], dc)) * dt;
                if (Lx) {
                    stnt(&A.m[M_SXX_X][c], p0); stnt(&A.m[M_SXY_Y][c], p1); stnt(&A.m[M_SXZ_Z][c], p2);
                    vx = n;
                    stnt(&A.v[0][c], vx);
                }
            }
            {   // vy: cpmlfdtd.f08:2055-2080
                T d1 = tplus5(CORR ? sxy[c + xm1] : T(0), sxy[c], sxy[c + xp1], sxy[c + xp2], CORR) * rx;
                T d2 = tplus5(CORR ? syy[c + ym1] : T(0), syy[c], syy[c + yp1], syy[c + yp2], CORR) * ry;
                T d3 = tminus5(syz[c + zm2], syz[c + zm1], syz[c], CORR ? syz[c + zp1] : T(0), CORR) * rz;
                d1 = cpml_upd(d1, &A.m[M_SXY_X][c], ah, bh, ikh, p0);
                d2 = cpml_upd(d2, &A.m[M_SYY_Y][c], ah, bh, ikh, p1);
                d3 = cpml_upd(d3, &A.m[M_SYZ_Z][c], a, b, ik, p2);
                const T n = vy + (d1 / scalar_mean(r, A.rho[c + xp1], dc) + d2 / ryy +
                                  d3 / scalar_mean(r, A.rho[c + zm1], dc)) * dt;
                if (Ly) {
                    stnt(&A.m[M_SXY_X][c], p0); stnt(&A.m[M_SYY_Y][c], p1); stnt(&A.m[M_SYZ_Z][c], p2);
                    vy = n;
                    stnt(&A.v[1][c], vy);
                }
            }
            {   // vz: cpmlfdtd.f08:2098-2127
                T d1 = tplus5(CORR ? sxz[c + xm1] : T(0), sxz[c], sxz[c + xp1], sxz[c + xp2], CORR) * rx;
                T d2 = tminus5(syz[c + ym2], syz[c + ym1], syz[c], CORR ? syz[c + yp1] : T(0), CORR) * ry;
                T d3 = tplus5(CORR ? szz[c + zm1] : T(0), szz[c], szz[c + zp1], szz[c + zp2], CORR) * rz;
                d1 = cpml_upd(d1, &A.m[M_SXZ_X][c], ah, bh, ikh, p0);
                d2 = cpml_upd(d2, &A.m[M_SYZ_Y][c], a, b, ik, p1);
                d3 = cpml_upd(d3, &A.m[M_SZZ_Z][c], ah, bh, ikh, p2);
                const T n = vz + (d1 / scalar_mean(r, A.rho[c + xp1], dc) + d2 / ryy +
                                  d3 / scalar_mean(r, A.rho[c + zp1], dc)) * dt;
                if (Lz) {
                    stnt(&A.m[M_SXZ_X][c], p0); stnt(&A.m[M_SYZ_Y][c], p1); stnt(&A.m[M_SZZ_Z][c], p2);
                    vz = n;
                    stnt(&A.v[2][c], vz);
                }
            }
        }
        const bool fix = c == A.vfix[0] || c == A.vfix[1] || c == A.vfix[2];
        const double a0 = vx, a1 = vy, a2 = vz;
        nrm = fix ? 0.0 : sqrt(a0 * a0 + a1 * a1 + a2 * a2);
    }
    block_max_to(&A.vnorm[*A.d_it - 1], nrm);
}

// ----------------------------------------------------------------- launchers
static ShellMap shell_map(const Device &d, int64_t &nshell)
{
    ShellMap S;
    for (int q = 0; q < 6; ++q) S.b[q] = d.box[q];
    const bool empty = d.box[1] < d.box[0] || d.box[3] < d.box[2] || d.box[5] < d.box[4];
    S.nxi = empty ? 0 : d.box[1] - d.box[0] + 1;
    S.nyi = empty ? 0 : d.box[3] - d.box[2] + 1;
    S.nzi = empty ? 0 : d.box[5] - d.box[4] + 1;
    if (empty) { S.b[4] = 1; S.b[2] = 1; S.b[0] = 1; }
    S.nzs = (int64_t)d.nx * d.ny * (d.nz - S.nzi);
    S.nys = (int64_t)d.nx * (d.ny - S.nyi) * S.nzi;
    S.nxs = (int64_t)(d.nx - S.nxi) * S.nyi * S.nzi;
    nshell = S.nzs + S.nys + S.nxs;
    return S;
}

static int el3_kc(const Device &d) { return d.zchunk > 0 ? d.zchunk : 16; }

static dim3 interior_grid(const Device &d, int kc)
{
    const int nxi = d.box[1] - d.box[0] + 1, nyi = d.box[3] - d.box[2] + 1, nzi = d.box[5] - d.box[4] + 1;
    return dim3((nxi + 63) / 64, (nyi + 3) / 4, (nzi + kc - 1) / kc);
}

static bool box_empty(const Device &d)
{
    return d.box[1] < d.box[0] || d.box[3] < d.box[2] || d.box[5] < d.box[4];
}

template <typename T>
void el3_fast_stress(Device &d, const El3Args<T> &A)
{
    const int kc = el3_kc(d);
    if (!box_empty(d))
        hipLaunchKernelGGL(el3_int_stress_kernel<T>, interior_grid(d, kc), dim3(64, 4, 1), 0, d.stream, A, kc);
    int64_t nshell;
    const ShellMap S = shell_map(d, nshell);
    if (nshell > 0)
        hipLaunchKernelGGL(el3_shell_stress_kernel<T>, dim3((unsigned)((nshell + 255) / 256)), dim3(256), 0,
                           d.stream, A, S, nshell);
}

template <typename T>
void el3_fast_velocity(Device &d, const El3Args<T> &A)
{
    const int kc = el3_kc(d);
    const bool corr = d.stencil == SDT_STENCIL_CORRECTED;
    if (!box_empty(d)) {
        if (corr)
            hipLaunchKernelGGL((el3_int_velocity_kernel<T, true>), interior_grid(d, kc), dim3(64, 4, 1), 0,
                               d.stream, A, kc);
        else
            hipLaunchKernelGGL((el3_int_velocity_kernel<T, false>), interior_grid(d, kc), dim3(64, 4, 1), 0,
                               d.stream, A, kc);
    }
    int64_t nshell;
    const ShellMap S = shell_map(d, nshell);
    if (nshell > 0) {
        const dim3 g((unsigned)((nshell + 255) / 256));
        if (corr) hipLaunchKernelGGL((el3_shell_velocity_kernel<T, true>), g, dim3(256), 0, d.stream, A, S, nshell);
        else hipLaunchKernelGGL((el3_shell_velocity_kernel<T, false>), g, dim3(256), 0, d.stream, A, S, nshell);
    }
}

template void el3_fast_stress<double>(Device &, const El3Args<double> &);
template void el3_fast_stress<float>(Device &, const El3Args<float> &);
template void el3_fast_velocity<double>(Device &, const El3Args<double> &);
template void el3_fast_velocity<float>(Device &, const El3Args<float> &);

}  // namespace sdt
