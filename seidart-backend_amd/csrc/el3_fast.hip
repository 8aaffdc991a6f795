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
//    Dirichlet faces), as a list of z-marching tiles of 64 cells per plane:
//    64 x 1 over the z slabs and the y strips, 16 x 4 over the (12-cell) x
//    strips, three cooperating wavefronts per tile.  Their memory variables
//    and CPML coefficients are stored compactly in tile order, so those
//    streams are contiguous per wavefront.  Neighbour indices are clamped
//    into the grid so every load is unconditional; nest ranges only mask the
//    stores.
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

// The six stress updates use 36 stiffness entries of 21 arrays; load each
// array once (non-temporal loads are not merged by the compiler).
template <typename T>
struct Stiff {
    T v[21];
};
template <typename T>
__device__ __forceinline__ void load_stiff(Stiff<T> &K, const T *const *C, int64_t c)
{
#pragma unroll
    for (int q = 0; q < 21; ++q) K.v[q] = ldnt(&C[q][c]);
}
// Material of cell c: 21 stiffness entries and the six attenuation
// divisors, from the dense per-cell arrays or from the material table.  When
// every lane of the wavefront holds the same material (the common case: a
// wavefront spans one x row) the row is read with scalar loads through the
// constant address space; otherwise each lane gathers its own row.
template <typename T>
using const_row = __attribute__((address_space(4))) const T *;

template <typename T, int MAT>
__device__ __forceinline__ void load_material(Stiff<T> &K, T (&ig)[6], const El3Args<T> &A, int64_t c)
{
    if (MAT == kMatTable) {
        const int m = A.mid[c];
        const int m0 = __builtin_amdgcn_readfirstlane(m);
        if (__all(m == m0)) {
            const_row<T> row = (const_row<T>)(A.tab) + m0 * kTabRow;
#pragma unroll
            for (int q = 0; q < 21; ++q) K.v[q] = row[q];
#pragma unroll
            for (int q = 0; q < 6; ++q) ig[q] = row[kTabIg + q];
        } else {
            const T *row = (const T *)__builtin_assume_aligned(A.tab + m * kTabRow, kTabRow * sizeof(T));
#pragma unroll
            for (int q = 0; q < 21; ++q) K.v[q] = row[q];
#pragma unroll
            for (int q = 0; q < 6; ++q) ig[q] = row[kTabIg + q];
        }
    } else {
        if (MAT == kMatIso) {
            const T c11 = ldnt(&A.C[0][c]), c12 = ldnt(&A.C[1][c]), c44 = ldnt(&A.C[15][c]);
#pragma unroll
            for (int q = 0; q < 21; ++q) K.v[q] = iso_entry(q, c11, c12, c44);
        } else {
            load_stiff(K, A.C, c);
        }
        if (A.ig_same) {   // wave-uniform branch
            const T g = ldnt(&A.ig[0][c]);
#pragma unroll
            for (int q = 0; q < 6; ++q) ig[q] = g;
        } else {
#pragma unroll
            for (int q = 0; q < 6; ++q) ig[q] = ldnt(&A.ig[q][c]);
        }
    }
}
template <typename T, bool TAB>
__device__ __forceinline__ T rho_at(const El3Args<T> &A, int64_t c)
{
    if (!TAB) return A.rho[c];
    const int m = A.mid[c];
    const int m0 = __builtin_amdgcn_readfirstlane(m);
    if (__all(m == m0)) return ((const_row<T>)(A.tab))[m0 * kTabRow + kTabRho];
    return A.tab[m * kTabRow + kTabRho];
}

// Density in two steps, so that the per-lane load (the material id, or the
// density itself without a table) can be issued a plane ahead of the cached
// table read.
template <typename T>
struct RhoPre {
    int m;
    T r;
};
template <typename T, bool TAB>
__device__ __forceinline__ RhoPre<T> rho_pre(const El3Args<T> &A, int64_t c)
{
    RhoPre<T> p;
    p.m = TAB ? (int)A.mid[c] : 0;
    p.r = TAB ? T(0) : A.rho[c];
    return p;
}
template <typename T, bool TAB>
__device__ __forceinline__ T rho_get(const El3Args<T> &A, const RhoPre<T> &p)
{
    if (!TAB) return p.r;
    const int m0 = __builtin_amdgcn_readfirstlane(p.m);
    if (__all(p.m == m0)) return ((const_row<T>)(A.tab))[m0 * kTabRow + kTabRho];
    return A.tab[p.m * kTabRow + kTabRho];
}

template <typename T>
__device__ __forceinline__ T supk(T s, const Stiff<T> &K, int i0, int i1, int i2, int i3, int i4, int i5, T dxx, T dyy,
                                  T dzz, T yz, T xz, T xy, T dt, T ig)
{
    return (s + (K.v[i0] * dxx + K.v[i1] * dyy + K.v[i2] * dzz + K.v[i3] * yz + K.v[i4] * xz + K.v[i5] * xy) * dt) * ig;
}

// ------------------------------------------------------------------ interior
template <typename T, int TY, int MAT>
__global__ void __launch_bounds__(64 * TY) el3_int_stress_kernel(El3Args<T> A, int kc)
{
    const int i0 = A.box[0], j0 = A.box[2];
    const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
    const unsigned b0 = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const unsigned L = (A.remap & 1) ? xcd_remap(b0, nb) : b0;
    const unsigned bx = L % gridDim.x, by = (L / gridDim.x) % gridDim.y, bz = L / (gridDim.x * gridDim.y);
    const int i = i0 + bx * 64 + threadIdx.x;
    const int j = j0 + by * TY + threadIdx.y;
    const int kb = A.box[4] + bz * kc;
    const int ke = min(kb + kc - 1, A.box[5]);
    const bool act = i <= A.box[1] && j <= A.box[3];
    const int64_t sy = A.pitch, sz = (int64_t)A.pitch * A.ny;
    int64_t c = A.xoff + (int64_t)((act ? i : i0) - 1) + sy * ((act ? j : j0) - 1) + sz * (kb - 1);
    const T rx = A.rd24[0], ry = A.rd24[1], rz = A.rd24[2], dt = A.dt;
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
        Stiff<T> K;
        T ig[6];
        load_material<T, MAT>(K, ig, A, c);
        T sxx, syy, szz, sxy, sxz, syz;
        {   // L1: D+x, D-y, D-z
            const T yz = dym[2] + dzm[1], xz = dzm[0] + dxp[2], xy = dxp[1] + dym[0];
            sxx = supk(ldnt(&A.s[0][c]), K, 0, 1, 2, 3, 4, 5, dxp[0], dym[1], dzm[2], yz, xz, xy, dt, ig[0]);
            syy = supk(ldnt(&A.s[1][c]), K, 1, 6, 7, 8, 9, 10, dxp[0], dym[1], dzm[2], yz, xz, xy, dt, ig[1]);
            szz = supk(ldnt(&A.s[2][c]), K, 2, 7, 11, 12, 13, 14, dxp[0], dym[1], dzm[2], yz, xz, xy, dt, ig[2]);
        }
        {   // L2: D-x, D+y, D-z
            const T yz = dzm[1] + dyp[2], xz = dzm[0] + dxm[2], xy = dxm[1] + dyp[0];
            sxy = supk(ldnt(&A.s[3][c]), K, 5, 10, 14, 17, 19, 20, dxm[0], dyp[1], dzm[2], yz, xz, xy, dt, ig[3]);
        }
        {   // L3: D-x, D-y, D+z
            const T yz = dzp[1] + dym[2], xz = dzp[0] + dxm[2], xy = dxm[1] + dym[0];
            sxz = supk(ldnt(&A.s[4][c]), K, 4, 9, 13, 16, 18, 19, dxm[0], dym[1], dzp[2], yz, xz, xy, dt, ig[4]);
        }
        {   // L4: D+x, D+y, D+z
            const T yz = dzp[1] + dyp[2], xz = dzp[0] + dxp[2], xy = dxp[1] + dyp[0];
            syz = supk(ldnt(&A.s[5][c]), K, 3, 8, 12, 15, 16, 17, dxp[0], dyp[1], dzp[2], yz, xz, xy, dt, ig[5]);
        }
        if (act) {
            stnt(&A.s[0][c], sxx); stnt(&A.s[1][c], syy); stnt(&A.s[2][c], szz);
            stnt(&A.s[3][c], sxy); stnt(&A.s[4][c], sxz); stnt(&A.s[5][c], syz);
            const bool fix = c == A.sfix[0] || c == A.sfix[1] || c == A.sfix[2] || c == A.sfix[3] ||
                             pw_face(A.pwf, i, j, k);
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

template <typename T, bool CORR, int TY, bool TAB>
__global__ void __launch_bounds__(64 * TY) el3_int_velocity_kernel(El3Args<T> A, int kc)
{
    const int i0 = A.box[0], j0 = A.box[2];
    const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
    const unsigned b0 = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const unsigned L = (A.remap & 1) ? xcd_remap(b0, nb) : b0;
    const unsigned bx = L % gridDim.x, by = (L / gridDim.x) % gridDim.y, bz = L / (gridDim.x * gridDim.y);
    const int i = i0 + bx * 64 + threadIdx.x;
    const int j = j0 + by * TY + threadIdx.y;
    const int kb = A.box[4] + bz * kc;
    const int ke = min(kb + kc - 1, A.box[5]);
    const bool act = i <= A.box[1] && j <= A.box[3];
    const int64_t sy = A.pitch, sz = (int64_t)A.pitch * A.ny;
    int64_t c = A.xoff + (int64_t)((act ? i : i0) - 1) + sy * ((act ? j : j0) - 1) + sz * (kb - 1);
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
    qr[0] = rho_at<T, TAB>(A, c - sz);
    qr[1] = rho_at<T, TAB>(A, c);
    // per-lane density loads of plane k (ids with the table), issued during
    // plane k-1 after its own loads; the cached table reads follow in plane k
    RhoPre<T> rz_next = rho_pre<T, TAB>(A, c + sz), rxm_next = rho_pre<T, TAB>(A, c - 1),
              rxp_next = rho_pre<T, TAB>(A, c + 1);
    double nrm = 0.0;
    for (int k = kb; k <= ke; ++k, c += sz) {
        qxz[4] = sxz[c + 2 * sz];
        qyz[4] = syz[c + 2 * sz];
        qzz[4] = szz[c + 2 * sz];
        qr[2] = rho_get<T, TAB>(A, rz_next);
        const T r = qr[1], rxm = rho_get<T, TAB>(A, rxm_next), rxp = rho_get<T, TAB>(A, rxp_next);
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
            const bool fix = c == A.vfix[0] || c == A.vfix[1] || c == A.vfix[2] || pw_face(A.pwf, i, j, k);
            const double a0 = vx, a1 = vy, a2 = vz;
            nrm = fix ? nrm : fmax_nan(nrm, sqrt(a0 * a0 + a1 * a1 + a2 * a2));
        }
        if (k < ke) {
            rz_next = rho_pre<T, TAB>(A, c + 2 * sz);
            rxm_next = rho_pre<T, TAB>(A, c + sz - 1);
            rxp_next = rho_pre<T, TAB>(A, c + sz + 1);
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
// A shell tile is kShellLanes (64) cells of a plane (64 x 1 over the z slabs
// and y strips, 16 x 4 over the x strips) marched over planes k0..k1.  Its
// workgroup is three wavefronts over the same 64 cells, so that no lane has to
// hold a whole cell's state:
//   stress pass: wavefront m computes the seven CPML-corrected derivatives of
//     velocity component m (updating its memory variables) into LDS; after a
//     barrier each wavefront updates two of the six stresses from them;
//   velocity pass: wavefront m updates velocity component m (its three
//     memory variables); the three meet in LDS only for the norm.
// Lanes outside the grid or the tile's region load from an active cell of the
// tile (no extra traffic) and store nothing.
__device__ __forceinline__ bool shell_lane(const El3Tile &t, const int (&b)[6], int i, int j, int nx, int ny)
{
    if (i < 1 || i > nx || j > ny) return false;
    if (t.kind == 0) return true;
    if (t.kind == 1) return j < b[2] || j > b[3];
    return j >= b[2] && j <= b[3] && (i < b[0] || i > b[1]);
}

struct ShellCol {
    bool on;
    int i, j;                 // cell (i, j) of the lane (the anchor cell when !on)
    int64_t col;              // element of (i, j, 1)
    int64_t xm2, xm1, xp1, xp2, ym2, ym1, yp1, yp2;   // neighbour offsets clamped into the grid
};

__device__ __forceinline__ ShellCol shell_col(const El3Tile &tl, const int (&box)[6], int lane, int nx, int ny,
                                              int64_t sy, int64_t xoff)
{
    ShellCol L;
    const int tx = lane & ((1 << tl.txl) - 1), ty = lane >> tl.txl;
    const int i0 = tl.x0 + tx, j0 = tl.y0 + ty;
    L.on = shell_lane(tl, box, i0, j0, nx, ny);
    L.i = L.on ? i0 : tl.ax;
    L.j = L.on ? j0 : tl.ay;
    const int i = L.i, j = L.j;
    L.xm2 = max(i - 2, 1) - i; L.xm1 = max(i - 1, 1) - i; L.xp1 = min(i + 1, nx) - i; L.xp2 = min(i + 2, nx) - i;
    L.ym2 = (max(j - 2, 1) - j) * sy; L.ym1 = (max(j - 1, 1) - j) * sy;
    L.yp1 = (min(j + 1, ny) - j) * sy; L.yp2 = (min(j + 2, ny) - j) * sy;
    L.col = xoff + (int64_t)(i - 1) + sy * (j - 1);
    return L;
}

// First element of a lane's streams in the tile's first plane; the next
// plane's block follows after S_NQ * kShellLanes elements.
template <typename T>
__device__ __forceinline__ T *shell_block(const El3Args<T> &A, const El3Tile &tl, int lane)
{
    return A.shell + (tl.base / kShellLanes) * (int64_t)(A.snq * kShellLanes) + lane;
}

// CPML coefficients of a shell slot, the source fixed at compile time (no
// branch, so the scheduler can issue the id load with the plane's other
// loads): from the dictionary (one-byte id, cached table) or from the
// block's coefficient streams.
template <typename T, bool CTAB>
__device__ __forceinline__ void shell_coefs_t(const El3Args<T> &A, const T *blk, int64_t slot, T (&cf)[6])
{
    if (CTAB) {
        const T *row = A.ctab + 8 * (int)A.cid[slot];
#pragma unroll
        for (int r = 0; r < 6; ++r) cf[r] = row[r];
    } else {
#pragma unroll
        for (int r = 0; r < 6; ++r) cf[r] = blk[(S_CF + r) * kShellLanes];
    }
}

// CPML coefficients of a shell slot: from the dictionary (one-byte id, cached
// table) or from the block's coefficient streams.
template <typename T>
__device__ __forceinline__ void shell_coefs(const El3Args<T> &A, const T *blk, int64_t slot, T (&cf)[6])
{
    if (A.ctab) {
        const T *row = A.ctab + 8 * (int)A.cid[slot];
#pragma unroll
        for (int r = 0; r < 6; ++r) cf[r] = row[r];
    } else {
#pragma unroll
        for (int r = 0; r < 6; ++r) cf[r] = blk[(S_CF + r) * kShellLanes];
    }
}

// The same in two steps (the velocity kernel loads the slot's dictionary id a
// plane ahead, then reads the cached row).
template <typename T>
__device__ __forceinline__ int coef_id(const El3Args<T> &A, int64_t slot)
{
    return A.ctab ? (int)A.cid[slot] : 0;
}
template <typename T>
__device__ __forceinline__ void shell_coefs_id(const El3Args<T> &A, const T *blk, int id, T (&cf)[6])
{
    if (A.ctab) {
        const T *row = A.ctab + 8 * id;
#pragma unroll
        for (int r = 0; r < 6; ++r) cf[r] = row[r];
    } else {
#pragma unroll
        for (int r = 0; r < 6; ++r) cf[r] = blk[(S_CF + r) * kShellLanes];
    }
}

template <typename T>
__device__ __forceinline__ T cpml_upd(T d, const T *psi_p, T a, T b, T ik, T &psi_out)
{
    psi_out = b * ldnt(psi_p) + a * d;
    return d * ik + psi_out;
}

// Stress pass, phase 1 for velocity component M: the four stress nests
// (cpmlfdtd.f08:1761-1992) use seven (direction, stencil, coefficient set)
// derivatives per component; each memory variable is updated once, on the
// union of the ranges of the nests that read it.
// INSIDE: read the CPML coefficients here, after the velocity loads and in
// their block, so the dictionary id's latency overlaps theirs (measured
// faster with the isotropic arrays, slower with the table or all 21 arrays);
// otherwise the caller passes them.
template <typename T, int M, bool CTAB, bool INSIDE>
__device__ __forceinline__ void shell_derivs(const El3Args<T> &A, const ShellCol &L, T (&q)[5], int64_t c,
                                             int64_t zp2, T *blk, bool L1, bool L2, bool L3, bool L4,
                                             int64_t slot, const T (&cf_in)[6], T *D, int lane)
{
    const T *v = A.v[M];
    q[4] = v[c + zp2];
    const T v0 = q[2];
    const T rx = A.rd24[0], ry = A.rd24[1], rz = A.rd24[2];
    const T dxp = dp4(v[c + L.xm1], v0, v[c + L.xp1], v[c + L.xp2], rx);
    const T dxm = dm4(v[c + L.xm2], v[c + L.xm1], v0, v[c + L.xp1], rx);
    const T dyp = dp4(v[c + L.ym1], v0, v[c + L.yp1], v[c + L.yp2], ry);
    const T dym = dm4(v[c + L.ym2], v[c + L.ym1], v0, v[c + L.yp1], ry);
    const T dzp = dp4(q[1], v0, q[3], q[4], rz);
    const T dzm = dm4(q[0], q[1], v0, q[3], rz);
    T cf[6];
    if (INSIDE) {
        shell_coefs_t<T, CTAB>(A, blk, slot, cf);
    } else {
#pragma unroll
        for (int r = 0; r < 6; ++r) cf[r] = cf_in[r];
    }
    const T a = cf[0], b = cf[1], ik = cf[2], ah = cf[3], bh = cf[4], ikh = cf[5];
    T *mm = blk + M * K_NKIND * kShellLanes;   // this component's seven streams
    T *Dm = D + M * K_NKIND * kShellLanes + lane;
    T p0, p1, p2, p3, p4, p5 = T(0), p6;
    Dm[K_XPH * kShellLanes] = cpml_upd(dxp, &mm[K_XPH * kShellLanes], ah, bh, ikh, p0);
    Dm[K_XMF * kShellLanes] = cpml_upd(dxm, &mm[K_XMF * kShellLanes], a, b, ik, p1);
    Dm[K_YMF * kShellLanes] = cpml_upd(dym, &mm[K_YMF * kShellLanes], a, b, ik, p2);
    Dm[K_YPH * kShellLanes] = cpml_upd(dyp, &mm[K_YPH * kShellLanes], ah, bh, ikh, p3);
    Dm[K_ZMF * kShellLanes] = cpml_upd(dzm, &mm[K_ZMF * kShellLanes], a, b, ik, p4);
    // vx's z derivative enters sigma_xy uncorrected (cpmlfdtd.f08:1858-1865)
    Dm[K_ZMH * kShellLanes] = M == 0 ? dzm : cpml_upd(dzm, &mm[K_ZMH * kShellLanes], ah, bh, ikh, p5);
    Dm[K_ZPH * kShellLanes] = cpml_upd(dzp, &mm[K_ZPH * kShellLanes], ah, bh, ikh, p6);
    if (L1 || L4) stnt(&mm[K_XPH * kShellLanes], p0);
    if (L2 || L3) stnt(&mm[K_XMF * kShellLanes], p1);
    if (L1 || L3) stnt(&mm[K_YMF * kShellLanes], p2);
    if (L2 || L4) stnt(&mm[K_YPH * kShellLanes], p3);
    if (L1) stnt(&mm[K_ZMF * kShellLanes], p4);
    if (M != 0 && L2) stnt(&mm[K_ZMH * kShellLanes], p5);
    if (L3 || L4) stnt(&mm[K_ZPH * kShellLanes], p6);
}

// Row accessors of the stiffness entries (0..20) and attenuation divisors
// (21..26) of one cell.
template <typename T>
struct ScalarRow {   // wave-uniform table row, scalar loads
    const_row<T> p;
    __device__ T operator[](int q) const { return p[q]; }
};
template <typename T>
struct LaneRow {     // per-lane table row
    const T *p;
    __device__ T operator[](int q) const { return p[q]; }
};
template <typename T>
struct DenseRow {    // per-cell arrays (cached loads: the three wavefronts share entries)
    const El3Args<T> *A;
    int64_t c;
    __device__ T operator[](int q) const { return q < 21 ? A->C[q][c] : A->ig[q - 21][c]; }
};
template <typename T>
struct IsoRow {      // isotropic-form per-cell arrays c11, c12, c44
    const El3Args<T> *A;
    int64_t c;
    T c11, c12, c44;
    __device__ T operator[](int q) const { return q < 21 ? iso_entry(q, c11, c12, c44) : A->ig[q - 21][c]; }
};

template <typename T, typename R>
__device__ __forceinline__ T supr(T s, const R &K, int i0, int i1, int i2, int i3, int i4, int i5, T dxx, T dyy,
                                  T dzz, T yz, T xz, T xy, T dt, int ig)
{
    return (s + (K[i0] * dxx + K[i1] * dyy + K[i2] * dzz + K[i3] * yz + K[i4] * xz + K[i5] * xy) * dt) * K[ig];
}

// Stress pass, phase 2 for wavefront W: W = 0 sigma_xx, sigma_yy (L1);
// W = 1 sigma_zz (L1), sigma_xy (L2); W = 2 sigma_xz (L3), sigma_yz (L4).
// Returns this wavefront's share of the stress norm's sum of squares.
template <typename T, int W, typename R>
__device__ __forceinline__ double shell_stresses(const El3Args<T> &A, const R &K, int64_t c, const T *D, int lane,
                                                 bool L1, bool L2, bool L3, bool L4, T sa, T sb)
{
    auto d = [&](int m, int kind) { return D[(m * K_NKIND + kind) * kShellLanes + lane]; };
    const T dt = A.dt;
    if (W == 0 || W == 1) {
        // L1 derivatives: D+x half, D-y full, D-z full
        const T dxx = d(0, K_XPH), dyy = d(1, K_YMF), dzz = d(2, K_ZMF);
        const T yz = d(1, K_ZMF) + d(2, K_YMF), xz = d(0, K_ZMF) + d(2, K_XPH), xy = d(1, K_XPH) + d(0, K_YMF);
        if (W == 0) {
            T sxx = sa, syy = sb;
            const T n0 = supr(sxx, K, 0, 1, 2, 3, 4, 5, dxx, dyy, dzz, yz, xz, xy, dt, kTabIg + 0);
            const T n1 = supr(syy, K, 1, 6, 7, 8, 9, 10, dxx, dyy, dzz, yz, xz, xy, dt, kTabIg + 1);
            if (L1) { sxx = n0; syy = n1; stnt(&A.s[0][c], sxx); stnt(&A.s[1][c], syy); }
            const double a0 = sxx, a1 = syy;
            return a0 * a0 + a1 * a1;
        }
        T szz = sa, sxy = sb;
        const T n2 = supr(szz, K, 2, 7, 11, 12, 13, 14, dxx, dyy, dzz, yz, xz, xy, dt, kTabIg + 2);
        if (L1) { szz = n2; stnt(&A.s[2][c], szz); }
        {   // L2 derivatives: D-x full, D+y half, D-z half (vx's uncorrected)
            const T ex = d(0, K_XMF), ey = d(1, K_YPH), ez = d(2, K_ZMH);
            const T eyz = d(1, K_ZMH) + d(2, K_YPH), exz = d(0, K_ZMH) + d(2, K_XMF), exy = d(1, K_XMF) + d(0, K_YPH);
            const T n3 = supr(sxy, K, 5, 10, 14, 17, 19, 20, ex, ey, ez, eyz, exz, exy, dt, kTabIg + 3);
            if (L2) { sxy = n3; stnt(&A.s[3][c], sxy); }
        }
        const double a2 = szz, a3 = sxy;
        return a2 * a2 + 2.0 * (a3 * a3);
    }
    T sxz = sa, syz = sb;
    {   // L3 derivatives: D-x full, D-y full, D+z half
        const T dxx = d(0, K_XMF), dyy = d(1, K_YMF), dzz = d(2, K_ZPH);
        const T yz = d(1, K_ZPH) + d(2, K_YMF), xz = d(0, K_ZPH) + d(2, K_XMF), xy = d(1, K_XMF) + d(0, K_YMF);
        const T n4 = supr(sxz, K, 4, 9, 13, 16, 18, 19, dxx, dyy, dzz, yz, xz, xy, dt, kTabIg + 4);
        if (L3) { sxz = n4; stnt(&A.s[4][c], sxz); }
    }
    {   // L4 derivatives: D+x half, D+y half, D+z half
        const T dxx = d(0, K_XPH), dyy = d(1, K_YPH), dzz = d(2, K_ZPH);
        const T yz = d(1, K_ZPH) + d(2, K_YPH), xz = d(0, K_ZPH) + d(2, K_XPH), xy = d(1, K_XPH) + d(0, K_YPH);
        const T n5 = supr(syz, K, 3, 8, 12, 15, 16, 17, dxx, dyy, dzz, yz, xz, xy, dt, kTabIg + 5);
        if (L4) { syz = n5; stnt(&A.s[5][c], syz); }
    }
    const double a4 = syz, a5 = sxz;
    return 2.0 * (a4 * a4 + a5 * a5);
}

// Phase-2 operands of one plane that do not depend on phase 1: this
// wavefront's two stresses and the handle of the cell's material (table id,
// or the isotropic entries).
template <typename T>
struct ShellPre {
    T sa, sb;
    int m;
    T c11, c12, c44;
};

template <typename T, int MAT>
__device__ __forceinline__ ShellPre<T> shell_pre(const El3Args<T> &A, int64_t c, int w)
{
    ShellPre<T> P;
    P.sa = ldnt(&A.s[2 * w][c]);
    P.sb = ldnt(&A.s[2 * w + 1][c]);
    P.m = 0;
    P.c11 = P.c12 = P.c44 = T(0);
    if (MAT == kMatTable) P.m = A.mid[c];
    if (MAT == kMatIso) {
        P.c11 = A.C[0][c];
        P.c12 = A.C[1][c];
        P.c44 = A.C[15][c];
    }
    return P;
}

template <typename T, int MAT, int W>
__device__ __forceinline__ double shell_stresses_at(const El3Args<T> &A, int64_t c, const T *D, int lane, bool L1,
                                                    bool L2, bool L3, bool L4, const ShellPre<T> &P)
{
    if (MAT == kMatIso) {
        const IsoRow<T> K{&A, c, P.c11, P.c12, P.c44};
        return shell_stresses<T, W>(A, K, c, D, lane, L1, L2, L3, L4, P.sa, P.sb);
    }
    if (MAT == kMatTable) {
        const int m = P.m;
        const int m0 = __builtin_amdgcn_readfirstlane(m);
        if (__all(m == m0)) {
            const ScalarRow<T> K{(const_row<T>)(A.tab) + m0 * kTabRow};
            return shell_stresses<T, W>(A, K, c, D, lane, L1, L2, L3, L4, P.sa, P.sb);
        }
        const LaneRow<T> K{A.tab + m * kTabRow};
        return shell_stresses<T, W>(A, K, c, D, lane, L1, L2, L3, L4, P.sa, P.sb);
    }
    const DenseRow<T> K{&A, c};
    return shell_stresses<T, W>(A, K, c, D, lane, L1, L2, L3, L4, P.sa, P.sb);
}

// The stress kernel runs its two phases as a pipeline over the planes: in one
// barrier interval a wavefront computes the derivatives of plane k+1 (into
// one LDS buffer) and two stresses of plane k (from the other), so the loads
// of both phases are in flight together.  Buffers alternate per plane; every
// buffer is rewritten only after the barrier that follows its last read.
template <typename T, int MAT, bool CTAB>
__global__ void __launch_bounds__(3 * kShellLanes) el3_shell_stress_kernel(El3Args<T> A, const El3Tile *tiles)
{
    __shared__ T D[2][3 * K_NKIND * kShellLanes];
    __shared__ double N[2][3 * kShellLanes];
    const El3Tile tl = tiles[(A.remap & 2) ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x];
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const int lane = threadIdx.x & (kShellLanes - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kShellLanes);
    const int64_t sy = A.pitch, sz = (int64_t)A.pitch * ny;
    const ShellCol L = shell_col(tl, A.box, lane, nx, ny, sy, A.xoff);
    const int i = L.i, j = L.j;
    auto nests = [&](int k, bool &L1, bool &L2, bool &L3, bool &L4) {
        L1 = L.on && i >= 2 && i <= nx - 2 && j >= 3 && j <= ny - 1 && k >= 3 && k <= nz - 1;
        L2 = L.on && i >= 3 && i <= nx - 1 && j >= 2 && j <= ny - 2 && k >= 3 && k <= nz - 1;
        L3 = L.on && i >= 3 && i <= nx - 1 && j >= 3 && j <= ny - 1 && k >= 2 && k <= nz - 2;
        L4 = L.on && i >= 2 && i <= nx - 2 && j >= 2 && j <= ny - 2 && k >= 2 && k <= nz - 2;
    };
    // z queue of this wavefront's velocity component
    const T *vq = w == 0 ? A.v[0] : (w == 1 ? A.v[1] : A.v[2]);
    T q[5];
#pragma unroll
    for (int s = 0; s < 4; ++s) q[s] = vq[L.col + sz * (min(max(tl.k0 - 2 + s, 1), nz) - 1)];
    const int lsel = L.on ? lane : tl.alane;
    T *const blk0 = shell_block(A, tl, lsel);
    const int64_t bstride = (int64_t)A.snq * kShellLanes;
    // phase 1 of plane k into buffer b
    auto phase1 = [&](int k, int b) {
        const int64_t c = L.col + sz * (k - 1);
        const int64_t zp2 = sz * (min(k + 2, nz) - k);
        T *const blk = blk0 + (int64_t)(k - tl.k0) * bstride;
        bool L1, L2, L3, L4;
        nests(k, L1, L2, L3, L4);
        const int64_t slot = tl.base + (int64_t)(k - tl.k0) * kShellLanes + lsel;
        // coefficients read by all three wavefronts: cached loads
        constexpr bool inside = MAT == kMatIso;
        T cf[6];
        if (!inside) shell_coefs_t<T, CTAB>(A, blk, slot, cf);
        if (w == 0) shell_derivs<T, 0, CTAB, inside>(A, L, q, c, zp2, blk, L1, L2, L3, L4, slot, cf, D[b], lane);
        else if (w == 1) shell_derivs<T, 1, CTAB, inside>(A, L, q, c, zp2, blk, L1, L2, L3, L4, slot, cf, D[b], lane);
        else shell_derivs<T, 2, CTAB, inside>(A, L, q, c, zp2, blk, L1, L2, L3, L4, slot, cf, D[b], lane);
        q[0] = q[1]; q[1] = q[2]; q[2] = q[3]; q[3] = q[4];
    };
    phase1(tl.k0, 0);
    __syncthreads();
    double nrm = 0.0;
    for (int k = tl.k0; k <= tl.k1; ++k) {
        const int b = (k - tl.k0) & 1;
        const int64_t c = L.col + sz * (k - 1);
        // per-cell arrays: load the phase-2 operands ahead of phase 1 (their
        // latency is exposed otherwise); with the table (scalar row loads)
        // the kernel measured faster loading them after it
        constexpr bool ahead = MAT != kMatTable;
        ShellPre<T> P;
        if (ahead) P = shell_pre<T, MAT>(A, c, w);
        if (k < tl.k1) phase1(k + 1, b ^ 1);
        if (!ahead) P = shell_pre<T, MAT>(A, c, w);
        bool L1, L2, L3, L4;
        nests(k, L1, L2, L3, L4);
        double part;
        if (w == 0) part = shell_stresses_at<T, MAT, 0>(A, c, D[b], lane, L1, L2, L3, L4, P);
        else if (w == 1) part = shell_stresses_at<T, MAT, 1>(A, c, D[b], lane, L1, L2, L3, L4, P);
        else part = shell_stresses_at<T, MAT, 2>(A, c, D[b], lane, L1, L2, L3, L4, P);
        N[b][w * kShellLanes + lane] = part;
        __syncthreads();
        if (w == 0 && L.on) {
            const bool fix = c == A.sfix[0] || c == A.sfix[1] || c == A.sfix[2] || c == A.sfix[3] ||
                             pw_face(A.pwf, L.i, L.j, k);
            const double n = sqrt(N[b][lane] + N[b][kShellLanes + lane] + N[b][2 * kShellLanes + lane]);
            nrm = fix ? nrm : fmax_nan(nrm, n);
        }
    }
    block_max_to(&A.snorm[*A.d_it - 1], nrm);
}

// Velocity pass for component M (cpmlfdtd.f08:2014-2038 vx, 2055-2080 vy,
// 2098-2127 vz): three stress derivatives, their memory variables, the
// averaged densities; Dirichlet faces (2226-2248) give 0.  Returns the new
// value (the old one outside the nest's range).
template <typename T, bool CORR, bool TAB, int M>
__device__ __forceinline__ T shell_velocity(const El3Args<T> &A, const ShellCol &L, T (&qs)[5], T (&qr)[3], int64_t c,
                                            T *blk, int k, const T (&cf)[6], T rx_nb)
{
    const int nx = A.nx, ny = A.ny, nz = A.nz, i = L.i, j = L.j;
    const T rx = A.rd24[0], ry = A.rd24[1], rz = A.rd24[2], dt = A.dt;
    const int dc = A.density_code;
    const T a = cf[0], b = cf[1], ik = cf[2], ah = cf[3], bh = cf[4], ikh = cf[5];
    const bool face = i == 1 || i == nx || j == 1 || j == ny || k == 1 || k == nz;
    const T r = qr[1];
    const T ryy = scalar_mean(r, r, dc);
    T v = ldnt(&A.v[M][c]);
    T p0, p1, p2;
    T d1, d2, d3;
    T *m1, *m2, *m3;
    T r1, r3;
    bool in;
    if (M == 0) {
        const T *sxx = A.s[0], *sxy = A.s[3];
        d1 = tminus5(sxx[c + L.xm2], sxx[c + L.xm1], sxx[c], CORR ? sxx[c + L.xp1] : T(0), CORR) * rx;
        d2 = tminus5(sxy[c + L.ym2], sxy[c + L.ym1], sxy[c], CORR ? sxy[c + L.yp1] : T(0), CORR) * ry;
        d3 = tminus5(qs[0], qs[1], qs[2], qs[3], CORR) * rz;
        m1 = blk + M_SXX_X * kShellLanes; m2 = blk + M_SXY_Y * kShellLanes; m3 = blk + M_SXZ_Z * kShellLanes;
        d1 = cpml_upd(d1, m1, ah, bh, ikh, p0);
        d2 = cpml_upd(d2, m2, ah, bh, ikh, p1);
        d3 = cpml_upd(d3, m3, ah, bh, ikh, p2);
        r1 = scalar_mean(r, rx_nb, dc);   // x neighbour i-1
        r3 = scalar_mean(r, qr[0], dc);
        in = i >= 3 && j >= 3 && k >= 3;
    } else if (M == 1) {
        const T *sxy = A.s[3], *syy = A.s[1];
        d1 = tplus5(CORR ? sxy[c + L.xm1] : T(0), sxy[c], sxy[c + L.xp1], sxy[c + L.xp2], CORR) * rx;
        d2 = tplus5(CORR ? syy[c + L.ym1] : T(0), syy[c], syy[c + L.yp1], syy[c + L.yp2], CORR) * ry;
        d3 = tminus5(qs[0], qs[1], qs[2], qs[3], CORR) * rz;
        m1 = blk + M_SXY_X * kShellLanes; m2 = blk + M_SYY_Y * kShellLanes; m3 = blk + M_SYZ_Z * kShellLanes;
        d1 = cpml_upd(d1, m1, ah, bh, ikh, p0);
        d2 = cpml_upd(d2, m2, ah, bh, ikh, p1);
        d3 = cpml_upd(d3, m3, a, b, ik, p2);
        r1 = scalar_mean(r, rx_nb, dc);   // x neighbour i+1
        r3 = scalar_mean(r, qr[0], dc);
        in = i <= nx - 2 && j <= ny - 2 && k >= 3;
    } else {
        const T *sxz = A.s[4], *syz = A.s[5];
        d1 = tplus5(CORR ? sxz[c + L.xm1] : T(0), sxz[c], sxz[c + L.xp1], sxz[c + L.xp2], CORR) * rx;
        d2 = tminus5(syz[c + L.ym2], syz[c + L.ym1], syz[c], CORR ? syz[c + L.yp1] : T(0), CORR) * ry;
        d3 = tplus5(qs[1], qs[2], qs[3], qs[4], CORR) * rz;
        m1 = blk + M_SXZ_X * kShellLanes; m2 = blk + M_SYZ_Y * kShellLanes; m3 = blk + M_SZZ_Z * kShellLanes;
        d1 = cpml_upd(d1, m1, ah, bh, ikh, p0);
        d2 = cpml_upd(d2, m2, a, b, ik, p1);
        d3 = cpml_upd(d3, m3, ah, bh, ikh, p2);
        r1 = scalar_mean(r, rx_nb, dc);   // x neighbour i+1
        r3 = scalar_mean(r, qr[2], dc);
        in = i <= nx - 2 && j >= 3 && k <= nz - 2;
    }
    const T n = v + (d1 / r1 + d2 / ryy + d3 / r3) * dt;
    if (L.on && !face && in) {
        stnt(m1, p0); stnt(m2, p1); stnt(m3, p2);
        v = n;
    }
    if (face) v = T(0);
    if (L.on) stnt(&A.v[M][c], v);
    return v;
}

// One barrier per plane: the norm of a plane needs all three components, so
// each wavefront leaves v^2 in LDS (double-buffered) for wavefront 0.
template <typename T, bool CORR, bool TAB>
__global__ void __launch_bounds__(3 * kShellLanes) el3_shell_velocity_kernel(El3Args<T> A, const El3Tile *tiles)
{
    __shared__ double V2[2][3 * kShellLanes];
    const El3Tile tl = tiles[(A.remap & 2) ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x];
    const int nz = A.nz;
    const int lane = threadIdx.x & (kShellLanes - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kShellLanes);
    const int64_t sy = A.pitch, sz = (int64_t)A.pitch * A.ny;
    const ShellCol L = shell_col(tl, A.box, lane, A.nx, A.ny, sy, A.xoff);
    // z queue: sigma_xz (vx), sigma_yz (vy), sigma_zz (vz), planes k-2..k+2
    const T *sq = w == 0 ? A.s[4] : (w == 1 ? A.s[5] : A.s[2]);
    T qs[5], qr[3];
#pragma unroll
    for (int s = 0; s < 4; ++s) qs[s] = sq[L.col + sz * (min(max(tl.k0 - 2 + s, 1), nz) - 1)];
    qr[0] = rho_at<T, TAB>(A, L.col + sz * (max(tl.k0 - 1, 1) - 1));
    qr[1] = rho_at<T, TAB>(A, L.col + sz * (tl.k0 - 1));
    double nrm = 0.0;
    const int lsel = L.on ? lane : tl.alane;
    T *blk = shell_block(A, tl, lsel);
    int64_t slot = tl.base + lsel;
    // x neighbour of the x-face density mean: i-1 for vx, i+1 for vy and vz
    const int64_t xo = w == 0 ? L.xm1 : L.xp1;
    // per-lane loads of plane k, issued during plane k-1 (after its own loads)
    auto zp1_of = [&](int k) { return sz * (min(k + 1, nz) - k); };
    auto zp2_of = [&](int k) { return sz * (min(k + 2, nz) - k); };
    const int64_t c0 = L.col + sz * (tl.k0 - 1);
    T qs_next = sq[c0 + zp2_of(tl.k0)];
    RhoPre<T> rz_next = rho_pre<T, TAB>(A, c0 + zp1_of(tl.k0)), rx_next = rho_pre<T, TAB>(A, c0 + xo);
    int cid_next = coef_id(A, slot);
    for (int k = tl.k0; k <= tl.k1; ++k, blk += (int64_t)A.snq * kShellLanes, slot += kShellLanes) {
        const int b = (k - tl.k0) & 1;
        const int64_t c = L.col + sz * (k - 1);
        qs[4] = qs_next;
        qr[2] = rho_get<T, TAB>(A, rz_next);
        const T rx_nb = rho_get<T, TAB>(A, rx_next);
        // read by all three wavefronts: cached loads
        T cf[6];
        shell_coefs_id(A, blk, cid_next, cf);
        T v;
        if (w == 0) v = shell_velocity<T, CORR, TAB, 0>(A, L, qs, qr, c, blk, k, cf, rx_nb);
        else if (w == 1) v = shell_velocity<T, CORR, TAB, 1>(A, L, qs, qr, c, blk, k, cf, rx_nb);
        else v = shell_velocity<T, CORR, TAB, 2>(A, L, qs, qr, c, blk, k, cf, rx_nb);
        if (k < tl.k1) {
            const int64_t cn = c + sz;
            qs_next = sq[cn + zp2_of(k + 1)];
            rz_next = rho_pre<T, TAB>(A, cn + zp1_of(k + 1));
            rx_next = rho_pre<T, TAB>(A, cn + xo);
            cid_next = coef_id(A, slot + kShellLanes);
        }
        const double dv = v;
        V2[b][w * kShellLanes + lane] = dv * dv;
        __syncthreads();
        if (w == 0 && L.on) {
            const bool fix = c == A.vfix[0] || c == A.vfix[1] || c == A.vfix[2] || pw_face(A.pwf, L.i, L.j, k);
            const double n = sqrt(V2[b][lane] + V2[b][kShellLanes + lane] + V2[b][2 * kShellLanes + lane]);
            nrm = fix ? nrm : fmax_nan(nrm, n);
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) qs[s] = qs[s + 1];
        qr[0] = qr[1]; qr[1] = qr[2];
    }
    block_max_to(&A.vnorm[*A.d_it - 1], nrm);
}

// ----------------------------------------------------------------- launchers
static int el3_kc(const Device &d) { return d.zchunk > 0 ? d.zchunk : 8; }

static int el3_ty(const Device &d) { return d.tile_rows == 8 || d.tile_rows == 16 ? d.tile_rows : 4; }

static dim3 interior_grid(const Device &d, int kc, int ty)
{
    const int nxi = d.box[1] - d.box[0] + 1, nyi = d.box[3] - d.box[2] + 1, nzi = d.box[5] - d.box[4] + 1;
    return dim3((nxi + 63) / 64, (nyi + ty - 1) / ty, (nzi + kc - 1) / kc);
}

static bool box_empty(const Device &d)
{
    return d.box[1] < d.box[0] || d.box[3] < d.box[2] || d.box[5] < d.box[4];
}

template <typename T, int TY, int MAT>
static void int_stress(Device &d, const El3Args<T> &A, int kc)
{
    hipLaunchKernelGGL((el3_int_stress_kernel<T, TY, MAT>), interior_grid(d, kc, TY), dim3(64, TY, 1), 0, d.stream,
                       A, kc);
}

template <typename T, bool CORR, int TY, bool TAB>
static void int_velocity(Device &d, const El3Args<T> &A, int kc)
{
    hipLaunchKernelGGL((el3_int_velocity_kernel<T, CORR, TY, TAB>), interior_grid(d, kc, TY), dim3(64, TY, 1), 0,
                       d.stream, A, kc);
}

// The interior and shell kernels of a pass write disjoint cells and read only
// the other half step's fields, so they may run concurrently: the shell kernel
// goes to the side stream between a fork and a join event.
static hipStream_t shell_stream_begin(Device &d, int pass)
{
    if (!d.concurrent || !d.side) return d.stream;
    (void)hipEventRecord(d.fork[pass], d.stream);
    (void)hipStreamWaitEvent(d.side, d.fork[pass], 0);
    return d.side;
}

static void shell_stream_end(Device &d, int pass)
{
    if (!d.concurrent || !d.side) return;
    (void)hipEventRecord(d.join[pass], d.side);
    (void)hipStreamWaitEvent(d.stream, d.join[pass], 0);
}

template <typename T, int MAT>
static void fast_stress(Device &d, const El3Args<T> &A)
{
    const int kc = el3_kc(d), ty = el3_ty(d);
    const hipStream_t ss = shell_stream_begin(d, 0);
    if (d.ntiles > 0) {
        if (A.ctab)
            hipLaunchKernelGGL((el3_shell_stress_kernel<T, MAT, true>), dim3(d.ntiles), dim3(3 * kShellLanes), 0, ss,
                               A, (const El3Tile *)d.tiles);
        else
            hipLaunchKernelGGL((el3_shell_stress_kernel<T, MAT, false>), dim3(d.ntiles), dim3(3 * kShellLanes), 0, ss,
                               A, (const El3Tile *)d.tiles);
    }
    if (!box_empty(d)) {
        if (ty == 16) int_stress<T, 16, MAT>(d, A, kc);
        else if (ty == 8) int_stress<T, 8, MAT>(d, A, kc);
        else int_stress<T, 4, MAT>(d, A, kc);
    }
    shell_stream_end(d, 0);
}

template <typename T, bool CORR, bool TAB>
static void fast_velocity(Device &d, const El3Args<T> &A)
{
    const int kc = el3_kc(d), ty = el3_ty(d);
    const hipStream_t ss = shell_stream_begin(d, 1);
    if (d.ntiles > 0)
        hipLaunchKernelGGL((el3_shell_velocity_kernel<T, CORR, TAB>), dim3(d.ntiles), dim3(3 * kShellLanes), 0, ss,
                           A, (const El3Tile *)d.tiles);
    if (!box_empty(d)) {
        if (ty == 16) int_velocity<T, CORR, 16, TAB>(d, A, kc);
        else if (ty == 8) int_velocity<T, CORR, 8, TAB>(d, A, kc);
        else int_velocity<T, CORR, 4, TAB>(d, A, kc);
    }
    shell_stream_end(d, 1);
}

template <typename T>
void el3_fast_stress(Device &d, const El3Args<T> &A)
{
    if (d.mat_table) fast_stress<T, kMatTable>(d, A);
    else if (d.mat_iso) fast_stress<T, kMatIso>(d, A);
    else fast_stress<T, kMatDense>(d, A);
}

template <typename T>
void el3_fast_velocity(Device &d, const El3Args<T> &A)
{
    const bool corr = d.stencil == SDT_STENCIL_CORRECTED;
    if (d.mat_table) {
        if (corr) fast_velocity<T, true, true>(d, A);
        else fast_velocity<T, false, true>(d, A);
    } else {
        if (corr) fast_velocity<T, true, false>(d, A);
        else fast_velocity<T, false, false>(d, A);
    }
}

template void el3_fast_stress<double>(Device &, const El3Args<double> &);
template void el3_fast_stress<float>(Device &, const El3Args<float> &);
template void el3_fast_velocity<double>(Device &, const El3Args<double> &);
template void el3_fast_velocity<float>(Device &, const El3Args<float> &);

}  // namespace sdt
