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
//    Dirichlet faces), as a list of z-marching tiles of 256 cells per plane,
//    one thread per cell like the interior kernels: 256 consecutive
//    elements of a plane over the z slabs and the y strips (rows follow one
//    another; only the row padding leaves threads idle), and over the x
//    strips 256 consecutive cells of the strip, numbered row by row.
//    Their memory
//    variables and CPML coefficient ids are stored compactly in tile order,
//    so those streams are contiguous per wavefront.  Neighbour indices are
//    clamped into the grid so every load is unconditional; nest ranges mask
//    the field stores.
//
// Whole-line stores.  A 128-byte line that a kernel writes only in part costs
// the memory a read of the line besides the write, so the kernels write whole
// lines where they can: every lane of a memory-variable line (16 slots in
// FP64) that holds a cell stores its memory variables (lanes without a cell,
// or outside the nest ranges that read a variable, store into slots no update
// reads; lines without a cell are skipped), and the interior kernels store
// the shell cell next to their last column again, with the value the shell
// kernel of the same pass gave it, so that the line holding a row's last
// interior cells is written by one store (280^3 FP64: 1.894 -> 1.758 ms per
// step; without the first, the 12-wide x-strip wavefronts wrote a partial
// line per memory variable).
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
    const bool edge = A.edge && i == A.box[1] + 1 && j <= A.box[3];   // the shell cell sharing the last line
    const bool use = act || edge;
    const int64_t sy = A.pitch, sz = (int64_t)A.pitch * A.ny;
    int64_t c = A.xoff + (int64_t)((use ? i : i0) - 1) + sy * ((use ? j : j0) - 1) + sz * (kb - 1);
    const T rx = A.rd24[0], ry = A.rd24[1], rz = A.rd24[2], dt = A.dt;
    T q[3][5];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int s = 0; s < 4; ++s) q[m][s] = A.v[m][c + (s - 2) * sz];
    double nrm = 0.0;   // largest squared magnitude (sqrt is monotone and correctly rounded)
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
        T so[6];
#pragma unroll
        for (int s = 0; s < 6; ++s) so[s] = ldnt(&A.s[s][c]);
        T sxx, syy, szz, sxy, sxz, syz;
        {   // L1: D+x, D-y, D-z
            const T yz = dym[2] + dzm[1], xz = dzm[0] + dxp[2], xy = dxp[1] + dym[0];
            sxx = supk(so[0], K, 0, 1, 2, 3, 4, 5, dxp[0], dym[1], dzm[2], yz, xz, xy, dt, ig[0]);
            syy = supk(so[1], K, 1, 6, 7, 8, 9, 10, dxp[0], dym[1], dzm[2], yz, xz, xy, dt, ig[1]);
            szz = supk(so[2], K, 2, 7, 11, 12, 13, 14, dxp[0], dym[1], dzm[2], yz, xz, xy, dt, ig[2]);
        }
        {   // L2: D-x, D+y, D-z
            const T yz = dzm[1] + dyp[2], xz = dzm[0] + dxm[2], xy = dxm[1] + dyp[0];
            sxy = supk(so[3], K, 5, 10, 14, 17, 19, 20, dxm[0], dyp[1], dzm[2], yz, xz, xy, dt, ig[3]);
        }
        {   // L3: D-x, D-y, D+z
            const T yz = dzp[1] + dym[2], xz = dzp[0] + dxm[2], xy = dxm[1] + dym[0];
            sxz = supk(so[4], K, 4, 9, 13, 16, 18, 19, dxm[0], dym[1], dzp[2], yz, xz, xy, dt, ig[4]);
        }
        {   // L4: D+x, D+y, D+z
            const T yz = dzp[1] + dyp[2], xz = dzp[0] + dxp[2], xy = dxp[1] + dyp[0];
            syz = supk(so[5], K, 3, 8, 12, 15, 16, 17, dxp[0], dyp[1], dzp[2], yz, xz, xy, dt, ig[5]);
        }
        if (use) {   // one store per stream, so the edge lane's value joins the row's line
            stnt(&A.s[0][c], act ? sxx : so[0]); stnt(&A.s[1][c], act ? syy : so[1]);
            stnt(&A.s[2][c], act ? szz : so[2]); stnt(&A.s[3][c], act ? sxy : so[3]);
            stnt(&A.s[4][c], act ? sxz : so[4]); stnt(&A.s[5][c], act ? syz : so[5]);
        }
        if (act) {
            const bool fix = c == A.sfix[0] || c == A.sfix[1] || c == A.sfix[2] || c == A.sfix[3] ||
                             pw_face(A.pwf, i, j, k);
            const double a0 = sxx, a1 = syy, a2 = szz, a3 = sxy, a4 = syz, a5 = sxz;
            const double n2 = a0 * a0 + a1 * a1 + a2 * a2 + 2.0 * (a3 * a3 + a4 * a4 + a5 * a5);
            nrm = fix ? nrm : fmax_nan(nrm, n2);
        }
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            q[m][0] = q[m][1]; q[m][1] = q[m][2]; q[m][2] = q[m][3]; q[m][3] = q[m][4];
        }
    }
    block_max_to(&A.snorm[*A.d_it - 1], sqrt(nrm));
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
    const bool edge = A.edge && i == A.box[1] + 1 && j <= A.box[3];   // the shell cell sharing the last line
    const bool use = act || edge;
    const int64_t sy = A.pitch, sz = (int64_t)A.pitch * A.ny;
    int64_t c = A.xoff + (int64_t)((use ? i : i0) - 1) + sy * ((use ? j : j0) - 1) + sz * (kb - 1);
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
    double nrm = 0.0;   // largest squared magnitude (sqrt is monotone and correctly rounded)
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
        const T vx0 = ldnt(&A.v[0][c]), vy0 = ldnt(&A.v[1][c]), vz0 = ldnt(&A.v[2][c]);
        T vx = vx0, vy = vy0, vz = vz0;
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
        if (use) {   // one store per component, so the edge lane's value joins the row's line
            stnt(&A.v[0][c], act ? vx : vx0); stnt(&A.v[1][c], act ? vy : vy0); stnt(&A.v[2][c], act ? vz : vz0);
        }
        if (act) {
            const bool fix = c == A.vfix[0] || c == A.vfix[1] || c == A.vfix[2] || pw_face(A.pwf, i, j, k);
            const double a0 = vx, a1 = vy, a2 = vz;
            nrm = fix ? nrm : fmax_nan(nrm, a0 * a0 + a1 * a1 + a2 * a2);
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
    block_max_to(&A.vnorm[*A.d_it - 1], sqrt(nrm));
}

// --------------------------------------------------------------------- shell
// The CPML shell (the cells outside the box: the z slabs, y strips and x
// strips, Dirichlet faces included) as a list of z-marching tiles of
// kShellThreads cells per plane, one thread per cell: 256 consecutive
// elements of a plane over the z slabs and y strips (every wavefront reads
// whole 128-byte lines, and only the row padding leaves threads idle: 280^3
// FP64 1.732 -> 1.705 ms per step against 64 x 4 tiles, whose last tile of a
// row was half empty), and over the x strips 256 consecutive cells of the
// strip numbered row by row (a wavefront on about five rows of a 12- or
// 13-cell strip, no thread idle: 1.697 -> 1.679 ms per step against tiles of
// whole rows per wavefront, which left 4 or 12 threads idle).  A thread keeps
// its column's z stencil planes in registers, issues every load of a plane
// before its first store, and updates all of its cell's memory variables and
// fields itself -- no LDS and
// no barriers, so each wavefront keeps a plane's loads in flight like the
// interior kernels.  Memory variables are stored in tile-plane blocks
// (stream q of slot s at ((s / 256) * snq + q) * 256 + s % 256, slot = tile
// base + (k - k0) * 256 + thread), so a wavefront's access of one stream is
// 64 consecutive values.
// Lanes outside the grid or the tile's region compute on clamped indices and
// store no field; neighbour indices are clamped into the grid, and the nest
// ranges of cpmlfdtd.f08 mask the field stores.
// the region test of host::shell_region (kinds 0 z slab, 1 / 3 y strip below
// / above the box, 2 / 4 x strip left / right of it)
__device__ __forceinline__ bool shell_lane(int kind, const int (&b)[6], int i, int j, int nx, int ny)
{
    if (i < 1 || i > nx || j < 1 || j > ny) return false;
    if (kind == 0) return true;
    if (kind == 1) return j < b[2];
    if (kind == 3) return j > b[3];
    const bool rows = j >= b[2] && j <= b[3];
    return kind == 2 ? rows && i < b[0] : rows && i > b[1];
}

struct ShellCell {
    bool on;
    int i, j;                                     // clamped into the grid
    int64_t col;                                  // element of (i, j, 1)
    int xm2, xm1, xp1, xp2, ym2, ym1, yp1, yp2;   // neighbour offsets clamped into the grid (elements)
};

__device__ __forceinline__ ShellCell shell_cell(const El3Tile &tl, const int (&box)[6], int nx, int ny, int64_t sy,
                                                int64_t xoff)
{
    ShellCell L;
    int i0, j0;   // host::tile_lane
    if (tl.tw == 0) {   // z slab / y strip: consecutive elements of the plane
        const int e = tl.x0 + (int)threadIdx.x;
        j0 = e / tl.pitch + 1;
        i0 = e - (j0 - 1) * tl.pitch - tl.xoff + 1;
        L.on = shell_lane(tl.kind, box, i0, j0, nx, ny);
    } else {            // x strip: the strip's cells row by row, every thread on one
        const int w = -tl.tw, n = tl.n0 + (int)threadIdx.x, r = n / w;
        i0 = tl.x0 + (n - r * w);
        j0 = tl.y0 + r;
        L.on = shell_lane(tl.kind, box, i0, j0, nx, ny);
    }
    const int i = min(max(i0, 1), nx), j = min(max(j0, 1), ny);
    L.i = i;
    L.j = j;
    L.xm2 = max(i - 2, 1) - i; L.xm1 = max(i - 1, 1) - i; L.xp1 = min(i + 1, nx) - i; L.xp2 = min(i + 2, nx) - i;
    L.ym2 = (max(j - 2, 1) - j) * (int)sy; L.ym1 = (max(j - 1, 1) - j) * (int)sy;
    L.yp1 = (min(j + 1, ny) - j) * (int)sy; L.yp2 = (min(j + 2, ny) - j) * (int)sy;
    L.col = xoff + (int64_t)(i - 1) + sy * (j - 1);
    return L;
}

// Element of stream q of shell slot `slot`: the snq streams of a tile-plane
// (256 slots) are stored together as one block, stream after stream, so that
// a tile's memory-variable traffic of a plane is one contiguous region.
template <typename T>
__device__ __forceinline__ int64_t shell_at(const El3Args<T> &A, int64_t slot, int q)
{
    static_assert(kShellThreads == 256, "tile-plane blocks of 256 slots");
    return ((slot >> 8) * A.snq + q) * kShellThreads + (slot & (kShellThreads - 1));
}

// CPML coefficients a, b, 1/kappa, a_half, b_half, 1/kappa_half of a shell
// slot: from the dictionary (one-byte id, cached rows of 8) or from the six
// coefficient streams after the memory variables.
template <typename T, bool CTAB>
__device__ __forceinline__ void shell_coefs(const El3Args<T> &A, int64_t slot, T (&cf)[6])
{
    if (CTAB) {
        const T *row = A.ctab + 8 * (int)A.cid[slot];
#pragma unroll
        for (int r = 0; r < 6; ++r) cf[r] = row[r];
    } else {
#pragma unroll
        for (int r = 0; r < 6; ++r) cf[r] = ldnt(&A.shell[shell_at(A, slot, S_CF + r)]);
    }
}

template <typename T>
__device__ __forceinline__ T cpml_upd(T d, T psi, T a, T b, T ik, T &psi_out)
{
    psi_out = b * psi + a * d;
    return d * ik + psi_out;
}

// Row accessors of a cell's stiffness entries (0..20) and attenuation
// divisors (21..26), read where they are used.
template <typename T>
struct ScalarRow {   // wave-uniform table row, scalar loads
    const_row<T> p;
    __device__ T operator[](int q) const { return p[q]; }
};
template <typename T>
struct LaneRow {     // per-lane table row (cached)
    const T *p;
    __device__ T operator[](int q) const { return p[q]; }
};
template <typename T>
struct DenseRow {    // per-cell arrays (C: the 21 stiffness arrays, ig: the six divisor arrays)
    const T *const *C;
    const T *const *ig;
    int64_t c;
    int ig_same;
    __device__ T operator[](int q) const { return q < 21 ? ldnt(&C[q][c]) : ldnt(&ig[ig_same ? 0 : q - 21][c]); }
};
template <typename T>
struct IsoRow {      // isotropic-form per-cell arrays c11, c12, c44
    const T *const *ig;
    int64_t c;
    int ig_same;
    T c11, c12, c44;
    __device__ T operator[](int q) const
    {
        return q < 21 ? iso_entry(q, c11, c12, c44) : ldnt(&ig[ig_same ? 0 : q - 21][c]);
    }
};

// (sigma + (sum_slot C[row][slot] d_slot) dt) * 1/(1 + gamma dt), slots in the
// reference's order dxx, dyy, dzz, yz, xz, xy
template <typename T, typename R>
__device__ __forceinline__ T supr(T s, const R &K, int i0, int i1, int i2, int i3, int i4, int i5, T dxx, T dyy,
                                  T dzz, T yz, T xz, T xy, T dt, int ig)
{
    return (s + (K[i0] * dxx + K[i1] * dyy + K[i2] * dzz + K[i3] * yz + K[i4] * xz + K[i5] * xy) * dt) * K[ig];
}

// The six stress updates of one shell cell from its 21 corrected
// derivatives D[component][kind]; masked by the nest ranges.
template <typename T, typename R>
__device__ __forceinline__ void shell_stress_update(const R &K, const T (&D)[3][K_NKIND], const T (&so)[6], T dt,
                                                    bool L1, bool L2, bool L3, bool L4, T (&sn)[6])
{
    {   // L1: D+x half, D-y full, D-z full
        const T dxx = D[0][K_XPH], dyy = D[1][K_YMF], dzz = D[2][K_ZMF];
        const T yz = D[1][K_ZMF] + D[2][K_YMF], xz = D[0][K_ZMF] + D[2][K_XPH], xy = D[1][K_XPH] + D[0][K_YMF];
        sn[0] = L1 ? supr(so[0], K, 0, 1, 2, 3, 4, 5, dxx, dyy, dzz, yz, xz, xy, dt, kTabIg + 0) : so[0];
        sn[1] = L1 ? supr(so[1], K, 1, 6, 7, 8, 9, 10, dxx, dyy, dzz, yz, xz, xy, dt, kTabIg + 1) : so[1];
        sn[2] = L1 ? supr(so[2], K, 2, 7, 11, 12, 13, 14, dxx, dyy, dzz, yz, xz, xy, dt, kTabIg + 2) : so[2];
    }
    {   // L2: D-x full, D+y half, D-z half (vx's uncorrected)
        const T dxx = D[0][K_XMF], dyy = D[1][K_YPH], dzz = D[2][K_ZMH];
        const T yz = D[1][K_ZMH] + D[2][K_YPH], xz = D[0][K_ZMH] + D[2][K_XMF], xy = D[1][K_XMF] + D[0][K_YPH];
        sn[3] = L2 ? supr(so[3], K, 5, 10, 14, 17, 19, 20, dxx, dyy, dzz, yz, xz, xy, dt, kTabIg + 3) : so[3];
    }
    {   // L3: D-x full, D-y full, D+z half
        const T dxx = D[0][K_XMF], dyy = D[1][K_YMF], dzz = D[2][K_ZPH];
        const T yz = D[1][K_ZPH] + D[2][K_YMF], xz = D[0][K_ZPH] + D[2][K_XMF], xy = D[1][K_XMF] + D[0][K_YMF];
        sn[4] = L3 ? supr(so[4], K, 4, 9, 13, 16, 18, 19, dxx, dyy, dzz, yz, xz, xy, dt, kTabIg + 4) : so[4];
    }
    {   // L4: D+x half, D+y half, D+z half
        const T dxx = D[0][K_XPH], dyy = D[1][K_YPH], dzz = D[2][K_ZPH];
        const T yz = D[1][K_ZPH] + D[2][K_YPH], xz = D[0][K_ZPH] + D[2][K_XPH], xy = D[1][K_XPH] + D[0][K_YPH];
        sn[5] = L4 ? supr(so[5], K, 3, 8, 12, 15, 16, 17, dxx, dyy, dzz, yz, xz, xy, dt, kTabIg + 5) : so[5];
    }
}

// Stress pass (cpmlfdtd.f08:1761-1992): per velocity component the seven
// (direction, stencil, coefficient set) derivatives the four nests read,
// each memory variable updated once (stored at once, by every lane: only the
// cells inside the ranges of the nests that read it ever use the stored
// value, cpmlfdtd.f08 updating it on those ranges), then the six stresses of
// the nests whose range holds the cell.  vx's z derivative enters sigma_xy
// uncorrected (1858-1865).
template <typename T, int MAT, bool CTAB>
__global__ void __launch_bounds__(kShellThreads) el3_shell_stress_kernel(El3Args<T> A, const El3Tile *tiles)
{
    const El3Tile tl = tiles[(A.remap & 2) ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x];
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const int64_t sy = A.pitch, sz = (int64_t)A.pitch * ny;
    const ShellCell L = shell_cell(tl, A.box, nx, ny, sy, A.xoff);
    const bool mst = line_any<T>(L.on);   // memory-variable lines holding a cell (stored whole)
    const int i = L.i, j = L.j;
    const T rx = A.rd24[0], ry = A.rd24[1], rz = A.rd24[2], dt = A.dt;
    T q[3][5];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int s = 0; s < 4; ++s) q[m][s] = A.v[m][L.col + sz * (min(max(tl.k0 - 2 + s, 1), nz) - 1)];
    constexpr int64_t ns = kShellThreads;   // stream stride inside a tile-plane block
    double nrm = 0.0;   // largest squared magnitude (sqrt is monotone and correctly rounded)
    int64_t slot = tl.base + threadIdx.x;
    for (int k = tl.k0; k <= tl.k1; ++k, slot += kShellThreads) {
        const int64_t c = L.col + sz * (k - 1);
        const int64_t zp2 = sz * (min(k + 2, nz) - k);
        const bool L1 = L.on && i >= 2 && i <= nx - 2 && j >= 3 && j <= ny - 1 && k >= 3 && k <= nz - 1;
        const bool L2 = L.on && i >= 3 && i <= nx - 1 && j >= 2 && j <= ny - 2 && k >= 3 && k <= nz - 1;
        const bool L3 = L.on && i >= 3 && i <= nx - 1 && j >= 3 && j <= ny - 1 && k >= 2 && k <= nz - 2;
        const bool L4 = L.on && i >= 2 && i <= nx - 2 && j >= 2 && j <= ny - 2 && k >= 2 && k <= nz - 2;
        T cf[6];
        shell_coefs<T, CTAB>(A, slot, cf);
        const T a = cf[0], b = cf[1], ik = cf[2], ah = cf[3], bh = cf[4], ikh = cf[5];
        T so[6];
#pragma unroll
        for (int s = 0; s < 6; ++s) so[s] = ldnt(&A.s[s][c]);
        const int mid = MAT == kMatTable ? (int)A.mid[c] : 0;
        T ic11 = T(0), ic12 = T(0), ic44 = T(0);
        if (MAT == kMatIso) { ic11 = ldnt(&A.C[0][c]); ic12 = ldnt(&A.C[1][c]); ic44 = ldnt(&A.C[15][c]); }
        T *ps = A.shell + shell_at(A, slot, 0);
        T D[3][K_NKIND];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const T *v = A.v[m];
            q[m][4] = v[c + zp2];
            const T v0 = q[m][2];
            const T xm2 = v[c + L.xm2], xm1 = v[c + L.xm1], xp1 = v[c + L.xp1], xp2 = v[c + L.xp2];
            const T ym2 = v[c + L.ym2], ym1 = v[c + L.ym1], yp1 = v[c + L.yp1], yp2 = v[c + L.yp2];
            T pin[K_NKIND];
#pragma unroll
            for (int kd = 0; kd < K_NKIND; ++kd)
                pin[kd] = (m == 0 && kd == K_ZMH) || !L.on ? T(0) : ldnt(&ps[(m * K_NKIND + kd) * ns]);
            const T dxp = dp4(xm1, v0, xp1, xp2, rx), dxm = dm4(xm2, xm1, v0, xp1, rx);
            const T dyp = dp4(ym1, v0, yp1, yp2, ry), dym = dm4(ym2, ym1, v0, yp1, ry);
            const T dzp = dp4(q[m][1], v0, q[m][3], q[m][4], rz), dzm = dm4(q[m][0], q[m][1], v0, q[m][3], rz);
            T P[K_NKIND];
            D[m][K_XPH] = cpml_upd(dxp, pin[K_XPH], ah, bh, ikh, P[K_XPH]);
            D[m][K_XMF] = cpml_upd(dxm, pin[K_XMF], a, b, ik, P[K_XMF]);
            D[m][K_YMF] = cpml_upd(dym, pin[K_YMF], a, b, ik, P[K_YMF]);
            D[m][K_YPH] = cpml_upd(dyp, pin[K_YPH], ah, bh, ikh, P[K_YPH]);
            D[m][K_ZMF] = cpml_upd(dzm, pin[K_ZMF], a, b, ik, P[K_ZMF]);
            if (m == 0) { D[m][K_ZMH] = dzm; P[K_ZMH] = T(0); }
            else D[m][K_ZMH] = cpml_upd(dzm, pin[K_ZMH], ah, bh, ikh, P[K_ZMH]);
            D[m][K_ZPH] = cpml_upd(dzp, pin[K_ZPH], ah, bh, ikh, P[K_ZPH]);
            if (mst) {
                stnt(&ps[(m * K_NKIND + K_XPH) * ns], P[K_XPH]);
                stnt(&ps[(m * K_NKIND + K_XMF) * ns], P[K_XMF]);
                stnt(&ps[(m * K_NKIND + K_YMF) * ns], P[K_YMF]);
                stnt(&ps[(m * K_NKIND + K_YPH) * ns], P[K_YPH]);
                stnt(&ps[(m * K_NKIND + K_ZMF) * ns], P[K_ZMF]);
                if (m != 0) stnt(&ps[(m * K_NKIND + K_ZMH) * ns], P[K_ZMH]);
                stnt(&ps[(m * K_NKIND + K_ZPH) * ns], P[K_ZPH]);
            }
            q[m][0] = q[m][1]; q[m][1] = q[m][2]; q[m][2] = q[m][3]; q[m][3] = q[m][4];
        }
        T sn[6];
        if constexpr (MAT == kMatTable) {
            const int m0 = __builtin_amdgcn_readfirstlane(mid);
            if (__all(mid == m0))
                shell_stress_update(ScalarRow<T>{(const_row<T>)(A.tab) + m0 * kTabRow}, D, so, dt, L1, L2, L3, L4, sn);
            else
                shell_stress_update(LaneRow<T>{A.tab + mid * kTabRow}, D, so, dt, L1, L2, L3, L4, sn);
        } else if constexpr (MAT == kMatIso) {
            shell_stress_update(IsoRow<T>{A.ig, c, A.ig_same, ic11, ic12, ic44}, D, so, dt, L1, L2, L3, L4, sn);
        } else {
            shell_stress_update(DenseRow<T>{A.C, A.ig, c, A.ig_same}, D, so, dt, L1, L2, L3, L4, sn);
        }
        if (tl.tw < 0) {
            // x strip (uniform per workgroup): ordinary stores, so the row pieces
            // of neighbouring rows may meet in L2 before they go to memory
            // (280^3 FP64 1.692 -> 1.682 ms per step against non-temporal ones)
            if (L1) { A.s[0][c] = sn[0]; A.s[1][c] = sn[1]; A.s[2][c] = sn[2]; }
            if (L2) A.s[3][c] = sn[3];
            if (L3) A.s[4][c] = sn[4];
            if (L4) A.s[5][c] = sn[5];
        } else {
            if (L1) { stnt(&A.s[0][c], sn[0]); stnt(&A.s[1][c], sn[1]); stnt(&A.s[2][c], sn[2]); }
            if (L2) stnt(&A.s[3][c], sn[3]);
            if (L3) stnt(&A.s[4][c], sn[4]);
            if (L4) stnt(&A.s[5][c], sn[5]);
        }
        if (L.on) {
            const bool fix = c == A.sfix[0] || c == A.sfix[1] || c == A.sfix[2] || c == A.sfix[3] ||
                             pw_face(A.pwf, i, j, k);
            const double a0 = sn[0], a1 = sn[1], a2 = sn[2], a3 = sn[3], a4 = sn[5], a5 = sn[4];
            const double n2 = a0 * a0 + a1 * a1 + a2 * a2 + 2.0 * (a3 * a3 + a4 * a4 + a5 * a5);
            nrm = fix ? nrm : fmax_nan(nrm, n2);
        }
    }
    block_max_to(&A.snorm[*A.d_it - 1], sqrt(nrm));
}

// Velocity pass (cpmlfdtd.f08:2014-2038 vx, 2055-2080 vy, 2098-2127 vz): the
// three components of the cell, each from three stress differences with
// their memory variables and the averaged densities; Dirichlet faces
// (2226-2248) give 0.
template <typename T, bool CORR, bool TAB, bool CTAB>
__global__ void __launch_bounds__(kShellThreads) el3_shell_velocity_kernel(El3Args<T> A, const El3Tile *tiles)
{
    const El3Tile tl = tiles[(A.remap & 2) ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x];
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const int64_t sy = A.pitch, sz = (int64_t)A.pitch * ny;
    const ShellCell L = shell_cell(tl, A.box, nx, ny, sy, A.xoff);
    const bool mst = line_any<T>(L.on);   // memory-variable lines holding a cell (stored whole)
    const int i = L.i, j = L.j;
    const T rx = A.rd24[0], ry = A.rd24[1], rz = A.rd24[2], dt = A.dt;
    const int dc = A.density_code;
    const T *sxx = A.s[0], *syy = A.s[1], *szz = A.s[2], *sxy = A.s[3], *sxz = A.s[4], *syz = A.s[5];
    auto zc = [&](int k) { return sz * (min(max(k, 1), nz) - 1); };
    T qxz[5], qyz[5], qzz[5], qr[3];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        qxz[s] = sxz[L.col + zc(tl.k0 - 2 + s)];
        qyz[s] = syz[L.col + zc(tl.k0 - 2 + s)];
        qzz[s] = szz[L.col + zc(tl.k0 - 2 + s)];
    }
    qr[0] = rho_at<T, TAB>(A, L.col + zc(tl.k0 - 1));
    qr[1] = rho_at<T, TAB>(A, L.col + zc(tl.k0));
    RhoPre<T> rz_next = rho_pre<T, TAB>(A, L.col + zc(tl.k0 + 1)),
              rxm_next = rho_pre<T, TAB>(A, L.col + zc(tl.k0) + L.xm1),
              rxp_next = rho_pre<T, TAB>(A, L.col + zc(tl.k0) + L.xp1);
    constexpr int64_t ns = kShellThreads;   // stream stride inside a tile-plane block
    double nrm = 0.0;   // largest squared magnitude (sqrt is monotone and correctly rounded)
    int64_t slot = tl.base + threadIdx.x;
    for (int k = tl.k0; k <= tl.k1; ++k, slot += kShellThreads) {
        const int64_t c = L.col + sz * (k - 1);
        qxz[4] = sxz[L.col + zc(k + 2)];
        qyz[4] = syz[L.col + zc(k + 2)];
        qzz[4] = szz[L.col + zc(k + 2)];
        qr[2] = rho_get<T, TAB>(A, rz_next);
        const T r = qr[1], rxm = rho_get<T, TAB>(A, rxm_next), rxp = rho_get<T, TAB>(A, rxp_next);
        const T sxx_m2 = sxx[c + L.xm2], sxx_m1 = sxx[c + L.xm1], sxx_0 = sxx[c], sxx_p1 = CORR ? sxx[c + L.xp1] : T(0);
        const T sxy_ym2 = sxy[c + L.ym2], sxy_ym1 = sxy[c + L.ym1], sxy_0 = sxy[c],
                sxy_yp1 = CORR ? sxy[c + L.yp1] : T(0);
        const T sxy_xm1 = CORR ? sxy[c + L.xm1] : T(0), sxy_xp1 = sxy[c + L.xp1], sxy_xp2 = sxy[c + L.xp2];
        const T syy_m1 = CORR ? syy[c + L.ym1] : T(0), syy_0 = syy[c], syy_p1 = syy[c + L.yp1],
                syy_p2 = syy[c + L.yp2];
        const T sxz_xm1 = CORR ? sxz[c + L.xm1] : T(0), sxz_xp1 = sxz[c + L.xp1], sxz_xp2 = sxz[c + L.xp2];
        const T syz_ym2 = syz[c + L.ym2], syz_ym1 = syz[c + L.ym1], syz_yp1 = CORR ? syz[c + L.yp1] : T(0);
        T cf[6];
        shell_coefs<T, CTAB>(A, slot, cf);
        const T a = cf[0], b = cf[1], ik = cf[2], ah = cf[3], bh = cf[4], ikh = cf[5];
        T *ps = A.shell + shell_at(A, slot, 0);
        T pin[9];
#pragma unroll
        for (int s = 0; s < 9; ++s) pin[s] = L.on ? ldnt(&ps[(M_SXX_X + s) * ns]) : T(0);
        T vx = ldnt(&A.v[0][c]), vy = ldnt(&A.v[1][c]), vz = ldnt(&A.v[2][c]);
        const bool face = i == 1 || i == nx || j == 1 || j == ny || k == 1 || k == nz;
        const T ryy = scalar_mean(r, r, dc);
        T p[9];
        T nvx, nvy, nvz;
        {   // vx: T-x sigma_xx, T-y sigma_xy, T-z sigma_xz, all half-set recursions
            const T d1 = cpml_upd(tminus5(sxx_m2, sxx_m1, sxx_0, sxx_p1, CORR) * rx, pin[0], ah, bh, ikh, p[0]);
            const T d2 = cpml_upd(tminus5(sxy_ym2, sxy_ym1, sxy_0, sxy_yp1, CORR) * ry, pin[1], ah, bh, ikh, p[1]);
            const T d3 = cpml_upd(tminus5(qxz[0], qxz[1], qxz[2], qxz[3], CORR) * rz, pin[2], ah, bh, ikh, p[2]);
            nvx = vx + (d1 / scalar_mean(r, rxm, dc) + d2 / ryy + d3 / scalar_mean(r, qr[0], dc)) * dt;
        }
        {   // vy: T+x sigma_xy (half), T+y sigma_yy (half), T-z sigma_yz (full)
            const T d1 = cpml_upd(tplus5(sxy_xm1, sxy_0, sxy_xp1, sxy_xp2, CORR) * rx, pin[3], ah, bh, ikh, p[3]);
            const T d2 = cpml_upd(tplus5(syy_m1, syy_0, syy_p1, syy_p2, CORR) * ry, pin[4], ah, bh, ikh, p[4]);
            const T d3 = cpml_upd(tminus5(qyz[0], qyz[1], qyz[2], qyz[3], CORR) * rz, pin[5], a, b, ik, p[5]);
            nvy = vy + (d1 / scalar_mean(r, rxp, dc) + d2 / ryy + d3 / scalar_mean(r, qr[0], dc)) * dt;
        }
        {   // vz: T+x sigma_xz (half), T-y sigma_yz (full), T+z sigma_zz (half)
            const T d1 = cpml_upd(tplus5(sxz_xm1, qxz[2], sxz_xp1, sxz_xp2, CORR) * rx, pin[6], ah, bh, ikh, p[6]);
            const T d2 = cpml_upd(tminus5(syz_ym2, syz_ym1, qyz[2], syz_yp1, CORR) * ry, pin[7], a, b, ik, p[7]);
            const T d3 = cpml_upd(tplus5(qzz[1], qzz[2], qzz[3], qzz[4], CORR) * rz, pin[8], ah, bh, ikh, p[8]);
            nvz = vz + (d1 / scalar_mean(r, rxp, dc) + d2 / ryy + d3 / scalar_mean(r, qr[2], dc)) * dt;
        }
        const bool inx = L.on && !face && i >= 3 && j >= 3 && k >= 3;
        const bool iny = L.on && !face && i <= nx - 2 && j <= ny - 2 && k >= 3;
        const bool inz = L.on && !face && i <= nx - 2 && j >= 3 && k <= nz - 2;
        // memory variables from every lane of a line holding a cell (whole
        // lines); only the cells of the update ranges below read theirs back
        if (mst) {
            stnt(&ps[M_SXX_X * ns], p[0]); stnt(&ps[M_SXY_Y * ns], p[1]); stnt(&ps[M_SXZ_Z * ns], p[2]);
            stnt(&ps[M_SXY_X * ns], p[3]); stnt(&ps[M_SYY_Y * ns], p[4]); stnt(&ps[M_SYZ_Z * ns], p[5]);
            stnt(&ps[M_SXZ_X * ns], p[6]); stnt(&ps[M_SYZ_Y * ns], p[7]); stnt(&ps[M_SZZ_Z * ns], p[8]);
        }
        if (inx) vx = nvx;
        if (iny) vy = nvy;
        if (inz) vz = nvz;
        if (face) { vx = T(0); vy = T(0); vz = T(0); }
        if (L.on && tl.tw < 0) {   // x strip: ordinary stores, as in the stress kernel
            A.v[0][c] = vx; A.v[1][c] = vy; A.v[2][c] = vz;
        } else if (L.on) {
            stnt(&A.v[0][c], vx); stnt(&A.v[1][c], vy); stnt(&A.v[2][c], vz);
        }
        if (L.on) {
            const bool fix = c == A.vfix[0] || c == A.vfix[1] || c == A.vfix[2] || pw_face(A.pwf, i, j, k);
            const double a0 = vx, a1 = vy, a2 = vz;
            nrm = fix ? nrm : fmax_nan(nrm, a0 * a0 + a1 * a1 + a2 * a2);
        }
        if (k < tl.k1) {
            rz_next = rho_pre<T, TAB>(A, L.col + zc(k + 2));
            rxm_next = rho_pre<T, TAB>(A, c + sz + L.xm1);
            rxp_next = rho_pre<T, TAB>(A, c + sz + L.xp1);
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            qxz[s] = qxz[s + 1]; qyz[s] = qyz[s + 1]; qzz[s] = qzz[s + 1];
        }
        qr[0] = qr[1]; qr[1] = qr[2];
    }
    block_max_to(&A.vnorm[*A.d_it - 1], sqrt(nrm));
}

// ----------------------------------------------------------------- launchers
// planes per interior workgroup when the z-chunk option is 0: 8, except the
// stress pass with per-cell material arrays (27 more streams per cell), which
// runs faster with the resident workgroups on fewer planes (280^3 FP64: stress
// pass 2.131 ms at 3 planes against 2.189 at 8; the velocity passes and the
// table layout's stress pass are fastest at 8)
static int el3_kc(const Device &d, bool dense_stress = false)
{
    return d.zchunk > 0 ? d.zchunk : dense_stress ? 3 : 8;
}

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
    const int kc = el3_kc(d, MAT == kMatDense), ty = el3_ty(d);
    const hipStream_t ss = shell_stream_begin(d, 0);
    if (d.ntiles > 0) {
        if (A.ctab)
            hipLaunchKernelGGL((el3_shell_stress_kernel<T, MAT, true>), dim3(d.ntiles), dim3(kShellThreads), 0, ss,
                               A, (const El3Tile *)d.tiles);
        else
            hipLaunchKernelGGL((el3_shell_stress_kernel<T, MAT, false>), dim3(d.ntiles), dim3(kShellThreads), 0, ss,
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
    if (d.ntiles > 0) {
        if (A.ctab)
            hipLaunchKernelGGL((el3_shell_velocity_kernel<T, CORR, TAB, true>), dim3(d.ntiles), dim3(kShellThreads), 0,
                               ss, A, (const El3Tile *)d.tiles);
        else
            hipLaunchKernelGGL((el3_shell_velocity_kernel<T, CORR, TAB, false>), dim3(d.ntiles), dim3(kShellThreads),
                               0, ss, A, (const El3Tile *)d.tiles);
    }
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
