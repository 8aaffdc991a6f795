// el3_common.h -- shared declarations of the 3D elastic kernels (el3.hip,
// el3_fast.hip).
#pragma once

#include "host_layout.h"
#include "sdt_device.h"
#include "sdt_internal.h"

namespace sdt {

// stress-side recursion slots: m[comp * 7 + kind]
enum { K_XPH = 0, K_XMF, K_YMF, K_YPH, K_ZMF, K_ZMH, K_ZPH, K_NKIND };
// velocity-side slots follow: 21 + q
enum { M_SXX_X = 21, M_SXY_Y, M_SXZ_Z, M_SXY_X, M_SYY_Y, M_SYZ_Z, M_SXZ_X, M_SYZ_Y, M_SZZ_Z, M_TOTAL };
// streams of a shell tile-plane block: the M_TOTAL memory variables, then the
// CPML coefficients a, b, 1/kappa, a_half, b_half, 1/kappa_half
enum { S_CF = M_TOTAL, S_NQ = M_TOTAL + 6 };
// planes per z-marching tile of the CPML shell when the z-chunk option is 0
constexpr int kEl3ShellChunk = 3;

template <typename T>
struct El3Args {
    T *v[3];        // vx vy vz
    T *s[6];        // sxx syy szz sxy sxz syz
    const T *C[21]; // c11 c12 c13 c14 c15 c16 c22 c23 c24 c25 c26 c33 c34 c35 c36 c44 c45 c46 c55 c56 c66
    const T *rho;
    const T *ig[6]; // 1/(1+gamma*dt): x y z xy xz yz
    // material table (when the model has at most 256 distinct materials):
    // mid[c] indexes rows of tab, row = c11..c66 (21), 1/(1+gamma*dt) x y z
    // xy xz yz (6), rho, 4 unused
    const uint8_t *mid;
    const T *tab;
    T *shell;       // CPML-shell storage: per tile-plane (256 slots) a block of snq streams of 256 values
    int snq;        // streams: M_TOTAL memory variables (+ 6 coefficient streams without ctab)
    int64_t nslots;
    // CPML coefficient dictionary of the shell (at most 256 distinct
    // (a, b, 1/kappa, a_half, b_half, 1/kappa_half) tuples): id per slot and
    // rows of 8 (6 used); ctab == nullptr keeps the coefficients in the block
    const uint8_t *cid;
    const T *ctab;
    const uint8_t *flags;
    int box[6];
    int nx, ny, nz, tiles_x, tiles_y, pitch;
    int64_t xoff;
    T rd24[3];      // 1/(24 dx), 1/(24 dy), 1/(24 dz)
    T dt;
    int density_code, corrected;
    int remap;      // bit 0: XCD-contiguous block order of the interior kernels, bit 1: of the shell kernels
    int ig_same;    // 1: the six ig arrays hold equal values (one array, loaded once per cell)
    int edge;       // 1: the shell kernels of a pass finish before its interior kernels start, and the
                    // interior kernels store column box[1] + 1 again (the shell's new value) so that
                    // the line holding their last column is written whole
    const int *d_it;
    unsigned long long *vnorm, *snorm;
    int64_t vfix[3], sfix[4];
    PwFaces pwf;    // plane-wave cells: skipped by the fused norms like vfix / sfix
};

// psi = b*psi + a*d ; d = d/kappa + psi  (e.g. cpmlfdtd.f08:1777, 1787)
template <typename T>
__device__ __forceinline__ T cpml_fix3(T d, T *psi_p, T a, T b, T ik)
{
    const T psi = b * ldnt(psi_p) + a * d;
    stnt(psi_p, psi);
    return d * ik + psi;
}

// forward / backward 4th-order differences (cpmlfdtd.f08:1765-1775)
template <typename T>
__device__ __forceinline__ T dplus(const T *f, int64_t c, int64_t s)
{
    return T(27) * (f[c + s] - f[c]) - (f[c + 2 * s] - f[c - s]);
}
template <typename T>
__device__ __forceinline__ T dminus(const T *f, int64_t c, int64_t s)
{
    return T(27) * (f[c] - f[c - s]) - (f[c + s] - f[c - 2 * s]);
}

template <typename T>
__device__ __forceinline__ T stress_update(T s, const T *const *C, int64_t c, int i0, int i1, int i2, int i3,
                                           int i4, int i5, T dx_, T dy_, T dz_, T yz, T xz, T xy, T dt,
                                           T ig)
{
    return (s + (ldnt(&C[i0][c]) * dx_ + ldnt(&C[i1][c]) * dy_ + ldnt(&C[i2][c]) * dz_ +
                 ldnt(&C[i3][c]) * yz + ldnt(&C[i4][c]) * xz + ldnt(&C[i5][c]) * xy) * dt) * ig;
}

// 3-term (reference) / 4-term (corrected) velocity differences from
// consecutive samples f(-2..+2) (cpmlfdtd.f08:2022-2024, 2063-2066, 2107-2111)
template <typename T>
__device__ __forceinline__ T tminus5(T fm2, T fm1, T f0, T fp1, int corr)
{
    return corr ? T(27) * (f0 - fm1) - (fp1 - fm2) : T(27) * f0 - T(27) * fm1 + fm2;
}
template <typename T>
__device__ __forceinline__ T tplus5(T fm1, T f0, T fp1, T fp2, int corr)
{
    return corr ? T(27) * (fp1 - f0) - (fp2 - fm1) : -T(27) * f0 + T(27) * fp1 - fp2;
}

// Where the stress kernels read a cell's stiffness from: all 21 per-cell
// arrays, the material table, or the isotropic arrays c11, c12, c44.
enum { kMatDense = 0, kMatTable = 1, kMatIso = 2 };

constexpr int kTabRow = kEl3TabRow;   // material-table row length (elements)
constexpr int kTabIg = 21, kTabRho = 27;
constexpr int kMaxMaterials = 256;

// Interior (branch-free, z-marching) and CPML-shell kernels, el3_fast.hip.
template <typename T>
void el3_fast_stress(Device &d, const El3Args<T> &A);
template <typename T>
void el3_fast_velocity(Device &d, const El3Args<T> &A);

}  // namespace sdt
