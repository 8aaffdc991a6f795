// el25.hip -- 2.5D anisotropic elastic CPML-FDTD (seismic25,
// cpmlfdtd.f08:600-1337) on gfx950.
//
// The grid is the (x, z) plane; the y dependence is exp(i ky y), so every
// field is complex and every y derivative is i*ky*f = (-ky*im, ky*re)
// (cpmlfdtd.f08:946-948).  Complex arrays are interleaved (re, im) pairs, one
// 16-byte (FP64) or 8-byte (FP32) load per cell.  Material arrays stay real.
//
// Layout and tiling as the 2D elastic solver: padded rows, one thread per cell,
// 64 x 4-cell tiles; tiles entirely inside the identity-CPML box run the body
// without memory variables.  The reference keeps one memory array per nest
// and derivative (30); recursions that share field, direction, stencil and
// coefficient set hold identical values, so 21 arrays are kept, each updated
// once per step on the union of the nests' ranges:
//   stress side  X+H (L1, L4), X-F (L2, L3), Z-F (L1), Z-H (L2), Z+H (L3, L4),
//                three velocity components each                    (15)
//   velocity side dsxx/dx H, dsxz/dz H, dsxy/dx H, dsyz/dz F, dsxz/dx H,
//                dszz/dz H                                          (6)
#include <math.h>

#include <vector>

#include "sdt_device.h"
#include "sdt_internal.h"

namespace sdt {

template <typename T>
struct alignas(2 * sizeof(T)) Cx {
    T re, im;
};

template <typename T>
__device__ __forceinline__ Cx<T> cmk(T re, T im)
{
    Cx<T> z;
    z.re = re;
    z.im = im;
    return z;
}
template <typename T>
__device__ __forceinline__ Cx<T> operator+(Cx<T> a, Cx<T> b) { return cmk(a.re + b.re, a.im + b.im); }
template <typename T>
__device__ __forceinline__ Cx<T> operator*(T s, Cx<T> a) { return cmk(s * a.re, s * a.im); }
// i*ky*f
template <typename T>
__device__ __forceinline__ Cx<T> iky(T ky, Cx<T> f) { return cmk(-ky * f.im, ky * f.re); }
// psi = b*psi + a*d ; d = d/kappa + psi (1/kappa precomputed)
template <typename T>
__device__ __forceinline__ Cx<T> cfix(Cx<T> d, Cx<T> &m, T a, T b, T ik)   // m: the memory variable, updated in place
{
    m.re = b * m.re + a * d.re;
    m.im = b * m.im + a * d.im;
    return cmk(d.re * ik + m.re, d.im * ik + m.im);
}
template <typename T>
__device__ __forceinline__ Cx<T> cfix_mem(Cx<T> d, Cx<T> *psi, int64_t c, T a, T b, T ik)   // psi[c] updated
{
    Cx<T> m = psi[c];
    const Cx<T> r = cfix(d, m, a, b, ik);
    psi[c] = m;
    return r;
}
// 4th-order staggered difference 27(f1 - f0) - (f2 - fm1), scaled
template <typename T>
__device__ __forceinline__ Cx<T> cd4(Cx<T> fm1, Cx<T> f0, Cx<T> f1, Cx<T> f2, T r)
{
    return cmk((T(27) * (f1.re - f0.re) - (f2.re - fm1.re)) * r, (T(27) * (f1.im - f0.im) - (f2.im - fm1.im)) * r);
}
// velocity-from-stress stencils: the reference's 3-term forms
// (cpmlfdtd.f08:1132-1134, 1161-1163, 1189-1191) or the 4-term ones
template <typename T>
__device__ __forceinline__ T tm1(T sm2, T sm1, T s0, T sp1, int corr)
{
    return corr ? T(27) * (s0 - sm1) - (sp1 - sm2) : T(27) * s0 - T(27) * sm1 + sm2;
}
template <typename T>
__device__ __forceinline__ T tp1(T sm1, T s0, T sp1, T sp2, int corr)
{
    return corr ? T(27) * (sp1 - s0) - (sp2 - sm1) : -T(27) * s0 + T(27) * sp1 - sp2;
}
template <typename T>
__device__ __forceinline__ Cx<T> ctm(const Cx<T> *f, int64_t c, int64_t s, int corr, T r)
{
    const Cx<T> a = f[c - 2 * s], b = f[c - s], z = f[c], d = f[c + s];
    return cmk(tm1(a.re, b.re, z.re, d.re, corr) * r, tm1(a.im, b.im, z.im, d.im, corr) * r);
}
template <typename T>
__device__ __forceinline__ Cx<T> ctp(const Cx<T> *f, int64_t c, int64_t s, int corr, T r)
{
    const Cx<T> a = f[c - s], b = f[c], z = f[c + s], d = f[c + 2 * s];
    return cmk(tp1(a.re, b.re, z.re, d.re, corr) * r, tp1(a.im, b.im, z.im, d.im, corr) * r);
}

// memory-variable slots
enum {
    P_XPH = 0,    // + m: D+x, half set        (L1, L4)
    P_XMF = 3,    // D-x, full set              (L2, L3)
    P_ZMF = 6,    // D-z, full set              (L1)
    P_ZMH = 9,    // D-z, half set              (L2)
    P_ZPH = 12,   // D+z, half set              (L3, L4)
    P_VX_X = 15, P_VX_Z, P_VY_X, P_VY_Z, P_VZ_X, P_VZ_Z,
    P_TOTAL = 21
};
// material slots: 21 stiffness, 6 attenuation divisors, 5 densities
enum { M_IG = 21, M_RHO = 27, M_RHO_PI, M_RHO_PJ, M_RHO_MI, M_RHO_MJ, M_TOTAL25 = 32 };

template <typename T>
struct El25Args {
    Cx<T> *v[3];          // vx, vy, vz
    Cx<T> *s[6];          // sxx, syy, szz, sxy, sxz, syz
    const T *mat[M_TOTAL25];
    const T *a, *b, *ik, *ah, *bh, *ikh;
    Cx<T> *m[P_TOTAL];
    const uint8_t *flags;
    int box[6];
    int nx, nz, tiles_x, pitch;
    int64_t xoff;
    T rdx24, rdz24, dt, ky;
    int corrected;
    const int *d_it;
    unsigned long long *vnorm;
    int64_t vfix[3];      // source cells excluded from the fused norm
    PwFaces pwf;          // plane-wave cells, likewise
    // material table (at most 256 distinct (c11..c66, attenuation) tuples):
    // id per padded cell, rows of kEl25TabRow (21 stiffness, 6 divisors)
    const uint8_t *mid;
    const T *tab;
};

// Where the stress update reads a cell's stiffness and attenuation divisors
// from: the 27 per-cell arrays, the material table, or the isotropic arrays
// c11, c12, c44 with per-cell divisors.
enum { kEl25Dense = 0, kEl25Table = 1, kEl25Iso = 2 };

// Row accessors: entry q < 21 is Voigt slot q of the stiffness, 21 + q the
// divisor of nest q (M_IG + q).
template <typename T>
using el25_const_row = __attribute__((address_space(4))) const T *;
template <typename T>
struct El25ScalarRow {   // wave-uniform table row, scalar loads
    el25_const_row<T> p;
    __device__ T operator[](int q) const { return p[q]; }
};
template <typename T>
struct El25LaneRow {     // per-lane table row (cached)
    const T *p;
    __device__ T operator[](int q) const { return p[q]; }
};
template <typename T>
struct El25DenseRow {    // per-cell arrays
    const T *const *M;
    int64_t c;
    __device__ T operator[](int q) const { return M[q][c]; }
};
template <typename T>
struct El25IsoRow {      // isotropic form: c11, c12, c44, per-cell divisors
    const T *const *M;
    int64_t c;
    T c11, c12, c44;
    __device__ T operator[](int q) const { return q < 21 ? iso_entry(q, c11, c12, c44) : M[q][c]; }
};

// sigma = (sigma + (K1 d1 + K2 d2 + K3 d3 + K4 yz + K5 xz + K6 xy) dt) * ig,
// cpmlfdtd.f08:974-987
template <typename T>
__device__ __forceinline__ Cx<T> cstress(Cx<T> s, T k1, T k2, T k3, T k4, T k5, T k6, Cx<T> d1, Cx<T> d2,
                                         Cx<T> d3, Cx<T> yz, Cx<T> xz, Cx<T> xy, T dt, T ig)
{
    const T re = k1 * d1.re + k2 * d2.re + k3 * d3.re + k4 * yz.re + k5 * xz.re + k6 * xy.re;
    const T im = k1 * d1.im + k2 * d2.im + k3 * d3.im + k4 * yz.im + k5 * xz.im + k6 * xy.im;
    return cmk((s.re + re * dt) * ig, (s.im + im * dt) * ig);
}

// The six stress updates of the four nests (cpmlfdtd.f08:974-1113) from the
// cell's derivatives; K is the row accessor of the cell's material.
// Voigt slots of the 21 stiffness entries: c11 0, c12 1, c13 2, c14 3, c15 4,
// c16 5, c22 6, c23 7, c24 8, c25 9, c26 10, c33 11, c34 12, c35 13, c36 14,
// c44 15, c45 16, c46 17, c55 18, c56 19, c66 20
template <typename T, typename R>
__device__ __forceinline__ void el25_stresses(const El25Args<T> &A, int64_t c, const R &K, bool L1, bool L2,
                                              bool L3, bool L4, const Cx<T> (&dy)[3], const Cx<T> (&dxp)[3],
                                              const Cx<T> (&dxm)[3], const Cx<T> (&dzmF)[3],
                                              const Cx<T> (&dzmH)[3], const Cx<T> (&dzp)[3])
{
    const T dt = A.dt;
    if (L1) {
        const Cx<T> yz = dzmF[1] + dy[2], xz = dzmF[0] + dxp[2], xy = dxp[1] + dy[0];
        A.s[0][c] = cstress(A.s[0][c], K[0], K[1], K[2], K[3], K[4], K[5], dxp[0], dy[1], dzmF[2],
                            yz, xz, xy, dt, K[M_IG + 0]);
        A.s[1][c] = cstress(A.s[1][c], K[1], K[6], K[7], K[8], K[9], K[10], dxp[0], dy[1], dzmF[2],
                            yz, xz, xy, dt, K[M_IG + 1]);
        A.s[2][c] = cstress(A.s[2][c], K[2], K[7], K[11], K[12], K[13], K[14], dxp[0], dy[1],
                            dzmF[2], yz, xz, xy, dt, K[M_IG + 2]);
    }
    if (L2) {
        const Cx<T> yz = dzmH[1] + dy[2], xz = dzmH[0] + dxm[2], xy = dxm[1] + dy[0];
        A.s[3][c] = cstress(A.s[3][c], K[5], K[10], K[14], K[17], K[19], K[20], dxm[0], dy[1],
                            dzmH[2], yz, xz, xy, dt, K[M_IG + 3]);
    }
    if (L3) {
        const Cx<T> yz = dzp[1] + dy[2], xz = dzp[0] + dxm[2], xy = dxm[1] + dy[0];
        A.s[4][c] = cstress(A.s[4][c], K[4], K[9], K[13], K[16], K[18], K[19], dxm[0], dy[1],
                            dzp[2], yz, xz, xy, dt, K[M_IG + 4]);
    }
    if (L4) {
        const Cx<T> yz = dzp[1] + dy[2], xz = dzp[0] + dxp[2], xy = dxp[1] + dy[0];
        A.s[5][c] = cstress(A.s[5][c], K[3], K[8], K[12], K[15], K[16], K[17], dxp[0], dy[1],
                            dzp[2], yz, xz, xy, dt, K[M_IG + 5]);
    }
}

// the four stress nests, cpmlfdtd.f08:941-1113.  (Loading every operand
// before the first store, as the 2D and velocity kernels do, needs about
// twice the registers here and measured slower.)
template <typename T, bool PML, int MAT>
__device__ __forceinline__ void el25_stress_cell(const El25Args<T> &A, int i, int k)
{
    const int nx = A.nx, nz = A.nz;
    const int64_t c = A.xoff + (int64_t)(i - 1) + (int64_t)A.pitch * (k - 1);
    const int64_t sz = A.pitch;
    const bool L1 = k >= 3 && k <= nz - 1 && i >= 2 && i <= nx - 2;
    const bool L2 = k >= 3 && k <= nz - 1 && i >= 3 && i <= nx - 1;
    const bool L3 = k >= 2 && k <= nz - 2 && i >= 3 && i <= nx - 1;
    const bool L4 = k >= 2 && k <= nz - 2 && i >= 2 && i <= nx - 2;
    if (!(L1 || L2 || L3 || L4)) return;
    // the material id first: its row loads wait on it
    const int m = MAT == kEl25Table ? (int)A.mid[c] : 0;
    const T rx = A.rdx24, rz = A.rdz24, ky = A.ky;
    T a = 0, b = 0, ik = 1, ah = 0, bh = 0, ikh = 1;
    if (PML) {
        a = A.a[c]; b = A.b[c]; ik = A.ik[c];
        ah = A.ah[c]; bh = A.bh[c]; ikh = A.ikh[c];
    }
    Cx<T> dy[3], dxp[3], dxm[3], dzmF[3], dzmH[3], dzp[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const Cx<T> *v = A.v[q];
        const Cx<T> v0 = v[c];
        dy[q] = iky(ky, v0);
        if (L1 || L4) {
            dxp[q] = cd4(v[c - 1], v0, v[c + 1], v[c + 2], rx);
            if (PML) dxp[q] = cfix_mem(dxp[q], A.m[P_XPH + q], c, ah, bh, ikh);
        }
        if (L2 || L3) {
            dxm[q] = cd4(v[c - 2], v[c - 1], v0, v[c + 1], rx);
            if (PML) dxm[q] = cfix_mem(dxm[q], A.m[P_XMF + q], c, a, b, ik);
        }
        if (L1 || L2) {
            const Cx<T> d = cd4(v[c - 2 * sz], v[c - sz], v0, v[c + sz], rz);
            dzmF[q] = d;
            dzmH[q] = d;
            if (PML && L1) dzmF[q] = cfix_mem(d, A.m[P_ZMF + q], c, a, b, ik);
            if (PML && L2) dzmH[q] = cfix_mem(d, A.m[P_ZMH + q], c, ah, bh, ikh);
        }
        if (L3 || L4) {
            dzp[q] = cd4(v[c - sz], v0, v[c + sz], v[c + 2 * sz], rz);
            if (PML) dzp[q] = cfix_mem(dzp[q], A.m[P_ZPH + q], c, ah, bh, ikh);
        }
    }
    if (MAT == kEl25Table) {
        const int m0 = __builtin_amdgcn_readfirstlane(m);
        if (__all(m == m0)) {
            const El25ScalarRow<T> K{(el25_const_row<T>)(A.tab) + m0 * kEl25TabRow};
            el25_stresses(A, c, K, L1, L2, L3, L4, dy, dxp, dxm, dzmF, dzmH, dzp);
        } else {
            const El25LaneRow<T> K{A.tab + m * kEl25TabRow};
            el25_stresses(A, c, K, L1, L2, L3, L4, dy, dxp, dxm, dzmF, dzmH, dzp);
        }
    } else if (MAT == kEl25Iso) {
        const T *const *M = A.mat;
        const El25IsoRow<T> K{M, c, M[0][c], M[1][c], M[15][c]};
        el25_stresses(A, c, K, L1, L2, L3, L4, dy, dxp, dxm, dzmF, dzmH, dzp);
    } else {
        const El25DenseRow<T> K{A.mat, c};
        el25_stresses(A, c, K, L1, L2, L3, L4, dy, dxp, dxm, dzmF, dzmH, dzp);
    }
}

template <typename T, int MAT>
__global__ void __launch_bounds__(256) el25_stress_kernel(El25Args<T> A)
{
    const int i = blockIdx.x * 64 + threadIdx.x + 1;
    const int k = blockIdx.y * 4 + threadIdx.y + 1;
    if (i > A.nx || k > A.nz) return;
    if (!in_box(A.box, i, 1, k) || A.flags[blockIdx.x + (size_t)A.tiles_x * blockIdx.y])
        el25_stress_cell<T, true, MAT>(A, i, k);
    else
        el25_stress_cell<T, false, MAT>(A, i, k);
}

// v += (d1 / rho_x + d2 / rho_y + d3 / rho_z) * dt
template <typename T>
__device__ __forceinline__ Cx<T> vupd(Cx<T> v, Cx<T> d1, Cx<T> d2, Cx<T> d3, T r1, T r2, T r3, T dt)
{
    return cmk(v.re + (d1.re / r1 + d2.re / r2 + d3.re / r3) * dt, v.im + (d1.im / r1 + d2.im / r2 + d3.im / r3) * dt);
}

// vx, vy, vz (cpmlfdtd.f08:1130-1200), Dirichlet edges (1280-1285) and the
// per-step velocity norm (1291) without the source cells; loads before
// stores as in the stress update
template <typename T, bool PML>
__device__ __forceinline__ double el25_velocity_cell(const El25Args<T> &A, int i, int k)
{
    const int nx = A.nx, nz = A.nz;
    const int64_t c = A.xoff + (int64_t)(i - 1) + (int64_t)A.pitch * (k - 1);
    const int64_t sz = A.pitch;
    Cx<T> vx = A.v[0][c], vy = A.v[1][c], vz = A.v[2][c];
    if (i == 1 || i == nx || k == 1 || k == nz) {
        vx = vy = vz = cmk(T(0), T(0));
        A.v[0][c] = vx;
        A.v[1][c] = vy;
        A.v[2][c] = vz;
    } else {
        const T rx = A.rdx24, rz = A.rdz24, ky = A.ky, dt = A.dt;
        const int corr = A.corrected;
        const T *const *M = A.mat;
        const T rho = M[M_RHO][c];
        T a = 0, b = 0, ik = 1, ah = 0, bh = 0, ikh = 1;
        if (PML) {
            a = A.a[c]; b = A.b[c]; ik = A.ik[c];
            ah = A.ah[c]; bh = A.bh[c]; ikh = A.ikh[c];
        }
        const bool LX = k >= 3 && k <= nz - 1 && i >= 3 && i <= nx - 1;
        const bool LY = k >= 3 && k <= nz - 1 && i >= 2 && i <= nx - 2;
        const bool LZ = k >= 2 && k <= nz - 2 && i >= 2 && i <= nx - 2;
        const Cx<T> z0 = cmk(T(0), T(0));
        Cx<T> mx1 = z0, mx3 = z0, my1 = z0, my3 = z0, mz1 = z0, mz3 = z0;
        if (PML) {
            if (LX) { mx1 = A.m[P_VX_X][c]; mx3 = A.m[P_VX_Z][c]; }
            if (LY) { my1 = A.m[P_VY_X][c]; my3 = A.m[P_VY_Z][c]; }
            if (LZ) { mz1 = A.m[P_VZ_X][c]; mz3 = A.m[P_VZ_Z][c]; }
        }
        Cx<T> nvx = vx, nvy = vy, nvz = vz;
        if (LX) {
            Cx<T> d1 = ctm(A.s[0], c, (int64_t)1, corr, rx);
            const Cx<T> d2 = iky(ky, A.s[3][c]);
            Cx<T> d3 = ctm(A.s[4], c, sz, corr, rz);
            if (PML) {
                d1 = cfix(d1, mx1, ah, bh, ikh);
                d3 = cfix(d3, mx3, ah, bh, ikh);
            }
            nvx = vupd(vx, d1, d2, d3, M[M_RHO_PI][c], rho, M[M_RHO_PJ][c], dt);
        }
        if (LY) {
            Cx<T> d1 = ctp(A.s[3], c, (int64_t)1, corr, rx);
            const Cx<T> d2 = iky(ky, A.s[1][c]);
            Cx<T> d3 = ctm(A.s[5], c, sz, corr, rz);
            if (PML) {
                d1 = cfix(d1, my1, ah, bh, ikh);
                d3 = cfix(d3, my3, a, b, ik);
            }
            nvy = vupd(vy, d1, d2, d3, M[M_RHO_MI][c], rho, M[M_RHO_PJ][c], dt);
        }
        if (LZ) {
            Cx<T> d1 = ctp(A.s[4], c, (int64_t)1, corr, rx);
            const Cx<T> d2 = iky(ky, A.s[5][c]);
            Cx<T> d3 = ctp(A.s[2], c, sz, corr, rz);
            if (PML) {
                d1 = cfix(d1, mz1, ah, bh, ikh);
                d3 = cfix(d3, mz3, ah, bh, ikh);
            }
            nvz = vupd(vz, d1, d2, d3, M[M_RHO_MI][c], rho, M[M_RHO_MJ][c], dt);
        }
        // stores
        if (LX) { vx = nvx; A.v[0][c] = vx; }
        if (LY) { vy = nvy; A.v[1][c] = vy; }
        if (LZ) { vz = nvz; A.v[2][c] = vz; }
        if (PML) {
            if (LX) { A.m[P_VX_X][c] = mx1; A.m[P_VX_Z][c] = mx3; }
            if (LY) { A.m[P_VY_X][c] = my1; A.m[P_VY_Z][c] = my3; }
            if (LZ) { A.m[P_VZ_X][c] = mz1; A.m[P_VZ_Z][c] = mz3; }
        }
    }
    if (c == A.vfix[0] || c == A.vfix[1] || c == A.vfix[2] || pw_face(A.pwf, i, 1, k)) return 0.0;
    const double s = ((double)vx.re * vx.re + (double)vx.im * vx.im) + ((double)vy.re * vy.re + (double)vy.im * vy.im) +
                     ((double)vz.re * vz.re + (double)vz.im * vz.im);
    return sqrt(s);
}

template <typename T>
__global__ void __launch_bounds__(256) el25_velocity_kernel(El25Args<T> A)
{
    const int i = blockIdx.x * 64 + threadIdx.x + 1;
    const int k = blockIdx.y * 4 + threadIdx.y + 1;
    double nrm = 0.0;
    if (i <= A.nx && k <= A.nz) {
        if (!in_box(A.box, i, 1, k) || A.flags[blockIdx.x + (size_t)A.tiles_x * blockIdx.y])
            nrm = el25_velocity_cell<T, true>(A, i, k);
        else
            nrm = el25_velocity_cell<T, false>(A, i, k);
    }
    block_max_to(&A.vnorm[*A.d_it - 1], nrm);
}

template <typename T>
static El25Args<T> el25_args(Device &d)
{
    El25Args<T> A;
    for (int q = 0; q < 3; ++q) A.v[q] = (Cx<T> *)d.field[SDT_F_VX + q];
    for (int q = 0; q < 6; ++q) A.s[q] = (Cx<T> *)d.field[SDT_F_SXX + q];
    for (int q = 0; q < M_TOTAL25; ++q) A.mat[q] = (const T *)d.mat[q];
    A.a = (const T *)d.cpml[0];
    A.b = (const T *)d.cpml[1];
    A.ik = (const T *)d.cpml[2];
    A.ah = (const T *)d.cpml[3];
    A.bh = (const T *)d.cpml[4];
    A.ikh = (const T *)d.cpml[5];
    for (int q = 0; q < P_TOTAL; ++q) A.m[q] = (Cx<T> *)d.mem[q];
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
    A.ky = (T)d.ky;
    A.corrected = d.stencil == SDT_STENCIL_CORRECTED;
    A.d_it = d.d_it;
    A.vnorm = d.norms;
    for (int q = 0; q < 3; ++q) A.vfix[q] = d.tail.vfix[q];
    A.pwf = d.pw.faces();
    A.mid = d.mid;
    A.tab = (const T *)d.tab;
    return A;
}

template <typename T>
static void el25_step_t(Device &d, bool timed)
{
    const El25Args<T> A = el25_args<T>(d);
    const dim3 grid(d.tiles_x, d.tiles_z), block(64, 4);
    if (timed) (void)hipEventRecord(d.ev[0], d.stream);
    if (d.mat_table) hipLaunchKernelGGL((el25_stress_kernel<T, kEl25Table>), grid, block, 0, d.stream, A);
    else if (d.mat_iso) hipLaunchKernelGGL((el25_stress_kernel<T, kEl25Iso>), grid, block, 0, d.stream, A);
    else hipLaunchKernelGGL((el25_stress_kernel<T, kEl25Dense>), grid, block, 0, d.stream, A);
    if (timed) (void)hipEventRecord(d.ev[1], d.stream);
    hipLaunchKernelGGL(el25_velocity_kernel<T>, grid, block, 0, d.stream, A);
    if (timed) (void)hipEventRecord(d.ev[2], d.stream);
    launch_plane_wave(d);
    launch_tail(d);
}

void el25_step(Device &d, bool timed)
{
    if (d.precision == SDT_FP64) el25_step_t<double>(d, timed);
    else el25_step_t<float>(d, timed);
}

// In-place density sweep of array_averaging2d (averaging.f08:27-80), run once
// at setup: the '+' directions average with the neighbour already overwritten
// in the same sweep, the '-' directions with the still-original neighbour.
static void smear(std::vector<double> &a, int nx, int nz, int code, int dir)
{
    auto A = [&](int i, int k) -> double & { return a[(size_t)(i - 1) + (size_t)nx * (size_t)(k - 1)]; };
    auto mean = [&](double x, double y) {
        switch (code) {
        case 1: return 2.0 * x * y / (x + y);
        case 2: return sqrt(x * y);
        case 3: return 0.5 * (x + y);
        default: return x;
        }
    };
    if (dir == 0)
        for (int k = 1; k <= nz; ++k)
            for (int i = 2; i <= nx; ++i) A(i, k) = mean(A(i, k), A(i - 1, k));
    else if (dir == 1)
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx - 1; ++i) A(i, k) = mean(A(i, k), A(i + 1, k));
    else if (dir == 2)
        for (int k = 2; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) A(i, k) = mean(A(i, k), A(i, k - 1));
    else
        for (int k = 1; k <= nz - 1; ++k)
            for (int i = 1; i <= nx; ++i) A(i, k) = mean(A(i, k), A(i, k + 1));
}

// Device material slots: c11..c66 (21; only c11, c12, c44 for an
// isotropic-form model), 1/(1 + gamma*dt) for x, y, z, xy, xz, yz, then rho
// and its four averaged copies ('+i' rhoxx, '+j' rhozx = rhozy, '-i' rhoxy =
// rhoxz, '-j' rhozz; cpmlfdtd.f08:889-898).  A model with at most 256
// distinct (stiffness, divisor) tuples keeps those 27 entries in a material
// table instead (one byte of id per cell); the smeared densities vary from
// row to row and stay per-cell arrays.
int el25_setup(Device &d, const double *const *in)
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
    std::vector<double> zero((size_t)n, 0.0);
    // isotropic form: only c11, c12, c44 (Voigt slots 0, 1, 15) are held
    d.mat_iso = d.iso_opt && stiffness_isotropic(in, n);
    const int gids[] = {SDT_GAMMA_X, SDT_GAMMA_Y, SDT_GAMMA_Z, SDT_GAMMA_XY, SDT_GAMMA_XZ, SDT_GAMMA_YZ};
    auto divisor = [&](int q, int64_t c) { return 1.0 / (1.0 + (in[gids[q]] ? in[gids[q]][c] : 0.0) * d.dt); };
    if (!d.mat_iso && d.mat_table_opt) {
        rc = build_material_table(d, kEl25TabRow, [&](int64_t c, double *r) {
            for (int q = 0; q < 21; ++q) r[q] = in[SDT_C11 + q] ? in[SDT_C11 + q][c] : 0.0;
            for (int q = 0; q < 6; ++q) r[M_IG + q] = divisor(q, c);
        });
        if (rc) return rc;
    }
    d.nmat = 0;
    std::vector<double> tmp((size_t)n);
    if (d.mat_table) {
        d.nmat = M_RHO;   // stiffness and divisors live in the table
    } else {
        for (int id = SDT_C11; id <= SDT_C66; ++id, ++d.nmat) {
            if (d.mat_iso && id != SDT_C11 && id != SDT_C12 && id != SDT_C44) continue;
            rc = upload_grid(d, in[id] ? in[id] : zero.data(), &d.mat[d.nmat]);
            if (rc) return rc;
        }
        for (int q = 0; q < 6; ++q) {
            for (int64_t c = 0; c < n; ++c) tmp[(size_t)c] = divisor(q, c);
            rc = upload_grid(d, tmp.data(), &d.mat[d.nmat++]);
            if (rc) return rc;
        }
    }
    const double *rho = in[SDT_RHO] ? in[SDT_RHO] : zero.data();
    rc = upload_grid(d, rho, &d.mat[d.nmat++]);
    if (rc) return rc;
    for (int dir : {0, 2, 1, 3}) {   // M_RHO_PI, M_RHO_PJ, M_RHO_MI, M_RHO_MJ
        tmp.assign(rho, rho + n);
        smear(tmp, d.nx, d.nz, d.density_code, dir);
        rc = upload_grid(d, tmp.data(), &d.mat[d.nmat++]);
        if (rc) return rc;
    }
    d.nmem = P_TOTAL;
    for (int q = 0; q < d.nmem; ++q) {
        rc = alloc_grid_zeroed(d, &d.mem[q], 2);
        if (rc) return rc;
    }
    for (int f = SDT_F_VX; f <= SDT_F_SYZ; ++f) {
        if (f <= SDT_F_VZ && in[SDT_IC_X + f]) rc = upload_grid(d, in[SDT_IC_X + f], &d.field[f], 2);
        else rc = alloc_grid_zeroed(d, &d.field[f], 2);
        if (rc) return rc;
    }

    // point sources, cpmlfdtd.f08:1268-1276 (isource = xind + cpml, ksource =
    // zind + cpml; the series are real and add to the real part)
    const int is = d.isrc, ks = d.ksrc, nx = d.nx, nz = d.nz;
    auto cell = [&](int i, int k) { return cell_index(d, i, 1, k); };
    auto R = [&](int i, int k) { return rho[(int64_t)(i - 1) + (int64_t)nx * (k - 1)]; };
    auto face = [&](int i, int k) { return (int)(i == 1 || i == nx || k == 1 || k == nz); };
    TailArgs &t = d.tail;
    t.nterms = 0;
    t.terms[t.nterms++] = {cell(is, ks), SDT_F_SXX, SDT_SRC_XX, 1.0, R(is, ks), 0};
    t.terms[t.nterms++] = {cell(is + 1, ks), SDT_F_SXY, SDT_SRC_XY, 1.0, R(is + 1, ks), 0};
    t.terms[t.nterms++] = {cell(is + 1, ks + 1), SDT_F_SXZ, SDT_SRC_XZ, 1.0, R(is + 1, ks), 0};
    t.terms[t.nterms++] = {cell(is, ks), SDT_F_SYY, SDT_SRC_YY, 1.0, R(is, ks), 0};
    t.terms[t.nterms++] = {cell(is, ks + 1), SDT_F_SYZ, SDT_SRC_YZ, 1.0, R(is, ks + 1), 0};
    t.terms[t.nterms++] = {cell(is, ks), SDT_F_SZZ, SDT_SRC_ZZ, 1.0, R(is, ks), 0};
    t.terms[t.nterms++] = {cell(is + 1, ks), SDT_F_VX, SDT_SRC_X, d.dt, R(is + 1, ks), face(is + 1, ks)};
    t.terms[t.nterms++] = {cell(is, ks), SDT_F_VY, SDT_SRC_Y, d.dt, R(is, ks), face(is, ks)};
    t.terms[t.nterms++] = {cell(is, ks + 1), SDT_F_VZ, SDT_SRC_Z, d.dt, R(is, ks + 1), face(is, ks + 1)};
    t.nvfix = 3;
    t.vfix[0] = cell(is + 1, ks);
    t.vfix[1] = cell(is, ks);
    t.vfix[2] = cell(is, ks + 1);
    t.nvcomp = 3;
    t.vcomp[0] = SDT_F_VX;
    t.vcomp[1] = SDT_F_VY;
    t.vcomp[2] = SDT_F_VZ;
    t.stress_norm = 0;
    t.nsfix = 0;
    t.rec_field[0] = SDT_F_VX;
    t.rec_field[1] = SDT_F_VY;
    t.rec_field[2] = SDT_F_VZ;
    t.cplx = 1;
    return upload_cpml_full(d, in);
}

}  // namespace sdt
