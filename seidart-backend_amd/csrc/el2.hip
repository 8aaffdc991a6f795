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
};

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
template <typename T, bool PML>
__device__ __forceinline__ void el2_stress_cell(const El2Args<T> &A, int i, int j)
{
    const int nx = A.nx, nz = A.nz;
    const bool r1 = i >= 2 && i <= nx - 2 && j >= 3 && j <= nz - 1;   // sigma_xx, sigma_zz
    const bool r2 = i >= 3 && i <= nx - 1 && j >= 2 && j <= nz - 2;   // sigma_xz
    if (!r1 && !r2) return;
    const int64_t c = A.xoff + (int64_t)(i - 1) + (int64_t)A.pitch * (j - 1);
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
        k11 = ldnt(A.c11 + c);
        k13 = ldnt(A.c13 + c);
        k33 = ldnt(A.c33 + c);
        gx = ldnt(A.igx + c);
        gz = ldnt(A.igz + c);
    }
    if (r2) {
        xl2 = vx[c - 2 * sx];
        xu2 = vx[c + 2 * sz];
        zl2 = vz[c - 2 * sx];
        zu2 = vz[c + 2 * sz];
        s_xz = A.sxz[c];
        k55 = ldnt(A.c55 + c);
        gxz = ldnt(A.igxz + c);
    }
    k15 = ldnt(A.c15 + c);
    k35 = ldnt(A.c35 + c);
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
        s_xx = (s_xx + (k11 * dvx_dx + k13 * dvz_dz + k15 * shear) * A.dt) * gx;
        s_zz = (s_zz + (k13 * dvx_dx + k33 * dvz_dz + k35 * shear) * A.dt) * gz;
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
        s_xz = (s_xz + (k15 * dvx_dx + k35 * dvz_dz + k55 * (dvz_dx + dvx_dz)) * A.dt) * gxz;
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
template <typename T, bool PML>
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
    if (rx || rz) r = ldnt(A.rho + c);
    if (rx) {
        rxl = A.rho[c - sx];
        rxd = A.rho[c - sz];
        f[0] = A.sxx[c];
        f[1] = A.sxx[c - sx];
        f[2] = A.sxx[c - 2 * sx];
        f[3] = corr ? A.sxx[c + sx] : T(0);
        f[4] = A.sxz[c];
        f[5] = A.sxz[c - sz];
        f[6] = A.sxz[c - 2 * sz];
        f[7] = corr ? A.sxz[c + sz] : T(0);
        if (PML) {
            a = ldnt(A.a + c);
            b = ldnt(A.b + c);
            ik = ldnt(A.ik + c);
            m[0] = ldnt(A.m[8] + c);
            m[1] = ldnt(A.m[9] + c);
        }
    }
    if (rz) {
        rzr = A.rho[c + sx];
        rzu = A.rho[c + sz];
        g[0] = A.sxz[c];
        g[1] = A.sxz[c + sx];
        g[2] = A.sxz[c + 2 * sx];
        g[3] = corr ? A.sxz[c - sx] : T(0);
        g[4] = A.szz[c];
        g[5] = A.szz[c + sz];
        g[6] = A.szz[c + 2 * sz];
        g[7] = corr ? A.szz[c - sz] : T(0);
        if (PML) {
            ah = ldnt(A.ah + c);
            bh = ldnt(A.bh + c);
            ikh = ldnt(A.ikh + c);
            m[2] = ldnt(A.m[10] + c);
            m[3] = ldnt(A.m[11] + c);
        }
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
        }
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
    if (c == A.src_cell) return 0.0;
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

template <typename T, int R>
__device__ __forceinline__ void el2_stress_tile(const El2Args<T> &A, int tx, int band)
{
    const int i = tx * 64 + (int)threadIdx.x - (int)A.xoff + 1;
    const bool valid = i >= 1 && i <= A.nx;
    const int j0 = (band * kTile2dWaves + (int)threadIdx.y) * R + 1;
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        if (j > A.nz) break;
        if (__any(valid && el2_pml(A, i, j))) {
            if (valid) el2_stress_cell<T, true>(A, i, j);
        } else if (valid) {
            el2_stress_cell<T, false>(A, i, j);
        }
    }
}

template <typename T, int R>
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
            if (valid) v = el2_velocity_cell<T, true>(A, i, j);
        } else if (valid) {
            v = el2_velocity_cell<T, false>(A, i, j);
        }
        nrm = (v > nrm || v != v) ? v : nrm;
    }
    return nrm;
}

template <typename T, int R>
__global__ void __launch_bounds__(64 * kTile2dWaves) el2_stress_kernel(El2Args<T> A)
{
    el2_stress_tile<T, R>(A, blockIdx.x, blockIdx.y);
}

template <typename T, int R>
__global__ void __launch_bounds__(64 * kTile2dWaves) el2_velocity_kernel(El2Args<T> A)
{
    const double nrm = sqrt(el2_velocity_tile<T, R>(A, blockIdx.x, blockIdx.y));
    block_max_to(&A.vnorm[*A.d_it - 1], nrm);
}

template <typename T, int R>
__global__ void __launch_bounds__(64 * kTile2dWaves) el2_stress_batch_kernel(const El2Args<T> *args,
                                                                             const int4 *tiles)
{
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;   // padding tile
    const El2Args<T> A = args[t.x];
    if (*A.d_it > A.nsteps) return;   // finished shot (uniform per workgroup)
    el2_stress_tile<T, R>(A, t.y, t.z);
}

template <typename T, int R>
__global__ void __launch_bounds__(64 * kTile2dWaves) el2_velocity_batch_kernel(const El2Args<T> *args,
                                                                               const int4 *tiles)
{
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;   // padding tile
    const El2Args<T> A = args[t.x];
    if (*A.d_it > A.nsteps) return;
    const double nrm = sqrt(el2_velocity_tile<T, R>(A, t.y, t.z));
    block_max_to(&A.vnorm[*A.d_it - 1], nrm);
}

template <typename T>
static int el2_batch_args_t(Batch &b, Device *const *shots, int n)
{
    std::vector<El2Args<T>> a((size_t)n);
    for (int s = 0; s < n; ++s) a[(size_t)s] = el2_args<T>(*shots[s]);
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
    hipLaunchKernelGGL((el2_stress_batch_kernel<T, R>), grid, block, 0, b.stream, a, b.tiles[which]);
    hipLaunchKernelGGL((el2_velocity_batch_kernel<T, R>), grid, block, 0, b.stream, a, b.tiles[which]);
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

template <typename T>
static void el2_step_t(Device &d, bool timed)
{
    constexpr int R = kSingleRows2d;
    const El2Args<T> A = el2_args<T>(d);
    const dim3 grid(tiles_x_2d(d), tiles_z_2d(d, R)), block(64, kTile2dWaves);
    if (timed) (void)hipEventRecord(d.ev[0], d.stream);
    hipLaunchKernelGGL((el2_stress_kernel<T, R>), grid, block, 0, d.stream, A);
    if (timed) (void)hipEventRecord(d.ev[1], d.stream);
    hipLaunchKernelGGL((el2_velocity_kernel<T, R>), grid, block, 0, d.stream, A);
    if (timed) (void)hipEventRecord(d.ev[2], d.stream);
    launch_tail(d);
}

int el2_adj_setup(Device &d)
{
    (void)d;
    return SDT_ERR_ARG;
}

void el2_adj_step(Device &d, int t)
{
    (void)d;
    (void)t;
}

void el2_adj_correlate(Device &d, int first, int n)
{
    (void)d;
    (void)first;
    (void)n;
}

int el2_batch_adj_args(Batch &b, Device *const *shots, int n)
{
    (void)b;
    (void)shots;
    (void)n;
    return SDT_ERR_ARG;
}

void el2_batch_adj_step(Batch &b, int which, int t, int first)
{
    (void)b;
    (void)which;
    (void)t;
    (void)first;
}

void el2_batch_adj_correlate(Batch &b, int which, int n)
{
    (void)b;
    (void)which;
    (void)n;
}

void el2_step(Device &d, bool timed)
{
    if (d.precision == SDT_FP64) el2_step_t<double>(d, timed);
    else el2_step_t<float>(d, timed);
}

// Material layout on the device: c11 c13 c15 c33 c35 c55 rho and the three
// attenuation divisors precomputed as 1/(1 + gamma*dt) in FP64.
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
    for (int id : ids) {
        rc = upload_grid(d, in[id] ? in[id] : zero.data(), &d.mat[d.nmat++]);
        if (rc) return rc;
    }
    const int gids[] = {SDT_GAMMA_X, SDT_GAMMA_Z, SDT_GAMMA_XZ};
    std::vector<double> inv((size_t)n);
    for (int id : gids) {
        for (int64_t c = 0; c < n; ++c) inv[(size_t)c] = 1.0 / (1.0 + (in[id] ? in[id][c] : 0.0) * d.dt);
        rc = upload_grid(d, inv.data(), &d.mat[d.nmat++]);
        if (rc) return rc;
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
    return upload_cpml_full(d, in);
}

}  // namespace sdt
