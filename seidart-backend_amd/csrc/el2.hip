// el2.hip -- 2D anisotropic elastic CPML-FDTD (seismic2, cpmlfdtd.f08:57-597)
// on gfx950.
//
// Layout: every array is (nx, nz) Fortran order, x fastest, in HBM.  One
// thread per cell, 64x4-cell tiles (one 64-wide wavefront per x row).  A tile
// whose cells all have identity CPML coefficients runs the branch-free
// interior body (no memory variables, no a/b/kappa reads); other tiles run the
// CPML body.  The choice is uniform per workgroup.
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

// sigma_xx, sigma_zz (cpmlfdtd.f08:324-353) and sigma_xz (362-386)
template <typename T, bool PML>
__device__ __forceinline__ void el2_stress_cell(const El2Args<T> &A, int i, int j)
{
    const int nx = A.nx, nz = A.nz;
    const int64_t c = A.xoff + (int64_t)(i - 1) + (int64_t)A.pitch * (j - 1);
    const T *vx = A.vx, *vz = A.vz;
    const int64_t sx = 1, sz = A.pitch;
    if (i >= 2 && i <= nx - 2 && j >= 3 && j <= nz - 1) {
        T dvx_dx = (T(27) * (vx[c + sx] - vx[c]) - (vx[c + 2 * sx] - vx[c - sx])) * A.rdx24;
        T dvz_dz = (T(27) * (vz[c] - vz[c - sz]) - (vz[c + sz] - vz[c - 2 * sz])) * A.rdz24;
        T dvz_dx = (T(27) * (vz[c + sx] - vz[c]) - (vz[c + 2 * sx] - vz[c - sx])) * A.rdx24;
        T dvx_dz = (T(27) * (vx[c] - vx[c - sz]) - (vx[c + sz] - vx[c - 2 * sz])) * A.rdz24;
        if (PML) {
            const T ah = A.ah[c], bh = A.bh[c], ikh = A.ikh[c];
            const T a = A.a[c], b = A.b[c], ik = A.ik[c];
            dvx_dx = cpml_fix(dvx_dx, A.m[0][c], ah, bh, ikh);
            dvx_dz = cpml_fix(dvx_dz, A.m[1][c], ah, bh, ikh);
            dvz_dz = cpml_fix(dvz_dz, A.m[2][c], a, b, ik);
            dvz_dx = cpml_fix(dvz_dx, A.m[3][c], a, b, ik);
        }
        const T shear = dvz_dx + dvx_dz;
        A.sxx[c] = (A.sxx[c] + (A.c11[c] * dvx_dx + A.c13[c] * dvz_dz + A.c15[c] * shear) * A.dt) * A.igx[c];
        A.szz[c] = (A.szz[c] + (A.c13[c] * dvx_dx + A.c33[c] * dvz_dz + A.c35[c] * shear) * A.dt) * A.igz[c];
    }
    if (i >= 3 && i <= nx - 1 && j >= 2 && j <= nz - 2) {
        T dvx_dx = (T(27) * (vx[c] - vx[c - sx]) - (vx[c + sx] - vx[c - 2 * sx])) * A.rdx24;
        T dvz_dz = (T(27) * (vz[c + sz] - vz[c]) - (vz[c + 2 * sz] - vz[c - sz])) * A.rdz24;
        T dvz_dx = (T(27) * (vz[c] - vz[c - sx]) - (vz[c + sx] - vz[c - 2 * sx])) * A.rdx24;
        T dvx_dz = (T(27) * (vx[c + sz] - vx[c]) - (vx[c + 2 * sz] - vx[c - sz])) * A.rdz24;
        if (PML) {
            const T ah = A.ah[c], bh = A.bh[c], ikh = A.ikh[c];
            dvx_dx = cpml_fix(dvx_dx, A.m[4][c], ah, bh, ikh);
            dvx_dz = cpml_fix(dvx_dz, A.m[5][c], ah, bh, ikh);
            dvz_dz = cpml_fix(dvz_dz, A.m[6][c], ah, bh, ikh);
            dvz_dx = cpml_fix(dvz_dx, A.m[7][c], ah, bh, ikh);
        }
        A.sxz[c] = (A.sxz[c] + (A.c15[c] * dvx_dx + A.c35[c] * dvz_dz + A.c55[c] * (dvz_dx + dvx_dz)) * A.dt) * A.igxz[c];
    }
}

template <typename T>
__global__ void __launch_bounds__(256) el2_stress_kernel(El2Args<T> A)
{
    const int i = blockIdx.x * 64 + threadIdx.x + 1;
    const int j = blockIdx.y * 4 + threadIdx.y + 1;
    if (i > A.nx || j > A.nz) return;
    if (!in_box(A.box, i, 1, j) || A.flags[blockIdx.x + (size_t)A.tiles_x * blockIdx.y]) el2_stress_cell<T, true>(A, i, j);
    else el2_stress_cell<T, false>(A, i, j);
}

// 3-term (reference) / 4-term (corrected) velocity stencils,
// cpmlfdtd.f08:405-406 and 434-435.
template <typename T>
__device__ __forceinline__ T tminus(const T *f, int64_t c, int64_t s, int corr)
{
    return corr ? T(27) * (f[c] - f[c - s]) - (f[c + s] - f[c - 2 * s])
                : T(27) * f[c] - T(27) * f[c - s] + f[c - 2 * s];
}
template <typename T>
__device__ __forceinline__ T tplus(const T *f, int64_t c, int64_t s, int corr)
{
    return corr ? T(27) * (f[c + s] - f[c]) - (f[c + 2 * s] - f[c - s])
                : -T(27) * f[c] + T(27) * f[c + s] - f[c + 2 * s];
}

// vx (cpmlfdtd.f08:399-416), vz (428-446), Dirichlet edges (534-549) and the
// per-step velocity norm (552-567) excluding the source cell.
template <typename T, bool PML>
__device__ __forceinline__ double el2_velocity_cell(const El2Args<T> &A, int i, int j)
{
    const int nx = A.nx, nz = A.nz;
    const int64_t c = A.xoff + (int64_t)(i - 1) + (int64_t)A.pitch * (j - 1);
    const int64_t sx = 1, sz = A.pitch;
    T vx = A.vx[c], vz = A.vz[c];
    if (i == 1 || i == nx || j == 1 || j == nz) {
        vx = T(0);
        vz = T(0);
        A.vx[c] = vx;
        A.vz[c] = vz;
    } else {
        const T r = A.rho[c];
        if (i >= 3 && i <= nx - 1 && j >= 3 && j <= nz - 1) {
            const T rhoxx = scalar_mean(r, A.rho[c - sx], A.density_code);
            const T rhozx = scalar_mean(r, A.rho[c - sz], A.density_code);
            T d1 = tminus(A.sxx, c, sx, A.corrected) * A.rdx24;
            T d2 = tminus(A.sxz, c, sz, A.corrected) * A.rdz24;
            if (PML) {
                const T a = A.a[c], b = A.b[c], ik = A.ik[c];
                d1 = cpml_fix(d1, A.m[8][c], a, b, ik);
                d2 = cpml_fix(d2, A.m[9][c], a, b, ik);
            }
            vx = vx + A.dt * (d1 / rhoxx + d2 / rhozx);
            A.vx[c] = vx;
        }
        if (i >= 2 && i <= nx - 2 && j >= 2 && j <= nz - 2) {
            const T rhoxz = scalar_mean(r, A.rho[c + sx], A.density_code);
            const T rhozz = scalar_mean(r, A.rho[c + sz], A.density_code);
            T d1 = tplus(A.sxz, c, sx, A.corrected) * A.rdx24;
            T d2 = tplus(A.szz, c, sz, A.corrected) * A.rdz24;
            if (PML) {
                const T ah = A.ah[c], bh = A.bh[c], ikh = A.ikh[c];
                d1 = cpml_fix(d1, A.m[10][c], ah, bh, ikh);
                d2 = cpml_fix(d2, A.m[11][c], ah, bh, ikh);
            }
            vz = vz + A.dt * (d1 / rhoxz + d2 / rhozz);
            A.vz[c] = vz;
        }
    }
    if (c == A.src_cell) return 0.0;
    const double dvx = (double)vx, dvz = (double)vz;
    return sqrt(dvx * dvx + dvz * dvz);
}

template <typename T>
__global__ void __launch_bounds__(256) el2_velocity_kernel(El2Args<T> A)
{
    const int i = blockIdx.x * 64 + threadIdx.x + 1;
    const int j = blockIdx.y * 4 + threadIdx.y + 1;
    double nrm = 0.0;
    if (i <= A.nx && j <= A.nz) {
        if (!in_box(A.box, i, 1, j) || A.flags[blockIdx.x + (size_t)A.tiles_x * blockIdx.y]) nrm = el2_velocity_cell<T, true>(A, i, j);
        else nrm = el2_velocity_cell<T, false>(A, i, j);
    }
    block_max_to(&A.vnorm[*A.d_it - 1], nrm);
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

// ---- batched shots: one launch covers the tiles of every shot --------------
// CPML classification per cell from the shot's 64 x 4 tile flags, so any
// tile height works.
template <typename T>
__device__ __forceinline__ bool el2_pml(const El2Args<T> &A, int i, int j)
{
    return !in_box(A.box, i, 1, j) || A.flags[((i - 1) >> 6) + (size_t)A.tiles_x * ((j - 1) >> 2)];
}

template <typename T, int TY>
__global__ void __launch_bounds__(64 * TY) el2_stress_batch_kernel(const El2Args<T> *args, const int4 *tiles)
{
    const int4 t = tiles[blockIdx.x];
    const El2Args<T> &A = args[t.x];
    if (*A.d_it > A.nsteps) return;
    const int i = t.y * 64 + threadIdx.x + 1;
    const int j = t.z * TY + threadIdx.y + 1;
    if (i > A.nx || j > A.nz) return;
    if (el2_pml(A, i, j)) el2_stress_cell<T, true>(A, i, j);
    else el2_stress_cell<T, false>(A, i, j);
}

template <typename T, int TY>
__global__ void __launch_bounds__(64 * TY) el2_velocity_batch_kernel(const El2Args<T> *args, const int4 *tiles)
{
    const int4 t = tiles[blockIdx.x];
    const El2Args<T> &A = args[t.x];
    if (*A.d_it > A.nsteps) return;   // uniform per workgroup
    const int i = t.y * 64 + threadIdx.x + 1;
    const int j = t.z * TY + threadIdx.y + 1;
    double nrm = 0.0;
    if (i <= A.nx && j <= A.nz)
        nrm = el2_pml(A, i, j) ? el2_velocity_cell<T, true>(A, i, j) : el2_velocity_cell<T, false>(A, i, j);
    block_max_to(&A.vnorm[*A.d_it - 1], nrm);
}

template <typename T>
static int el2_batch_args_t(Batch &b, Device *const *shots, int n)
{
    std::vector<El2Args<T>> a((size_t)n);
    for (int s = 0; s < n; ++s) a[(size_t)s] = el2_args<T>(*shots[s]);
    if (hipMalloc(&b.args, a.size() * sizeof(El2Args<T>)) != hipSuccess) return SDT_ERR_HIP;
    if (hipMemcpy(b.args, a.data(), a.size() * sizeof(El2Args<T>), hipMemcpyHostToDevice) != hipSuccess)
        return SDT_ERR_HIP;
    return SDT_OK;
}

int el2_batch_args(Batch &b, Device *const *shots, int n)
{
    return b.precision == SDT_FP64 ? el2_batch_args_t<double>(b, shots, n) : el2_batch_args_t<float>(b, shots, n);
}

template <typename T, int TY>
static void el2_batch_launch_t(Batch &b, int which)
{
    const El2Args<T> *a = (const El2Args<T> *)b.args;
    const dim3 grid(b.ntiles[which]), block(64, TY);
    hipLaunchKernelGGL((el2_stress_batch_kernel<T, TY>), grid, block, 0, b.stream, a, b.tiles[which]);
    hipLaunchKernelGGL((el2_velocity_batch_kernel<T, TY>), grid, block, 0, b.stream, a, b.tiles[which]);
}

template <typename T>
static void el2_batch_launch_p(Batch &b, int which)
{
    switch (kBatchTileRows[which]) {
    case 2: el2_batch_launch_t<T, 2>(b, which); break;
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
    const El2Args<T> A = el2_args<T>(d);
    const dim3 grid(d.tiles_x, d.tiles_z), block(64, 4);
    if (timed) (void)hipEventRecord(d.ev[0], d.stream);
    hipLaunchKernelGGL(el2_stress_kernel<T>, grid, block, 0, d.stream, A);
    if (timed) (void)hipEventRecord(d.ev[1], d.stream);
    hipLaunchKernelGGL(el2_velocity_kernel<T>, grid, block, 0, d.stream, A);
    if (timed) (void)hipEventRecord(d.ev[2], d.stream);
    launch_tail(d);
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
