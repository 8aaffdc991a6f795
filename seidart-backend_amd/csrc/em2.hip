// em2.hip -- 2D anisotropic electromagnetic CPML-FDTD (electromag2,
// cpmlfdtd.f08:2317-2760) on gfx950: Hy from (Ex, Ez), then (Ex, Ez) from Hy,
// 2nd-order Yee differences, full 2x2 permittivity and conductivity tensors.
//
// E_old == E at the start of every step (the reference copies E into E_old
// after the Dirichlet step, 2716-2721), and each E cell reads only its own
// E_old, so the E update runs in place without the E_old arrays.
//
// Dirichlet faces of Hy (2692-2711) are applied lazily: the H kernel treats a
// face cell's previous Hy as 0 and stores the fresh value, which the E kernel
// of the same step reads exactly as the reference does (its zeroing follows
// the E update); field read-back masks the faces.
#include <math.h>

#include <algorithm>
#include <vector>

#include "sdt_device.h"
#include "sdt_internal.h"

namespace sdt {

template <typename T>
struct Em2Args {
    T *Ex, *Ez, *Hy;
    const T *aEx, *bEx, *bEz, *dEx, *eEx, *dEz, *eEz;   // aEz == bEx
    const T *a, *b, *ik, *ah, *bh, *ikh;
    T *m_dEx_dz, *m_dEz_dx, *m_dHy_dx, *m_dHy_dz;
    const uint8_t *flags;
    int box[6];
    int nx, nz, tiles_x, pitch;
    int64_t xoff;
    T rdx, rdz, dt, dbHy;
    const int *d_it;
    unsigned long long *vnorm;
    int64_t src_cell;
    PwFaces pwf;        // plane-wave face cells, excluded from the fused norm
    int nsteps;         // time_steps of the run (batched shots stop there)
    // segment recomputation for the adjoint: per step (0-based t), slot
    // t - store_first of `store` holds E_x, E_z before the update and the
    // CPML-corrected dHy/dz, dHy/dx of every cell ([slot][4][stride])
    T *store;
    int64_t stride;
    int store_first;
    // material table (a model with at most 256 distinct coefficient tuples
    // and no adjoint): id per padded cell, rows of kEm2TabRow holding aEx bEx
    // bEz dEx eEx dEz eEz; null: per-cell arrays
    const uint8_t *mid;
    const T *tab;
};

template <typename T>
using em2_const_row = __attribute__((address_space(4))) const T *;

// The seven update coefficients of a cell from its table row: scalar loads
// when the wavefront's cells share one material, per-lane cached loads
// otherwise.
template <typename T>
__device__ __forceinline__ void em2_row(const T *tab, int m, T (&co)[7])
{
    const int m0 = __builtin_amdgcn_readfirstlane(m);
    if (__all(m == m0)) {
        const em2_const_row<T> p = (em2_const_row<T>)tab + m0 * kEm2TabRow;
#pragma unroll
        for (int q = 0; q < 7; ++q) co[q] = p[q];
    } else {
        const T *p = tab + m * kEm2TabRow;
#pragma unroll
        for (int q = 0; q < 7; ++q) co[q] = p[q];
    }
}

// Hy (cpmlfdtd.f08:2576-2593); every operand is read before the first store
template <typename T, bool PML>
__device__ __forceinline__ void em2_h_cell(const Em2Args<T> &A, int i, int j)
{
    const int nx = A.nx, nz = A.nz;
    const int64_t c = A.xoff + (int64_t)(i - 1) + (int64_t)A.pitch * (j - 1);
    if (i == nx || j == nz) { A.Hy[c] = T(0); return; }
    const T ex0 = A.Ex[c], exu = A.Ex[c + A.pitch], ez0 = A.Ez[c], ezr = A.Ez[c + 1];
    const T hold = (i == 1 || j == 1) ? T(0) : A.Hy[c];
    T mx = 0, mz = 0, a = 0, b = 0, ik = 0, ah = 0, bh = 0, ikh = 0;
    if (PML) {
        mx = ldnt(A.m_dEx_dz + c);
        mz = ldnt(A.m_dEz_dx + c);
        a = ldnt(A.a + c);
        b = ldnt(A.b + c);
        ik = ldnt(A.ik + c);
        ah = ldnt(A.ah + c);
        bh = ldnt(A.bh + c);
        ikh = ldnt(A.ikh + c);
    }
    T dEx_dz = (exu - ex0) * A.rdz;
    T dEz_dx = (ezr - ez0) * A.rdx;
    if (PML) {
        mx = b * mx + a * dEx_dz;
        mz = bh * mz + ah * dEz_dx;
        dEz_dx = dEz_dx * ikh + mz;
        dEx_dz = dEx_dz * ik + mx;
        A.m_dEx_dz[c] = mx;
        A.m_dEz_dx[c] = mz;
    }
    A.Hy[c] = hold + A.dbHy * (dEz_dx - dEx_dz);
}

// Ex, Ez (cpmlfdtd.f08:2605-2620), zero edges, and the squared field
// magnitude for the per-step norm, 0 at the source cell and the plane-wave
// face cells (their values change after this kernel; the norm is the
// square root of the maximum, which equals the maximum of the magnitudes).
// STORE: also write the values the adjoint gradient needs (see Em2Args).
template <typename T, bool PML, bool STORE, bool TAB>
__device__ __forceinline__ double em2_e_cell(const Em2Args<T> &A, int i, int j)
{
    const int nx = A.nx, nz = A.nz;
    const int64_t c = A.xoff + (int64_t)(i - 1) + (int64_t)A.pitch * (j - 1);
    T ex, ez;
    T *st = nullptr;
    if (STORE) st = A.store + (int64_t)(*A.d_it - 1 - A.store_first) * 4 * A.stride + c;
    if (i == 1 || i == nx || j == 1 || j == nz) {
        ex = ez = T(0);
        if (STORE) st[0] = st[A.stride] = st[2 * A.stride] = st[3 * A.stride] = T(0);
    } else {
        // TAB: the launch may meet a table (compiled out otherwise); uniform
        // over the launch's shot; recording runs (the adjoint's) never hold
        // a table
        const bool tab = TAB && !STORE && A.mid != nullptr;
        const int mat = tab ? (int)A.mid[c] : 0;
        const T h0 = A.Hy[c], hd = A.Hy[c - A.pitch], hl = A.Hy[c - 1];
        const T exo = A.Ex[c], ezo = A.Ez[c];
        T co[7];   // aEx bEx bEz dEx eEx dEz eEz
        if (!tab) {
            co[1] = ldnt(A.bEx + c); co[2] = ldnt(A.bEz + c); co[5] = ldnt(A.dEz + c); co[6] = ldnt(A.eEz + c);
            co[0] = ldnt(A.aEx + c); co[3] = ldnt(A.dEx + c); co[4] = ldnt(A.eEx + c);
        }
        T mx = 0, mz = 0, a = 0, b = 0, ik = 0, ah = 0, bh = 0, ikh = 0;
        if (PML) {
            mx = ldnt(A.m_dHy_dx + c);
            mz = ldnt(A.m_dHy_dz + c);
            a = ldnt(A.a + c);
            b = ldnt(A.b + c);
            ik = ldnt(A.ik + c);
            ah = ldnt(A.ah + c);
            bh = ldnt(A.bh + c);
            ikh = ldnt(A.ikh + c);
        }
        if (tab) em2_row(A.tab, mat, co);
        const T aEx = co[0], bEx = co[1], bEz = co[2], dEx = co[3], eEx = co[4], dEz = co[5], eEz = co[6];
        T dHy_dz = (h0 - hd) * A.rdz;
        T dHy_dx = (h0 - hl) * A.rdx;
        if (PML) {
            mx = bh * mx + ah * dHy_dx;
            mz = b * mz + a * dHy_dz;
            dHy_dz = dHy_dz * ik + mz;
            dHy_dx = dHy_dx * ikh + mx;
            A.m_dHy_dx[c] = mx;
            A.m_dHy_dz[c] = mz;
        }
        ez = ezo + (-bEx * dHy_dz + bEz * dHy_dx - dEz * exo - eEz * ezo) * A.dt;
        ex = exo + (-aEx * dHy_dz + bEx * dHy_dx - dEx * exo - eEx * ezo) * A.dt;
        if (STORE) {
            st[0] = exo;
            st[A.stride] = ezo;
            st[2 * A.stride] = dHy_dz;
            st[3 * A.stride] = dHy_dx;
        }
    }
    A.Ex[c] = ex;
    A.Ez[c] = ez;
    if (c == A.src_cell || pw_face(A.pwf, i, 1, j)) return 0.0;
    const double x = ex, z = ez;
    return x * x + z * z;
}

template <typename T>
static Em2Args<T> em2_args(Device &d)

{
    Em2Args<T> A;
    A.Ex = (T *)d.field[SDT_F_EX];
    A.Ez = (T *)d.field[SDT_F_EZ];
    A.Hy = (T *)d.field[SDT_F_HY];
    const T **mp[] = {&A.aEx, &A.bEx, &A.bEz, &A.dEx, &A.eEx, &A.dEz, &A.eEz};
    for (int q = 0; q < 7; ++q) *mp[q] = (const T *)d.mat[q];
    A.a = (const T *)d.cpml[0];
    A.b = (const T *)d.cpml[1];
    A.ik = (const T *)d.cpml[2];
    A.ah = (const T *)d.cpml[3];
    A.bh = (const T *)d.cpml[4];
    A.ikh = (const T *)d.cpml[5];
    A.m_dEx_dz = (T *)d.mem[0];
    A.m_dEz_dx = (T *)d.mem[1];
    A.m_dHy_dx = (T *)d.mem[2];
    A.m_dHy_dz = (T *)d.mem[3];
    A.flags = d.tile_flags;
    for (int q = 0; q < 6; ++q) A.box[q] = d.box[q];
    A.nx = d.nx;
    A.nz = d.nz;
    A.pitch = d.pitch;
    A.xoff = d.xoff;
    A.tiles_x = d.tiles_x;
    A.rdx = (T)(1.0 / d.dx);
    A.rdz = (T)(1.0 / d.dz);
    A.dt = (T)d.dt;
    A.dbHy = (T)(d.dt / kMu0);   // cpmlfdtd.f08:2516
    A.d_it = d.d_it;
    A.vnorm = d.norms;
    A.src_cell = d.tail.vfix[0];
    A.pwf = d.pw.faces();
    A.nsteps = d.time_steps;
    A.store = (T *)d.adj.seg_fwd;
    A.stride = d.ncell_dev;
    A.store_first = d.adj.store_first;
    A.mid = d.mat_table ? d.mid : nullptr;
    A.tab = d.mat_table ? (const T *)d.tab : nullptr;
    return A;
}

// ---- tiles ------------------------------------------------------------------
// Same tiling as the 2D elastic solver: 64 line-aligned columns x four
// wavefronts marching R rows each; the CPML body runs for a wavefront row
// when any of its cells needs it (identity cells give the same values
// through either body).
template <typename T>
__device__ __forceinline__ bool em2_pml(const Em2Args<T> &A, int i, int j)
{
    return !in_box(A.box, i, 1, j) || A.flags[((i - 1) >> 6) + (size_t)A.tiles_x * ((j - 1) >> 2)];
}

template <typename T, int R>
__device__ __forceinline__ void em2_h_tile(const Em2Args<T> &A, int tx, int band)
{
    const int i = tx * 64 + (int)threadIdx.x - (int)A.xoff + 1;
    const bool valid = i >= 1 && i <= A.nx;
    const int j0 = (band * kTile2dWaves + (int)threadIdx.y) * R + 1;
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        if (j > A.nz) break;
        if (__any(valid && em2_pml(A, i, j))) {
            if (valid) em2_h_cell<T, true>(A, i, j);
        } else if (valid) {
            em2_h_cell<T, false>(A, i, j);
        }
    }
}

template <typename T, int R, bool STORE, bool TAB>
__device__ __forceinline__ double em2_e_tile(const Em2Args<T> &A, int tx, int band)
{
    const int i = tx * 64 + (int)threadIdx.x - (int)A.xoff + 1;
    const bool valid = i >= 1 && i <= A.nx;
    const int j0 = (band * kTile2dWaves + (int)threadIdx.y) * R + 1;
    double nrm = 0.0;
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        if (j > A.nz) break;
        double v = 0.0;
        if (__any(valid && em2_pml(A, i, j))) {
            if (valid) v = em2_e_cell<T, true, STORE, TAB>(A, i, j);
        } else if (valid) {
            v = em2_e_cell<T, false, STORE, TAB>(A, i, j);
        }
        nrm = (v > nrm || v != v) ? v : nrm;
    }
    return nrm;
}

template <typename T, int R>
__global__ void __launch_bounds__(64 * kTile2dWaves) em2_h_kernel(Em2Args<T> A)
{
    em2_h_tile<T, R>(A, blockIdx.x, blockIdx.y);
}

template <typename T, int R, bool STORE, bool TAB>
__global__ void __launch_bounds__(64 * kTile2dWaves) em2_e_kernel(Em2Args<T> A)
{
    const double nrm = sqrt(em2_e_tile<T, R, STORE, TAB>(A, blockIdx.x, blockIdx.y));
    block_max_to(&A.vnorm[*A.d_it - 1], nrm);
}

template <typename T, int R>
__global__ void __launch_bounds__(64 * kTile2dWaves) em2_h_batch_kernel(const Em2Args<T> *args, const int4 *tiles)
{
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;   // padding tile
    const Em2Args<T> A = args[t.x];
    if (*A.d_it > A.nsteps) return;   // finished shot (uniform per workgroup)
    em2_h_tile<T, R>(A, t.y, t.z);
}

template <typename T, int R, bool STORE, bool TAB>
__global__ void __launch_bounds__(64 * kTile2dWaves) em2_e_batch_kernel(const Em2Args<T> *args, const int4 *tiles)
{
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;   // padding tile
    const Em2Args<T> A = args[t.x];
    if (*A.d_it > A.nsteps) return;
    const double nrm = sqrt(em2_e_tile<T, R, STORE, TAB>(A, t.y, t.z));
    block_max_to(&A.vnorm[*A.d_it - 1], nrm);
}

template <typename T>
static int em2_batch_args_t(Batch &b, Device *const *shots, int n)
{
    std::vector<Em2Args<T>> a((size_t)n);
    b.any_table = false;
    for (int s = 0; s < n; ++s) {
        a[(size_t)s] = em2_args<T>(*shots[s]);
        b.any_table = b.any_table || a[(size_t)s].mid != nullptr;
    }
    if (!b.args && hipMalloc(&b.args, a.size() * sizeof(Em2Args<T>)) != hipSuccess) return SDT_ERR_HIP;
    if (hipMemcpy(b.args, a.data(), a.size() * sizeof(Em2Args<T>), hipMemcpyHostToDevice) != hipSuccess)
        return SDT_ERR_HIP;
    return SDT_OK;
}

int em2_batch_args(Batch &b, Device *const *shots, int n)
{
    return b.precision == SDT_FP64 ? em2_batch_args_t<double>(b, shots, n) : em2_batch_args_t<float>(b, shots, n);
}

template <typename T, int R>
static void em2_batch_launch_t(Batch &b, int which)
{
    const Em2Args<T> *a = (const Em2Args<T> *)b.args;
    const dim3 grid(b.ntiles[which]), block(64, kTile2dWaves);
    hipLaunchKernelGGL((em2_h_batch_kernel<T, R>), grid, block, 0, b.stream, a, b.tiles[which]);
    if (b.store)
        hipLaunchKernelGGL((em2_e_batch_kernel<T, R, true, false>), grid, block, 0, b.stream, a, b.tiles[which]);
    else if (b.any_table)
        hipLaunchKernelGGL((em2_e_batch_kernel<T, R, false, true>), grid, block, 0, b.stream, a, b.tiles[which]);
    else
        hipLaunchKernelGGL((em2_e_batch_kernel<T, R, false, false>), grid, block, 0, b.stream, a, b.tiles[which]);
}

template <typename T>
static void em2_batch_launch_p(Batch &b, int which)
{
    switch (kBatchShapes[which].rows) {
    case 1: em2_batch_launch_t<T, 1>(b, which); break;
    case 8: em2_batch_launch_t<T, 8>(b, which); break;
    default: em2_batch_launch_t<T, 4>(b, which); break;
    }
}

void em2_batch_launch(Batch &b, int which)
{
    if (b.precision == SDT_FP64) em2_batch_launch_p<double>(b, which);
    else em2_batch_launch_p<float>(b, which);
}

template <typename T>
static void em2_step_t(Device &d, bool timed)
{
    constexpr int R = kSingleRows2d;
    const Em2Args<T> A = em2_args<T>(d);
    const dim3 grid(tiles_x_2d(d), tiles_z_2d(d, R)), block(64, kTile2dWaves);
    if (timed) (void)hipEventRecord(d.ev[0], d.stream);
    hipLaunchKernelGGL((em2_h_kernel<T, R>), grid, block, 0, d.stream, A);
    if (timed) (void)hipEventRecord(d.ev[1], d.stream);
    if (d.adj.store)
        hipLaunchKernelGGL((em2_e_kernel<T, R, true, false>), grid, block, 0, d.stream, A);
    else if (d.mat_table)
        hipLaunchKernelGGL((em2_e_kernel<T, R, false, true>), grid, block, 0, d.stream, A);
    else
        hipLaunchKernelGGL((em2_e_kernel<T, R, false, false>), grid, block, 0, d.stream, A);
    if (timed) (void)hipEventRecord(d.ev[2], d.stream);
    launch_plane_wave(d);
    launch_tail(d);
}

// ---- adjoint state (discrete adjoint of the step above) ----------------------
// Step t maps (E^t, Hy^{t-1/2}, psi^t) to (E^{t+1}, Hy^{t+1/2}, psi^{t+1}):
// H update (reads E^t), E update (reads E^t, Hy^{t+1/2}), point sources,
// receivers (read E^{t+1}).  Its transpose runs per step, latest first:
//   tail   the trace-gradient row t is added to lam(E^{t+1}) at the receivers;
//   E pass lam(E^{t+1}) is completed with the H-pass transfer terms of step
//          t+1 (lgz, lgx), recorded at the source cell and in the segment
//          store, then pushed through the E update: lam(E^t) (own-cell part),
//          lam(psi_E), and the transfer terms lfz, lfx to lam(Hy^{t+1/2});
//   H pass lam(Hy^{t+1/2}) is completed with lfz, lfx and pushed through the
//          H update: lam(Hy^{t-1/2}), lam(psi_H), transfer terms lgz, lgx to
//          lam(E^t), which the next E pass gathers.
// Every cell gathers from its neighbours' transfer terms, so no two threads
// write one element.  Cells outside an update's domain hold zero transfer
// terms.  Gradients: E^{t+1} = E^t + dt * (-aEx fz + bEx fx - dEx Ex - eEx Ez,
// -bEx fz + bEz fx - dEz Ex - eEz Ez) gives, per cell,
//   d/daEx = -dt lx fz, d/dbEx = dt (lx fx - lz fz), d/dbEz = dt lz fx,
//   d/ddEx = -dt lx Ex, d/deEx = -dt lx Ez, d/ddEz = -dt lz Ex, d/deEz = -dt lz Ez
// with (lx, lz) = lam(E^{t+1}) and (Ex, Ez, fz, fx) the stored forward values.
template <typename T>
struct Em2AdjArgs {
    Em2Args<T> F;
    T *lex, *lez, *lhy;        // lam(Ex), lam(Ez) (own-cell part), lam(Hy)
    T *lm[4];                  // lam of m_dEx_dz, m_dEz_dx, m_dHy_dx, m_dHy_dz
    T *lfz, *lfx, *lgz, *lgx;  // transfer terms
    T *adj;                    // segment store: [slot][2][stride] lam(Ex^{t+1}), lam(Ez^{t+1})
    double *src_lam;           // [time_steps][kAdjSrcSlots]
    double *grad;              // [7][stride] gradient accumulators
};

template <typename T, bool PML>
__device__ __forceinline__ void em2_adj_e_cell(const Em2AdjArgs<T> &B, int i, int j, int t, int first)
{
    const Em2Args<T> &A = B.F;
    const int nx = A.nx, nz = A.nz;
    const int64_t c = A.xoff + (int64_t)(i - 1) + (int64_t)A.pitch * (j - 1);
    // unconditional loads (the own element at the low faces), so they issue
    // with the cell's other loads
    const T gzd_v = B.lgz[j >= 2 ? c - A.pitch : c], gxl_v = B.lgx[i >= 2 ? c - 1 : c];
    const T gzd = j >= 2 ? gzd_v : T(0);
    const T gxl = i >= 2 ? gxl_v : T(0);
    const T lx = B.lex[c] + (gzd - B.lgz[c]) * A.rdz;
    const T lz = B.lez[c] + (gxl - B.lgx[c]) * A.rdx;
    if (c == A.src_cell) {
        B.src_lam[kAdjSrcSlots * (int64_t)t] = (double)lx;
        B.src_lam[kAdjSrcSlots * (int64_t)t + 1] = (double)lz;
    }
    T *as = B.adj + (int64_t)(t - first) * 2 * A.stride + c;
    if (i == 1 || i == nx || j == 1 || j == nz) {   // E^{t+1} = 0 there: no dependence on step t
        B.lex[c] = T(0);
        B.lez[c] = T(0);
        B.lfz[c] = T(0);
        B.lfx[c] = T(0);
        as[0] = as[A.stride] = T(0);
        return;
    }
    const T aEx = ldnt(A.aEx + c), bEx = ldnt(A.bEx + c), bEz = ldnt(A.bEz + c);
    const T dEx = ldnt(A.dEx + c), eEx = ldnt(A.eEx + c), dEz = ldnt(A.dEz + c), eEz = ldnt(A.eEz + c);
    T lfz = A.dt * (-aEx * lx - bEx * lz);
    T lfx = A.dt * (bEx * lx + bEz * lz);
    if (PML) {
        // fz' = fz ik + psi_z', psi_z' = b psi_z + a fz (and the half-cell set for x)
        const T L4 = B.lm[3][c] + lfz, L3 = B.lm[2][c] + lfx;
        lfz = lfz * ldnt(A.ik + c) + ldnt(A.a + c) * L4;
        lfx = lfx * ldnt(A.ikh + c) + ldnt(A.ah + c) * L3;
        B.lm[3][c] = ldnt(A.b + c) * L4;
        B.lm[2][c] = ldnt(A.bh + c) * L3;
    }
    as[0] = lx;
    as[A.stride] = lz;
    B.lex[c] = lx - A.dt * (dEx * lx + dEz * lz);
    B.lez[c] = lz - A.dt * (eEx * lx + eEz * lz);
    B.lfz[c] = lfz;
    B.lfx[c] = lfx;
}

template <typename T, bool PML>
__device__ __forceinline__ void em2_adj_h_cell(const Em2AdjArgs<T> &B, int i, int j)
{
    const Em2Args<T> &A = B.F;
    const int nx = A.nx, nz = A.nz;
    const int64_t c = A.xoff + (int64_t)(i - 1) + (int64_t)A.pitch * (j - 1);
    if (i == nx || j == nz) {   // Hy = 0 there: no dependence
        B.lhy[c] = T(0);
        B.lgz[c] = T(0);
        B.lgx[c] = T(0);
        return;
    }
    const T lh = B.lhy[c] + (B.lfz[c] - B.lfz[c + A.pitch]) * A.rdz + (B.lfx[c] - B.lfx[c + 1]) * A.rdx;
    T lgx = A.dbHy * lh, lgz = -A.dbHy * lh;
    if (PML) {
        const T L2 = B.lm[1][c] + lgx, L1 = B.lm[0][c] + lgz;
        lgx = lgx * ldnt(A.ikh + c) + ldnt(A.ah + c) * L2;
        lgz = lgz * ldnt(A.ik + c) + ldnt(A.a + c) * L1;
        B.lm[1][c] = ldnt(A.bh + c) * L2;
        B.lm[0][c] = ldnt(A.b + c) * L1;
    }
    B.lhy[c] = (i == 1 || j == 1) ? T(0) : lh;
    B.lgz[c] = lgz;
    B.lgx[c] = lgx;
}

template <typename T, int R>
__device__ __forceinline__ void em2_adj_e_tile(const Em2AdjArgs<T> &B, int tx, int band, int t, int first)
{
    const Em2Args<T> &A = B.F;
    const int i = tx * 64 + (int)threadIdx.x - (int)A.xoff + 1;
    const bool valid = i >= 1 && i <= A.nx;
    const int j0 = (band * kTile2dWaves + (int)threadIdx.y) * R + 1;
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        if (j > A.nz) break;
        if (__any(valid && em2_pml(A, i, j))) {
            if (valid) em2_adj_e_cell<T, true>(B, i, j, t, first);
        } else if (valid) {
            em2_adj_e_cell<T, false>(B, i, j, t, first);
        }
    }
}

template <typename T, int R>
__device__ __forceinline__ void em2_adj_h_tile(const Em2AdjArgs<T> &B, int tx, int band)
{
    const Em2Args<T> &A = B.F;
    const int i = tx * 64 + (int)threadIdx.x - (int)A.xoff + 1;
    const bool valid = i >= 1 && i <= A.nx;
    const int j0 = (band * kTile2dWaves + (int)threadIdx.y) * R + 1;
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        if (j > A.nz) break;
        if (__any(valid && em2_pml(A, i, j))) {
            if (valid) em2_adj_h_cell<T, true>(B, i, j);
        } else if (valid) {
            em2_adj_h_cell<T, false>(B, i, j);
        }
    }
}

// gradient of segment [first, first + n): per cell, sums over the stored
// steps in FP64, scaled by dt once
template <typename T, int R>
__device__ __forceinline__ void em2_correlate_tile(const Em2AdjArgs<T> &B, int tx, int band, int n)
{
    double *grad = B.grad;
    const Em2Args<T> &A = B.F;
    const int i = tx * 64 + (int)threadIdx.x - (int)A.xoff + 1;
    if (i < 1 || i > A.nx) return;
    const int j0 = (band * kTile2dWaves + (int)threadIdx.y) * R + 1;
    const int64_t st = A.stride;
    for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        if (j > A.nz) break;
        const int64_t c = A.xoff + (int64_t)(i - 1) + (int64_t)A.pitch * (j - 1);
        double g[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int s = 0; s < n; ++s) {
            const T *f = A.store + (int64_t)s * 4 * st + c;
            const T *l = B.adj + (int64_t)s * 2 * st + c;
            const double lx = ldnt(l), lz = ldnt(l + st);
            const double ex = ldnt(f), ez = ldnt(f + st), fz = ldnt(f + 2 * st), fx = ldnt(f + 3 * st);
            g[0] -= lx * fz;
            g[1] += lx * fx - lz * fz;
            g[2] += lz * fx;
            g[3] -= lx * ex;
            g[4] -= lx * ez;
            g[5] -= lz * ex;
            g[6] -= lz * ez;
        }
        const double dt = (double)A.dt;
        for (int q = 0; q < 7; ++q) grad[q * st + c] += dt * g[q];
    }
}

template <typename T, int R>
__global__ void __launch_bounds__(64 * kTile2dWaves) em2_adj_e_kernel(Em2AdjArgs<T> B, int t, int first)
{
    em2_adj_e_tile<T, R>(B, blockIdx.x, blockIdx.y, t, first);
}

template <typename T, int R>
__global__ void __launch_bounds__(64 * kTile2dWaves) em2_adj_h_kernel(Em2AdjArgs<T> B)
{
    em2_adj_h_tile<T, R>(B, blockIdx.x, blockIdx.y);
}

template <typename T, int R>
__global__ void __launch_bounds__(64 * kTile2dWaves) em2_correlate_kernel(Em2AdjArgs<T> B, int n)
{
    em2_correlate_tile<T, R>(B, blockIdx.x, blockIdx.y, n);
}

// batched forms: tile lists name the shot (all shots reverse the same step)
template <typename T, int R>
__global__ void __launch_bounds__(64 * kTile2dWaves) em2_adj_e_batch_kernel(const Em2AdjArgs<T> *args,
                                                                            const int4 *tiles, int t, int first)
{
    const int4 q = tiles[blockIdx.x];
    if (q.x < 0) return;
    em2_adj_e_tile<T, R>(args[q.x], q.y, q.z, t, first);
}

template <typename T, int R>
__global__ void __launch_bounds__(64 * kTile2dWaves) em2_adj_h_batch_kernel(const Em2AdjArgs<T> *args,
                                                                            const int4 *tiles)
{
    const int4 q = tiles[blockIdx.x];
    if (q.x < 0) return;
    em2_adj_h_tile<T, R>(args[q.x], q.y, q.z);
}

template <typename T, int R>
__global__ void __launch_bounds__(64 * kTile2dWaves) em2_correlate_batch_kernel(const Em2AdjArgs<T> *args,
                                                                                const int4 *tiles, int n)
{
    const int4 q = tiles[blockIdx.x];
    if (q.x < 0) return;
    em2_correlate_tile<T, R>(args[q.x], q.y, q.z, n);
}

template <typename T>
static Em2AdjArgs<T> em2_adj_args(Device &d)
{
    Em2AdjArgs<T> B;
    B.F = em2_args<T>(d);
    T **l = (T **)d.adj.lam;
    B.lex = l[0];
    B.lez = l[1];
    B.lhy = l[2];
    for (int q = 0; q < 4; ++q) B.lm[q] = l[3 + q];
    B.lfz = l[7];
    B.lfx = l[8];
    B.lgz = l[9];
    B.lgx = l[10];
    B.adj = (T *)d.adj.seg_adj;
    B.src_lam = d.adj.src_lam;
    B.grad = d.adj.grad;
    return B;
}

int em2_adj_setup(Device &d)
{
    d.adj.nlam = 11;
    for (int q = 0; q < d.adj.nlam; ++q)
        if (!d.adj.lam[q] && alloc_grid_zeroed(d, &d.adj.lam[q]) != SDT_OK) return SDT_ERR_HIP;
    return SDT_OK;
}

template <typename T>
static void em2_adj_step_t(Device &d, int t)
{
    constexpr int R = kSingleRows2d;
    const Em2AdjArgs<T> B = em2_adj_args<T>(d);
    void *lam_field[3] = {B.lex, nullptr, B.lez};
    launch_adj_tail(d, t, lam_field);
    const dim3 grid(tiles_x_2d(d), tiles_z_2d(d, R)), block(64, kTile2dWaves);
    hipLaunchKernelGGL((em2_adj_e_kernel<T, R>), grid, block, 0, d.stream, B, t, d.adj.store_first);
    hipLaunchKernelGGL((em2_adj_h_kernel<T, R>), grid, block, 0, d.stream, B);
}

void em2_adj_step(Device &d, int t)
{
    if (d.precision == SDT_FP64) em2_adj_step_t<double>(d, t);
    else em2_adj_step_t<float>(d, t);
}

template <typename T>
static void em2_adj_correlate_t(Device &d, int n)
{
    constexpr int R = kSingleRows2d;
    const Em2AdjArgs<T> B = em2_adj_args<T>(d);
    const dim3 grid(tiles_x_2d(d), tiles_z_2d(d, R)), block(64, kTile2dWaves);
    hipLaunchKernelGGL((em2_correlate_kernel<T, R>), grid, block, 0, d.stream, B, n);
}

void em2_adj_correlate(Device &d, int first, int n)
{
    (void)first;   // the store slots are relative to the segment start
    if (d.precision == SDT_FP64) em2_adj_correlate_t<double>(d, n);
    else em2_adj_correlate_t<float>(d, n);
}

template <typename T>
static int em2_batch_adj_args_t(Batch &b, Device *const *shots, int n)
{
    std::vector<Em2AdjArgs<T>> a((size_t)n);
    std::vector<AdjTailArgs> tails((size_t)n);
    for (int s = 0; s < n; ++s) {
        a[(size_t)s] = em2_adj_args<T>(*shots[s]);
        AdjTailArgs &t = tails[(size_t)s];
        t.lam[0] = a[(size_t)s].lex;
        t.lam[1] = nullptr;
        t.lam[2] = a[(size_t)s].lez;
        t.rec_cell = shots[s]->rec_cell;
        t.tg = shots[s]->adj.trace_grad;
        t.nrec = shots[s]->nrec;
    }
    if (!b.adj_args && hipMalloc(&b.adj_args, a.size() * sizeof(Em2AdjArgs<T>)) != hipSuccess) return SDT_ERR_HIP;
    if (!b.adj_tails && hipMalloc((void **)&b.adj_tails, tails.size() * sizeof(AdjTailArgs)) != hipSuccess)
        return SDT_ERR_HIP;
    if (hipMemcpy(b.adj_args, a.data(), a.size() * sizeof(Em2AdjArgs<T>), hipMemcpyHostToDevice) != hipSuccess)
        return SDT_ERR_HIP;
    if (hipMemcpy(b.adj_tails, tails.data(), tails.size() * sizeof(AdjTailArgs), hipMemcpyHostToDevice) != hipSuccess)
        return SDT_ERR_HIP;
    return SDT_OK;
}

int em2_batch_adj_args(Batch &b, Device *const *shots, int n)
{
    return b.precision == SDT_FP64 ? em2_batch_adj_args_t<double>(b, shots, n)
                                   : em2_batch_adj_args_t<float>(b, shots, n);
}

template <typename T, int R>
static void em2_batch_adj_launch(Batch &b, int which, int t, int first, int n, bool correlate)
{
    const Em2AdjArgs<T> *a = (const Em2AdjArgs<T> *)b.adj_args;
    const dim3 grid(b.ntiles[which]), block(64, kTile2dWaves);
    if (correlate) {
        hipLaunchKernelGGL((em2_correlate_batch_kernel<T, R>), grid, block, 0, b.stream, a, b.tiles[which], n);
        return;
    }
    hipLaunchKernelGGL((em2_adj_e_batch_kernel<T, R>), grid, block, 0, b.stream, a, b.tiles[which], t, first);
    hipLaunchKernelGGL((em2_adj_h_batch_kernel<T, R>), grid, block, 0, b.stream, a, b.tiles[which]);
}

template <typename T>
static void em2_batch_adj_p(Batch &b, int which, int t, int first, int n, bool correlate)
{
    switch (kBatchShapes[which].rows) {
    case 1: em2_batch_adj_launch<T, 1>(b, which, t, first, n, correlate); break;
    case 8: em2_batch_adj_launch<T, 8>(b, which, t, first, n, correlate); break;
    default: em2_batch_adj_launch<T, 4>(b, which, t, first, n, correlate); break;
    }
}

void em2_batch_adj_step(Batch &b, int which, int t, int first)
{
    if (b.precision == SDT_FP64) em2_batch_adj_p<double>(b, which, t, first, 0, false);
    else em2_batch_adj_p<float>(b, which, t, first, 0, false);
}

void em2_batch_adj_correlate(Batch &b, int which, int n)
{
    if (b.precision == SDT_FP64) em2_batch_adj_p<double>(b, which, 0, 0, n, true);
    else em2_batch_adj_p<float>(b, which, 0, 0, n, true);
}

void em2_step(Device &d, bool timed)
{
    if (d.precision == SDT_FP64) em2_step_t<double>(d, timed);
    else em2_step_t<float>(d, timed);
}

// Coefficients of cpmlfdtd.f08:2519-2540 (permittivity scaled by eps0,
// determinant clamped at 1e-18*eps0^2), computed in FP64 on the host and
// stored as aEx, bEx (= aEz), bEz, dEx, eEx, dEz, eEz: per-cell arrays, or,
// without an adjoint and with at most 256 distinct tuples, table rows with
// one byte of id per cell.
size_t src_window_bytes(const SrcWindow &w) { return 3 * (size_t)(w.ni + 1) * (w.nj + 1) * sizeof(double); }

// the window's (and ring's) values of field f from an unpadded FP64 host
// field in Fortran order (nx, nz); cells past the grid (none for E) stay
int src_window_load(Device &d, int f, const double *host)
{
    const SrcWindow &w = d.tail.win;
    if (!w.v || (f != SDT_F_EX && f != SDT_F_EZ && f != SDT_F_HY)) return SDT_OK;
    const int m = w.ni + 1, mm = m * (w.nj + 1);
    std::vector<double> v((size_t)mm, 0.0);
    for (int jj = 0; jj <= w.nj; ++jj)
        for (int ii = 0; ii <= w.ni; ++ii) {
            const int i = w.i0 + ii, j = w.j0 + jj;
            if (i <= d.nx && j <= d.nz) v[(size_t)(ii + m * jj)] = host[(int64_t)(i - 1) + (int64_t)d.nx * (j - 1)];
        }
    double *dst = w.v + (size_t)(f == SDT_F_EX ? 0 : f == SDT_F_EZ ? 1 : 2) * mm;
    return hipMemcpy(dst, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess ? SDT_OK
                                                                                                   : SDT_ERR_HIP;
}

int em2_setup(Device &d, const double *const *in)
{
    d.tile_x = 64;
    d.tile_y = 1;
    d.tile_z = 4;
    int rc = compute_tile_flags(d, in);
    if (rc) return rc;
    set_layout(d);
    const int64_t n = d.ncell;
    std::vector<double> co[7];
    for (auto &v : co) v.resize((size_t)n);
    auto val = [&](int id, int64_t c) { return in[id] ? in[id][c] : 0.0; };
    for (int64_t c = 0; c < n && d.coef_in; ++c)   // coefficients given directly
        for (int q = 0; q < 7; ++q) co[q][(size_t)c] = d.coef_in[(size_t)q * n + c];
    for (int64_t c = 0; c < n && !d.coef_in; ++c) {
        const double e11 = val(SDT_E11, c) * kEps0, e13 = val(SDT_E13, c) * kEps0, e33 = val(SDT_E33, c) * kEps0;
        const double s11 = val(SDT_S11, c), s13 = val(SDT_S13, c), s33 = val(SDT_S33, c);
        double det = e11 * e33 - e13 * e13;
        if (det <= 1.0e-18 * kEps0 * kEps0) det = 1.0e-18 * kEps0 * kEps0;
        co[0][(size_t)c] = e33 / det;
        co[1][(size_t)c] = -e13 / det;
        co[2][(size_t)c] = e11 / det;
        co[3][(size_t)c] = (e33 * s11 - e13 * s13) / det;
        co[4][(size_t)c] = (e33 * s13 - e13 * s33) / det;
        co[5][(size_t)c] = (e11 * s13 - e13 * s11) / det;
        co[6][(size_t)c] = (e11 * s33 - e13 * s13) / det;
    }
    if (!d.coef_in && d.mat_table_opt && d.adj.every == 0) {
        rc = build_material_table(d, kEm2TabRow, [&](int64_t c, double *r) {
            for (int q = 0; q < 7; ++q) r[q] = co[q][(size_t)c];
        });
        if (rc) return rc;
    }
    d.nmat = 0;
    for (int q = 0; q < 7; ++q, ++d.nmat) {
        if (d.mat_table) continue;   // the rows hold them
        rc = upload_grid(d, co[q].data(), &d.mat[d.nmat]);
        if (rc) return rc;
    }
    d.nmem = 4;
    for (int q = 0; q < d.nmem; ++q) {
        rc = alloc_grid_zeroed(d, &d.mem[q]);
        if (rc) return rc;
    }
    const int fids[] = {SDT_F_EX, SDT_F_EZ, SDT_F_HY};
    for (int f : fids) {
        if (f == SDT_F_EX && in[SDT_IC_X]) rc = upload_grid(d, in[SDT_IC_X], &d.field[f]);
        else if (f == SDT_F_EZ && in[SDT_IC_Z]) rc = upload_grid(d, in[SDT_IC_Z], &d.field[f]);
        else rc = alloc_grid_zeroed(d, &d.field[f]);
        if (rc) return rc;
    }

    // point source, cpmlfdtd.f08:2678-2681
    const int is = d.isrc, js = d.ksrc, nx = d.nx, nz = d.nz;
    const int64_t hc = (int64_t)(is - 1) + (int64_t)nx * (js - 1);
    const int64_t c = cell_index(d, is, 1, js);
    const int face = (is == 1 || is == nx || js == 1 || js == nz);
    TailArgs &t = d.tail;
    t.nterms = 0;
    t.terms[t.nterms++] = {c, SDT_F_EX, SDT_SRC_X, d.dt, val(SDT_E11, hc) * kEps0, face};
    t.terms[t.nterms++] = {c, SDT_F_EZ, SDT_SRC_Z, d.dt, val(SDT_E33, hc) * kEps0, face};
    t.nvfix = 1;
    t.vfix[0] = c;
    t.nvcomp = 2;
    t.vcomp[0] = SDT_F_EX;
    t.vcomp[1] = SDT_F_EZ;
    t.stress_norm = 0;
    t.nsfix = 0;
    t.rec_field[0] = SDT_F_EX;
    t.rec_field[1] = -1;
    t.rec_field[2] = SDT_F_EZ;

    // FP64 source window (FP32 runs): up to kSrcWinHalf cells on each side
    // of the source, inside the CPML-free box and off the faces
    if (d.precision == SDT_FP32 && d.src_window_opt && !face) {
        const int ilo = std::max({is - kSrcWinHalf, d.box[0], 2}), ihi = std::min({is + kSrcWinHalf, d.box[1], nx - 1});
        const int jlo = std::max({js - kSrcWinHalf, d.box[4], 2}), jhi = std::min({js + kSrcWinHalf, d.box[5], nz - 1});
        if (ilo <= is && is <= ihi && jlo <= js && js <= jhi) {
            SrcWindow &w = t.win;
            w.i0 = ilo;
            w.j0 = jlo;
            w.ni = ihi - ilo + 1;
            w.nj = jhi - jlo + 1;
            w.pitch = d.pitch;
            w.base = cell_index(d, ilo, 1, jlo);
            w.rdx = 1.0 / d.dx;
            w.rdz = 1.0 / d.dz;
            w.dt = d.dt;
            w.dbHy = d.dt / kMu0;
            const size_t nv = src_window_bytes(w) / sizeof(double), nc = (size_t)w.ni * w.nj;
            std::vector<double> hv(nv + 7 * nc, 0.0);
            for (int jj = 0; jj < w.nj; ++jj)
                for (int ii = 0; ii < w.ni; ++ii) {
                    const int64_t h = (int64_t)(ilo + ii - 1) + (int64_t)nx * (jlo + jj - 1);
                    for (int q = 0; q < 7; ++q) hv[nv + q * nc + (size_t)(ii + w.ni * jj)] = co[q][(size_t)h];
                }
            void *p = nullptr;
            if (hipMalloc(&p, hv.size() * sizeof(double)) != hipSuccess) return SDT_ERR_HIP;
            w.v = (double *)p;
            w.co = w.v + nv;
            if (hipMemcpy(p, hv.data(), hv.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
                return SDT_ERR_HIP;
            // initial E in FP64 (Hy starts at zero)
            if (in[SDT_IC_X] && (rc = src_window_load(d, SDT_F_EX, in[SDT_IC_X])) != SDT_OK) return rc;
            if (in[SDT_IC_Z] && (rc = src_window_load(d, SDT_F_EZ, in[SDT_IC_Z])) != SDT_OK) return rc;
        }
    }

    // adjoint layout: checkpoint = fields + memory variables; 7 coefficients;
    // per step 4 forward values (Ex, Ez, fz, fx) and 2 adjoint values
    Adjoint &a = d.adj;
    a.nstate = 0;
    for (int f : fids) a.state[a.nstate++] = d.field[f];
    for (int q = 0; q < d.nmem; ++q) a.state[a.nstate++] = d.mem[q];
    a.ncoef = 7;
    a.ngrad = 7;
    a.nfwd = 4;
    a.nadj = 2;
    return upload_cpml_full(d, in);
}

}  // namespace sdt
